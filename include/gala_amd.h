/*
 * gala_amd.h -- C ABI of the MI355X orbit-integration engine.
 *
 * Plain C: no torch or HIP types.  Array arguments are DEVICE pointers
 * (HBM-resident, caller-owned) unless a comment says "host"; `stream` is a
 * hipStream_t passed as void* (NULL = the default stream).  Every call is
 * stream-ordered and returns 0 on success or a negative gamd_status.  Per-orbit
 * integration outcomes go in status[] arrays using gala's DOP853 codes
 * (1 success, -2 nmax exceeded, -3 step size too small, -4 stiff).
 *
 * Layouts follow gala's C layer: positions are structure-of-arrays (3,N)
 * (q[k*N+i], potential/src/vectorization.h:18-25), phase-space states (6,N),
 * saved trajectories (6, ntimes, N) as leapfrog.pyx:91-95 returns them.
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef GALA_AMD_H
#define GALA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum gamd_status {
    GAMD_OK = 0,
    GAMD_ERR_INVALID = -1,   /* bad argument (shape, count, null pointer) */
    GAMD_ERR_HIP = -2,       /* a HIP runtime call failed */
    GAMD_ERR_NOMEM = -3,     /* device allocation failed */
    GAMD_ERR_UNSUPPORTED = -4
};

/* status[] of an orbit that an integration never started (gamd_dop853 with
 * an order that does not name it); gala's own codes are 1, -2, -3, -4 */
#define GAMD_STATUS_NOT_RUN 0

/* ---- potentials -------------------------------------------------------- */
enum gamd_component_type {
    GAMD_NULL = 0,           /* NullWrapper, cybuiltin.pyx:362-372 */
    GAMD_HERNQUIST = 1,      /* pars [G, m, c]            cybuiltin.pyx:165-174 */
    GAMD_PLUMMER = 2,        /* pars [G, m, b]            cybuiltin.pyx:177-186 */
    GAMD_MIYAMOTONAGAI = 3,  /* pars [G, m, a, b]         cybuiltin.pyx:264-273 */
    GAMD_NFW = 4,            /* pars [G, m, r_s, a, b, c] SphericalNFWWrapper, cybuiltin.pyx:292-301 */
    GAMD_SCF = 5             /* pars [G, m, r_s] + S/T    SCFWrapper, scf/bfe_class.pyx:38-47 */
};

#define GAMD_MAX_COMPONENTS 8
#define GAMD_SCF_MAX_NMAX 20
#define GAMD_SCF_MAX_LMAX 10

/* One component of a (possibly composite) potential: one slot of gala's
 * struct CPotential (potential/potential/src/cpotential.h:10-36).
 * pars[0] = G, then the PotentialParameter values in declaration order
 * (cpotential.pyx:281-316).  shift_rotate = 0 means q0 = 0 and R = identity
 * (cpotential.pyx:97-102); otherwise the component is evaluated at
 * q' = R (q - q0) and its gradient rotated back with R^T (cpotential.cpp:246-279). */
typedef struct {
    int32_t type;            /* enum gamd_component_type */
    int32_t shift_rotate;
    int32_t nmax, lmax;      /* SCF only */
    double pars[8];
    double q0[3];
    double R[9];             /* row-major */
    const double *S;         /* SCF only, HOST (nmax+1)(lmax+1)(lmax+1) C order, copied */
    const double *T;         /* SCF only, HOST, same shape */
} gamd_component;

typedef struct gamd_potential gamd_potential;

/* Replaces CPotentialWrapper.init + CCompositePotentialWrapper assembly
 * (cpotential.pyx:57-102, ccompositepotential.pyx:27-70): copies the
 * component table (and SCF coefficients) to the device once. */
int gamd_potential_create(const gamd_component *components, int ncomponents,
                          gamd_potential **out);
int gamd_potential_destroy(gamd_potential *pot);

/* Batched evaluation on (3,N) positions.  Replace CPotentialWrapper.energy /
 * density / gradient / hessian (cpotential.pyx:104-185) over c_potential /
 * c_density / c_gradient / c_hessian (cpotential.cpp:170-314).
 * value, density: out (N).  gradient: grad (3,N), OVERWRITTEN (the zeroing
 * of cpotential.cpp:238-240 is folded in).  hessian: hess (9,N) row-major
 * 3x3 per point, rotation not applied (as cpotential.cpp:309-311). */
int gamd_potential_value(const gamd_potential *pot, double t, const double *q, size_t N,
                         double *out, void *stream);
int gamd_potential_density(const gamd_potential *pot, double t, const double *q, size_t N,
                           double *out, void *stream);
int gamd_potential_gradient(const gamd_potential *pot, double t, const double *q, size_t N,
                            double *grad, void *stream);
int gamd_potential_hessian(const gamd_potential *pot, double t, const double *q, size_t N,
                           double *hess, void *stream);
/* c_d2_dr2 (cpotential.cpp:346-371): radial second difference, h = 1e-2.
 * q (3,N) -> out (N). Used by the stream DF (df.pyx:84-86). */
int gamd_potential_d2_dr2(const gamd_potential *pot, double t, const double *q, size_t N,
                          double *out, void *stream);
/* Hamiltonian.energy for the static frame (chamiltonian.cpp:7-19,
 * builtin_frames.cpp:8-16): w (6,N) -> E (N) = Phi(q) + |p|^2/2. */
int gamd_hamiltonian_energy(const gamd_potential *pot, double t, const double *w, size_t N,
                            double *out, void *stream);

/* ---- Leapfrog ----------------------------------------------------------- */
/* Replaces leapfrog_integrate_hamiltonian (leapfrog.pyx:54-121).
 * w0 (6,N); fixed dt (leapfrog.pyx:78 uses t[1]-t[0]); n_steps steps for
 * every orbit, or n_steps_per_orbit[i] (device, nullable) steps for orbit i.
 * Saved trajectory w_out (6, nsave, N) at steps 0, s, 2s, ... (s =
 * save_stride, nsave = n_steps/s + 1; s = 1 is gala's save_all=True); w_out
 * may be NULL.  w_final (6,N) receives the state after the orbit's last step
 * (gala's save_all=False result). */
int gamd_leapfrog(const gamd_potential *pot, const double *w0, size_t N, double t0, double dt,
                  int64_t n_steps, const int64_t *n_steps_per_orbit, int64_t save_stride,
                  double *w_out, double *w_final, void *stream);

/* Resume a leapfrog integration from the state its loop carries between
 * steps (leapfrog.pyx:106-116): positions and HALF-STEP velocities w_half
 * (6,N), so no initial half kick.  n_steps steps, or n_steps_per_orbit[i]
 * (device, nullable); w_final (6,N) receives the positions and the
 * synchronised velocities after the last step (an orbit with 0 steps keeps
 * its input).  Arithmetic as gamd_leapfrog's loop, so resuming from the
 * state after step j of a run continues that run bit for bit.  No gala
 * counterpart (gala cannot resume); the learned surrogate hands orbits on
 * the exact force to it. */
int gamd_leapfrog_resume(const gamd_potential *pot, const double *w_half, size_t N, double dt,
                         int64_t n_steps, const int64_t *n_steps_per_orbit, double *w_final,
                         void *stream);

/* Leapfrog whose saved trajectory is streamed to HOST memory while the
 * integration runs (gala's leapfrog_integrate_hamiltonian returns it on the
 * host, leapfrog.pyx:91-121).  The nsave = n_steps / save_stride + 1 saved
 * rows go to host_out (HOST, ideally pinned, (6, nsave, N)) in chunks of
 * chunk_rows rows: each chunk is one kernel writing a slot of a two-slot
 * device ring, then copied to the host on a second stream while the next
 * chunk computes.  Results equal gamd_leapfrog's w_out and w_final (device,
 * (6,N)) bit for bit.  use_graph != 0 captures the chunk kernels and copies
 * as one HIP graph and launches that (no memset nodes).  workspace: device,
 * gamd_leapfrog_stream_workspace_bytes(N, chunk_rows) bytes.  The call
 * returns when host_out is complete (it synchronises `stream`).  No gala
 * counterpart for the streaming itself; the chunk size is a speed knob
 * (gala_amd.autotune picks it). */
int64_t gamd_leapfrog_stream_workspace_bytes(size_t N, int64_t chunk_rows);
int gamd_leapfrog_stream(const gamd_potential *pot, const double *w0, size_t N, double t0,
                         double dt, int64_t n_steps, int64_t save_stride, int64_t chunk_rows,
                         int use_graph, double *host_out, double *w_final, void *workspace,
                         size_t workspace_bytes, void *stream);

/* ---- Ruth4 ------------------------------------------------------------- */
/* Replaces ruth4_integrate_hamiltonian (ruth4.pyx:24-106): fixed dt, per step
 * four kick-drift substeps with Ruth's coefficients (:67-80).  Arguments and
 * the saved layout as gamd_leapfrog (no per-orbit step counts); w_final (6,N)
 * is the state after n_steps steps.  The first substep's kick has d_0 = 0 and
 * is skipped (identical results wherever the gradient is finite). */
int gamd_ruth4(const gamd_potential *pot, const double *w0, size_t N, double t0, double dt,
               int64_t n_steps, int64_t save_stride, double *w_out, double *w_final,
               void *stream);

/* ---- DOP853 ------------------------------------------------------------- */
/* Replaces dop853_integrate_hamiltonian / dop853_helper / dop853 /
 * Fwrapper_T (dop853.pyx:90-250, dop853.cpp:89-866, 979-988).  Each orbit is
 * an independent 6-dimensional system with its own adaptive step (gala's
 * nbatch=1), integrated from t0[i] to t1[i] with initial step h0[i] (0 =
 * Hairer's hinit, dop853.cpp:18-86) and hmax = |t1-t0| (dop853.cpp:787-788)
 * when hmax <= 0.  nmax 0 -> 1e6, nstiff 0 -> 1000, < 0 -> off
 * (dop853.cpp:673-694).  Dense output (dop853.cpp:492-612, 869-904): when
 * w_out is non-NULL, every orbit is sampled at the n_out shared times t_out
 * (device) into w_out (6, n_out, N).  w_final (6,N), status (N) and stats
 * (N,4) = nstep, naccpt, nrejct, nfcn (nullable) are per orbit.
 * order (device, nullable): the order in which the persistent kernel's work
 * queue hands orbits to lanes (a permutation of 0..N-1; longest first keeps
 * the tail short).  With an order, status is first set to
 * GAMD_STATUS_NOT_RUN for every orbit; entries outside 0..N-1 are skipped,
 * so an orbit that the order does not name keeps that status (the device
 * array is not read back to check it; gala_amd.integrate.dop853 checks that
 * a caller's order is a permutation before the call).  n_tail >= 0: the
 * first n_tail orbits of that order are integrated from the start by
 * dedicated waves, four orbits per wave, each spread over 16 lanes (capped
 * at half the grid); results, dense output included, are identical for
 * every n_tail.  block: threads per workgroup,
 * 64, 128, 192 or 256 (0 = 256); results are identical for every block.
 * Both are speed knobs only: gala_amd.autotune times them on the caller's
 * workload. */
int gamd_dop853(const gamd_potential *pot, const double *w0, size_t N, const double *t0,
                const double *t1, const double *h0, double rtol, double atol, int64_t nmax,
                int64_t nstiff, double hmax, const double *t_out, size_t n_out, double *w_out,
                double *w_final, int32_t *status, int64_t *stats, const int32_t *order,
                int64_t n_tail, int block, void *stream);

/* gamd_dop853 for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Arguments, results and errors are
 * gamd_dop853's; the work-queue counter, which gamd_dop853 allocates per
 * call, lives instead in a fixed device array of the library
 * (GAMD_DOP853_SCRATCH_BYTES per GPU, held from the library's load on that
 * GPU until the process ends).  No input size is limited by it: the counter
 * does not grow with N.  The call makes no hipMalloc, hipFree or other
 * allocating call; it clears the counter with a kernel (no memset node).
 * Every call on one GPU shares that GPU's array, so calls that overlap on one
 * GPU (concurrent streams, or threads) race: serialise them.  The GPU is the
 * current device of the calling thread. */
#define GAMD_DOP853_SCRATCH_BYTES 256
int gamd_dop853_noalloc(const gamd_potential *pot, const double *w0, size_t N,
                        const double *t0, const double *t1, const double *h0, double rtol,
                        double atol, int64_t nmax, int64_t nstiff, double hmax,
                        const double *t_out, size_t n_out, double *w_out, double *w_final,
                        int32_t *status, int64_t *stats, const int32_t *order, int64_t n_tail,
                        int block, void *stream);

/* gamd_dop853's dense output streamed to HOST memory while the integration
 * runs (gala's dop853_integrate_hamiltonian returns its dense output on the
 * host, dop853.pyx:228-250; the samples are those of dop853.cpp:584-612,
 * 869-904).  The n_out shared times t_out (device) are taken in windows of
 * window_rows samples: one launch per window carries every orbit to the
 * first step boundary past the window's last sample (its state kept in the
 * workspace between windows; a step that also covers the next window's first
 * sample keeps its dense-output polynomial for that window) and writes the
 * window's samples to a slot of a two-slot device ring, copied to host_out
 * (HOST, ideally pinned, (6, n_out, N)) on a second stream while the next
 * window runs.  Steps, samples, w_final, status and stats equal gamd_dop853's
 * with w_out bit for bit; a sample outside an orbit's [t0, t1] is not
 * written (gamd_dop853 leaves it unwritten too).  The launch layouts are
 * gamd_dop853's (order, n_tail dedicated rows, quads and rows at the queue's
 * end; a row or quad stops at a window's end as a lane does).  use_graph != 0
 * captures the window kernels and copies as one HIP graph and launches that
 * (no memset nodes).  workspace: device,
 * gamd_dop853_dense_workspace_bytes(N, window_rows) bytes (the work-queue
 * counter included: the call allocates no device memory).  Returns when
 * host_out is complete (it synchronises `stream`).  Other arguments as
 * gamd_dop853.  No gala counterpart for the windows themselves; window_rows
 * is a speed knob (gala_amd.autotune picks it). */
int64_t gamd_dop853_dense_workspace_bytes(size_t N, int64_t window_rows);
int gamd_dop853_dense_stream(const gamd_potential *pot, const double *w0, size_t N,
                             const double *t0, const double *t1, const double *h0, double rtol,
                             double atol, int64_t nmax, int64_t nstiff, double hmax,
                             const double *t_out, size_t n_out, int64_t window_rows,
                             int use_graph, double *host_out, double *w_final, int32_t *status,
                             int64_t *stats, const int32_t *order, int64_t n_tail, int block,
                             void *workspace, size_t workspace_bytes, void *stream);

#define GAMD_DOP853_MAX_NBATCH 1024  /* orbits of one batched system (one workgroup) */

/* Replaces dop853_integrate_hamiltonian with gala's default batching
 * (dop853.pyx:196-250; Hamiltonian.integrate_orbit passes nbatch=100,
 * chamiltonian.pyx:348): orbits [b nbatch, min((b+1) nbatch, N)) form one
 * (6, nb) system with ONE shared adaptive step (dop853_helper, :90-193),
 * integrated from t0 to t1 (host scalars) with initial step h0 (gala passes
 * t[1]-t[0], dop853.pyx:171; 0 = hinit over the batch) and hmax = |t1-t0|
 * when hmax <= 0.  The error norm and the stiffness test sum over the
 * system's 6 nb components in the reference's order (dop853.cpp:416-444,
 * 463-486), so one orbit's error shapes the whole batch's steps as in gala.
 * nmax / nstiff defaults as gamd_dop853.  Dense output at t_out (device,
 * n_out) into w_out (6, n_out, N) when w_out != NULL.  w_final (6, N);
 * status (N) and stats (N, 4) (nullable) hold each orbit's batch's idid and
 * nstep, naccpt, nrejct, nfcn.  1 <= nbatch <= GAMD_DOP853_MAX_NBATCH
 * (GAMD_ERR_UNSUPPORTED above).  With more batches than the GPU holds at
 * once, the batches run in two launches: every batch takes 192 attempts
 * (GAMD_DOP853_BATCH_PILOT) and is kept, then the batches resume in order of
 * their predicted remaining attempts, longest first; the kept states (29 N
 * doubles) are allocated stream-ordered for the call.  Same results as one
 * launch (GAMD_DOP853_BATCH_ORDER=0, which allocates nothing; =1 forces the
 * two launches). */
int gamd_dop853_batched(const gamd_potential *pot, const double *w0, size_t N, double t0,
                        double t1, double h0, double rtol, double atol, int64_t nmax,
                        int64_t nstiff, double hmax, int nbatch, const double *t_out,
                        size_t n_out, double *w_out, double *w_final, int32_t *status,
                        int64_t *stats, void *stream);

/* ---- N-body systems ------------------------------------------------------ */
#define GAMD_MAX_BODIES 64   /* bodies of one DOP853 N-body system (one wave) */
#define GAMD_MAX_MASSIVE 8   /* bodies carried by every lane of the Leapfrog N-body kernels */

/* Body potentials: `bodies` is a HOST array of nbody potential handles for
 * bodies 0..nbody-1; body j's potential acts on every other body, centred on
 * body j's position (c_nbody_acceleration / c_nbody_gradient_symplectic set
 * do_shift_rotate = 1 and q0 = the body's position, keeping each
 * component's R, cpotential.cpp:389-442).  A NULL entry is a massless body
 * (NullPotential, skipped as the reference skips pots[j]->null).  Bodies
 * from index nbody on are massless test particles (DirectNBody puts the
 * massive bodies first, dynamics/nbody/core.py:212-227). */

/* Replaces leapfrog_integrate_nbody (leapfrog.pyx:161-257): n bodies, w0
 * (6, n), fixed dt, n_steps steps.  Within a step bodies move in index order
 * (c_leapfrog_step_nbody, leapfrog.pyx:140-158, :241-252): body i's kick sees
 * bodies below i already drifted and bodies above i not yet; test particles
 * see every massive body drifted.  nbody <= GAMD_MAX_MASSIVE.  w_out (6,
 * nsave, n) and w_final (6, n) as gamd_leapfrog. */
int gamd_nbody_leapfrog(const gamd_potential *ext, const gamd_potential *const *bodies, int nbody,
                        const double *w0, size_t n, double t0, double dt, int64_t n_steps,
                        int64_t save_stride, double *w_out, double *w_final, void *stream);

/* Replaces ruth4_integrate_nbody (ruth4.pyx:138-261): arguments and layouts
 * as gamd_nbody_leapfrog.  Within a step each body takes all four
 * kick-then-drift substeps (c_ruth4_step_nbody, :119-136) before the next
 * body moves, in index order; test particles move last. */
int gamd_nbody_ruth4(const gamd_potential *ext, const gamd_potential *const *bodies, int nbody,
                     const double *w0, size_t n, double t0, double dt, int64_t n_steps,
                     int64_t save_stride, double *w_out, double *w_final, void *stream);

/* Replaces direct_nbody_dop853 (dynamics/nbody/nbody.pyx:30-115): the n
 * bodies form ONE 6n-dimensional system (body-major state, the reference's
 * (n, 6) rows) with one shared adaptive step and Fwrapper_direct_nbody
 * forces (dop853.cpp:990-1006), integrated from t0 to t1 with initial step
 * h0 (nonzero; the reference passes t[1] - t[0]), uround = machine epsilon,
 * nmax / nstiff / hmax as gamd_dop853 (direct_nbody_dop853 passes nstiff
 * -1: off).  n <= GAMD_MAX_BODIES, nbody <= GAMD_MAX_MASSIVE.  Dense output
 * at the n_out times t_out into w_out (6, n_out, n) when w_out is non-NULL;
 * w_final (6, n); status (1) gala's idid; stats (4) nullable. */
int gamd_nbody_dop853(const gamd_potential *ext, const gamd_potential *const *bodies, int nbody,
                      const double *w0, size_t n, double t0, double t1, double h0, double rtol,
                      double atol, int64_t nmax, int64_t nstiff, double hmax, const double *t_out,
                      size_t n_out, double *w_out, double *w_final, int32_t *status,
                      int64_t *stats, void *stream);

/* ---- mock streams ------------------------------------------------------- */
/* Replaces mockstream_leapfrog's release-and-integrate loop
 * (mockstream.pyx:551-599): the system bodies 0..nbody-1 (body 0 the
 * progenitor, then any perturbers; potentials as above) and massless stream
 * particles.
 *   body_traj (6, nfull, nbody): the bodies on the forward grid t0 + k dt
 *     (the leapfrog_integrate_nbody trajectory of mockstream.pyx:536-549);
 *   w0 (6, n): particle states at release; rel_t (n): release times;
 *   out (6, n): particle states at tfinal.
 * Every particle integrates the bodies again from their states at its
 * release alongside itself (bodies first, in index order), exactly as each
 * release batch does in the reference.  nbody <= GAMD_MAX_MASSIVE. */
int gamd_mockstream_leapfrog(const gamd_potential *ext, const gamd_potential *const *bodies,
                             int nbody, const double *body_traj, size_t nfull, double t0,
                             double dt, const double *w0, const double *rel_t, size_t n,
                             double tfinal, double *out, void *stream);

/* Replaces the release loop of mockstream_dop853 (mockstream.pyx:286-303),
 * gala's default mock-stream integrator: release i's system is the bodies
 * 0..nbody-1 at body_w[i] (nbody, 6) (their states at the release time from
 * the dense output of the bodies' own DOP853 integration,
 * mockstream.pyx:274-282) followed by the nstream[i] particles in rows
 * offset[i].. of stream_w0 (ntot, 6); the (nbody + k, 6) system is
 * integrated with one shared adaptive step from rel_t[i] to tfinal with the
 * arguments dop853_step passes (dop853.pyx:29-79: h = dt0, uround 2.3e-16,
 * nstiff 1, hmax |tfinal - t|) and Fwrapper_direct_nbody forces
 * (dop853.cpp:990-1006).  All arrays are device arrays; nbody +
 * nstream[i] <= GAMD_MAX_BODIES (status GAMD_ERR_UNSUPPORTED otherwise),
 * nbody <= GAMD_MAX_MASSIVE.  out (ntot, 6): particle states at tfinal;
 * body_final (nrel, nbody, 6): the bodies of each system at tfinal; status
 * (nrel): gala's idid (0 for releases without particles; such a release
 * leaves its rows of body_final and stats untouched); stats (nrel, 4)
 * nullable.  Systems of at most 4 bodies run one per 16-lane row (up to
 * 2 x 64 x 4 x CUs / 16 systems) or one per 4-lane quad (more systems);
 * gamd_nbody_dop853 likewise runs at most 4 bodies on one row, more on one
 * wave.  Environment GAMD_DOP853_BODY_ROWS=0 / =1 forces one body per lane /
 * rows, with identical results.  GAMD_DOP853_WIDE=0 / =1 likewise switches
 * off gamd_dop853's 16-lane / 4-lane tail stages (identical results). */
int gamd_mockstream_dop853(const gamd_potential *ext, const gamd_potential *const *bodies,
                           int nbody, const double *body_w, size_t nrel, const double *rel_t,
                           const int32_t *nstream, const int64_t *offset, const double *stream_w0,
                           double tfinal, double dt0, double rtol, double atol, int64_t nmax,
                           double *out, double *body_final, int32_t *status, int64_t *stats,
                           void *stream);

/* ---- learned SCF-force surrogate --------------------------------------- */
/* No counterpart in gala: a 4 -> 256 x 4 (SiLU) -> 4 MLP standing in for the
 * SCF force of scf/src/bfe.cpp:114-189, on BF16 matrix cores with FP32
 * accumulation.  Parameters: one flat FP32 master buffer theta of
 * gamd_mlp_param_count() values, layout W1[4][256] b1[256] W2[256][256] b2
 * W3 b3 W4 b4 W5[256][4] b5[4] (W[in][out]); BF16 copies wbf of
 * gamd_mlp_bf16_count() values made by gamd_mlp_sync_bf16.
 * Features of a point q: (x/r, y/r, z/r, (ln r - mu)/sigma).  The network's
 * potential is Phi_net(q) = Phi_b(r) (1 + phi_c o0) with the Hernquist
 * monopole Phi_b = -gm / (r + rs) of the expansion, and the orbit force is
 * -grad Phi_net (exact reverse-mode derivative of that potential, so
 * surrogate orbits are a conservative system).  Outputs o1..o3 are auxiliary
 * training targets, grad Phi / s_grad(r), with s_grad a radial scale profile
 * (typically the RMS of one exact gradient component over shells),
 * piecewise linear in ln r between GAMD_MLP_KNOTS knots at ln r = lnr0 +
 * k / inv_dlnr and constant beyond the end knots. */
#define GAMD_MLP_KNOTS 32
typedef struct {
    double mu, inv_sigma;    /* ln r feature: (ln r - mu) * inv_sigma */
    double lnr0, inv_dlnr;   /* knot positions of the s_grad profile */
    double pad_r;            /* any positive radius, used for padding rows */
    const double *scale;     /* DEVICE [GAMD_MLP_KNOTS]: s_grad knots */
    double gm, rs;           /* the monopole Phi_b = -gm / (r + rs), gm = G M */
    double phi_c;            /* scale of o0: Phi_net = Phi_b (1 + phi_c o0) */
} gamd_mlp_norm;

int64_t gamd_mlp_param_count(void);
int64_t gamd_mlp_bf16_count(void);
/* bytes of workspace gamd_mlp_gradients needs for a batch of M points */
int64_t gamd_mlp_workspace_bytes(int64_t M);
/* BF16 copies of the hidden weights (high and low parts in matrix-core
 * fragment order, their transposes, and row-major high parts) */
int gamd_mlp_sync_bf16(const float *theta, uint16_t *wbf, void *stream);
/* inference: q (3,M) FP64 -> Phi_net (M) and grad Phi_net (3,M) FP64, the
 * hidden products carried to ~16 significant bits (three BF16 products of
 * a high/low split) */
int gamd_mlp_forward(const float *theta, const uint16_t *wbf, const gamd_mlp_norm *norm,
                     const double *q, size_t M, double *phi, double *grad, void *stream);
/* training data: exact phi (M) and grad (3,M) at q (3,M) -> features
 * (M,4) and scaled targets (M,4), FP32 */
int gamd_mlp_make_targets(const gamd_mlp_norm *norm, const double *q, const double *phi,
                          const double *grad, size_t M, float *feat, float *target, void *stream);
/* forward + backward of the mean squared error over a batch (M a multiple
 * of 64): grad (param_count FP32) overwritten; loss[0] = sum of squared
 * errors over the batch (nullable) */
int gamd_mlp_gradients(const float *theta, const uint16_t *wbf, const float *feat,
                       const float *target, size_t M, float *grad, float *loss, void *workspace,
                       size_t workspace_bytes, void *stream);
/* one Adam step on the FP32 masters (bias-corrected, step >= 1), then the
 * BF16 copies are refreshed */
int gamd_mlp_adam(float *theta, const float *grad, float *m, float *v, int64_t step, float lr,
                  float beta1, float beta2, float eps, uint16_t *wbf, void *stream);
/* One leapfrog step of N orbits with surrogate forces (the loop of
 * leapfrog.pyx:106-116 with the network in place of c_gradient).  w (6,N)
 * holds positions and half-step velocities between calls; vsync (3,N) the
 * synchronised velocity after the step; E0 (N) the initial energy (written
 * by mode 0, the initial half kick); fallback_step (N) is -1 while an orbit
 * uses the surrogate force -grad Phi_net and the step at which it switched
 * to the exact SCF force of `scf` (a single unshifted SCF component) when its
 * surrogate energy Phi_net + |v|^2/2 drifted by more than tol |E0|; from
 * that step on the orbit follows gamd_leapfrog's arithmetic and SCF
 * evaluator exactly, and w_switch (6,N, nullable) receives its state before
 * the step (position and half-step velocity, gamd_leapfrog_resume's input).  mode 0 is the initial half
 * kick (and sets E0), mode 1 one drift + kick. */
int gamd_surrogate_leapfrog_step(const float *theta, const uint16_t *wbf, const gamd_mlp_norm *norm,
                                 const gamd_potential *scf, double *w, double *vsync, double *E0,
                                 int32_t *fallback_step, double *w_switch, size_t N, double dt,
                                 int mode, int32_t step, double tol, void *stream);

/* Library version string (host). */
const char *gamd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GALA_AMD_H */
