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
 * the tail short).  block: threads per workgroup (0 = autotuned). */
int gamd_dop853(const gamd_potential *pot, const double *w0, size_t N, const double *t0,
                const double *t1, const double *h0, double rtol, double atol, int64_t nmax,
                int64_t nstiff, double hmax, const double *t_out, size_t n_out, double *w_out,
                double *w_final, int32_t *status, int64_t *stats, const int32_t *order,
                int block, void *stream);

/* ---- mock streams ------------------------------------------------------- */
/* Replaces mockstream_leapfrog's release-and-integrate loop
 * (mockstream.pyx:551-599) for one massive body (the progenitor, whose
 * potential is prog, evaluated around the progenitor's position,
 * cpotential.cpp:418-442) and massless stream particles.
 *   prog_traj (6, nfull): progenitor states on the forward grid t0 + k dt
 *     (the leapfrog_integrate_nbody trajectory of mockstream.pyx:536-549);
 *   w0 (6, n): particle states at release; rel_t (n): release times;
 *   out (6, n): particle states at tfinal.
 * Every particle integrates the progenitor again from its release state
 * alongside itself, exactly as each release batch does in the reference. */
int gamd_mockstream_leapfrog(const gamd_potential *ext, const gamd_potential *prog,
                             const double *prog_traj, size_t nfull, double t0, double dt,
                             const double *w0, const double *rel_t, size_t n, double tfinal,
                             double *out, void *stream);

/* ---- SCF surrogate (no reference counterpart) --------------------------- */
/* A 4 -> 256 x 4 (SiLU) -> 4 MLP standing in for the SCF force; see the
 * gamd_mlp_* entry points.  Weights are BF16 copies of FP32 masters. */
typedef struct {
    int32_t hidden;          /* 256 */
    int32_t layers;          /* 4 hidden layers */
    double in_scale[4];      /* feature scaling: x/L, y/L, z/L, (log r - mu)/sig */
    double in_shift[4];
    double out_scale[4];     /* Phi, dPhi/dx, dPhi/dy, dPhi/dz de-normalisation */
} gamd_mlp_desc;

/* Library version string (host). */
const char *gamd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GALA_AMD_H */
