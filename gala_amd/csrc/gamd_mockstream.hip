// gamd_mockstream.hip -- Leapfrog N-body systems and batched mock-stream
// release-and-integrate.
//
// C ABI: gamd_nbody_leapfrog, gamd_nbody_ruth4, gamd_mockstream_leapfrog
// (include/gala_amd.h).
//
// gamd_nbody_leapfrog restates leapfrog_integrate_nbody (integrate/
// cyintegrators/leapfrog.pyx:161-257): half kicks for every body
// (c_init_velocity_nbody, :126-137), then per step every body in index order
// drifts, evaluates the external gradient plus the potentials of the other
// massive bodies centred on their current positions (c_nbody_gradient_
// symplectic, potential/src/cpotential.cpp:418-442) and kicks
// (c_leapfrog_step_nbody, :140-158).  Because the bodies move one after the
// other, body i sees bodies below i already drifted and bodies above i not
// yet; massless test particles (after the massive bodies) see them all
// drifted and act on nothing.
//
// gamd_mockstream_leapfrog restates the per-release loop of
// mockstream_leapfrog (dynamics/mockstream/mockstream.pyx:551-599): each
// release batch starts from the bodies' states on the forward grid
// (idx = int((t_i - t_0)/dt + 0.5), :546) and the particles' release states
// and takes n_steps = int((tfinal - t_i)/dt + 0.5) steps (:581) of the same
// leapfrog.
//
// GPU structure, both kernels: one lane per test particle, and every lane
// carries its own copy of the massive bodies (at most GAMD_MAX_MASSIVE; up to
// four stay in registers, more spill to scratch).  The bodies' arithmetic is
// identical in every lane, so each lane reproduces the reference's bodies
// exactly, and the uneven per-release lengths (0 .. nfull-1 steps) become
// independent lanes of one launch.  The external potential is evaluated by
// the straight-line Milky-Way form when its table has that layout (the DOP853
// N-body kernels choose the same way), else by the general component loop.
#include <math.h>
#include <string.h>

#include "gamd_host.hpp"

using namespace gamd;

namespace {

// Gradient felt at (x, y, z) by body `self` (-1 for a test particle): the
// external potential, then every other body with a potential, centred on that
// body's position, in body order (c_gradient, then c_nbody_gradient_symplectic
// adding each body's f2).
template <bool SCF, int LAY, int KMAX>
__device__ __forceinline__ void nbody_grad(const DevPot &E, const BodyTable &bt, int self,
                                           const double (&bx)[KMAX][3], double x, double y,
                                           double z, double &gx, double &gy, double &gz) {
    pot_grad_l<SCF, LAY>(E, x, y, z, gx, gy, gz);
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        if (j < bt.n && j != self && bt.p[j]) {
            double fx, fy, fz;
            body_grad<SCF>(*bt.p[j], x, y, z, bx[j][0], bx[j][1], bx[j][2], fx, fy, fz);
            gx += fx;
            gy += fy;
            gz += fz;
        }
    }
}

// The massive bodies of one lane: positions, half-step velocities and (when
// the caller needs them) the synchronised velocities of the last step.
template <int KMAX>
struct Bodies {
    double x[KMAX][3], vh[KMAX][3], v[KMAX][3];
};

// c_init_velocity_nbody for bodies 0..K-1 (positions unchanged, so the order
// does not matter here)
template <bool SCF, int LAY, int KMAX>
__device__ __forceinline__ void bodies_init(const DevPot &E, const BodyTable &bt, int K,
                                            double dt, Bodies<KMAX> &B) {
#pragma unroll
    for (int i = 0; i < KMAX; i++) {
        if (i < K) {
            double gx, gy, gz;
            nbody_grad<SCF, LAY, KMAX>(E, bt, i, B.x, B.x[i][0], B.x[i][1], B.x[i][2], gx, gy, gz);
            B.vh[i][0] = B.v[i][0] - gx * dt / 2.0;
            B.vh[i][1] = B.v[i][1] - gy * dt / 2.0;
            B.vh[i][2] = B.v[i][2] - gz * dt / 2.0;
        }
    }
}

// one c_leapfrog_step_nbody for bodies 0..K-1, in index order
template <bool SCF, int LAY, int KMAX>
__device__ __forceinline__ void bodies_step(const DevPot &E, const BodyTable &bt, int K,
                                            double dt, Bodies<KMAX> &B) {
#pragma unroll
    for (int i = 0; i < KMAX; i++) {
        if (i < K) {
            B.x[i][0] = B.x[i][0] + B.vh[i][0] * dt;
            B.x[i][1] = B.x[i][1] + B.vh[i][1] * dt;
            B.x[i][2] = B.x[i][2] + B.vh[i][2] * dt;
            double gx, gy, gz;
            nbody_grad<SCF, LAY, KMAX>(E, bt, i, B.x, B.x[i][0], B.x[i][1], B.x[i][2], gx, gy, gz);
            B.v[i][0] = B.vh[i][0] - gx * dt / 2.0;
            B.v[i][1] = B.vh[i][1] - gy * dt / 2.0;
            B.v[i][2] = B.vh[i][2] - gz * dt / 2.0;
            B.vh[i][0] = B.vh[i][0] - gx * dt;
            B.vh[i][1] = B.vh[i][1] - gy * dt;
            B.vh[i][2] = B.vh[i][2] - gz * dt;
        }
    }
}

// a test particle's state and its init / step against the lane's bodies
struct Particle {
    double x, y, z, vx, vy, vz, hx, hy, hz;
};

template <bool SCF, int LAY, int KMAX>
__device__ __forceinline__ void particle_init(const DevPot &E, const BodyTable &bt,
                                              const Bodies<KMAX> &B, double dt, Particle &p) {
    double gx, gy, gz;
    nbody_grad<SCF, LAY, KMAX>(E, bt, -1, B.x, p.x, p.y, p.z, gx, gy, gz);
    p.hx = p.vx - gx * dt / 2.0;
    p.hy = p.vy - gy * dt / 2.0;
    p.hz = p.vz - gz * dt / 2.0;
}

template <bool SCF, int LAY, int KMAX>
__device__ __forceinline__ void particle_step(const DevPot &E, const BodyTable &bt,
                                              const Bodies<KMAX> &B, double dt, Particle &p) {
    p.x = p.x + p.hx * dt;
    p.y = p.y + p.hy * dt;
    p.z = p.z + p.hz * dt;
    double gx, gy, gz;
    nbody_grad<SCF, LAY, KMAX>(E, bt, -1, B.x, p.x, p.y, p.z, gx, gy, gz);
    p.vx = p.hx - gx * dt / 2.0;
    p.vy = p.hy - gy * dt / 2.0;
    p.vz = p.hz - gz * dt / 2.0;
    p.hx = p.hx - gx * dt;
    p.hy = p.hy - gy * dt;
    p.hz = p.hz - gz * dt;
}

// leapfrog_integrate_nbody: lane l integrates the K massive bodies and test
// particle K + l (when there is one); lane 0 writes the bodies.
template <bool SCF, int LAY, bool SAVE, int KMAX>
__global__ __launch_bounds__(64) void k_nbody_leapfrog(const DevPot *__restrict__ Ep, BodyTable bt,
                                                       const double *__restrict__ w0, int64_t n,
                                                       double dt, int64_t n_steps, int64_t stride,
                                                       int64_t nsave, double *__restrict__ wout,
                                                       double *__restrict__ wfinal) {
    const DevPot &E = *Ep;
    const int K = bt.n;
    const int64_t ntest = n - K;
    const int64_t lane = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (lane >= (ntest > 0 ? ntest : 1)) return;
    const bool lead = lane == 0;
    const bool has_p = ntest > 0;
    const int64_t col = K + lane;
    Bodies<KMAX> B;
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            B.x[j][c] = j < K ? w0[c * n + j] : 0.0;
            B.v[j][c] = j < K ? w0[(3 + c) * n + j] : 0.0;
            B.vh[j][c] = 0.0;
        }
    }
    Particle p{};
    if (has_p) {
        p.x = w0[col];
        p.y = w0[n + col];
        p.z = w0[2 * n + col];
        p.vx = w0[3 * n + col];
        p.vy = w0[4 * n + col];
        p.vz = w0[5 * n + col];
    }
    auto store = [&](double *dst, int64_t so, int64_t row) {
        // (6, rows, n) with row stride n: dst[(c * so + row) * n + col]
        if (lead) {
#pragma unroll
            for (int j = 0; j < KMAX; j++) {
                if (j < K) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        dst[(c * so + row) * n + j] = B.x[j][c];
                        dst[((3 + c) * so + row) * n + j] = B.v[j][c];
                    }
                }
            }
        }
        if (has_p) {
            dst[(0 * so + row) * n + col] = p.x;
            dst[(1 * so + row) * n + col] = p.y;
            dst[(2 * so + row) * n + col] = p.z;
            dst[(3 * so + row) * n + col] = p.vx;
            dst[(4 * so + row) * n + col] = p.vy;
            dst[(5 * so + row) * n + col] = p.vz;
        }
    };
    if (SAVE) store(wout, nsave, 0);
    bodies_init<SCF, LAY, KMAX>(E, bt, K, dt, B);
    if (has_p) particle_init<SCF, LAY, KMAX>(E, bt, B, dt, p);
    // the saved row s / stride as a running count (no 64-bit divide per step)
    int64_t next_save = stride, row = 0;
    for (int64_t s = 1; s <= n_steps; s++) {
        bodies_step<SCF, LAY, KMAX>(E, bt, K, dt, B);
        if (has_p) particle_step<SCF, LAY, KMAX>(E, bt, B, dt, p);
        if (SAVE && s == next_save) {
            next_save += stride;
            store(wout, nsave, ++row);
        }
    }
    if (wfinal) store(wfinal, 1, 0);
}

// Ruth4 coefficients (ruth4.pyx:67-80, the same in ruth4_integrate_nbody)
struct Ruth4Coef {
    double c[4], d[4];
};

// one c_ruth4_step_nbody (ruth4.pyx:119-136) of the position x, velocity v
// against the bodies' current positions: four kick-then-drift substeps.  The
// first kick has d_0 = 0 and adds exactly zero wherever the force is finite,
// so its force evaluation is skipped.
template <bool SCF, int LAY, int KMAX>
__device__ __forceinline__ void ruth4_substeps(const DevPot &E, const BodyTable &bt, int self,
                                               const double (&bx)[KMAX][3], double dt,
                                               const Ruth4Coef &cf, double (&x)[3],
                                               double (&v)[3]) {
#pragma unroll
    for (int c = 0; c < 3; c++) x[c] = x[c] + cf.c[0] * v[c] * dt;
#pragma unroll
    for (int k = 1; k < 4; k++) {
        double g[3];
        nbody_grad<SCF, LAY, KMAX>(E, bt, self, bx, x[0], x[1], x[2], g[0], g[1], g[2]);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            v[c] = v[c] - cf.d[k] * g[c] * dt;
            x[c] = x[c] + cf.c[k] * v[c] * dt;
        }
    }
}

// ruth4_integrate_nbody (ruth4.pyx:138-261): per step, body by body in index
// order, all four substeps of one body before the next (bodies below i have
// already moved a full step when body i moves); lane l carries the bodies
// and test particle K + l.  Lane 0 writes the bodies.
template <bool SCF, int LAY, bool SAVE, int KMAX>
__global__ __launch_bounds__(64) void k_nbody_ruth4(const DevPot *__restrict__ Ep, BodyTable bt,
                                                    const double *__restrict__ w0, int64_t n,
                                                    double dt, int64_t n_steps, int64_t stride,
                                                    int64_t nsave, Ruth4Coef cf,
                                                    double *__restrict__ wout,
                                                    double *__restrict__ wfinal) {
    const DevPot &E = *Ep;
    const int K = bt.n;
    const int64_t ntest = n - K;
    const int64_t lane = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (lane >= (ntest > 0 ? ntest : 1)) return;
    const bool lead = lane == 0;
    const bool has_p = ntest > 0;
    const int64_t col = K + lane;
    double bx[KMAX][3], bv[KMAX][3];
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            bx[j][c] = j < K ? w0[c * n + j] : 0.0;
            bv[j][c] = j < K ? w0[(3 + c) * n + j] : 0.0;
        }
    }
    double px[3] = {0.0, 0.0, 0.0}, pv[3] = {0.0, 0.0, 0.0};
    if (has_p) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            px[c] = w0[c * n + col];
            pv[c] = w0[(3 + c) * n + col];
        }
    }
    auto store = [&](double *dst, int64_t so, int64_t row) {
        if (lead) {
#pragma unroll
            for (int j = 0; j < KMAX; j++) {
                if (j < K) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        dst[(c * so + row) * n + j] = bx[j][c];
                        dst[((3 + c) * so + row) * n + j] = bv[j][c];
                    }
                }
            }
        }
        if (has_p) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                dst[(c * so + row) * n + col] = px[c];
                dst[((3 + c) * so + row) * n + col] = pv[c];
            }
        }
    };
    if (SAVE) store(wout, nsave, 0);
    int64_t next_save = stride, row = 0;  // saved row s / stride without a divide
    for (int64_t s = 1; s <= n_steps; s++) {
#pragma unroll
        for (int i = 0; i < KMAX; i++)
            if (i < K) ruth4_substeps<SCF, LAY, KMAX>(E, bt, i, bx, dt, cf, bx[i], bv[i]);
        if (has_p) ruth4_substeps<SCF, LAY, KMAX>(E, bt, -1, bx, dt, cf, px, pv);
        if (SAVE && s == next_save) {
            next_save += stride;
            store(wout, nsave, ++row);
        }
    }
    if (wfinal) store(wfinal, 1, 0);
}

// mockstream_leapfrog's release batches: lane i = particle i with its own
// copy of the bodies from their states at its release.
template <bool SCF, int LAY, int KMAX>
__global__ __launch_bounds__(64) void k_mockstream(const DevPot *__restrict__ Ep, BodyTable bt,
                                                   const double *__restrict__ traj, int64_t nfull,
                                                   double t0, double dt,
                                                   const double *__restrict__ w0,
                                                   const double *__restrict__ rel_t, int64_t n,
                                                   double tfinal, double *__restrict__ out) {
    const DevPot &E = *Ep;
    const int K = bt.n;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t idx = (int64_t)((rel_t[i] - t0) / dt + 0.5);
    if (idx < 0) idx = 0;
    if (idx >= nfull) idx = nfull - 1;
    const int64_t nsteps = (int64_t)((tfinal - rel_t[i]) / dt + 0.5);
    // bodies at release: traj (6, nfull, K)
    Bodies<KMAX> B;
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            B.x[j][c] = j < K ? traj[(c * nfull + idx) * K + j] : 0.0;
            B.v[j][c] = j < K ? traj[((3 + c) * nfull + idx) * K + j] : 0.0;
            B.vh[j][c] = 0.0;
        }
    }
    Particle p{};
    p.x = w0[i];
    p.y = w0[n + i];
    p.z = w0[2 * n + i];
    p.vx = w0[3 * n + i];
    p.vy = w0[4 * n + i];
    p.vz = w0[5 * n + i];
    // c_init_velocity_nbody: the bodies, then the particle
    bodies_init<SCF, LAY, KMAX>(E, bt, K, dt, B);
    particle_init<SCF, LAY, KMAX>(E, bt, B, dt, p);
    for (int64_t s = 0; s < nsteps; s++) {
        bodies_step<SCF, LAY, KMAX>(E, bt, K, dt, B);
        particle_step<SCF, LAY, KMAX>(E, bt, B, dt, p);
    }
    out[i] = p.x;
    out[n + i] = p.y;
    out[2 * n + i] = p.z;
    out[3 * n + i] = p.vx;
    out[4 * n + i] = p.vy;
    out[5 * n + i] = p.vz;
}

// The one-body stream (the progenitor alone, gala's common case) on lane
// pairs: lane 2i carries particle i's copy of the progenitor, lane 2i + 1
// the particle.  Both lanes drift, evaluate the external gradient at their
// own positions and kick with the same instructions, so the two external
// evaluations of a step run side by side instead of one after the other;
// the particle lane takes the progenitor's drifted position from its
// partner and adds the progenitor's own potential.  Each lane's arithmetic
// is that of bodies_step / particle_step (same expressions, same order), so
// the results equal k_mockstream's bit for bit.
template <bool SCF, int LAY>
__global__ __launch_bounds__(64) void k_mockstream_pair(const DevPot *__restrict__ Ep, BodyTable bt,
                                                        const double *__restrict__ traj,
                                                        int64_t nfull, double t0, double dt,
                                                        const double *__restrict__ w0,
                                                        const double *__restrict__ rel_t,
                                                        int64_t n, double tfinal,
                                                        double *__restrict__ out) {
    const DevPot &E = *Ep;
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t i = g >> 1;             // the particle of this pair
    const bool part = (g & 1) != 0;       // the particle's lane, else the progenitor's
    const int64_t ii = i < n ? i : n - 1;  // lanes past the last pair mirror it (no stores)
    int64_t idx = (int64_t)((rel_t[ii] - t0) / dt + 0.5);
    if (idx < 0) idx = 0;
    if (idx >= nfull) idx = nfull - 1;
    const int64_t nsteps = (int64_t)((tfinal - rel_t[ii]) / dt + 0.5);
    double X[3], V[3], H[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        X[c] = part ? w0[c * n + ii] : traj[(c * nfull + idx) * 1];
        V[c] = part ? w0[(3 + c) * n + ii] : traj[((3 + c) * nfull + idx) * 1];
    }
    const bool body_pot = bt.p[0] != nullptr;
    // c_init_velocity_nbody: the progenitor at its release state, the
    // particle against the progenitor there
    {
        double gx, gy, gz;
        pot_grad_l<SCF, LAY>(E, X[0], X[1], X[2], gx, gy, gz);
        const double bx = __shfl_xor(X[0], 1), by = __shfl_xor(X[1], 1), bz = __shfl_xor(X[2], 1);
        if (part && body_pot) {
            double fx, fy, fz;
            body_grad<SCF>(*bt.p[0], X[0], X[1], X[2], bx, by, bz, fx, fy, fz);
            gx += fx;
            gy += fy;
            gz += fz;
        }
        H[0] = V[0] - gx * dt / 2.0;
        H[1] = V[1] - gy * dt / 2.0;
        H[2] = V[2] - gz * dt / 2.0;
    }
    for (int64_t s = 0; s < nsteps; s++) {
        X[0] = X[0] + H[0] * dt;
        X[1] = X[1] + H[1] * dt;
        X[2] = X[2] + H[2] * dt;
        double gx, gy, gz;
        pot_grad_l<SCF, LAY>(E, X[0], X[1], X[2], gx, gy, gz);
        // the progenitor's position after this step's drift
        const double bx = __shfl_xor(X[0], 1), by = __shfl_xor(X[1], 1), bz = __shfl_xor(X[2], 1);
        if (part && body_pot) {
            double fx, fy, fz;
            body_grad<SCF>(*bt.p[0], X[0], X[1], X[2], bx, by, bz, fx, fy, fz);
            gx += fx;
            gy += fy;
            gz += fz;
        }
        V[0] = H[0] - gx * dt / 2.0;
        V[1] = H[1] - gy * dt / 2.0;
        V[2] = H[2] - gz * dt / 2.0;
        H[0] = H[0] - gx * dt;
        H[1] = H[1] - gy * dt;
        H[2] = H[2] - gz * dt;
    }
    if (part && i < n) {
        out[i] = X[0];
        out[n + i] = X[1];
        out[2 * n + i] = X[2];
        out[3 * n + i] = V[0];
        out[4 * n + i] = V[1];
        out[5 * n + i] = V[2];
    }
}

// The external potential's evaluator: the straight-line Milky-Way form when
// the table has that layout and no body is an SCF expansion (as the DOP853
// N-body kernels choose), else the general component loop, with the SCF
// branch compiled in only when some potential needs it.  Bodies per lane:
// 1, up to 4 (registers), up to 8 (the body arrays spill to scratch).
enum { EV_GENERIC, EV_SCF, EV_MW };

int evaluator(const gamd_potential *ext, const BodyTable &bt) {
    if (ext->dev.has_scf || bt.scf) return EV_SCF;
    return ext->dev.layout == LAYOUT_MW ? EV_MW : EV_GENERIC;
}

template <bool SCF, int LAY, bool SAVE>
auto leapfrog_kernel(int nbody) {
    return nbody <= 1   ? k_nbody_leapfrog<SCF, LAY, SAVE, 1>
           : nbody <= 4 ? k_nbody_leapfrog<SCF, LAY, SAVE, 4>
                        : k_nbody_leapfrog<SCF, LAY, SAVE, 8>;
}

template <bool SAVE>
auto leapfrog_kernel(int ev, int nbody) {
    return ev == EV_SCF  ? leapfrog_kernel<true, LAYOUT_GENERIC, SAVE>(nbody)
           : ev == EV_MW ? leapfrog_kernel<false, LAYOUT_MW, SAVE>(nbody)
                         : leapfrog_kernel<false, LAYOUT_GENERIC, SAVE>(nbody);
}

template <bool SCF, int LAY, bool SAVE>
auto ruth4_kernel(int nbody) {
    return nbody <= 1   ? k_nbody_ruth4<SCF, LAY, SAVE, 1>
           : nbody <= 4 ? k_nbody_ruth4<SCF, LAY, SAVE, 4>
                        : k_nbody_ruth4<SCF, LAY, SAVE, 8>;
}

template <bool SAVE>
auto ruth4_kernel(int ev, int nbody) {
    return ev == EV_SCF  ? ruth4_kernel<true, LAYOUT_GENERIC, SAVE>(nbody)
           : ev == EV_MW ? ruth4_kernel<false, LAYOUT_MW, SAVE>(nbody)
                         : ruth4_kernel<false, LAYOUT_GENERIC, SAVE>(nbody);
}

template <bool SCF, int LAY>
auto mockstream_kernel(int nbody) {
    return nbody <= 1   ? k_mockstream<SCF, LAY, 1>
           : nbody <= 4 ? k_mockstream<SCF, LAY, 4>
                        : k_mockstream<SCF, LAY, 8>;
}

}  // namespace

extern "C" int gamd_nbody_leapfrog(const gamd_potential *ext, const gamd_potential *const *bodies,
                                   int nbody, const double *w0, size_t n, double t0, double dt,
                                   int64_t n_steps, int64_t save_stride, double *w_out,
                                   double *w_final, void *stream) {
    (void)t0;  // static potentials only: the time argument of c_gradient is unused
    BodyTable bt;
    if (make_body_table(bodies, nbody, bt) != GAMD_OK) return GAMD_ERR_INVALID;
    if (!ext || !w0 || n_steps < 0 || (w_out && save_stride <= 0) || (!w_out && !w_final) ||
        (size_t)nbody > n)
        return GAMD_ERR_INVALID;
    if (n == 0) return GAMD_OK;
    const int64_t nsave = w_out ? n_steps / save_stride + 1 : 0;
    const int64_t lanes = (int64_t)n > nbody ? (int64_t)n - nbody : 1;
    const unsigned block = 64;
    const int ev = evaluator(ext, bt);
    auto kern = w_out ? leapfrog_kernel<true>(ev, nbody) : leapfrog_kernel<false>(ev, nbody);
    hipLaunchKernelGGL(kern, dim3(grid_for((size_t)lanes, block)), dim3(block), 0, as_stream(stream),
                       ext->ddev, bt, w0, (int64_t)n, dt, n_steps, w_out ? save_stride : 1, nsave,
                       w_out, w_final);
    return launch_status();
}

extern "C" int gamd_nbody_ruth4(const gamd_potential *ext, const gamd_potential *const *bodies,
                                int nbody, const double *w0, size_t n, double t0, double dt,
                                int64_t n_steps, int64_t save_stride, double *w_out,
                                double *w_final, void *stream) {
    (void)t0;  // static potentials only: the time argument of c_gradient is unused
    BodyTable bt;
    if (make_body_table(bodies, nbody, bt) != GAMD_OK) return GAMD_ERR_INVALID;
    if (!ext || !w0 || n_steps < 0 || (w_out && save_stride <= 0) || (!w_out && !w_final) ||
        (size_t)nbody > n)
        return GAMD_ERR_INVALID;
    if (n == 0) return GAMD_OK;
    // ruth4.pyx:174-188
    const double two_13 = pow(2.0, 1.0 / 3.0);
    Ruth4Coef cf;
    cf.c[0] = 1.0 / (2.0 * (2.0 - two_13));
    cf.c[1] = (1.0 - two_13) / (2.0 * (2.0 - two_13));
    cf.c[2] = (1.0 - two_13) / (2.0 * (2.0 - two_13));
    cf.c[3] = 1.0 / (2.0 * (2.0 - two_13));
    cf.d[0] = 0.0;
    cf.d[1] = 1.0 / (2.0 - two_13);
    cf.d[2] = -two_13 / (2.0 - two_13);
    cf.d[3] = 1.0 / (2.0 - two_13);
    const int64_t nsave = w_out ? n_steps / save_stride + 1 : 0;
    const int64_t lanes = (int64_t)n > nbody ? (int64_t)n - nbody : 1;
    const unsigned block = 64;
    const int ev = evaluator(ext, bt);
    auto kern = w_out ? ruth4_kernel<true>(ev, nbody) : ruth4_kernel<false>(ev, nbody);
    hipLaunchKernelGGL(kern, dim3(grid_for((size_t)lanes, block)), dim3(block), 0, as_stream(stream),
                       ext->ddev, bt, w0, (int64_t)n, dt, n_steps, w_out ? save_stride : 1, nsave,
                       cf, w_out, w_final);
    return launch_status();
}

extern "C" int gamd_mockstream_leapfrog(const gamd_potential *ext,
                                        const gamd_potential *const *bodies, int nbody,
                                        const double *body_traj, size_t nfull, double t0,
                                        double dt, const double *w0, const double *rel_t, size_t n,
                                        double tfinal, double *out, void *stream) {
    BodyTable bt;
    if (make_body_table(bodies, nbody, bt) != GAMD_OK) return GAMD_ERR_INVALID;
    if (!ext || nbody < 1 || !body_traj || nfull == 0 || !w0 || !rel_t || !out || dt == 0.0)
        return GAMD_ERR_INVALID;
    if (n == 0) return GAMD_OK;
    const int ev = evaluator(ext, bt);
    const unsigned block = 64;
    if (nbody == 1) {
        // the progenitor alone: lane pairs (k_mockstream_pair)
        auto pair = ev == EV_SCF  ? k_mockstream_pair<true, LAYOUT_GENERIC>
                    : ev == EV_MW ? k_mockstream_pair<false, LAYOUT_MW>
                                  : k_mockstream_pair<false, LAYOUT_GENERIC>;
        hipLaunchKernelGGL(pair, dim3(grid_for(2 * n, block)), dim3(block), 0,
                           as_stream(stream), ext->ddev, bt, body_traj, (int64_t)nfull, t0, dt,
                           w0, rel_t, (int64_t)n, tfinal, out);
        return launch_status();
    }
    auto kern = ev == EV_SCF  ? mockstream_kernel<true, LAYOUT_GENERIC>(nbody)
                : ev == EV_MW ? mockstream_kernel<false, LAYOUT_MW>(nbody)
                              : mockstream_kernel<false, LAYOUT_GENERIC>(nbody);
    hipLaunchKernelGGL(kern, dim3(grid_for(n, block)), dim3(block), 0, as_stream(stream),
                       ext->ddev, bt, body_traj, (int64_t)nfull, t0, dt, w0, rel_t, (int64_t)n,
                       tfinal, out);
    return launch_status();
}
