// gamd_leapfrog.hip -- one-lane-per-orbit FP64 symplectic integrators:
// leapfrog and Ruth4.
//
// C ABI: gamd_leapfrog, gamd_ruth4 (include/gala_amd.h).  gamd_leapfrog restates
// leapfrog_integrate_hamiltonian (integrate/cyintegrators/leapfrog.pyx:54-121):
// a half kick from the initial gradient (c_init_velocity, :24-32), then per
// step drift x += v_half dt, gradient at the new x, synchronised
// v = v_half - g dt/2 and v_half -= g dt (c_leapfrog_step, :35-51).
//
// Each lane keeps its orbit's 9 doubles (x, v, v_half) in VGPRs for the
// whole integration and touches HBM only to read w0 and write saved
// states: the (6, nsave, N) layout makes every store a coalesced 512-byte
// wave store.  The time loop runs inside the kernel, so there is one launch
// per integration, not one per step.
//
// gamd_ruth4 restates ruth4_integrate_hamiltonian (ruth4.pyx:24-106): per step
// four substeps j = 0..3 of a kick v -= d_j g(x) dt and a drift x += c_j v dt.
// d_0 = 0, so the first substep's kick is the identity for every finite
// gradient and its force evaluation is skipped (three evaluations per step
// instead of the reference's four; results differ only where the gradient
// itself is not finite, e.g. at r = 0 of an NFW halo, where the reference's
// 0 * NaN turns the velocity into NaN).
#include <math.h>

#include "gamd_host.hpp"

using namespace gamd;

namespace {

// RESUME: w0 holds positions and HALF-STEP velocities (the state the loop
// carries between steps), so the initial half kick is skipped
template <bool SCF, bool SAVE, int LAY, bool RESUME = false>
__global__ __launch_bounds__(256) void k_leapfrog(const DevPot *__restrict__ Pp, const double *__restrict__ w0,
                                                  int64_t N, double dt, int64_t n_steps,
                                                  const int64_t *__restrict__ nsteps_i,
                                                  int64_t stride, int64_t nsave,
                                                  double *__restrict__ wout,
                                                  double *__restrict__ wfinal) {
    const DevPot &P = *Pp;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= N) return;
    double x = w0[i], y = w0[N + i], z = w0[2 * N + i];
    double vx = w0[3 * N + i], vy = w0[4 * N + i], vz = w0[5 * N + i];
    const int64_t ns = nsteps_i ? nsteps_i[i] : n_steps;
    if (SAVE) {
        const int64_t s = (int64_t)nsave * N;
        wout[i] = x;
        wout[s + i] = y;
        wout[2 * s + i] = z;
        wout[3 * s + i] = vx;
        wout[4 * s + i] = vy;
        wout[5 * s + i] = vz;
    }
    double gx, gy, gz;
    double hx = vx, hy = vy, hz = vz;
    if (!RESUME) {
        pot_grad_l<SCF, LAY>(P, x, y, z, gx, gy, gz);
        hx = vx - gx * dt / 2.0;
        hy = vy - gy * dt / 2.0;
        hz = vz - gz * dt / 2.0;
    }
    int64_t next_save = stride, o = i;
    for (int64_t j = 1; j <= ns; j++) {
        x = x + hx * dt;
        y = y + hy * dt;
        z = z + hz * dt;
        pot_grad_l<SCF, LAY>(P, x, y, z, gx, gy, gz);
        vx = hx - gx * dt / 2.0;
        vy = hy - gy * dt / 2.0;
        vz = hz - gz * dt / 2.0;
        hx = hx - gx * dt;
        hy = hy - gy * dt;
        hz = hz - gz * dt;
        if (SAVE && j == next_save) {
            next_save += stride;
            o += N;  // saved row j / stride, kept as a running offset (no 64-bit divide)
            const int64_t s = (int64_t)nsave * N;
            wout[o] = x;
            wout[s + o] = y;
            wout[2 * s + o] = z;
            wout[3 * s + o] = vx;
            wout[4 * s + o] = vy;
            wout[5 * s + o] = vz;
        }
    }
    if (wfinal) {
        wfinal[i] = x;
        wfinal[N + i] = y;
        wfinal[2 * N + i] = z;
        wfinal[3 * N + i] = vx;
        wfinal[4 * N + i] = vy;
        wfinal[5 * N + i] = vz;
    }
}

struct Ruth4Coef {
    double c[4], d[4];
};

template <bool SCF, bool SAVE, int LAY>
__global__ __launch_bounds__(256) void k_ruth4(const DevPot *__restrict__ Pp, const double *__restrict__ w0,
                                               int64_t N, double dt, int64_t n_steps, int64_t stride,
                                               int64_t nsave, Ruth4Coef cf, double *__restrict__ wout,
                                               double *__restrict__ wfinal) {
    const DevPot &P = *Pp;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= N) return;
    double x = w0[i], y = w0[N + i], z = w0[2 * N + i];
    double vx = w0[3 * N + i], vy = w0[4 * N + i], vz = w0[5 * N + i];
    if (SAVE) {
        const int64_t s = (int64_t)nsave * N;
        wout[i] = x;
        wout[s + i] = y;
        wout[2 * s + i] = z;
        wout[3 * s + i] = vx;
        wout[4 * s + i] = vy;
        wout[5 * s + i] = vz;
    }
    int64_t next_save = stride, o = i;
    for (int64_t j = 1; j <= n_steps; j++) {
        // substep 0: d_0 = 0, drift only
        x = x + cf.c[0] * vx * dt;
        y = y + cf.c[0] * vy * dt;
        z = z + cf.c[0] * vz * dt;
#pragma unroll
        for (int k = 1; k < 4; k++) {
            double gx, gy, gz;
            pot_grad_l<SCF, LAY>(P, x, y, z, gx, gy, gz);
            vx = vx - cf.d[k] * gx * dt;
            vy = vy - cf.d[k] * gy * dt;
            vz = vz - cf.d[k] * gz * dt;
            x = x + cf.c[k] * vx * dt;
            y = y + cf.c[k] * vy * dt;
            z = z + cf.c[k] * vz * dt;
        }
        if (SAVE && j == next_save) {
            next_save += stride;
            o += N;  // saved row j / stride, kept as a running offset (no 64-bit divide)
            const int64_t s = (int64_t)nsave * N;
            wout[o] = x;
            wout[s + o] = y;
            wout[2 * s + o] = z;
            wout[3 * s + o] = vx;
            wout[4 * s + o] = vy;
            wout[5 * s + o] = vz;
        }
    }
    if (wfinal) {
        wfinal[i] = x;
        wfinal[N + i] = y;
        wfinal[2 * N + i] = z;
        wfinal[3 * N + i] = vx;
        wfinal[4 * N + i] = vy;
        wfinal[5 * N + i] = vz;
    }
}

constexpr unsigned LF_BLOCK = 64;  // one wave per workgroup spreads small N over all CUs

}  // namespace

extern "C" int gamd_leapfrog(const gamd_potential *pot, const double *w0, size_t N, double t0,
                             double dt, int64_t n_steps, const int64_t *n_steps_per_orbit,
                             int64_t save_stride, double *w_out, double *w_final, void *stream) {
    if (!pot || !w0 || n_steps < 0 || (w_out && save_stride <= 0) || (!w_out && !w_final))
        return GAMD_ERR_INVALID;
    if (w_out && n_steps_per_orbit) return GAMD_ERR_INVALID;  // saved grids need one length
    if (N == 0) return GAMD_OK;
    const int64_t nsave = w_out ? n_steps / save_stride + 1 : 0;
    const unsigned g = grid_for(N, LF_BLOCK);
    hipStream_t s = as_stream(stream);
    const bool scf = pot->dev.has_scf;
    const bool mw = pot->dev.layout == LAYOUT_MW;
    const bool scf106 = pot->dev.layout == LAYOUT_SCF106;
    auto kern = w_out ? (scf106 ? k_leapfrog<true, true, LAYOUT_SCF106>
                         : scf  ? k_leapfrog<true, true, LAYOUT_GENERIC>
                         : mw   ? k_leapfrog<false, true, LAYOUT_MW>
                                : k_leapfrog<false, true, LAYOUT_GENERIC>)
                      : (scf106 ? k_leapfrog<true, false, LAYOUT_SCF106>
                         : scf  ? k_leapfrog<true, false, LAYOUT_GENERIC>
                         : mw   ? k_leapfrog<false, false, LAYOUT_MW>
                                : k_leapfrog<false, false, LAYOUT_GENERIC>);
    hipLaunchKernelGGL(kern, dim3(g), dim3(LF_BLOCK), 0, s, pot->ddev, w0, (int64_t)N, dt, n_steps,
                       n_steps_per_orbit, save_stride, nsave, w_out, w_final);
    return launch_status();
}

extern "C" int gamd_leapfrog_resume(const gamd_potential *pot, const double *w_half, size_t N,
                                    double dt, int64_t n_steps, const int64_t *n_steps_per_orbit,
                                    double *w_final, void *stream) {
    if (!pot || !w_half || !w_final || n_steps < 0) return GAMD_ERR_INVALID;
    if (N == 0) return GAMD_OK;
    const unsigned g = grid_for(N, LF_BLOCK);
    const bool scf = pot->dev.has_scf;
    const bool mw = pot->dev.layout == LAYOUT_MW;
    const bool scf106 = pot->dev.layout == LAYOUT_SCF106;
    auto kern = scf106 ? k_leapfrog<true, false, LAYOUT_SCF106, true>
                : scf  ? k_leapfrog<true, false, LAYOUT_GENERIC, true>
                : mw   ? k_leapfrog<false, false, LAYOUT_MW, true>
                       : k_leapfrog<false, false, LAYOUT_GENERIC, true>;
    hipLaunchKernelGGL(kern, dim3(g), dim3(LF_BLOCK), 0, as_stream(stream), pot->ddev, w_half,
                       (int64_t)N, dt, n_steps, n_steps_per_orbit, (int64_t)1, (int64_t)0,
                       (double *)nullptr, w_final);
    return launch_status();
}

extern "C" int gamd_ruth4(const gamd_potential *pot, const double *w0, size_t N, double t0,
                          double dt, int64_t n_steps, int64_t save_stride, double *w_out,
                          double *w_final, void *stream) {
    (void)t0;  // static potentials only: the time argument of c_gradient is unused
    if (!pot || !w0 || n_steps < 0 || (w_out && save_stride <= 0) || (!w_out && !w_final))
        return GAMD_ERR_INVALID;
    if (N == 0) return GAMD_OK;
    // ruth4.pyx:67-80
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
    const unsigned g = grid_for(N, LF_BLOCK);
    hipStream_t s = as_stream(stream);
    const bool scf = pot->dev.has_scf;
    const bool mw = pot->dev.layout == LAYOUT_MW;
    const bool scf106 = pot->dev.layout == LAYOUT_SCF106;
    auto kern = w_out ? (scf106 ? k_ruth4<true, true, LAYOUT_SCF106>
                         : scf  ? k_ruth4<true, true, LAYOUT_GENERIC>
                         : mw   ? k_ruth4<false, true, LAYOUT_MW>
                                : k_ruth4<false, true, LAYOUT_GENERIC>)
                      : (scf106 ? k_ruth4<true, false, LAYOUT_SCF106>
                         : scf  ? k_ruth4<true, false, LAYOUT_GENERIC>
                         : mw   ? k_ruth4<false, false, LAYOUT_MW>
                                : k_ruth4<false, false, LAYOUT_GENERIC>);
    hipLaunchKernelGGL(kern, dim3(g), dim3(LF_BLOCK), 0, s, pot->ddev, w0, (int64_t)N, dt, n_steps,
                       save_stride, nsave, cf, w_out, w_final);
    return launch_status();
}
