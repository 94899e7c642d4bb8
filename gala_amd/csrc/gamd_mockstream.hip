// gamd_mockstream.hip -- batched release-and-integrate of mock-stream particles.
//
// C ABI: gamd_mockstream_leapfrog (include/gala_amd.h).  Restates the
// per-release loop of mockstream_leapfrog (dynamics/mockstream/mockstream.pyx:
// 551-599) for one massive body, the progenitor, plus massless particles:
//   * the release batch starts from the progenitor state on the forward grid
//     (idx = int((t_i - t_0)/dt + 0.5), :546) and the particles' release
//     states;
//   * half kicks for every body (c_init_velocity_nbody, :571-578), then
//     n_steps = int((tfinal - t_i)/dt + 0.5) leapfrog steps (:581-590) in
//     which the progenitor moves first and each particle's kick adds the
//     progenitor potential evaluated around the progenitor's drifted position
//     (c_leapfrog_step_nbody / c_nbody_gradient_symplectic, leapfrog.pyx:
//     140-158, cpotential.cpp:418-442);
//   * the progenitor does not feel the particles (NullPotential bodies).
// GPU structure: one lane per particle, and that lane carries its own copy of
// the progenitor from the particle's release state, so the uneven per-release
// lengths (0 .. nfull-1 steps) become independent lanes in one launch and each
// lane reproduces the reference's per-release progenitor restart exactly.
#include <string.h>

#include "gamd_host.hpp"

using namespace gamd;

namespace {

// gradient of the progenitor potential at q, centred on xp, added to g
template <bool SCF>
__device__ __forceinline__ void prog_grad(const DevPot &PP, double x, double y, double z,
                                          double px, double py, double pz, double &gx,
                                          double &gy, double &gz) {
    double fx, fy, fz;
    body_grad<SCF>(PP, x, y, z, px, py, pz, fx, fy, fz);
    gx += fx;
    gy += fy;
    gz += fz;
}

template <bool SCF>
__global__ __launch_bounds__(64) void k_mockstream(const DevPot *__restrict__ Ep, const DevPot *__restrict__ PPp, int has_prog,
                                                   const double *__restrict__ traj, int64_t nfull,
                                                   double t0, double dt,
                                                   const double *__restrict__ w0,
                                                   const double *__restrict__ rel_t, int64_t n,
                                                   double tfinal, double *__restrict__ out) {
    const DevPot &E = *Ep;
    const DevPot &PP = *PPp;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t idx = (int64_t)((rel_t[i] - t0) / dt + 0.5);
    if (idx < 0) idx = 0;
    if (idx >= nfull) idx = nfull - 1;
    const int64_t nsteps = (int64_t)((tfinal - rel_t[i]) / dt + 0.5);
    // progenitor at release
    double px = traj[idx], py = traj[nfull + idx], pz = traj[2 * nfull + idx];
    double pvx = traj[3 * nfull + idx], pvy = traj[4 * nfull + idx], pvz = traj[5 * nfull + idx];
    // particle
    double x = w0[i], y = w0[n + i], z = w0[2 * n + i];
    double vx = w0[3 * n + i], vy = w0[4 * n + i], vz = w0[5 * n + i];
    double gx, gy, gz;
    // c_init_velocity_nbody: progenitor (body 0), then the particle
    pot_grad<SCF>(E, px, py, pz, gx, gy, gz);
    double phx = pvx - gx * dt / 2.0, phy = pvy - gy * dt / 2.0, phz = pvz - gz * dt / 2.0;
    pot_grad<SCF>(E, x, y, z, gx, gy, gz);
    if (has_prog) prog_grad<SCF>(PP, x, y, z, px, py, pz, gx, gy, gz);
    double hx = vx - gx * dt / 2.0, hy = vy - gy * dt / 2.0, hz = vz - gz * dt / 2.0;
    for (int64_t s = 0; s < nsteps; s++) {
        px = px + phx * dt;
        py = py + phy * dt;
        pz = pz + phz * dt;
        pot_grad<SCF>(E, px, py, pz, gx, gy, gz);
        phx = phx - gx * dt;
        phy = phy - gy * dt;
        phz = phz - gz * dt;
        x = x + hx * dt;
        y = y + hy * dt;
        z = z + hz * dt;
        pot_grad<SCF>(E, x, y, z, gx, gy, gz);
        if (has_prog) prog_grad<SCF>(PP, x, y, z, px, py, pz, gx, gy, gz);
        vx = hx - gx * dt / 2.0;
        vy = hy - gy * dt / 2.0;
        vz = hz - gz * dt / 2.0;
        hx = hx - gx * dt;
        hy = hy - gy * dt;
        hz = hz - gz * dt;
    }
    out[i] = x;
    out[n + i] = y;
    out[2 * n + i] = z;
    out[3 * n + i] = vx;
    out[4 * n + i] = vy;
    out[5 * n + i] = vz;
}

}  // namespace

extern "C" int gamd_mockstream_leapfrog(const gamd_potential *ext, const gamd_potential *prog,
                                        const double *prog_traj, size_t nfull, double t0, double dt,
                                        const double *w0, const double *rel_t, size_t n,
                                        double tfinal, double *out, void *stream) {
    if (!ext || !prog_traj || nfull == 0 || !w0 || !rel_t || !out || dt == 0.0)
        return GAMD_ERR_INVALID;
    if (n == 0) return GAMD_OK;
    const bool scf = ext->dev.has_scf || (prog && prog->dev.has_scf);
    const DevPot *pp = prog ? prog->ddev : ext->ddev;  // not read when there is no progenitor
    const int has_prog = prog ? 1 : 0;
    const unsigned block = 64;
    if (scf)
        hipLaunchKernelGGL(k_mockstream<true>, dim3(grid_for(n, block)), dim3(block), 0,
                           as_stream(stream), ext->ddev, pp, has_prog, prog_traj, (int64_t)nfull, t0, dt, w0,
                           rel_t, (int64_t)n, tfinal, out);
    else
        hipLaunchKernelGGL(k_mockstream<false>, dim3(grid_for(n, block)), dim3(block), 0,
                           as_stream(stream), ext->ddev, pp, has_prog, prog_traj, (int64_t)nfull, t0, dt, w0,
                           rel_t, (int64_t)n, tfinal, out);
    return launch_status();
}
