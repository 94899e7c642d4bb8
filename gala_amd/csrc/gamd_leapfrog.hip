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

// ---------------------------------------------------------------------------
// Streamed saved trajectory: one chunk of `rows` saved rows [row0, row0 +
// rows) of the (6, nsave, N) trajectory, written to a device ring slot laid
// out (6, rows_max, N); the arithmetic is k_leapfrog's loop, so the streamed
// rows equal gamd_leapfrog's saved rows bit for bit.  The first chunk starts
// from w0 (row 0 is the initial state, then the initial half kick); later
// chunks continue from the carried state (positions and half-step
// velocities) in `state`.  The last chunk runs the steps after the last
// saved row (n_steps % stride) and writes w_final; the others update state.
struct ChunkArgs {
    const double *w0;
    double *state;     // (6, N): x, then the half-step velocity
    double *slot;      // (6, rows_max, N)
    double *w_final;   // (6, N), written by the last chunk
    int64_t N, rows_max, row0, rows, stride, n_steps;
    double dt;
    int first, last;
};

template <bool SCF, int LAY>
__global__ __launch_bounds__(256) void k_leapfrog_chunk(const DevPot *__restrict__ Pp, ChunkArgs a) {
    const DevPot &P = *Pp;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    const int64_t N = a.N, so = a.rows_max * N;
    double x, y, z, vx, vy, vz, hx, hy, hz, gx, gy, gz;
    int64_t r = 0;  // rows of this chunk written so far
    if (a.first) {
        x = a.w0[i];
        y = a.w0[N + i];
        z = a.w0[2 * N + i];
        vx = a.w0[3 * N + i];
        vy = a.w0[4 * N + i];
        vz = a.w0[5 * N + i];
        a.slot[i] = x;
        a.slot[so + i] = y;
        a.slot[2 * so + i] = z;
        a.slot[3 * so + i] = vx;
        a.slot[4 * so + i] = vy;
        a.slot[5 * so + i] = vz;
        r = 1;
        pot_grad_l<SCF, LAY>(P, x, y, z, gx, gy, gz);
        hx = vx - gx * a.dt / 2.0;
        hy = vy - gy * a.dt / 2.0;
        hz = vz - gz * a.dt / 2.0;
    } else {
        x = a.state[i];
        y = a.state[N + i];
        z = a.state[2 * N + i];
        hx = a.state[3 * N + i];
        hy = a.state[4 * N + i];
        hz = a.state[5 * N + i];
        vx = hx;  // replaced at the first step; every later chunk takes >= 1 step
        vy = hy;
        vz = hz;
    }
    const double dt = a.dt;
    // steps to the chunk's last saved row, then (last chunk) to n_steps
    const int64_t save_steps = (a.rows - r) * a.stride;
    const int64_t tail = a.last ? a.n_steps - (a.row0 + a.rows - 1) * a.stride : 0;
    int64_t next_save = a.stride;
    for (int64_t j = 1; j <= save_steps + tail; j++) {
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
        if (j <= save_steps && j == next_save) {
            next_save += a.stride;
            const int64_t o = r * N + i;
            a.slot[o] = x;
            a.slot[so + o] = y;
            a.slot[2 * so + o] = z;
            a.slot[3 * so + o] = vx;
            a.slot[4 * so + o] = vy;
            a.slot[5 * so + o] = vz;
            r++;
        }
    }
    if (a.last) {
        a.w_final[i] = x;
        a.w_final[N + i] = y;
        a.w_final[2 * N + i] = z;
        a.w_final[3 * N + i] = vx;
        a.w_final[4 * N + i] = vy;
        a.w_final[5 * N + i] = vz;
    } else {
        a.state[i] = x;
        a.state[N + i] = y;
        a.state[2 * N + i] = z;
        a.state[3 * N + i] = hx;
        a.state[4 * N + i] = hy;
        a.state[5 * N + i] = hz;
    }
}

// The chunk sequence on the compute stream s, with each chunk's rows copied
// to the host on stream c while the next chunk computes (two ring slots; a
// slot is rewritten only after its copy has finished).  Both streams end
// joined to s.
template <bool SCF, int LAY>
int stream_chunks(const DevPot *P, const double *w0, size_t N, double dt, int64_t n_steps,
                  int64_t stride, int64_t C, double *host_out, double *w_final, double *ring,
                  double *state, hipStream_t s, hipStream_t c, hipEvent_t comp,
                  hipEvent_t copied[2]) {
    const int64_t nsave = n_steps / stride + 1;
    const int64_t nchunk = (nsave + C - 1) / C;
    const unsigned g = grid_for(N, LF_BLOCK);
    for (int64_t k = 0; k < nchunk; k++) {
        const int sl = (int)(k & 1);
        double *slot = ring + (size_t)sl * 6 * C * N;
        if (k >= 2) GAMD_HIP_CHECK(hipStreamWaitEvent(s, copied[sl], 0));
        ChunkArgs a;
        a.w0 = w0;
        a.state = state;
        a.slot = slot;
        a.w_final = w_final;
        a.N = (int64_t)N;
        a.rows_max = C;
        a.row0 = k * C;
        a.rows = (k + 1) * C <= nsave ? C : nsave - k * C;
        a.stride = stride;
        a.n_steps = n_steps;
        a.dt = dt;
        a.first = k == 0;
        a.last = k == nchunk - 1;
        hipLaunchKernelGGL((k_leapfrog_chunk<SCF, LAY>), dim3(g), dim3(LF_BLOCK), 0, s, P, a);
        const int rc = launch_status();
        if (rc != GAMD_OK) return rc;
        GAMD_HIP_CHECK(hipEventRecord(comp, s));
        GAMD_HIP_CHECK(hipStreamWaitEvent(c, comp, 0));
        for (int q = 0; q < 6; q++)
            GAMD_HIP_CHECK(hipMemcpyAsync(host_out + ((size_t)q * nsave + a.row0) * N,
                                          slot + (size_t)q * C * N,
                                          (size_t)a.rows * N * sizeof(double),
                                          hipMemcpyDeviceToHost, c));
        GAMD_HIP_CHECK(hipEventRecord(copied[sl], c));
    }
    // join: the compute stream waits for the last copies
    GAMD_HIP_CHECK(hipStreamWaitEvent(s, copied[(nchunk - 1) & 1], 0));
    if (nchunk >= 2) GAMD_HIP_CHECK(hipStreamWaitEvent(s, copied[(nchunk - 2) & 1], 0));
    return GAMD_OK;
}

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

extern "C" int64_t gamd_leapfrog_stream_workspace_bytes(size_t N, int64_t chunk_rows) {
    if (chunk_rows < 1) return -1;
    return (int64_t)((2 * 6 * (size_t)chunk_rows * N + 6 * N) * sizeof(double));
}

extern "C" int gamd_leapfrog_stream(const gamd_potential *pot, const double *w0, size_t N,
                                    double t0, double dt, int64_t n_steps, int64_t save_stride,
                                    int64_t chunk_rows, int use_graph, double *host_out,
                                    double *w_final, void *workspace, size_t workspace_bytes,
                                    void *stream) {
    (void)t0;  // static potentials only
    if (!pot || !w0 || !host_out || !w_final || !workspace || n_steps < 0 || save_stride < 1 ||
        chunk_rows < 1)
        return GAMD_ERR_INVALID;
    if ((int64_t)workspace_bytes < gamd_leapfrog_stream_workspace_bytes(N, chunk_rows))
        return GAMD_ERR_INVALID;
    if (N == 0) return GAMD_OK;
    double *ring = static_cast<double *>(workspace);
    double *state = ring + 2 * 6 * (size_t)chunk_rows * N;
    hipStream_t s = as_stream(stream);
    hipStream_t c = nullptr, cs = nullptr;
    hipEvent_t comp = nullptr, copied[2] = {nullptr, nullptr};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int rc = GAMD_ERR_HIP;
    const bool scf = pot->dev.has_scf;
    const bool mw = pot->dev.layout == LAYOUT_MW;
    const bool scf106 = pot->dev.layout == LAYOUT_SCF106;
    auto run = scf106 ? stream_chunks<true, LAYOUT_SCF106>
               : scf  ? stream_chunks<true, LAYOUT_GENERIC>
               : mw   ? stream_chunks<false, LAYOUT_MW>
                      : stream_chunks<false, LAYOUT_GENERIC>;
    if (hipStreamCreateWithFlags(&c, hipStreamNonBlocking) != hipSuccess) goto done;
    if (hipEventCreateWithFlags(&comp, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&copied[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&copied[1], hipEventDisableTiming) != hipSuccess)
        goto done;
    if (!use_graph) {
        rc = run(pot->ddev, w0, N, dt, n_steps, save_stride, chunk_rows, host_out, w_final, ring,
                 state, s, c, comp, copied);
        // the events and the copy stream are released below: finish first
        if (rc == GAMD_OK && hipStreamSynchronize(s) != hipSuccess) rc = GAMD_ERR_HIP;
    } else {
        // the same sequence captured once on a private stream (the caller's
        // may be the legacy default stream, which cannot capture): kernel and
        // copy nodes, no memset; then launched as one graph on the caller's
        // stream, after its earlier work
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) goto done;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) goto done;
        rc = run(pot->ddev, w0, N, dt, n_steps, save_stride, chunk_rows, host_out, w_final, ring,
                 state, cs, c, comp, copied);
        const hipError_t ce = hipStreamEndCapture(cs, &graph);
        if (rc != GAMD_OK || ce != hipSuccess || !graph) {
            rc = GAMD_ERR_HIP;
            goto done;
        }
        rc = GAMD_ERR_HIP;
        if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) goto done;
        if (hipGraphLaunch(exec, s) != hipSuccess) goto done;
        if (hipStreamSynchronize(s) != hipSuccess) goto done;
        rc = GAMD_OK;
    }
done:
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (comp) (void)hipEventDestroy(comp);
    if (copied[0]) (void)hipEventDestroy(copied[0]);
    if (copied[1]) (void)hipEventDestroy(copied[1]);
    if (c) (void)hipStreamDestroy(c);
    if (cs) (void)hipStreamDestroy(cs);
    return rc;
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
