// gamd_dop853.hip -- persistent one-lane-per-orbit DOP853 with a work queue.
//
// C ABI: gamd_dop853 (include/gala_amd.h).  Restates Hairer's DOP853 as gala
// runs it (integrate/cyintegrators/dopri/dop853.cpp: dopcor :89-650, front
// end defaults :653-866, dense output :492-612 and contd8 :869-904), for
// systems of one orbit each (gala's nbatch=1, dop853.pyx:228-245), with the
// right-hand side of hamiltonian_gradient_T for the static frame
// (hamiltonian/src/chamiltonian.cpp:38-57): f(q, p) = (p, -grad Phi(q)).
//
// GPU structure:
//   * one lane owns one orbit; y, k1..k10 and the stage vector live in VGPRs
//     (about 80 doubles), so an accepted step touches HBM not at all;
//   * the grid is persistent (a few workgroups per CU) and lanes pull orbits
//     from a global queue: when a lane's orbit ends, the wave claims new
//     orbits for all its idle lanes with ONE atomicAdd (ballot + popcount +
//     mbcnt), so lanes never sit idle while the wave still has work and the
//     uneven per-orbit step counts do not serialise whole waves;
//   * the queue hands out orbits in a caller-chosen order (longest first
//     keeps the tail short); the loop ends when the queue is exhausted for
//     every lane of the wave, which every wave reaches.
//
// The same step-attempt body also integrates N-body systems with
// Fwrapper_direct_nbody forces (dop853.cpp:990-1006): mockstream_dop853's
// release systems (gamd_mockstream_dop853) and direct_nbody_dop853
// (gamd_nbody_dop853), one body per lane of a quad or a wave.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <type_traits>

#include "dop853_tableau.h"
#include "gamd_host.hpp"

using namespace gamd;
namespace T = gamd::dop;

namespace {

struct DopArgs {
    const double *w0;
    int64_t N;
    const double *t0;
    const double *t1;
    const double *h0;
    double rtol, atol;
    int64_t nmax, nstiff;
    double hmax;           // <= 0: |t1 - t0| per orbit
    const double *t_out;
    int64_t n_out;
    double *w_out;
    double *w_final;
    int32_t *status;
    int64_t *stats;
    const int32_t *order;
    unsigned long long *queue;
    int wide;              // 0: one lane per orbit to the end; 1: the last <= 4 orbits of a
                           // wave on 16-lane rows; 2: the last <= 16 on 4-lane quads first
    double uround;         // the "step size too small" threshold (dop853.cpp:349-354)
    int64_t n_tail;        // queue slots [0, n_tail) go to dedicated wide waves
    // dense output in windows of the t_out grid: w_out holds samples
    // [out_base, out_base + out_rows), and a step samples up to out_end
    // (the whole grid: 0, n_out, n_out).  pend: where a one-lane orbit keeps
    // the polynomial of a step that also covers sample out_end (nullptr: no
    // windows)
    int64_t out_base, out_rows, out_end;
    double *pend;
};

#define FOR6 _Pragma("unroll") for (int c = 0; c < 6; c++)

// f = (p, -grad Phi(q))
template <bool SCF, int LAY>
__device__ __forceinline__ void orbit_rhs(const DevPot &P, const double *y, double *f) {
    double gx, gy, gz;
    pot_grad_l<SCF, LAY>(P, y[0], y[1], y[2], gx, gy, gz);
    f[0] = y[3];
    f[1] = y[4];
    f[2] = y[5];
    f[3] = -gx;
    f[4] = -gy;
    f[5] = -gz;
}

__device__ __forceinline__ double sign_d(double a, double b) { return (b < 0.0) ? -fabs(a) : fabs(a); }
// comparison-based min/max as dop853.cpp:14-16: with a NaN argument they
// return the second operand, which decides how a NaN orbit ends (-2 after
// nmax attempts, as in the reference) where fmin/fmax would drop the NaN
__device__ __forceinline__ double min_d(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double max_d(double a, double b) { return (a > b) ? a : b; }

// Hairer's initial step guess, dop853.cpp:18-86 (scalar tolerances).
template <bool SCF, int LAY>
__device__ double hinit(const DevPot &P, double x, const double *y, double posneg,
                        const double *f0, double hmax, double rtol, double atol) {
    double dnf = 0.0, dny = 0.0;
    FOR6 {
        const double sk = atol + rtol * fabs(y[c]);
        double sqr = f0[c] / sk;
        dnf += sqr * sqr;
        sqr = y[c] / sk;
        dny += sqr * sqr;
    }
    double h = ((dnf <= 1.0e-10) || (dny <= 1.0e-10)) ? 1.0e-6 : sqrt(dny / dnf) * 0.01;
    h = min_d(h, hmax);
    h = sign_d(h, posneg);
    double yy[6], f1[6];
    FOR6 yy[c] = y[c] + h * f0[c];
    orbit_rhs<SCF, LAY>(P, yy, f1);
    double der2 = 0.0;
    FOR6 {
        const double sk = atol + rtol * fabs(y[c]);
        const double sqr = (f1[c] - f0[c]) / sk;
        der2 += sqr * sqr;
    }
    der2 = sqrt(der2) / h;
    const double der12 = max_d(fabs(der2), sqrt(dnf));
    const double h1 = (der12 <= 1.0e-15) ? max_d(1.0e-6, fabs(h) * 1.0e-3) : sqrt(sqrt(sqrt(0.01 / der12)));
    h = min_d(100.0 * fabs(h), min_d(h1, hmax));
    return sign_d(h, posneg);
}

// ---------------------------------------------------------------------------
// Two lane layouts share one step-attempt body, templated on NC, the number
// of state components a lane holds:
//   NC = 6: one lane owns one orbit (the bulk of the run);
//   NC = 1: the "wide tail": once the queue is drained and a wave holds at
//           most four unfinished orbits, each orbit is spread over one DPP
//           row of 16 lanes, lane c of the row holding component c % 6 (lanes
//           12..15 duplicate components 3..5).  The stage combinations, error
//           terms and dense updates then cost one component's arithmetic
//           instead of six; every lane of the row evaluates the force at the
//           row's broadcast position.  Sums over components are formed from
//           broadcast terms in the one-lane order, so both layouts produce
//           bit-identical results.
// The tail matters because the uneven workload's longest orbits need ~5x the
// average number of steps and end up running alone on their SIMDs.

constexpr int WIDE_ROW = 16;   // lanes per orbit in the wide tail (one DPP row)
constexpr int WIDE_ORBITS = 4; // orbits per wave in the wide tail
constexpr int QUAD_ORBITS = 16; // orbits per wave in the quad tail (one per 4-lane quad)

template <int SRC>
__device__ __forceinline__ double row_bcast(double v) {
    // DPP row_newbcast:SRC: every lane of a 16-lane row reads lane SRC of the
    // row (every lane has a source in its row, so no "old" value is needed)
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x150 + SRC, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x150 + SRC, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// DPP row_shl:3: lane i of a row reads lane i + 3 (lanes 13..15 keep their own)
__device__ __forceinline__ double row_shl3(double v) {
    const int lo0 = __double2loint(v), hi0 = __double2hiint(v);
    const int lo = __builtin_amdgcn_update_dpp(lo0, lo0, 0x103, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(hi0, hi0, 0x103, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// DPP quad_perm [B,B,B,B]: every lane of a 4-lane quad reads lane B of the quad
template <int B>
__device__ __forceinline__ double quad_bcast(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x55 * B, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x55 * B, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// per-orbit scalars of the integration (wave-uniform within a wide row)
struct DopState {
    int64_t orb;
    double x, xend, h, hmax, posneg, facold, hlamb;
    int64_t nstep, naccpt, nrejct, nfcn, oidx;
    int64_t stiff;  // naccpt % nstiff, kept incrementally (no 64-bit modulo per step)
    int reject, iasti, nonsti;
    int pend;  // dense output in windows: the last step's polynomial is kept
};

// Dense output in windows of the t_out grid (gamd_dop853_dense_stream).  The
// launch for window [out_base, out_end) takes every orbit from where the
// previous window left it, samples it into the window's slot and stops it at
// the first step boundary past the window's last sample; its state is kept
// and the slot is copied to host memory while the next window runs.  A step
// that also covers the next window's first sample keeps its polynomial
// (r1..r8, x0, h), and the next window samples the rest of it before the
// orbit steps on.  Steps, step sizes and samples are those of one gamd_dop853
// launch with w_out.
//
// Kept state, one row of N words per field:
enum : int {
    WS_X, WS_XEND, WS_H, WS_HMAX, WS_POSNEG, WS_FACOLD, WS_HLAMB,  // doubles
    WS_NSTEP, WS_NACCPT, WS_NREJCT, WS_NFCN, WS_OIDX, WS_STIFF,    // int64
    WS_REJECT, WS_IASTI, WS_NONSTI, WS_FLAGS,                       // int64
    WS_Y, WS_K1 = WS_Y + 6, WS_WORDS = WS_K1 + 6
};
constexpr int WS_PEND_WORDS = 8 * 6 + 2;    // the kept step: r1..r8, x0, h
constexpr int64_t WF_RUN = 1, WF_PEND = 2;  // still stepping; a kept step to sample

struct WinArgs {
    double *st;  // (WS_WORDS, N); nullptr: no windows
    int first;   // the first window: orbits start from w0
};

__device__ __forceinline__ void win_save_scalars(const WinArgs &w, int64_t N, const DopState &s,
                                                 int64_t flags) {
    double *d = w.st + s.orb;
    int64_t *q = reinterpret_cast<int64_t *>(w.st) + s.orb;
    d[WS_X * N] = s.x;
    d[WS_XEND * N] = s.xend;
    d[WS_H * N] = s.h;
    d[WS_HMAX * N] = s.hmax;
    d[WS_POSNEG * N] = s.posneg;
    d[WS_FACOLD * N] = s.facold;
    d[WS_HLAMB * N] = s.hlamb;
    q[WS_NSTEP * N] = s.nstep;
    q[WS_NACCPT * N] = s.naccpt;
    q[WS_NREJCT * N] = s.nrejct;
    q[WS_NFCN * N] = s.nfcn;
    q[WS_OIDX * N] = s.oidx;
    q[WS_STIFF * N] = s.stiff;
    q[WS_REJECT * N] = s.reject;
    q[WS_IASTI * N] = s.iasti;
    q[WS_NONSTI * N] = s.nonsti;
    q[WS_FLAGS * N] = flags;
}

// component c (0..5) of the state and of its first derivative
__device__ __forceinline__ void win_save_comp(const WinArgs &w, int64_t N, int64_t orb, int c,
                                              double y, double k1) {
    w.st[(WS_Y + c) * N + orb] = y;
    w.st[(WS_K1 + c) * N + orb] = k1;
}

// the kept state of orbit orb into one lane; returns its flags
__device__ __forceinline__ int64_t win_load(const WinArgs &w, int64_t N, int64_t orb,
                                            DopState &s, double (&y)[6], double (&k1)[6]) {
    const double *d = w.st + orb;
    const int64_t *q = reinterpret_cast<const int64_t *>(w.st) + orb;
    s.orb = orb;
    s.x = d[WS_X * N];
    s.xend = d[WS_XEND * N];
    s.h = d[WS_H * N];
    s.hmax = d[WS_HMAX * N];
    s.posneg = d[WS_POSNEG * N];
    s.facold = d[WS_FACOLD * N];
    s.hlamb = d[WS_HLAMB * N];
    s.nstep = q[WS_NSTEP * N];
    s.naccpt = q[WS_NACCPT * N];
    s.nrejct = q[WS_NREJCT * N];
    s.nfcn = q[WS_NFCN * N];
    s.oidx = q[WS_OIDX * N];
    s.stiff = q[WS_STIFF * N];
    s.reject = (int)q[WS_REJECT * N];
    s.iasti = (int)q[WS_IASTI * N];
    s.nonsti = (int)q[WS_NONSTI * N];
    s.pend = 0;
#pragma unroll
    for (int c = 0; c < 6; c++) {
        y[c] = d[(WS_Y + c) * N];
        k1[c] = d[(WS_K1 + c) * N];
    }
    return q[WS_FLAGS * N];
}

// component c of a kept step's polynomial, and the step's x0 and h
__device__ __forceinline__ void keep_comp(const DopArgs &a, int64_t orb, int c, double r1,
                                          double r2, double r3, double r4, double r5, double r6,
                                          double r7, double r8) {
    double *pp = a.pend + orb;
    const int64_t N = a.N;
    pp[(0 * 6 + c) * N] = r1;
    pp[(1 * 6 + c) * N] = r2;
    pp[(2 * 6 + c) * N] = r3;
    pp[(3 * 6 + c) * N] = r4;
    pp[(4 * 6 + c) * N] = r5;
    pp[(5 * 6 + c) * N] = r6;
    pp[(6 * 6 + c) * N] = r7;
    pp[(7 * 6 + c) * N] = r8;
}
__device__ __forceinline__ void keep_x0h(const DopArgs &a, int64_t orb, double x0, double h) {
    a.pend[(8 * 6) * a.N + orb] = x0;
    a.pend[(8 * 6 + 1) * a.N + orb] = h;
}

// Lane layouts of one DOP853 system (the policy of dop_attempt):
//   rhs   the lane's share of f(y);
//   sums  sa += a_i^2, sb += b_i^2 over the system's n components in the
//         order of its state vector (the error norm and stiffness sums);
//   n     the system's number of components.

// one orbit per lane: f = (p, -grad Phi(q))
template <bool SCF, int LAY>
struct OneLane {
    static constexpr int NC = 6;
    static constexpr bool WINDOWS = true;  // dense output in windows of the grid
    __device__ __forceinline__ void rhs(const DevPot &P, const double (&y)[6], double (&f)[6]) const {
        orbit_rhs<SCF, LAY>(P, y, f);
    }
    __device__ __forceinline__ void sums(const double (&a)[6], const double (&b)[6], double &sa,
                                         double &sb) const {
        FOR6 {
            sa += a[c] * a[c];
            sb += b[c] * b[c];
        }
    }
    __device__ __forceinline__ double n() const { return 6.0; }
    // dense sample s.oidx of orbit s.orb into w_out (6, out_rows, N)
    __device__ __forceinline__ void store_dense(const DopArgs &a, const DopState &s,
                                                const double (&v)[6]) const {
        const int64_t so = a.out_rows * a.N;
        FOR6 a.w_out[c * so + (s.oidx - a.out_base) * a.N + s.orb] = v[c];
    }
    // the step that covers the next window's first sample
    __device__ __forceinline__ void keep_step(const DopArgs &a, const DopState &s,
                                              const double (&r1)[6], const double (&r2)[6],
                                              const double (&r3)[6], const double (&r4)[6],
                                              const double (&r5)[6], const double (&r6)[6],
                                              const double (&r7)[6], const double (&r8)[6],
                                              double x0, double h) const {
        FOR6 keep_comp(a, s.orb, c, r1[c], r2[c], r3[c], r4[c], r5[c], r6[c], r7[c], r8[c]);
        keep_x0h(a, s.orb, x0, h);
    }
    __device__ __forceinline__ void save_state(const WinArgs &w, int64_t N, const DopState &s,
                                               const double (&y)[6], const double (&k1)[6],
                                               int64_t flags) const {
        win_save_scalars(w, N, s, flags);
        FOR6 win_save_comp(w, N, s.orb, c, y[c], k1[c]);
    }
};

// OneLane reading the tableau from an LDS copy (the queue loop's attempts)
template <bool SCF, int LAY>
struct OneLaneTab : OneLane<SCF, LAY> {
    static constexpr bool LDS_TABLEAU = true;
    const double *tab;
};

// one orbit per 16-lane row, lane c holding component comp
template <bool SCF, int LAY>
struct WideRow {
    static constexpr int NC = 1;
    int comp;
    bool writes;  // lanes 0..5 of the row store the dense samples
    __device__ __forceinline__ void rhs(const DevPot &P, const double (&y)[1], double (&f)[1]) const {
        const double qx = row_bcast<0>(y[0]), qy = row_bcast<1>(y[0]), qz = row_bcast<2>(y[0]);
        const double vp = row_shl3(y[0]);  // the velocity partner of components 0..2
        double gx, gy, gz;
        pot_grad_l<SCF, LAY>(P, qx, qy, qz, gx, gy, gz);
        const double g = comp == 3 ? gx : (comp == 4 ? gy : gz);
        f[0] = comp < 3 ? vp : -g;
    }
    __device__ __forceinline__ static void all6(const double (&v)[1], double (&o)[6]) {
        o[0] = row_bcast<0>(v[0]);
        o[1] = row_bcast<1>(v[0]);
        o[2] = row_bcast<2>(v[0]);
        o[3] = row_bcast<3>(v[0]);
        o[4] = row_bcast<4>(v[0]);
        o[5] = row_bcast<5>(v[0]);
    }
    __device__ __forceinline__ void sums(const double (&a)[1], const double (&b)[1], double &sa,
                                         double &sb) const {
        double qa[6], qb[6];
        all6(a, qa);
        all6(b, qb);
        FOR6 {
            sa += qa[c] * qa[c];
            sb += qb[c] * qb[c];
        }
    }
    __device__ __forceinline__ double n() const { return 6.0; }
    // dense sample s.oidx of orbit s.orb into w_out (6, out_rows, N)
    __device__ __forceinline__ void store_dense(const DopArgs &a, const DopState &s,
                                                const double (&v)[1]) const {
        if (writes) a.w_out[comp * a.out_rows * a.N + (s.oidx - a.out_base) * a.N + s.orb] = v[0];
    }
    static constexpr bool WINDOWS = true;
    __device__ __forceinline__ void keep_step(const DopArgs &a, const DopState &s,
                                              const double (&r1)[1], const double (&r2)[1],
                                              const double (&r3)[1], const double (&r4)[1],
                                              const double (&r5)[1], const double (&r6)[1],
                                              const double (&r7)[1], const double (&r8)[1],
                                              double x0, double h) const {
        if (writes) {
            keep_comp(a, s.orb, comp, r1[0], r2[0], r3[0], r4[0], r5[0], r6[0], r7[0], r8[0]);
            if (comp == 0) keep_x0h(a, s.orb, x0, h);
        }
    }
    __device__ __forceinline__ void save_state(const WinArgs &w, int64_t N, const DopState &s,
                                               const double (&y)[1], const double (&k1)[1],
                                               int64_t flags) const {
        if (writes) {
            win_save_comp(w, N, s.orb, comp, y[0], k1[0]);
            if (comp == 0) win_save_scalars(w, N, s, flags);
        }
    }
};

// one orbit per 4-lane quad, lane c < 3 holding components c and c + 3 (the
// position and velocity along one axis), lane 3 mirroring lane 0.  The stage
// combinations cost two components' arithmetic instead of six; every lane
// evaluates the force at the quad's broadcast position.  Sums over
// components are formed in the one-lane order, so results are bit-identical
// to one lane per orbit.
template <bool SCF, int LAY>
struct QuadPair {
    static constexpr int NC = 2;
    int axis;
    bool writes;  // lanes 0..2 of the quad store the dense samples
    __device__ __forceinline__ void rhs(const DevPot &P, const double (&y)[2], double (&f)[2]) const {
        const double qx = quad_bcast<0>(y[0]), qy = quad_bcast<1>(y[0]), qz = quad_bcast<2>(y[0]);
        double gx, gy, gz;
        pot_grad_l<SCF, LAY>(P, qx, qy, qz, gx, gy, gz);
        f[0] = y[1];
        f[1] = -(axis == 0 ? gx : (axis == 1 ? gy : gz));
    }
    __device__ __forceinline__ void sums(const double (&a)[2], const double (&b)[2], double &sa,
                                         double &sb) const {
        const double qa[6] = {quad_bcast<0>(a[0]), quad_bcast<1>(a[0]), quad_bcast<2>(a[0]),
                              quad_bcast<0>(a[1]), quad_bcast<1>(a[1]), quad_bcast<2>(a[1])};
        const double qb[6] = {quad_bcast<0>(b[0]), quad_bcast<1>(b[0]), quad_bcast<2>(b[0]),
                              quad_bcast<0>(b[1]), quad_bcast<1>(b[1]), quad_bcast<2>(b[1])};
        FOR6 {
            sa += qa[c] * qa[c];
            sb += qb[c] * qb[c];
        }
    }
    __device__ __forceinline__ double n() const { return 6.0; }
    // dense sample s.oidx of orbit s.orb into w_out (6, out_rows, N)
    __device__ __forceinline__ void store_dense(const DopArgs &a, const DopState &s,
                                                const double (&v)[2]) const {
        if (writes) {
            const int64_t so = a.out_rows * a.N, at = (s.oidx - a.out_base) * a.N + s.orb;
            a.w_out[axis * so + at] = v[0];
            a.w_out[(axis + 3) * so + at] = v[1];
        }
    }
    static constexpr bool WINDOWS = true;
    __device__ __forceinline__ void keep_step(const DopArgs &a, const DopState &s,
                                              const double (&r1)[2], const double (&r2)[2],
                                              const double (&r3)[2], const double (&r4)[2],
                                              const double (&r5)[2], const double (&r6)[2],
                                              const double (&r7)[2], const double (&r8)[2],
                                              double x0, double h) const {
        if (writes) {
            keep_comp(a, s.orb, axis, r1[0], r2[0], r3[0], r4[0], r5[0], r6[0], r7[0], r8[0]);
            keep_comp(a, s.orb, axis + 3, r1[1], r2[1], r3[1], r4[1], r5[1], r6[1], r7[1],
                      r8[1]);
            if (axis == 0) keep_x0h(a, s.orb, x0, h);
        }
    }
    __device__ __forceinline__ void save_state(const WinArgs &w, int64_t N, const DopState &s,
                                               const double (&y)[2], const double (&k1)[2],
                                               int64_t flags) const {
        if (writes) {
            win_save_comp(w, N, s.orb, axis, y[0], k1[0]);
            win_save_comp(w, N, s.orb, axis + 3, y[1], k1[1]);
            if (axis == 0) win_save_scalars(w, N, s, flags);
        }
    }
};

// N-body systems (Fwrapper_direct_nbody, dop853.cpp:990-1006): lane b of a
// group holds body b's six components; the state vector is body-major, as the
// reference's (nb, 6) rows.  Every body feels the external potential, then
// the potential of every other body j < bt.n that has one, centred on body j
// (c_nbody_acceleration, cpotential.cpp:389-415), accumulated in body order;
// f = -(g_ext + f_0 + f_1 + ...) equals the reference's
// (-g_ext) + (-f_0) + (-f_1) + ... exactly.  Lanes b >= nb mirror body 0 and
// are left out of every sum.
//   QuadBodies: one system of at most 4 bodies per 4-lane quad (several
//               systems per wave), DPP quad_perm broadcasts;
//   RowBodies:  one system of at most 4 bodies per 16-lane row, body b on
//               the row's quad b and lane c < 3 of the quad holding the
//               position and velocity along axis c (lane 3 mirrors lane 0):
//               the stage combinations cost two components' arithmetic
//               instead of six.  Body positions are broadcast with DPP
//               quad_perm (own body) and row_newbcast (other bodies); sums
//               follow the body-major order, so results are bit-identical to
//               QuadBodies;
//   WaveBodies: one system of at most 64 bodies per wave, readlane
//               broadcasts (wave-uniform source lane).
template <bool SCF, int LAY>
struct QuadBodies {
    static constexpr int NC = 6;
    int body, nb;
    const BodyTable &bt;  // bt.n <= nb <= 4
    template <int J>
    __device__ __forceinline__ void add_body(const double (&y)[6], double &gx, double &gy,
                                             double &gz) const {
        if (J < bt.n) {  // wave-uniform
            const double px = quad_bcast<J>(y[0]), py = quad_bcast<J>(y[1]),
                         pz = quad_bcast<J>(y[2]);
            if (J != body && body < nb && bt.p[J]) {
                double fx, fy, fz;
                body_grad<SCF>(*bt.p[J], y[0], y[1], y[2], px, py, pz, fx, fy, fz);
                gx += fx;
                gy += fy;
                gz += fz;
            }
        }
    }
    __device__ __forceinline__ void rhs(const DevPot &P, const double (&y)[6], double (&f)[6]) const {
        double gx, gy, gz;
        pot_grad_l<SCF, LAY>(P, y[0], y[1], y[2], gx, gy, gz);
        add_body<0>(y, gx, gy, gz);
        add_body<1>(y, gx, gy, gz);
        add_body<2>(y, gx, gy, gz);
        add_body<3>(y, gx, gy, gz);
        f[0] = y[3];
        f[1] = y[4];
        f[2] = y[5];
        f[3] = -gx;
        f[4] = -gy;
        f[5] = -gz;
    }
    template <int B>
    __device__ __forceinline__ void sums_body(const double (&a)[6], const double (&b)[6],
                                              double &sa, double &sb) const {
        double qa[6], qb[6];
        FOR6 {
            qa[c] = quad_bcast<B>(a[c]);
            qb[c] = quad_bcast<B>(b[c]);
        }
        if (B < nb) {
            FOR6 {
                sa += qa[c] * qa[c];
                sb += qb[c] * qb[c];
            }
        }
    }
    __device__ __forceinline__ void sums(const double (&a)[6], const double (&b)[6], double &sa,
                                         double &sb) const {
        sums_body<0>(a, b, sa, sb);
        sums_body<1>(a, b, sa, sb);
        sums_body<2>(a, b, sa, sb);
        sums_body<3>(a, b, sa, sb);
    }
    __device__ __forceinline__ double n() const { return 6.0 * (double)nb; }
};

// the RowBodies layout of the list above
template <bool SCF, int LAY>
struct RowBodies {
    static constexpr int NC = 2;
    int body, axis, lane4, nb;  // axis 0..2 (lane4 = 3, the quad's last lane, runs axis 0 again)
    const BodyTable &bt;        // bt.n <= nb <= 4
    template <int J>
    __device__ __forceinline__ void add_body(const double (&y)[2], double qx, double qy, double qz,
                                             double &gx, double &gy, double &gz) const {
        if (J < bt.n) {  // wave-uniform
            const double px = row_bcast<4 * J>(y[0]), py = row_bcast<4 * J + 1>(y[0]),
                         pz = row_bcast<4 * J + 2>(y[0]);
            if (J != body && body < nb && bt.p[J]) {
                double fx, fy, fz;
                body_grad<SCF>(*bt.p[J], qx, qy, qz, px, py, pz, fx, fy, fz);
                gx += fx;
                gy += fy;
                gz += fz;
            }
        }
    }
    __device__ __forceinline__ void rhs(const DevPot &P, const double (&y)[2], double (&f)[2]) const {
        const double qx = quad_bcast<0>(y[0]), qy = quad_bcast<1>(y[0]), qz = quad_bcast<2>(y[0]);
        double gx, gy, gz;
        pot_grad_l<SCF, LAY>(P, qx, qy, qz, gx, gy, gz);
        add_body<0>(y, qx, qy, qz, gx, gy, gz);
        add_body<1>(y, qx, qy, qz, gx, gy, gz);
        add_body<2>(y, qx, qy, qz, gx, gy, gz);
        add_body<3>(y, qx, qy, qz, gx, gy, gz);
        f[0] = y[1];
        f[1] = -(axis == 0 ? gx : (axis == 1 ? gy : gz));
    }
    template <int B>
    __device__ __forceinline__ void sums_body(const double (&a)[2], const double (&b)[2],
                                              double &sa, double &sb) const {
        const double qa[6] = {row_bcast<4 * B>(a[0]), row_bcast<4 * B + 1>(a[0]),
                              row_bcast<4 * B + 2>(a[0]), row_bcast<4 * B>(a[1]),
                              row_bcast<4 * B + 1>(a[1]), row_bcast<4 * B + 2>(a[1])};
        const double qb[6] = {row_bcast<4 * B>(b[0]), row_bcast<4 * B + 1>(b[0]),
                              row_bcast<4 * B + 2>(b[0]), row_bcast<4 * B>(b[1]),
                              row_bcast<4 * B + 1>(b[1]), row_bcast<4 * B + 2>(b[1])};
        if (B < nb) {
            FOR6 {
                sa += qa[c] * qa[c];
                sb += qb[c] * qb[c];
            }
        }
    }
    __device__ __forceinline__ void sums(const double (&a)[2], const double (&b)[2], double &sa,
                                         double &sb) const {
        sums_body<0>(a, b, sa, sb);
        sums_body<1>(a, b, sa, sb);
        sums_body<2>(a, b, sa, sb);
        sums_body<3>(a, b, sa, sb);
    }
    __device__ __forceinline__ double n() const { return 6.0 * (double)nb; }
    // dense sample s.oidx of body `body` into w_out (6, n_out, nb)
    __device__ __forceinline__ void store_dense(const DopArgs &a, const DopState &s,
                                                const double (&v)[2]) const {
        const int64_t so = a.out_rows * a.N, row = (s.oidx - a.out_base) * a.N;
        if (body < nb && lane4 < 3) {
            a.w_out[axis * so + row + body] = v[0];
            a.w_out[(axis + 3) * so + row + body] = v[1];
        }
    }
};

// v of lane src (wave-uniform) in every lane
__device__ __forceinline__ double lane_bcast(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

template <bool SCF, int LAY>
struct WaveBodies {
    static constexpr int NC = 6;
    int body, nb;
    const BodyTable &bt;  // bt.n <= nb <= 64
    __device__ __forceinline__ void rhs(const DevPot &P, const double (&y)[6], double (&f)[6]) const {
        double gx, gy, gz;
        pot_grad_l<SCF, LAY>(P, y[0], y[1], y[2], gx, gy, gz);
        for (int j = 0; j < bt.n; j++) {
            const double px = lane_bcast(y[0], j), py = lane_bcast(y[1], j),
                         pz = lane_bcast(y[2], j);
            if (j != body && body < nb && bt.p[j]) {
                double fx, fy, fz;
                body_grad<SCF>(*bt.p[j], y[0], y[1], y[2], px, py, pz, fx, fy, fz);
                gx += fx;
                gy += fy;
                gz += fz;
            }
        }
        f[0] = y[3];
        f[1] = y[4];
        f[2] = y[5];
        f[3] = -gx;
        f[4] = -gy;
        f[5] = -gz;
    }
    __device__ __forceinline__ void sums(const double (&a)[6], const double (&b)[6], double &sa,
                                         double &sb) const {
        for (int j = 0; j < nb; j++) {
            FOR6 {
                const double qa = lane_bcast(a[c], j), qb = lane_bcast(b[c], j);
                sa += qa * qa;
                sb += qb * qb;
            }
        }
    }
    __device__ __forceinline__ double n() const { return 6.0 * (double)nb; }
    // dense sample s.oidx of body `body` into w_out (6, n_out, nb)
    __device__ __forceinline__ void store_dense(const DopArgs &a, const DopState &s,
                                                const double (&v)[6]) const {
        const int64_t so = a.out_rows * a.N;
        if (body < nb) FOR6 a.w_out[c * so + (s.oidx - a.out_base) * a.N + body] = v[c];
    }
};

// gala's default batching (dop853_integrate_hamiltonian, dop853.pyx:228-245):
// the nb orbits of a batch form one (6, nb) system with one shared adaptive
// step.  Orbit j of the batch lives on thread j of one workgroup (its six
// components in VGPRs, as OneLane).  The error-norm and stiffness sums run
// over the system's 6 nb components in the order of the (6, nb) state
// vector, component-major, as dop853.cpp:416-444 and :463-486 loop over it:
// every lane stages its six squared terms in LDS, rounded as the reference
// rounds sqr*sqr, and the 6 nb terms of each sum are then added in that order
// by one wave (a workgroup of two or more waves gives the two sums to two
// waves, which halves the serial chain of dependent adds each wave waits on;
// a one-wave batch adds both), so the workgroup's lanes reach the same sums
// and take the same step decision.  Lanes past nb mirror orbit 0 of the batch
// and stay out of the sums.
template <bool SCF, int LAY>
struct BatchLanes {
    static constexpr int NC = 6;
    static constexpr bool LDS_TABLEAU = true;
    int j, nb;
    int64_t base;       // the batch's first orbit
    double *lds;        // 12 nb + 2 doubles: the a-terms, the b-terms, the two sums
    const double *tab;  // the tableau's LDS copy (TAB_WORDS doubles)
    __device__ __forceinline__ void rhs(const DevPot &P, const double (&y)[6], double (&f)[6]) const {
        orbit_rhs<SCF, LAY>(P, y, f);
    }
    __device__ __forceinline__ void sums(const double (&a)[6], const double (&b)[6], double &sa,
                                         double &sb) const {
        const int n = 6 * nb;
        __syncthreads();  // every lane has read the previous sums' terms
        if (j < nb) {
            FOR6 {
                const double pa = a[c] * a[c], pb = b[c] * b[c];
                lds[c * nb + j] = pa;
                lds[n + c * nb + j] = pb;
            }
        }
        __syncthreads();
        if (nb > 64) {
            // two or more waves: wave 0 adds the a-terms, wave 1 the b-terms,
            // and they trade results through LDS
            const int w = j >> 6;
            if (w < 2) {
                const double r = chain(w ? sb : sa, lds + w * n, n);
                if ((j & 63) == 0) lds[2 * n + w] = r;
            }
            __syncthreads();
            sa = lds[2 * n];
            sb = lds[2 * n + 1];
            return;
        }
        double xa = sa, xb = sb;
        int i = 0;
        // eight terms of each sum read ahead of their adds (the adds keep the
        // reference's order; the reads of a block overlap one LDS latency)
        constexpr int U = 8;
        for (; i + U <= n; i += U) {
            double ta[U], tb[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                ta[u] = lds[i + u];
                tb[u] = lds[n + i + u];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                xa += ta[u];
                xb += tb[u];
            }
        }
        for (; i < n; i++) {
            xa += lds[i];
            xb += lds[n + i];
        }
        sa = xa;
        sb = xb;
    }
    // x + t[0] + t[1] + ... + t[n - 1], added in that order
    __device__ __forceinline__ static double chain(double x, const double *t, int n) {
        int i = 0;
        constexpr int U = 8;
        for (; i + U <= n; i += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = t[i + u];
#pragma unroll
            for (int u = 0; u < U; u++) x += v[u];
        }
        for (; i < n; i++) x += t[i];
        return x;
    }
    __device__ __forceinline__ double n() const { return 6.0 * (double)nb; }
    // dense sample s.oidx of the lane's orbit into w_out (6, n_out, N)
    __device__ __forceinline__ void store_dense(const DopArgs &a, const DopState &s,
                                                const double (&v)[6]) const {
        const int64_t so = a.out_rows * a.N;
        if (j < nb) FOR6 a.w_out[c * so + (s.oidx - a.out_base) * a.N + base + j] = v[c];
    }
};

// Hairer's initial step guess (dop853.cpp:18-86) for a batched system: the
// norms dnf, dny and der2 are sums over the system's components in its
// state order (BatchLanes::sums).
template <bool SCF, int LAY>
__device__ double hinit_batch(const DevPot &P, const BatchLanes<SCF, LAY> &sys,
                              const double (&y)[6], double posneg, const double (&f0)[6],
                              double hmax, double rtol, double atol) {
    double a[6], b[6];
    FOR6 {
        const double sk = atol + rtol * fabs(y[c]);
        a[c] = f0[c] / sk;
        b[c] = y[c] / sk;
    }
    double dnf = 0.0, dny = 0.0;
    sys.sums(a, b, dnf, dny);
    double h = ((dnf <= 1.0e-10) || (dny <= 1.0e-10)) ? 1.0e-6 : sqrt(dny / dnf) * 0.01;
    h = min_d(h, hmax);
    h = sign_d(h, posneg);
    double yy[6], f1[6];
    FOR6 yy[c] = y[c] + h * f0[c];
    sys.rhs(P, yy, f1);
    FOR6 {
        const double sk = atol + rtol * fabs(y[c]);
        a[c] = (f1[c] - f0[c]) / sk;
        b[c] = 0.0;
    }
    double der2 = 0.0, unused = 0.0;
    sys.sums(a, b, der2, unused);
    der2 = sqrt(der2) / h;
    const double der12 = max_d(fabs(der2), sqrt(dnf));
    const double h1 = (der12 <= 1.0e-15) ? max_d(1.0e-6, fabs(h) * 1.0e-3)
                                         : pow(0.01 / der12, 1.0 / 8.0);
    h = min_d(100.0 * fabs(h), min_d(h1, hmax));
    return sign_d(h, posneg);
}

#define FORN _Pragma("unroll") for (int c = 0; c < NC; c++)

// t_out sample `to` inside the step from x0 to x1 (either direction)
__device__ __forceinline__ bool in_step(double x0, double x1, double to) {
    return (x0 <= to && to <= x1) || (x1 <= to && to <= x0);
}

// the dense-output polynomial at `to` of the step (x0, h) (dop853.cpp:869-904)
template <int NC>
__device__ __forceinline__ void dense_value(double to, double x0, double h, const double *r1,
                                            const double *r2, const double *r3, const double *r4,
                                            const double *r5, const double *r6, const double *r7,
                                            const double *r8, double (&v)[NC]) {
    const double sv = (to - x0) / h, s1 = 1.0 - sv;
    FORN v[c] = r1[c] + sv * (r2[c] + s1 * (r3[c] + sv * (r4[c] + s1 * (r5[c] +
                sv * (r6[c] + s1 * (r7[c] + sv * r8[c]))))));
}

// layouts that support dense output in windows declare WINDOWS = true
template <class S, class = void>
struct has_windows {
    static constexpr bool value = false;
};
template <class S>
struct has_windows<S, decltype((void)S::WINDOWS)> {
    static constexpr bool value = S::WINDOWS;
};

// The tableau as one array (A 16x16, D 4x16, B, E5, BHH), for layouts that
// read it from LDS (a layout declares LDS_TABLEAU = true and a member `tab`
// pointing at an LDS copy of kTableau).
constexpr int TAB_OFS_A = 0, TAB_OFS_D = 256, TAB_OFS_B = 320, TAB_OFS_E5 = 332,
              TAB_OFS_BHH = 345, TAB_WORDS = 358;
struct TableauWords {
    double w[TAB_WORDS];
};
constexpr TableauWords make_tableau() {
    TableauWords t{};
    for (int i = 0; i < 16; i++)
        for (int j = 0; j < 16; j++) t.w[TAB_OFS_A + i * 16 + j] = T::A[i][j];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 16; j++) t.w[TAB_OFS_D + i * 16 + j] = T::D[i][j];
    for (int k = 0; k < 12; k++) t.w[TAB_OFS_B + k] = T::B[k];
    for (int k = 0; k < 13; k++) t.w[TAB_OFS_E5 + k] = T::E5[k];
    for (int k = 0; k < 13; k++) t.w[TAB_OFS_BHH + k] = T::BHH[k];
    return t;
}
__constant__ TableauWords kTableau = make_tableau();

template <class S, class = void>
struct lds_tableau {
    static constexpr bool value = false;
};
template <class S>
struct lds_tableau<S, decltype((void)S::LDS_TABLEAU)> {
    static constexpr bool value = S::LDS_TABLEAU;
};

template <class Sys>
__device__ __forceinline__ double tab_entry(const Sys &sys, int ofs, int k, double c) {
    if constexpr (lds_tableau<Sys>::value)
        return sys.tab[ofs + k];
    else
        return c;
}

// One step attempt (dopcor basic step, dop853.cpp:339-648) for the lane's NC
// components.  Returns 0 to go on, else the orbit's final idid.
template <bool SCF, bool DENSE, int LAY, class Sys>
__device__ __forceinline__ int dop_attempt(const DevPot &P, const DopArgs &a, const Sys &sys,
                                           DopState &s, double (&y)[Sys::NC],
                                           double (&k1)[Sys::NC]) {
    constexpr int NC = Sys::NC;
    // tableau entries: compile-time constants, or for a layout that keeps the
    // tableau in LDS, reads at an offset the compiler cannot move out of the
    // attempt loop (held across the loop, the constants crowd the scalar
    // registers and spill)
    int tofs = 0;
    if constexpr (lds_tableau<Sys>::value) asm volatile("" : "+v"(tofs));
#define TAB_A(i, j) tab_entry(sys, tofs, TAB_OFS_A + (i) * 16 + (j), T::A[i][j])
#define TAB_D(i, j) tab_entry(sys, tofs, TAB_OFS_D + (i) * 16 + (j), T::D[i][j])
#define TAB_B(k) tab_entry(sys, tofs, TAB_OFS_B + (k), T::B[k])
#define TAB_E5(k) tab_entry(sys, tofs, TAB_OFS_E5 + (k), T::E5[k])
#define TAB_BHH(k) tab_entry(sys, tofs, TAB_OFS_BHH + (k), T::BHH[k])
    const double safe = 0.9, facc1 = 1.0 / 0.333, facc2 = 1.0 / 6.0;  // expo1 = 1/8
    if (s.nstep > a.nmax) return -2;
    if (0.1 * fabs(s.h) <= fabs(s.x) * a.uround) return -3;
    const bool last = (s.x + 1.01 * s.h - s.xend) * s.posneg > 0.0;
    const double h = last ? s.xend - s.x : s.h;
    s.h = h;
    s.nstep++;
    double k2[NC], k3[NC], k4[NC], k5[NC], k6[NC], k7[NC], k8[NC], k9[NC], k10[NC], yy[NC];
    FORN yy[c] = y[c] + h * TAB_A(1, 0) * k1[c];
    sys.rhs(P, yy, k2);
    FORN yy[c] = y[c] + h * (TAB_A(2, 0) * k1[c] + TAB_A(2, 1) * k2[c]);
    sys.rhs(P, yy, k3);
    FORN yy[c] = y[c] + h * (TAB_A(3, 0) * k1[c] + TAB_A(3, 2) * k3[c]);
    sys.rhs(P, yy, k4);
    FORN yy[c] = y[c] + h * (TAB_A(4, 0) * k1[c] + TAB_A(4, 2) * k3[c] + TAB_A(4, 3) * k4[c]);
    sys.rhs(P, yy, k5);
    FORN yy[c] = y[c] + h * (TAB_A(5, 0) * k1[c] + TAB_A(5, 3) * k4[c] + TAB_A(5, 4) * k5[c]);
    sys.rhs(P, yy, k6);
    FORN yy[c] = y[c] + h * (TAB_A(6, 0) * k1[c] + TAB_A(6, 3) * k4[c] + TAB_A(6, 4) * k5[c] +
                             TAB_A(6, 5) * k6[c]);
    sys.rhs(P, yy, k7);
    FORN yy[c] = y[c] + h * (TAB_A(7, 0) * k1[c] + TAB_A(7, 3) * k4[c] + TAB_A(7, 4) * k5[c] +
                             TAB_A(7, 5) * k6[c] + TAB_A(7, 6) * k7[c]);
    sys.rhs(P, yy, k8);
    FORN yy[c] = y[c] + h * (TAB_A(8, 0) * k1[c] + TAB_A(8, 3) * k4[c] + TAB_A(8, 4) * k5[c] +
                             TAB_A(8, 5) * k6[c] + TAB_A(8, 6) * k7[c] + TAB_A(8, 7) * k8[c]);
    sys.rhs(P, yy, k9);
    FORN yy[c] = y[c] + h * (TAB_A(9, 0) * k1[c] + TAB_A(9, 3) * k4[c] + TAB_A(9, 4) * k5[c] +
                             TAB_A(9, 5) * k6[c] + TAB_A(9, 6) * k7[c] + TAB_A(9, 7) * k8[c] +
                             TAB_A(9, 8) * k9[c]);
    sys.rhs(P, yy, k10);
    FORN yy[c] = y[c] + h * (TAB_A(10, 0) * k1[c] + TAB_A(10, 3) * k4[c] + TAB_A(10, 4) * k5[c] +
                             TAB_A(10, 5) * k6[c] + TAB_A(10, 6) * k7[c] + TAB_A(10, 7) * k8[c] +
                             TAB_A(10, 8) * k9[c] + TAB_A(10, 9) * k10[c]);
    sys.rhs(P, yy, k2);  // stage 11 reuses k2 (dop853.cpp:402)
    const double xph = s.x + h;
    FORN yy[c] = y[c] + h * (TAB_A(11, 0) * k1[c] + TAB_A(11, 3) * k4[c] + TAB_A(11, 4) * k5[c] +
                             TAB_A(11, 5) * k6[c] + TAB_A(11, 6) * k7[c] + TAB_A(11, 7) * k8[c] +
                             TAB_A(11, 8) * k9[c] + TAB_A(11, 9) * k10[c] + TAB_A(11, 10) * k2[c]);
    sys.rhs(P, yy, k3);  // stage 12 reuses k3 (:408)
    s.nfcn += 11;
    FORN {
        k4[c] = TAB_B(0) * k1[c] + TAB_B(5) * k6[c] + TAB_B(6) * k7[c] + TAB_B(7) * k8[c] +
                TAB_B(8) * k9[c] + TAB_B(9) * k10[c] + TAB_B(10) * k2[c] + TAB_B(11) * k3[c];
        k5[c] = y[c] + h * k4[c];
    }
    // error estimation (:416-444), n = 6; one divide serves both estimates
    double e2[NC], e1[NC];
    FORN {
        const double sk = a.atol + a.rtol * max_d(fabs(y[c]), fabs(k5[c]));
        const double isk = 1.0 / sk;
        double erri = k4[c] - TAB_BHH(0) * k1[c] - TAB_BHH(8) * k9[c] - TAB_BHH(11) * k3[c];
        e2[c] = erri * isk;
        erri = TAB_E5(0) * k1[c] + TAB_E5(5) * k6[c] + TAB_E5(6) * k7[c] + TAB_E5(7) * k8[c] +
               TAB_E5(8) * k9[c] + TAB_E5(9) * k10[c] + TAB_E5(10) * k2[c] + TAB_E5(11) * k3[c];
        e1[c] = erri * isk;
    }
    double err = 0.0, err2 = 0.0;
    sys.sums(e2, e1, err2, err);
    double deno = err + 0.01 * err2;
    if (deno <= 0.0) deno = 1.0;
    err = fabs(h) * err * sqrt(1.0 / (deno * sys.n()));
    // step size factor (:447-452), beta = 0: err^(1/8) as three square
    // roots (~330 cycles of latency where pow's log/exp pair is ~1000)
    const double fac11 = sqrt(sqrt(sqrt(err)));
    const double fac = max_d(facc2, min_d(facc1, fac11 / safe));
    double hnew = h / fac;
    int code = 0;
    if (err <= 1.0) {
        // accepted (:454-637)
        s.facold = max_d(err, 1.0e-4);
        s.naccpt++;
        if (++s.stiff == a.nstiff) s.stiff = 0;  // never 0 again when nstiff <= 0
        sys.rhs(P, k5, k4);  // FSAL
        s.nfcn++;
        // stiffness test (:463-486); returns -4 on the first detection
        if (a.nstiff > 0 && (s.stiff == 0 || s.iasti > 0)) {
            double dn[NC], dd[NC];
            FORN {
                dn[c] = k4[c] - k3[c];
                dd[c] = k5[c] - yy[c];
            }
            double stnum = 0.0, stden = 0.0;
            sys.sums(dn, dd, stnum, stden);
            if (stden > 0.0) s.hlamb = h * sqrt(stnum / stden);
            if (s.hlamb > 6.1) {
                s.nonsti = 0;
                s.iasti++;
                code = -4;
            } else {
                s.nonsti++;
                if (s.nonsti == 6) s.iasti = 0;
            }
        }
        if (code == 0) {
            if constexpr (DENSE) {
                // dense output coefficients (:492-582) and sampling (:584-612)
                double r1[NC], r2[NC], r3[NC], r4[NC], r5[NC], r6[NC], r7[NC], r8[NC];
                FORN {
                    r1[c] = y[c];
                    const double ydiff = k5[c] - y[c];
                    r2[c] = ydiff;
                    const double bspl = h * k1[c] - ydiff;
                    r3[c] = bspl;
                    r4[c] = ydiff - h * k4[c] - bspl;
                    r5[c] = TAB_D(0, 0) * k1[c] + TAB_D(0, 5) * k6[c] + TAB_D(0, 6) * k7[c] +
                            TAB_D(0, 7) * k8[c] + TAB_D(0, 8) * k9[c] + TAB_D(0, 9) * k10[c] +
                            TAB_D(0, 10) * k2[c] + TAB_D(0, 11) * k3[c];
                    r6[c] = TAB_D(1, 0) * k1[c] + TAB_D(1, 5) * k6[c] + TAB_D(1, 6) * k7[c] +
                            TAB_D(1, 7) * k8[c] + TAB_D(1, 8) * k9[c] + TAB_D(1, 9) * k10[c] +
                            TAB_D(1, 10) * k2[c] + TAB_D(1, 11) * k3[c];
                    r7[c] = TAB_D(2, 0) * k1[c] + TAB_D(2, 5) * k6[c] + TAB_D(2, 6) * k7[c] +
                            TAB_D(2, 7) * k8[c] + TAB_D(2, 8) * k9[c] + TAB_D(2, 9) * k10[c] +
                            TAB_D(2, 10) * k2[c] + TAB_D(2, 11) * k3[c];
                    r8[c] = TAB_D(3, 0) * k1[c] + TAB_D(3, 5) * k6[c] + TAB_D(3, 6) * k7[c] +
                            TAB_D(3, 7) * k8[c] + TAB_D(3, 8) * k9[c] + TAB_D(3, 9) * k10[c] +
                            TAB_D(3, 10) * k2[c] + TAB_D(3, 11) * k3[c];
                }
                FORN yy[c] = y[c] + h * (TAB_A(13, 0) * k1[c] + TAB_A(13, 6) * k7[c] +
                                         TAB_A(13, 7) * k8[c] + TAB_A(13, 8) * k9[c] +
                                         TAB_A(13, 9) * k10[c] + TAB_A(13, 10) * k2[c] +
                                         TAB_A(13, 11) * k3[c] + TAB_A(13, 12) * k4[c]);
                sys.rhs(P, yy, k10);
                FORN yy[c] = y[c] + h * (TAB_A(14, 0) * k1[c] + TAB_A(14, 5) * k6[c] +
                                         TAB_A(14, 6) * k7[c] + TAB_A(14, 7) * k8[c] +
                                         TAB_A(14, 10) * k2[c] + TAB_A(14, 11) * k3[c] +
                                         TAB_A(14, 12) * k4[c] + TAB_A(14, 13) * k10[c]);
                sys.rhs(P, yy, k2);
                FORN yy[c] = y[c] + h * (TAB_A(15, 0) * k1[c] + TAB_A(15, 5) * k6[c] +
                                         TAB_A(15, 6) * k7[c] + TAB_A(15, 7) * k8[c] +
                                         TAB_A(15, 8) * k9[c] + TAB_A(15, 12) * k4[c] +
                                         TAB_A(15, 13) * k10[c] + TAB_A(15, 14) * k2[c]);
                sys.rhs(P, yy, k3);
                s.nfcn += 3;
                FORN {
                    r5[c] = h * (r5[c] + TAB_D(0, 12) * k4[c] + TAB_D(0, 13) * k10[c] +
                                 TAB_D(0, 14) * k2[c] + TAB_D(0, 15) * k3[c]);
                    r6[c] = h * (r6[c] + TAB_D(1, 12) * k4[c] + TAB_D(1, 13) * k10[c] +
                                 TAB_D(1, 14) * k2[c] + TAB_D(1, 15) * k3[c]);
                    r7[c] = h * (r7[c] + TAB_D(2, 12) * k4[c] + TAB_D(2, 13) * k10[c] +
                                 TAB_D(2, 14) * k2[c] + TAB_D(2, 15) * k3[c]);
                    r8[c] = h * (r8[c] + TAB_D(3, 12) * k4[c] + TAB_D(3, 13) * k10[c] +
                                 TAB_D(3, 14) * k2[c] + TAB_D(3, 15) * k3[c]);
                }
                const double x0 = s.x, x1 = s.x + h;
                while (s.oidx < a.out_end) {
                    const double to = a.t_out[s.oidx];
                    if (in_step(x0, x1, to)) {
                        double v[NC];
                        dense_value<NC>(to, x0, h, r1, r2, r3, r4, r5, r6, r7, r8, v);
                        sys.store_dense(a, s, v);
                        s.oidx++;
                    } else {
                        break;
                    }
                }
                if constexpr (has_windows<Sys>::value) {
                    // the window ends inside this step: the next window
                    // samples the rest of it from the same polynomial
                    if (a.pend && s.oidx == a.out_end && s.oidx < a.n_out &&
                        in_step(x0, x1, a.t_out[s.oidx])) {
                        sys.keep_step(a, s, r1, r2, r3, r4, r5, r6, r7, r8, x0, h);
                        s.pend = 1;
                    }
                }
            }
            FORN {
                k1[c] = k4[c];
                y[c] = k5[c];
            }
            s.x = xph;
            if (last) {
                code = 1;
            } else {
                if (fabs(hnew) > s.hmax) hnew = s.posneg * s.hmax;
                if (s.reject) hnew = s.posneg * min_d(fabs(hnew), fabs(h));
                s.reject = 0;
            }
        }
    } else {
        // rejected (:638-645)
        hnew = h / min_d(facc1, fac11 / safe);
        s.reject = 1;
        if (s.naccpt >= 1) s.nrejct++;
    }
    if (code == 0) s.h = hnew;
    return code;
#undef TAB_A
#undef TAB_D
#undef TAB_B
#undef TAB_E5
#undef TAB_BHH
}

__device__ __forceinline__ void finish_stats(const DopArgs &a, const DopState &s, int code) {
    if (a.status) a.status[s.orb] = code;
    if (a.stats) {
        a.stats[4 * s.orb + 0] = s.nstep;
        a.stats[4 * s.orb + 1] = s.naccpt;
        a.stats[4 * s.orb + 2] = s.nrejct;
        a.stats[4 * s.orb + 3] = s.nfcn;
    }
}

template <typename V>
__device__ __forceinline__ V from_lane(V v, int src) {
    return __shfl(v, src);
}

// Dense output in windows, after an attempt that returned `code`: keep the
// state of an orbit that finished with a kept step, or that reached the
// window's end (it then leaves the launch: returns true).  Nothing in the
// last window or without windows.
template <bool DENSE, class Sys, int NC>
__device__ __forceinline__ bool window_pause(const DopArgs &a, const WinArgs &w, const Sys &sys,
                                             const DopState &s, const double (&y)[NC],
                                             const double (&k1)[NC], int code) {
    if constexpr (DENSE) {
        if (!a.pend || a.out_end >= a.n_out) return false;
        if (code != 0) {
            // finished: later windows skip it, or sample its kept step
            sys.save_state(w, a.N, s, y, k1, s.pend ? WF_PEND : 0);
            return false;
        }
        if (s.oidx >= a.out_end) {
            sys.save_state(w, a.N, s, y, k1, WF_RUN | (s.pend ? WF_PEND : 0));
            return true;
        }
    }
    return false;
}

// The wide tail of one wave: the (at most WIDE_ORBITS) orbits still active
// move to one 16-lane row each and are integrated to the end (or, in a
// window of the dense output, to the window's end).  Called with every lane
// of the wave converged; the queue is drained.
template <bool SCF, bool DENSE, int LAY>
__device__ __forceinline__ void wide_tail(const DevPot &P, const DopArgs &a, const WinArgs &w,
                                       bool active, const DopState &s6, const double (&y6)[6],
                                       const double (&k16)[6]) {
    const int lane = threadIdx.x & 63;
    const int row = lane / WIDE_ROW, c = lane % WIDE_ROW;
    const int comp = c < 12 ? c % 6 : (c - 9 < 5 ? c - 9 : 5);
    unsigned long long m = __ballot(active);
    for (int i = 0; i < row; i++) m &= m - 1ull;  // the row-th active lane
    bool ract = m != 0ull;
    const int src = ract ? __ffsll((long long)m) - 1 : lane;
    double y[1] = {0.0}, k1[1] = {0.0};
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const double vy = from_lane(y6[j], src), vk = from_lane(k16[j], src);
        if (comp == j) {
            y[0] = vy;
            k1[0] = vk;
        }
    }
    DopState s;
    s.orb = from_lane(s6.orb, src);
    s.x = from_lane(s6.x, src);
    s.xend = from_lane(s6.xend, src);
    s.h = from_lane(s6.h, src);
    s.hmax = from_lane(s6.hmax, src);
    s.posneg = from_lane(s6.posneg, src);
    s.facold = from_lane(s6.facold, src);
    s.hlamb = from_lane(s6.hlamb, src);
    s.nstep = from_lane(s6.nstep, src);
    s.naccpt = from_lane(s6.naccpt, src);
    s.stiff = from_lane(s6.stiff, src);
    s.nrejct = from_lane(s6.nrejct, src);
    s.nfcn = from_lane(s6.nfcn, src);
    s.oidx = from_lane(s6.oidx, src);
    s.reject = from_lane(s6.reject, src);
    s.iasti = from_lane(s6.iasti, src);
    s.nonsti = from_lane(s6.nonsti, src);
    s.pend = 0;
    const WideRow<SCF, LAY> row_sys{comp, c < 6};
    // every row ends: each attempt advances nstep, and nmax bounds it
    while (__any(ract)) {
        if (ract) {
            const int code = dop_attempt<SCF, DENSE, LAY>(P, a, row_sys, s, y, k1);
            if (code != 0) {
                if (c < 6) a.w_final[c * a.N + s.orb] = y[0];
                if (c == 0) finish_stats(a, s, code);
                ract = false;
            }
            if (window_pause<DENSE>(a, w, row_sys, s, y, k1, code)) ract = false;
        }
    }
}

// The quad tail of one wave: the (at most QUAD_ORBITS) orbits still active
// move to one 4-lane quad each; once at most WIDE_ORBITS remain they move on
// to 16-lane rows (wide_tail).  Called with every lane of the wave converged;
// the queue is drained.
template <bool SCF, bool DENSE, int LAY>
__device__ __forceinline__ void quad_tail(const DevPot &P, const DopArgs &a, const WinArgs &w,
                                       bool active, const DopState &s6, const double (&y6)[6],
                                       const double (&k16)[6]) {
    const int lane = threadIdx.x & 63;
    const int quad = lane >> 2, c = lane & 3;
    const int axis = c < 3 ? c : 0;
    unsigned long long m = __ballot(active);
    for (int i = 0; i < quad; i++) m &= m - 1ull;  // the quad-th active lane
    bool qact = m != 0ull;
    const int src = qact ? __ffsll((long long)m) - 1 : lane;
    double y[2] = {0.0, 0.0}, k1[2] = {0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const double vy = from_lane(y6[j], src), vk = from_lane(k16[j], src);
        if (j == axis) {
            y[0] = vy;
            k1[0] = vk;
        }
        if (j == axis + 3) {
            y[1] = vy;
            k1[1] = vk;
        }
    }
    DopState s;
    s.orb = from_lane(s6.orb, src);
    s.x = from_lane(s6.x, src);
    s.xend = from_lane(s6.xend, src);
    s.h = from_lane(s6.h, src);
    s.hmax = from_lane(s6.hmax, src);
    s.posneg = from_lane(s6.posneg, src);
    s.facold = from_lane(s6.facold, src);
    s.hlamb = from_lane(s6.hlamb, src);
    s.nstep = from_lane(s6.nstep, src);
    s.naccpt = from_lane(s6.naccpt, src);
    s.stiff = from_lane(s6.stiff, src);
    s.nrejct = from_lane(s6.nrejct, src);
    s.nfcn = from_lane(s6.nfcn, src);
    s.oidx = from_lane(s6.oidx, src);
    s.reject = from_lane(s6.reject, src);
    s.iasti = from_lane(s6.iasti, src);
    s.nonsti = from_lane(s6.nonsti, src);
    s.pend = 0;
    const QuadPair<SCF, LAY> qsys{axis, c < 3};
    // every quad ends: each attempt advances nstep, and nmax bounds it
    while (__popcll(__ballot(qact && c == 0)) > WIDE_ORBITS) {
        if (qact) {
            const int code = dop_attempt<SCF, DENSE, LAY>(P, a, qsys, s, y, k1);
            if (code != 0) {
                if (c < 3) {
                    a.w_final[axis * a.N + s.orb] = y[0];
                    a.w_final[(axis + 3) * a.N + s.orb] = y[1];
                }
                if (c == 0) finish_stats(a, s, code);
                qact = false;
            }
            if (window_pause<DENSE>(a, w, qsys, s, y, k1, code)) qact = false;
        }
    }
    // the remaining orbits back in one lane each (lane 0 of their quad), then rows
    double yb[6], kb[6];
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const int from = 4 * quad + (j % 3);
        yb[j] = from_lane(j < 3 ? y[0] : y[1], from);
        kb[j] = from_lane(j < 3 ? k1[0] : k1[1], from);
    }
    wide_tail<SCF, DENSE, LAY>(P, a, w, qact && c == 0, s, yb, kb);
}

// start the orbit at queue slot `slot`: state, first derivative and, for
// h0 = 0, Hairer's initial step (dopcor's set-up, dop853.cpp:300-320).
// Returns false for an order entry outside 0..N-1, which names no orbit: the
// slot is skipped (an orbit that no slot names keeps the GAMD_STATUS_NOT_RUN
// written before the launch).
template <bool SCF, int LAY>
__device__ __forceinline__ bool load_orbit(const DevPot &P, const DopArgs &a, int64_t slot,
                                           DopState &s, double (&y)[6], double (&k1)[6]) {
    const int64_t orb = a.order ? (int64_t)a.order[slot] : slot;
    if (orb < 0 || orb >= a.N) return false;
    s.orb = orb;
    FOR6 y[c] = a.w0[c * a.N + orb];
    s.x = a.t0[orb];
    s.xend = a.t1[orb];
    s.h = a.h0 ? a.h0[orb] : 0.0;
    s.hmax = a.hmax > 0.0 ? a.hmax : fabs(s.xend - s.x);  // dop853.cpp:787-788, 312
    s.posneg = sign_d(1.0, s.xend - s.x);
    s.facold = 1.0e-4;
    s.hlamb = 0.0;
    s.nstep = s.naccpt = s.stiff = s.nrejct = 0;
    s.nfcn = 2;
    s.oidx = 0;
    s.reject = s.iasti = s.nonsti = s.pend = 0;
    orbit_rhs<SCF, LAY>(P, y, k1);
    if (s.h == 0.0) s.h = hinit<SCF, LAY>(P, s.x, y, s.posneg, k1, s.hmax, a.rtol, a.atol);
    return true;
}

// samples [s.oidx, out_end) of the step kept by the previous window; returns
// true when the step still covers the sample after this window's last
template <bool SCF, int LAY>
__device__ __forceinline__ bool sample_kept(const DopArgs &a, DopState &s) {
    constexpr int NC = 6;
    const double *pp = a.pend + s.orb;
    const int64_t N = a.N;
    double r[8][NC];
#pragma unroll
    for (int k = 0; k < 8; k++) FORN r[k][c] = pp[(k * NC + c) * N];
    const double x0 = pp[(8 * NC) * N], h = pp[(8 * NC + 1) * N], x1 = x0 + h;
    const OneLane<SCF, LAY> sys;
    while (s.oidx < a.out_end) {
        const double to = a.t_out[s.oidx];
        if (!in_step(x0, x1, to)) return false;
        double v[NC];
        dense_value<NC>(to, x0, h, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], v);
        sys.store_dense(a, s, v);
        s.oidx++;
    }
    return s.oidx < a.n_out && in_step(x0, x1, a.t_out[s.oidx]);
}

// the orbit at queue slot `slot` into one lane: from w0 (load_orbit), or, in
// a later window of the dense output, from its kept state, after sampling
// what its kept step covers of this window.  Returns whether it steps on in
// this launch.
template <bool SCF, bool DENSE, int LAY>
__device__ __forceinline__ bool take_orbit(const DevPot &P, const DopArgs &a, const WinArgs &w,
                                           int64_t slot, DopState &s, double (&y)[6],
                                           double (&k1)[6]) {
    if (!DENSE || !a.pend || w.first) return load_orbit<SCF, LAY>(P, a, slot, s, y, k1);
    const int64_t orb = a.order ? (int64_t)a.order[slot] : slot;
    // slots that name no orbit were skipped by the first window too
    int64_t flags = 0;
    if (orb >= 0 && orb < a.N) flags = win_load(w, a.N, orb, s, y, k1);
    // an orbit behind this window (one a window did not visit) is not run:
    // its samples would fall before the window's slot
    if (s.oidx < a.out_base) flags = 0;
    if (flags & WF_PEND) {
        const bool more = sample_kept<SCF, LAY>(a, s);
        flags = (flags & WF_RUN) | (more ? WF_PEND : 0);
        if (more || !(flags & WF_RUN)) OneLane<SCF, LAY>{}.save_state(w, a.N, s, y, k1, flags);
        if (more) flags = 0;  // no steps before its samples are out
    }
    return (flags & WF_RUN) != 0;
}

template <bool SCF, bool DENSE, int LAY>
__global__ __launch_bounds__(256) void k_dop853(const DevPot *__restrict__ Pp, DopArgs a,
                                                WinArgs w) {
    const DevPot &P = *Pp;
    const int lane = threadIdx.x & 63;

    bool active = false, exhausted = false;
    DopState s;
    s.orb = 0;
    s.x = s.xend = s.h = s.hmax = 0.0;
    s.posneg = 1.0;
    s.facold = 1e-4;
    s.hlamb = 0.0;
    s.nstep = s.naccpt = s.stiff = s.nrejct = s.nfcn = s.oidx = 0;
    s.reject = s.iasti = s.nonsti = s.pend = 0;
    double y[6], k1[6];
    FOR6 {
        y[c] = 0.0;
        k1[c] = 0.0;
    }
    __shared__ double tab[TAB_WORDS];
    for (int i = threadIdx.x; i < TAB_WORDS; i += blockDim.x) tab[i] = kTableau.w[i];
    __syncthreads();

    if (a.n_tail > 0) {
        // the first waves of the grid take the queue order's first n_tail
        // (longest) orbits, four per wave, and integrate them in 16-lane rows
        // from the start; the other waves share the rest of the queue
        const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
        if (wave * WIDE_ORBITS < a.n_tail) {
            const int64_t slot = wave * WIDE_ORBITS + lane;
            if (lane < WIDE_ORBITS && slot < a.n_tail)
                active = take_orbit<SCF, DENSE, LAY>(P, a, w, slot, s, y, k1);
            wide_tail<SCF, DENSE, LAY>(P, a, w, active, s, y, k1);
            return;
        }
    }

    while (true) {
        // ---- refill idle lanes from the queue (one atomic per wave) ----
        const bool need = !active && !exhausted;
        const unsigned long long want = __ballot(need);
        if (want) {
            const int leader = __ffsll((long long)want) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(a.queue, (unsigned long long)__popcll(want));
            base = __shfl(base, leader);
            if (need) {
                const unsigned long long below = want & ((1ull << lane) - 1ull);
                const unsigned long long slot = base + (unsigned long long)__popcll(below);
                const unsigned long long qslot = slot + (unsigned long long)a.n_tail;
                if (qslot < (unsigned long long)a.N) {
                    active = take_orbit<SCF, DENSE, LAY>(P, a, w, (int64_t)qslot, s, y, k1);
                } else {
                    exhausted = true;
                }
            }
        }
        // A slot can name an orbit with nothing to do (one that has finished,
        // in a window of the dense output, or an order entry that names no
        // orbit): idle lanes that have not yet found the queue empty take
        // slots again before the wave leaves or spreads its last orbits, so
        // that every slot's orbit is visited.  Without dense output the test
        // stays off the attempt loop's common path (a ballot per attempt cost
        // ~10 % of the full launch); the dense kernel takes it every pass.
        if constexpr (DENSE) {
            const bool searching = __any(!active && !exhausted);
            if (!__any(active)) {
                if (!searching) break;
                continue;
            }
            if (a.wide && !searching) {
                // every idle lane has found the queue empty; with few orbits
                // left, spread them over rows and finish them there
                const int left = __popcll(__ballot(active));
                if (left <= WIDE_ORBITS) {
                    wide_tail<SCF, DENSE, LAY>(P, a, w, active, s, y, k1);
                    break;
                }
                if (left <= QUAD_ORBITS && a.wide > 1) {
                    quad_tail<SCF, DENSE, LAY>(P, a, w, active, s, y, k1);
                    break;
                }
            }
        } else {
            if (!__any(active)) {
                if (__any(!exhausted)) continue;
                break;
            }
            const int left = a.wide ? __popcll(__ballot(active)) : 64;
            if (left <= QUAD_ORBITS && !__any(!active && !exhausted)) {
                // every idle lane has found the queue empty; with few orbits
                // left, spread them over rows and finish them there
                if (left <= WIDE_ORBITS) {
                    wide_tail<SCF, DENSE, LAY>(P, a, w, active, s, y, k1);
                    break;
                }
                if (a.wide > 1) {
                    quad_tail<SCF, DENSE, LAY>(P, a, w, active, s, y, k1);
                    break;
                }
            }
        }
        if (!active) continue;

        const int code = dop_attempt<SCF, DENSE, LAY>(P, a, OneLaneTab<SCF, LAY>{{}, tab}, s, y, k1);
        if (code != 0) {
            FOR6 a.w_final[c * a.N + s.orb] = y[c];
            finish_stats(a, s, code);
            active = false;
        }
        if (window_pause<DENSE>(a, w, OneLane<SCF, LAY>{}, s, y, k1, code)) active = false;
    }
}

// ---------------------------------------------------------------------------
// mockstream_dop853's release systems (mockstream.pyx:286-303): system i is
// the bodies (progenitor first, then any perturbers) at their states at the
// release plus the release's particles, one (K + k, 6) system integrated by
// dop853_step (dop853.pyx:29-79) from rel_t[i] to tfinal; all systems run
// concurrently.  Systems of at most 4 bodies run on 4-lane quads (16 per
// wave); larger ones (more perturbers or particles per release) take a wave
// each, in a second launch.
struct StreamArgs {
    const double *body_w;     // (nrel, K, 6) bodies at each release
    const double *rel_t;      // (nrel,)
    const int32_t *nstream;   // (nrel,) particles released
    const int64_t *offset;    // (nrel,) first particle row of each release
    const double *stream_w0;  // (ntot, 6)
    int64_t nrel;
    double tfinal, dt0;
    double *out;              // (ntot, 6)
    double *body_final;       // (nrel, K, 6)
    int32_t *status;          // (nrel,)
    int64_t *stats;           // (nrel, 4) or nullptr
};

// G = 4: QuadBodies; G = 16: RowBodies; G = 64: WaveBodies.  quads: the
// quad or row launch takes the systems of at most 4 bodies (so the wave
// launch skips them).
template <bool SCF, int LAY, int G>
__global__ __launch_bounds__(256) void k_dop853_stream(const DevPot *__restrict__ Pp, BodyTable bt,
                                                       int quads, StreamArgs sa, DopArgs a) {
    const DevPot &P = *Pp;
    const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t rel = gid / G;  // uniform over the group
    const int body = G == 16 ? (int)(gid % 16) / 4 : (int)(gid % G);
    const int lane4 = (int)(gid % 4);
    const int axis = lane4 < 3 ? lane4 : 0;
    if (rel >= sa.nrel) return;
    const int K = bt.n;
    const int k = sa.nstream[rel];
    const bool lead = G == 16 ? (gid % 16) == 0 : body == 0;
    if (k <= 0) {
        if (lead && (G != 64 || !quads)) sa.status[rel] = 0;
        return;
    }
    const int nb = K + k;
    if (G != 64 && nb > 4) return;             // a wave's system
    if (G == 64 && quads && nb <= 4) return;   // a quad's or row's system
    if (nb > (G == 16 ? 4 : G)) {
        if (lead) sa.status[rel] = GAMD_ERR_UNSUPPORTED;
        return;
    }
    using Sys = typename std::conditional<
        G == 4, QuadBodies<SCF, LAY>,
        typename std::conditional<G == 16, RowBodies<SCF, LAY>, WaveBodies<SCF, LAY>>::type>::type;
    constexpr int NC = Sys::NC;
    auto make_sys = [&]() {
        if constexpr (G == 16)
            return Sys{body, axis, lane4, nb, bt};
        else
            return Sys{body, nb, bt};
    };
    const Sys sys = make_sys();
    double w[6];
    if (body >= K && body < nb) {
        const double *src = sa.stream_w0 + 6 * (sa.offset[rel] + body - K);
        FOR6 w[c] = src[c];
    } else {
        const int b = body < K ? body : 0;
        FOR6 w[c] = sa.body_w[6 * (rel * K + b) + c];
    }
    double y[NC], k1[NC];
    if constexpr (NC == 6) {
        FOR6 y[c] = w[c];
    } else {
        y[0] = axis == 0 ? w[0] : (axis == 1 ? w[1] : w[2]);
        y[1] = axis == 0 ? w[3] : (axis == 1 ? w[4] : w[5]);
    }
    DopState s;
    s.orb = rel;
    s.x = sa.rel_t[rel];
    s.xend = sa.tfinal;
    s.h = sa.dt0;
    s.hmax = fabs(sa.tfinal - s.x);  // hmax 0 -> xend - x (dop853.cpp:787-788), |.| in dopcor
    s.posneg = sign_d(1.0, s.xend - s.x);
    s.facold = 1.0e-4;
    s.hlamb = 0.0;
    s.nstep = s.naccpt = s.stiff = s.nrejct = 0;
    s.nfcn = 2;
    s.oidx = 0;
    s.reject = s.iasti = s.nonsti = s.pend = 0;
    sys.rhs(P, y, k1);
    int code = 0;
    // every group ends: each attempt advances nstep, and nmax bounds it
    while (code == 0) code = dop_attempt<SCF, false, LAY>(P, a, sys, s, y, k1);
    double *dst = nullptr;
    if (body >= K && body < nb) dst = sa.out + 6 * (sa.offset[rel] + body - K);
    if (body < K) dst = sa.body_final + 6 * (rel * K + body);
    if (dst) {
        if constexpr (NC == 6) {
            FOR6 dst[c] = y[c];
        } else if (lane4 < 3) {
            dst[axis] = y[0];
            dst[axis + 3] = y[1];
        }
    }
    if (lead) {
        sa.status[rel] = code;
        if (sa.stats) {
            sa.stats[4 * rel + 0] = s.nstep;
            sa.stats[4 * rel + 1] = s.naccpt;
            sa.stats[4 * rel + 2] = s.nrejct;
            sa.stats[4 * rel + 3] = s.nfcn;
        }
    }
}

// direct_nbody_dop853 (nbody.pyx:30-115): the a.N bodies of w0 (6, N) as one
// system, dense output at t_out (dop853_helper, dop853.pyx:90-193).  G = 64:
// one wave, body b on lane b (WaveBodies); G = 16: at most 4 bodies on one
// 16-lane row (RowBodies, the other rows of the wave exit at once).
template <bool SCF, bool DENSE, int LAY, int G>
__global__ __launch_bounds__(64) void k_dop853_nbody(const DevPot *__restrict__ Pp, BodyTable bt,
                                                     double t0, double t1, double h0, DopArgs a) {
    const DevPot &P = *Pp;
    const int t = threadIdx.x;
    if (t >= G) return;
    const int body = G == 16 ? t / 4 : t;
    const int lane4 = t % 4;
    const int axis = lane4 < 3 ? lane4 : 0;
    const int nb = (int)a.N;
    using Sys = typename std::conditional<G == 16, RowBodies<SCF, LAY>, WaveBodies<SCF, LAY>>::type;
    constexpr int NC = Sys::NC;
    auto make_sys = [&]() {
        if constexpr (G == 16)
            return Sys{body, axis, lane4, nb, bt};
        else
            return Sys{body, nb, bt};
    };
    const Sys sys = make_sys();
    const int src = body < nb ? body : 0;
    double y[NC], k1[NC];
    if constexpr (NC == 6) {
        FOR6 y[c] = a.w0[c * a.N + src];
    } else {
        y[0] = a.w0[axis * a.N + src];
        y[1] = a.w0[(axis + 3) * a.N + src];
    }
    DopState s;
    s.orb = 0;
    s.x = t0;
    s.xend = t1;
    s.h = h0;
    s.hmax = a.hmax > 0.0 ? a.hmax : fabs(t1 - t0);  // dop853.cpp:787-788
    s.posneg = sign_d(1.0, t1 - t0);
    s.facold = 1.0e-4;
    s.hlamb = 0.0;
    s.nstep = s.naccpt = s.stiff = s.nrejct = 0;
    s.nfcn = 2;
    s.oidx = 0;
    s.reject = s.iasti = s.nonsti = s.pend = 0;
    sys.rhs(P, y, k1);
    int code = 0;
    // the group ends: each attempt advances nstep, and nmax bounds it
    while (code == 0) code = dop_attempt<SCF, DENSE, LAY>(P, a, sys, s, y, k1);
    if (body < nb) {
        if constexpr (NC == 6) {
            FOR6 a.w_final[c * a.N + body] = y[c];
        } else if (lane4 < 3) {
            a.w_final[axis * a.N + body] = y[0];
            a.w_final[(axis + 3) * a.N + body] = y[1];
        }
    }
    if (t == 0) finish_stats(a, s, code);
}

// One launch of the batched kernel.  order (nullable): workgroup i takes
// batch order[i] (entries outside [0, nbatches) are skipped).  st (nullable):
// each batch's kept state, in the per-orbit window layout (WS_WORDS, N), every
// orbit's column holding its own y and k1 and the batch's scalars; first: the
// batches start from w0, else from st (a batch kept as finished is skipped).
// pause > 0: a batch still stepping after that many attempts in this launch is
// kept in st and left for the next launch.  Pausing between attempts changes
// no step: a batch resumed from st takes the steps the one launch would.
struct BatchRun {
    const int64_t *order;
    int64_t nbatches;
    int64_t pause;
    WinArgs w;
};

// dop853_integrate_hamiltonian with nbatch orbits per system (dop853.pyx:
// 196-250): batch b, the orbits [b nbatch, min((b + 1) nbatch, N)), is one
// (6, nb) system integrated by one workgroup (BatchLanes) from t0 to t1 with
// initial step h0 (0: hinit over the batch) and dense output at t_out
// (dop853_helper, :90-193).  The batch's status and step counts go to each of
// its orbits.  Workgroups are independent: the hardware dispatcher hands
// each free slot the next batch.
template <bool SCF, bool DENSE, int LAY, int MAXT>
__global__ __launch_bounds__(MAXT) void k_dop853_batch(const DevPot *__restrict__ Pp, int nbatch,
                                                      double t0, double t1, double h0, DopArgs a,
                                                      BatchRun br) {
    extern __shared__ double lds[];
    const DevPot &P = *Pp;
    const int j = threadIdx.x;
    int64_t b = blockIdx.x;
    if (br.order) {
        b = br.order[b];
        if (b < 0 || b >= br.nbatches) return;
    }
    const int64_t base = b * nbatch;
    const int64_t left = a.N - base;
    const int nb = left < (int64_t)nbatch ? (int)left : nbatch;
    double *tab = lds + 12 * nbatch + 2;
    for (int i = j; i < TAB_WORDS; i += blockDim.x) tab[i] = kTableau.w[i];
    __syncthreads();
    const BatchLanes<SCF, LAY> sys{j, nb, base, lds, tab};
    const int64_t orb = base + (j < nb ? j : 0);
    double y[6], k1[6];
    DopState s;
    if (br.w.st && !br.w.first) {
        // every column of the batch was kept with the same flags
        if (win_load(br.w, a.N, orb, s, y, k1) == 0) return;
    } else {
        FOR6 y[c] = a.w0[c * a.N + orb];
        s.orb = orb;
        s.x = t0;
        s.xend = t1;
        s.h = h0;
        s.hmax = a.hmax > 0.0 ? a.hmax : fabs(t1 - t0);  // dop853.cpp:787-788, 312
        s.posneg = sign_d(1.0, t1 - t0);
        s.facold = 1.0e-4;
        s.hlamb = 0.0;
        s.nstep = s.naccpt = s.stiff = s.nrejct = 0;
        s.nfcn = 2;
        s.oidx = 0;
        s.reject = s.iasti = s.nonsti = s.pend = 0;
        sys.rhs(P, y, k1);
        if (s.h == 0.0)
            s.h = hinit_batch<SCF, LAY>(P, sys, y, s.posneg, k1, s.hmax, a.rtol, a.atol);
    }
    int code = 0;
    // every attempt advances nstep and nmax bounds it; the step decisions are
    // uniform over the workgroup, so its lanes leave together
    for (int64_t n = br.pause > 0 ? br.pause : -1; code == 0 && n != 0; n--)
        code = dop_attempt<SCF, DENSE, LAY>(P, a, sys, s, y, k1);
    if (j >= nb) return;
    if (code == 0) {  // paused: kept for the next launch
        win_save_scalars(br.w, a.N, s, WF_RUN);
        FOR6 win_save_comp(br.w, a.N, orb, c, y[c], k1[c]);
        return;
    }
    FOR6 a.w_final[c * a.N + orb] = y[c];
    finish_stats(a, s, code);
    if (br.w.st) win_save_scalars(br.w, a.N, s, 0);
}

// Predicted remaining attempts of each batch after a pause of `pause`
// attempts from t0: pause (t1 - x) / (x - t0), +inf without progress, -1 for
// a finished batch.
__global__ void k_batch_keys(const double *__restrict__ st, int64_t N, int nbatch,
                             int64_t nbatches, double t0, double t1, int64_t pause,
                             double *__restrict__ key) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbatches) return;
    const int64_t col = b * nbatch;
    const int64_t flags = reinterpret_cast<const int64_t *>(st)[WS_FLAGS * N + col];
    const double x = st[WS_X * N + col];
    double k = -1.0;
    if (flags != 0) {
        const double done = fabs(x - t0), rest = fabs(t1 - x);
        k = done > 0.0 ? (double)pause * (rest / done) : INFINITY;
        if (!(k >= 0.0)) k = INFINITY;  // NaN state: run it early, it will stop
    }
    key[b] = k;
}

// order = the batches by descending key, ties by index: batch b goes to
// position #{keys above key[b]} + #{equal keys at lower index}.  Each
// workgroup stages tiles of the keys in LDS; n^2 comparisons in all, which
// for the batch counts it serves (one batch per 1 to 1,024 orbits, each
// batch thousands of attempts) stays far below the integration's time.
constexpr int RANK_TILE = 256;
__global__ __launch_bounds__(RANK_TILE) void k_batch_rank(const double *__restrict__ key,
                                                          int64_t n, int64_t *__restrict__ order) {
    __shared__ double tile[RANK_TILE];
    const int64_t b = (int64_t)blockIdx.x * RANK_TILE + threadIdx.x;
    const double kb = b < n ? key[b] : 0.0;
    int64_t rank = 0;
    for (int64_t t = 0; t < n; t += RANK_TILE) {
        __syncthreads();
        if (t + threadIdx.x < n) tile[threadIdx.x] = key[t + threadIdx.x];
        __syncthreads();
        const int m = n - t < RANK_TILE ? (int)(n - t) : RANK_TILE;
        for (int i = 0; i < m; i++) {
            const double ki = tile[i];
            rank += (ki > kb) | ((ki == kb) & (t + i < b));
        }
    }
    if (b < n) order[rank] = b;
}

// Batches in two launches: the first takes each batch through a few attempts
// and keeps it, the kept states rank the batches (longest predicted remainder
// first, so the dispatcher does not leave the costliest batches to the end)
// and the second resumes them in that order.  GAMD_DOP853_BATCH_ORDER: 0 =
// one launch in index order, 1 = two launches whatever the batch count;
// unset: two launches when there are more batches than resident workgroups.
// GAMD_DOP853_BATCH_PILOT: the first launch's attempts per batch (default
// 192).  Same results every way.
int batch_order_mode() {
    const char *env = getenv("GAMD_DOP853_BATCH_ORDER");
    if (env && env[0] == '0') return 0;
    if (env && env[0] == '1') return 1;
    return -1;
}

int64_t batch_pilot() {
    const char *env = getenv("GAMD_DOP853_BATCH_PILOT");
    const int64_t v = env && env[0] ? atoll(env) : 0;
    return v > 0 ? v : 192;
}

template <bool SCF, bool DENSE, int LAY>
int launch_batch(const DevPot *P, int nbatch, double t0, double t1, double h0, const DopArgs &a,
                 hipStream_t s) {
    const int threads = (nbatch + 63) / 64 * 64;
    const int64_t grid = (a.N + nbatch - 1) / nbatch;
    const size_t lds = (size_t)(12 * nbatch + 2 + TAB_WORDS) * sizeof(double);
    const auto kern = threads <= 256 ? k_dop853_batch<SCF, DENSE, LAY, 256>
                                     : k_dop853_batch<SCF, DENSE, LAY, GAMD_DOP853_MAX_NBATCH>;
    BatchRun br{nullptr, grid, 0, WinArgs{nullptr, 1}};
    const int mode = batch_order_mode();
    const int64_t pilot = batch_pilot();
    bool two = mode == 1;
    if (mode < 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds) !=
            hipSuccess)
            per_cu = 0;
        // the ranking compares every pair of batches: at most 2^22 batches
        // (~2 s of ranking there; integrating that many batches takes tens of
        // seconds even at one orbit per batch)
        two = per_cu > 0 && grid > (int64_t)per_cu * cu_count() && grid <= ((int64_t)1 << 22);
    }
    double *st = nullptr;
    if (two) {
        // kept states, then keys and order
        const size_t bytes = ((size_t)WS_WORDS * a.N + 2 * (size_t)grid) * sizeof(double);
        if (hipMallocAsync((void **)&st, bytes, s) != hipSuccess) {
            (void)hipGetLastError();
            st = nullptr;
        }
    }
    if (!st) {
        hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(threads), lds, s, P, nbatch, t0, t1,
                           h0, a, br);
        return launch_status();
    }
    double *key = st + (size_t)WS_WORDS * a.N;
    int64_t *order = reinterpret_cast<int64_t *>(key + grid);
    br.pause = pilot;
    br.w.st = st;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(threads), lds, s, P, nbatch, t0, t1, h0,
                       a, br);
    const unsigned kb = (unsigned)((grid + 255) / 256);
    hipLaunchKernelGGL(k_batch_keys, dim3(kb), dim3(256), 0, s, st, a.N, nbatch, grid, t0, t1,
                       pilot, key);
    const unsigned rb = (unsigned)((grid + RANK_TILE - 1) / RANK_TILE);
    hipLaunchKernelGGL(k_batch_rank, dim3(rb), dim3(RANK_TILE), 0, s, key, grid, order);
    br.order = order;
    br.pause = 0;
    br.w.first = 0;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(threads), lds, s, P, nbatch, t0, t1, h0,
                       a, br);
    const int rc = launch_status();
    const hipError_t fe = hipFreeAsync(st, s);
    return rc != GAMD_OK ? rc : (fe == hipSuccess ? GAMD_OK : GAMD_ERR_HIP);
}

// Systems of at most 4 bodies on one 16-lane row (two components per lane)
// unless GAMD_DOP853_BODY_ROWS=0; =1 forces rows wherever they fit.  Same
// results either way.
bool body_rows(int64_t systems, int nb) {
    if (nb > 4) return false;
    const char *env = getenv("GAMD_DOP853_BODY_ROWS");
    if (env && env[0] == '0') return false;
    if (env && env[0] == '1') return true;
    // rows while they fit about two waves per SIMD (the launch is then bound
    // by its longest systems, whose attempts rows make shorter)
    return 16 * systems <= 2 * 64 * 4 * (int64_t)cu_count();
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

template <bool SCF, bool DENSE, int LAY>
int64_t dop_grid(int block, int64_t N) {
    // the persistent grid: every workgroup that fits at once, at most one
    // lane per orbit (queried before any stream capture begins)
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dop853<SCF, DENSE, LAY>, block,
                                                     0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    int64_t blocks = (int64_t)per_cu * cu_count();
    const int64_t needed = (N + block - 1) / block;
    if (blocks > needed) blocks = needed;
    return blocks < 1 ? 1 : blocks;
}

// dedicated wide waves: at most half the grid's waves, never all of it
int64_t cap_tail(const DopArgs &a, int64_t blocks, int block) {
    const int64_t waves = blocks * (block / 64);
    const int64_t max_tail = a.wide ? WIDE_ORBITS * (waves / 2) : 0;
    const int64_t t = a.n_tail < max_tail ? a.n_tail : max_tail;
    return t < a.N ? t : a.N;
}

template <bool SCF, bool DENSE, int LAY>
int launch(const DevPot *P, const DopArgs &a, int block, hipStream_t s) {
    const int64_t blocks = dop_grid<SCF, DENSE, LAY>(block, a.N);
    DopArgs b = a;
    b.n_tail = cap_tail(a, blocks, block);
    WinArgs w;
    w.st = nullptr;
    w.first = 1;
    hipLaunchKernelGGL((k_dop853<SCF, DENSE, LAY>), dim3((unsigned)blocks), dim3(block), 0, s, P,
                       b, w);
    return launch_status();
}

}  // namespace

namespace {

// gamd_dop853's argument checks and gala's defaults (dop853.cpp:673-694)
int dop853_check(const gamd_potential *pot, const double *w0, const double *t0, const double *t1,
                 const double *w_final, const double *w_out, const double *t_out, size_t n_out,
                 int64_t n_tail, int &block, int64_t &nmax, int64_t &nstiff) {
    if (!pot || !w0 || !t0 || !t1 || !w_final || (w_out && (!t_out || n_out == 0)))
        return GAMD_ERR_INVALID;
    if (n_tail < 0) return GAMD_ERR_INVALID;
    if (block == 0) block = 256;
    // whole waves, at most the kernel's launch bound
    if (block < 64 || block > 256 || block % 64) return GAMD_ERR_INVALID;
    if (nmax == 0) nmax = 1000000;                   // dop853.cpp:673-674
    else if (nmax < 0) return GAMD_ERR_INVALID;      // :675-679
    if (nstiff == 0) nstiff = 1000;                  // :691-692
    else if (nstiff < 0) nstiff = 0;                 // :693-694 (never tested)
    return GAMD_OK;
}

// gamd_dop853's kernel arguments
DopArgs dop_args(const double *w0, size_t N, const double *t0, const double *t1, const double *h0,
                 double rtol, double atol, int64_t nmax, int64_t nstiff, double hmax,
                 const double *t_out, size_t n_out, double *w_out, double *w_final,
                 int32_t *status, int64_t *stats, const int32_t *order, int64_t n_tail,
                 unsigned long long *queue) {
    DopArgs a;
    a.w0 = w0;
    a.N = (int64_t)N;
    a.t0 = t0;
    a.t1 = t1;
    a.h0 = h0;
    a.rtol = rtol;
    a.atol = atol;
    a.nmax = nmax;
    a.nstiff = nstiff;
    a.hmax = hmax;
    a.t_out = t_out;
    a.n_out = (int64_t)n_out;
    a.out_base = 0;
    a.out_rows = a.out_end = a.n_out;
    a.pend = nullptr;
    a.w_out = w_out;
    a.w_final = w_final;
    a.status = status;
    a.stats = stats;
    a.order = order;
    a.n_tail = n_tail;
    a.queue = queue;
    // GAMD_DOP853_WIDE=0 keeps one lane per orbit to the end, =1 skips the
    // quad stage (same results every way)
    const char *wide_env = getenv("GAMD_DOP853_WIDE");
    a.wide = !wide_env ? 2 : (wide_env[0] == '0' ? 0 : (wide_env[0] == '1' ? 1 : 2));
    a.uround = 2.220446049250313e-16;  // np.finfo(float).eps, dop853.pyx:165
    return a;
}

// The launches of gamd_dop853 after its checks, with the work-queue counter
// in `queue` (device memory the caller provides; cleared here by a kernel,
// so that the sequence holds no memset node).
int dop853_run(const gamd_potential *pot, const double *w0, size_t N, const double *t0,
               const double *t1, const double *h0, double rtol, double atol, int64_t nmax,
               int64_t nstiff, double hmax, const double *t_out, size_t n_out, double *w_out,
               double *w_final, int32_t *status, int64_t *stats, const int32_t *order,
               int64_t n_tail, int block, unsigned long long *queue, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, queue);
    int rc = launch_status();
    if (rc != GAMD_OK) return rc;
    DopArgs a = dop_args(w0, N, t0, t1, h0, rtol, atol, nmax, nstiff, hmax, t_out, n_out, w_out,
                         w_final, status, stats, order, n_tail, queue);
    if (order && status) {
        // an orbit that the order does not name keeps this status
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, status,
                           (int64_t)N, (int32_t)GAMD_STATUS_NOT_RUN);
        rc = launch_status();
        if (rc != GAMD_OK) return rc;
    }
    const bool scf = pot->dev.has_scf;
    const bool mw = pot->dev.layout == LAYOUT_MW;
    if (w_out)
        return scf  ? launch<true, true, LAYOUT_GENERIC>(pot->ddev, a, block, s)
               : mw ? launch<false, true, LAYOUT_MW>(pot->ddev, a, block, s)
                    : launch<false, true, LAYOUT_GENERIC>(pot->ddev, a, block, s);
    return scf  ? launch<true, false, LAYOUT_GENERIC>(pot->ddev, a, block, s)
           : mw ? launch<false, false, LAYOUT_MW>(pot->ddev, a, block, s)
                : launch<false, false, LAYOUT_GENERIC>(pot->ddev, a, block, s);
}

}  // namespace

extern "C" int gamd_dop853(const gamd_potential *pot, const double *w0, size_t N, const double *t0,
                           const double *t1, const double *h0, double rtol, double atol,
                           int64_t nmax, int64_t nstiff, double hmax, const double *t_out,
                           size_t n_out, double *w_out, double *w_final, int32_t *status,
                           int64_t *stats, const int32_t *order, int64_t n_tail, int block,
                           void *stream) {
    int rc = dop853_check(pot, w0, t0, t1, w_final, w_out, t_out, n_out, n_tail, block, nmax,
                          nstiff);
    if (rc != GAMD_OK || N == 0) return rc;
    hipStream_t s = as_stream(stream);
    unsigned long long *queue = nullptr;
    GAMD_HIP_CHECK(hipMallocAsync((void **)&queue, sizeof(unsigned long long), s));
    rc = dop853_run(pot, w0, N, t0, t1, h0, rtol, atol, nmax, nstiff, hmax, t_out, n_out, w_out,
                    w_final, status, stats, order, n_tail, block, queue, s);
    const hipError_t fe = hipFreeAsync(queue, s);
    if (rc == GAMD_OK && fe != hipSuccess) rc = GAMD_ERR_HIP;
    return rc;
}

// ---------------------------------------------------------------------------
// gamd_dop853_noalloc: gamd_dop853 with its work-queue counter in `scratch`,
// a fixed device array of this library, instead of a per-call stream-ordered
// allocation.  Each GPU holds its own copy of the array, at its own address,
// from the time the library's code object is loaded there until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// returns with that GPU current (the array's name in host code is a
// host-side stand-in, not a device address).
static __device__ __align__(256) unsigned long long scratch[GAMD_DOP853_SCRATCH_BYTES / 8];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex scratch_mu;
unsigned long long *scratch_dev[SCRATCH_MAX_DEVICES] = {};

// The current GPU's copy of `scratch`, looked up once per GPU.
int scratch_address(unsigned long long **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return GAMD_ERR_HIP;
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return GAMD_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "gala_amd: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return GAMD_ERR_HIP;
        }
        scratch_dev[dev] = static_cast<unsigned long long *>(p);
    }
    *out = scratch_dev[dev];
    return GAMD_OK;
}

}  // namespace

extern "C" int gamd_dop853_noalloc(const gamd_potential *pot, const double *w0, size_t N,
                                   const double *t0, const double *t1, const double *h0,
                                   double rtol, double atol, int64_t nmax, int64_t nstiff,
                                   double hmax, const double *t_out, size_t n_out, double *w_out,
                                   double *w_final, int32_t *status, int64_t *stats,
                                   const int32_t *order, int64_t n_tail, int block,
                                   void *stream) {
    int rc = dop853_check(pot, w0, t0, t1, w_final, w_out, t_out, n_out, n_tail, block, nmax,
                          nstiff);
    if (rc != GAMD_OK || N == 0) return rc;
    unsigned long long *queue = nullptr;
    rc = scratch_address(&queue);
    if (rc != GAMD_OK) return rc;
    // the counter is the array's first word; the kernels read no other
    return dop853_run(pot, w0, N, t0, t1, h0, rtol, atol, nmax, nstiff, hmax, t_out, n_out, w_out,
                      w_final, status, stats, order, n_tail, block, queue, as_stream(stream));
}

// ---------------------------------------------------------------------------
// gamd_dop853_dense_stream: gamd_dop853's dense output in windows of
// `window_rows` samples, each window's slot copied to host memory on a
// second stream while the next window runs (two slots in turn).

namespace {

// the window launches and their copies; kernels on s, copies on c
template <bool SCF, int LAY>
int dense_windows(const DevPot *P, DopArgs a, int block, int64_t blocks, int64_t W,
                  double *host_out, double *ring, WinArgs w, hipStream_t s, hipStream_t c,
                  hipEvent_t comp, hipEvent_t copied[2]) {
    auto kern = k_dop853<SCF, true, LAY>;
    const int64_t n_out = a.n_out, N = a.N;
    const int64_t nwin = (n_out + W - 1) / W;
    if (a.order && a.status) {
        // an orbit that the order does not name keeps this status
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s,
                           a.status, N, (int32_t)GAMD_STATUS_NOT_RUN);
        const int rc = launch_status();
        if (rc != GAMD_OK) return rc;
    }
    for (int64_t k = 0; k < nwin; k++) {
        const int sl = (int)(k & 1);
        double *slot = ring + (size_t)sl * 6 * W * N;
        if (k >= 2) GAMD_HIP_CHECK(hipStreamWaitEvent(s, copied[sl], 0));
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, a.queue);
        int rc = launch_status();
        if (rc != GAMD_OK) return rc;
        DopArgs b = a;
        b.w_out = slot;
        b.out_base = k * W;
        b.out_rows = W;
        b.out_end = (k + 1) * W < n_out ? (k + 1) * W : n_out;
        WinArgs wk = w;
        wk.first = k == 0;
        hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(block), 0, s, P, b, wk);
        rc = launch_status();
        if (rc != GAMD_OK) return rc;
        GAMD_HIP_CHECK(hipEventRecord(comp, s));
        GAMD_HIP_CHECK(hipStreamWaitEvent(c, comp, 0));
        const int64_t rows = b.out_end - b.out_base;
        for (int q = 0; q < 6; q++)
            GAMD_HIP_CHECK(hipMemcpyAsync(host_out + ((size_t)q * n_out + b.out_base) * N,
                                          slot + (size_t)q * W * N,
                                          (size_t)rows * N * sizeof(double),
                                          hipMemcpyDeviceToHost, c));
        GAMD_HIP_CHECK(hipEventRecord(copied[sl], c));
    }
    // join: the compute stream waits for the last copies
    GAMD_HIP_CHECK(hipStreamWaitEvent(s, copied[(nwin - 1) & 1], 0));
    if (nwin >= 2) GAMD_HIP_CHECK(hipStreamWaitEvent(s, copied[(nwin - 2) & 1], 0));
    return GAMD_OK;
}

}  // namespace

extern "C" int64_t gamd_dop853_dense_workspace_bytes(size_t N, int64_t window_rows) {
    if (window_rows < 1) return -1;
    return (int64_t)(256 + (2 * 6 * (size_t)window_rows + WS_WORDS + WS_PEND_WORDS) * N *
                               sizeof(double));
}

extern "C" int gamd_dop853_dense_stream(const gamd_potential *pot, const double *w0, size_t N,
                                        const double *t0, const double *t1, const double *h0,
                                        double rtol, double atol, int64_t nmax, int64_t nstiff,
                                        double hmax, const double *t_out, size_t n_out,
                                        int64_t window_rows, int use_graph, double *host_out,
                                        double *w_final, int32_t *status, int64_t *stats,
                                        const int32_t *order, int64_t n_tail, int block,
                                        void *workspace, size_t workspace_bytes,
                                        void *stream) {
    int rc = dop853_check(pot, w0, t0, t1, w_final, host_out, t_out, n_out, n_tail, block, nmax,
                          nstiff);
    if (rc != GAMD_OK) return rc;
    if (!host_out || !t_out || n_out == 0 || window_rows < 1 || !workspace)
        return GAMD_ERR_INVALID;
    if ((int64_t)workspace_bytes < gamd_dop853_dense_workspace_bytes(N, window_rows))
        return GAMD_ERR_INVALID;
    if (N == 0) return GAMD_OK;
    const int64_t W = window_rows < (int64_t)n_out ? window_rows : (int64_t)n_out;
    unsigned long long *queue = static_cast<unsigned long long *>(workspace);
    double *ring = reinterpret_cast<double *>(static_cast<char *>(workspace) + 256);
    double *st = ring + 2 * 6 * (size_t)window_rows * N;
    DopArgs a = dop_args(w0, N, t0, t1, h0, rtol, atol, nmax, nstiff, hmax, t_out, n_out, nullptr,
                         w_final, status, stats, order, n_tail, queue);
    a.pend = st + (size_t)WS_WORDS * N;
    WinArgs w;
    w.st = st;
    w.first = 1;
    hipStream_t s = as_stream(stream);
    hipStream_t c = nullptr, cs = nullptr;
    hipEvent_t comp = nullptr, copied[2] = {nullptr, nullptr};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const bool scf = pot->dev.has_scf;
    const bool mw = pot->dev.layout == LAYOUT_MW;
    auto run = scf  ? dense_windows<true, LAYOUT_GENERIC>
               : mw ? dense_windows<false, LAYOUT_MW>
                    : dense_windows<false, LAYOUT_GENERIC>;
    const int64_t blocks = scf  ? dop_grid<true, true, LAYOUT_GENERIC>(block, (int64_t)N)
                           : mw ? dop_grid<false, true, LAYOUT_MW>(block, (int64_t)N)
                                : dop_grid<false, true, LAYOUT_GENERIC>(block, (int64_t)N);
    a.n_tail = cap_tail(a, blocks, block);
    rc = GAMD_ERR_HIP;
    if (hipStreamCreateWithFlags(&c, hipStreamNonBlocking) != hipSuccess) goto done;
    if (hipEventCreateWithFlags(&comp, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&copied[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&copied[1], hipEventDisableTiming) != hipSuccess)
        goto done;
    if (!use_graph) {
        rc = run(pot->ddev, a, block, blocks, W, host_out, ring, w, s, c, comp, copied);
        // the events and the copy stream are released below: finish first
        if (rc == GAMD_OK && hipStreamSynchronize(s) != hipSuccess) rc = GAMD_ERR_HIP;
    } else {
        // the same sequence captured once on a private stream (the caller's
        // may be the legacy default stream, which cannot capture): kernel and
        // copy nodes, no memset; then launched as one graph on the caller's
        // stream, after its earlier work
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) goto done;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) goto done;
        rc = run(pot->ddev, a, block, blocks, W, host_out, ring, w, cs, c, comp, copied);
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

extern "C" int gamd_mockstream_dop853(const gamd_potential *ext,
                                      const gamd_potential *const *bodies, int nbody,
                                      const double *body_w, size_t nrel, const double *rel_t,
                                      const int32_t *nstream, const int64_t *offset,
                                      const double *stream_w0, double tfinal, double dt0,
                                      double rtol, double atol, int64_t nmax, double *out,
                                      double *body_final, int32_t *status, int64_t *stats,
                                      void *stream) {
    BodyTable bt;
    if (make_body_table(bodies, nbody, bt) != GAMD_OK) return GAMD_ERR_INVALID;
    if (!ext || nbody < 1 || !body_w || !rel_t || !nstream || !offset || !body_final ||
        !status || dt0 == 0.0)
        return GAMD_ERR_INVALID;
    if (nmax == 0) nmax = 1000000;  // dop853.cpp:673-674
    else if (nmax < 0) return GAMD_ERR_INVALID;
    if (nrel == 0) return GAMD_OK;
    DopArgs a;
    memset(&a, 0, sizeof(a));
    a.rtol = rtol;
    a.atol = atol;
    a.nmax = nmax;
    a.nstiff = 1;          // dop853_step passes nstiff = 1 (dop853.pyx:72)
    a.uround = 2.3e-16;    // dop853_step passes uround = 0: dop853() picks 2.3e-16 (:733-734)
    StreamArgs sa;
    sa.body_w = body_w;
    sa.rel_t = rel_t;
    sa.nstream = nstream;
    sa.offset = offset;
    sa.stream_w0 = stream_w0;
    sa.nrel = (int64_t)nrel;
    sa.tfinal = tfinal;
    sa.dt0 = dt0;
    sa.out = out;
    sa.body_final = body_final;
    sa.status = status;
    sa.stats = stats;
    const bool scf = ext->dev.has_scf || bt.scf;
    const bool mw = ext->dev.layout == LAYOUT_MW && !bt.scf;
    const int quads = nbody < 4 ? 1 : 0;  // a quad holds the bodies and at least one particle
    const int block = 256;
    hipStream_t s = as_stream(stream);
    if (quads) {
        // Systems of at most 4 bodies: one per 16-lane row (body_rows) or one
        // per quad (a quarter of the lanes per system).
        const bool rows = body_rows((int64_t)nrel, 4);
        const int G = rows ? 16 : 4;
        const unsigned grid = (unsigned)((G * (int64_t)nrel + block - 1) / block);
        auto kern = rows ? (scf  ? k_dop853_stream<true, LAYOUT_GENERIC, 16>
                            : mw ? k_dop853_stream<false, LAYOUT_MW, 16>
                                 : k_dop853_stream<false, LAYOUT_GENERIC, 16>)
                         : (scf  ? k_dop853_stream<true, LAYOUT_GENERIC, 4>
                            : mw ? k_dop853_stream<false, LAYOUT_MW, 4>
                                 : k_dop853_stream<false, LAYOUT_GENERIC, 4>);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, ext->ddev, bt, quads, sa, a);
        const int rc = launch_status();
        if (rc != GAMD_OK) return rc;
    }
    // systems of more than 4 bodies (every system when the bodies alone fill a quad)
    const unsigned grid = (unsigned)((64 * (int64_t)nrel + block - 1) / block);
    if (scf)
        hipLaunchKernelGGL((k_dop853_stream<true, LAYOUT_GENERIC, 64>), dim3(grid), dim3(block), 0,
                           s, ext->ddev, bt, quads, sa, a);
    else if (mw)
        hipLaunchKernelGGL((k_dop853_stream<false, LAYOUT_MW, 64>), dim3(grid), dim3(block), 0, s,
                           ext->ddev, bt, quads, sa, a);
    else
        hipLaunchKernelGGL((k_dop853_stream<false, LAYOUT_GENERIC, 64>), dim3(grid), dim3(block),
                           0, s, ext->ddev, bt, quads, sa, a);
    return launch_status();
}

extern "C" int gamd_nbody_dop853(const gamd_potential *ext, const gamd_potential *const *bodies,
                                 int nbody, const double *w0, size_t n, double t0, double t1,
                                 double h0, double rtol, double atol, int64_t nmax, int64_t nstiff,
                                 double hmax, const double *t_out, size_t n_out, double *w_out,
                                 double *w_final, int32_t *status, int64_t *stats, void *stream) {
    BodyTable bt;
    if (make_body_table(bodies, nbody, bt) != GAMD_OK) return GAMD_ERR_INVALID;
    if (!ext || !w0 || !w_final || !status || (size_t)nbody > n || n > GAMD_MAX_BODIES ||
        h0 == 0.0 || (w_out && (!t_out || n_out == 0)))
        return GAMD_ERR_INVALID;
    if (nmax == 0) nmax = 1000000;                   // dop853.cpp:673-674
    else if (nmax < 0) return GAMD_ERR_INVALID;
    if (nstiff == 0) nstiff = 1000;                  // :691-692
    else if (nstiff < 0) nstiff = 0;                 // :693-694 (never tested)
    if (n == 0) return GAMD_OK;
    DopArgs a;
    memset(&a, 0, sizeof(a));
    a.w0 = w0;
    a.N = (int64_t)n;
    a.rtol = rtol;
    a.atol = atol;
    a.nmax = nmax;
    a.nstiff = nstiff;
    a.hmax = hmax;
    a.t_out = t_out;
    a.n_out = w_out ? (int64_t)n_out : 0;
    a.out_rows = a.out_end = a.n_out;
    a.w_out = w_out;
    a.w_final = w_final;
    a.status = status;
    a.stats = stats;
    a.uround = 2.220446049250313e-16;  // np.finfo(float).eps, dop853.pyx:165
    const bool scf = ext->dev.has_scf || bt.scf;
    const bool mw = ext->dev.layout == LAYOUT_MW && !bt.scf;
    hipStream_t s = as_stream(stream);
    auto kern = body_rows(1, (int)n)
                    ? (w_out ? (scf  ? k_dop853_nbody<true, true, LAYOUT_GENERIC, 16>
                                : mw ? k_dop853_nbody<false, true, LAYOUT_MW, 16>
                                     : k_dop853_nbody<false, true, LAYOUT_GENERIC, 16>)
                             : (scf  ? k_dop853_nbody<true, false, LAYOUT_GENERIC, 16>
                                : mw ? k_dop853_nbody<false, false, LAYOUT_MW, 16>
                                     : k_dop853_nbody<false, false, LAYOUT_GENERIC, 16>))
                    : (w_out ? (scf  ? k_dop853_nbody<true, true, LAYOUT_GENERIC, 64>
                                : mw ? k_dop853_nbody<false, true, LAYOUT_MW, 64>
                                     : k_dop853_nbody<false, true, LAYOUT_GENERIC, 64>)
                             : (scf  ? k_dop853_nbody<true, false, LAYOUT_GENERIC, 64>
                                : mw ? k_dop853_nbody<false, false, LAYOUT_MW, 64>
                                     : k_dop853_nbody<false, false, LAYOUT_GENERIC, 64>));
    hipLaunchKernelGGL(kern, dim3(1), dim3(64), 0, s, ext->ddev, bt, t0, t1, h0, a);
    return launch_status();
}

extern "C" int gamd_dop853_batched(const gamd_potential *pot, const double *w0, size_t N,
                                   double t0, double t1, double h0, double rtol, double atol,
                                   int64_t nmax, int64_t nstiff, double hmax, int nbatch,
                                   const double *t_out, size_t n_out, double *w_out,
                                   double *w_final, int32_t *status, int64_t *stats,
                                   void *stream) {
    if (!pot || !w0 || !w_final || (w_out && (!t_out || n_out == 0))) return GAMD_ERR_INVALID;
    if (nbatch < 1) return GAMD_ERR_INVALID;
    if (nbatch > GAMD_DOP853_MAX_NBATCH) return GAMD_ERR_UNSUPPORTED;
    if (nmax == 0) nmax = 1000000;                   // dop853.cpp:673-674
    else if (nmax < 0) return GAMD_ERR_INVALID;      // :675-679
    if (nstiff == 0) nstiff = 1000;                  // :691-692
    else if (nstiff < 0) nstiff = 0;                 // :693-694 (never tested)
    if (N == 0) return GAMD_OK;
    if ((N + (size_t)nbatch - 1) / (size_t)nbatch > 0x7fffffffu) return GAMD_ERR_UNSUPPORTED;
    DopArgs a;
    memset(&a, 0, sizeof(a));
    a.w0 = w0;
    a.N = (int64_t)N;
    a.rtol = rtol;
    a.atol = atol;
    a.nmax = nmax;
    a.nstiff = nstiff;
    a.hmax = hmax;
    a.t_out = t_out;
    a.n_out = w_out ? (int64_t)n_out : 0;
    a.out_rows = a.out_end = a.n_out;
    a.w_out = w_out;
    a.w_final = w_final;
    a.status = status;
    a.stats = stats;
    a.uround = 2.220446049250313e-16;  // np.finfo(float).eps, dop853.pyx:165
    const bool scf = pot->dev.has_scf;
    const bool mw = pot->dev.layout == LAYOUT_MW;
    hipStream_t s = as_stream(stream);
    if (w_out)
        return scf  ? launch_batch<true, true, LAYOUT_GENERIC>(pot->ddev, nbatch, t0, t1, h0, a, s)
               : mw ? launch_batch<false, true, LAYOUT_MW>(pot->ddev, nbatch, t0, t1, h0, a, s)
                    : launch_batch<false, true, LAYOUT_GENERIC>(pot->ddev, nbatch, t0, t1, h0, a, s);
    return scf  ? launch_batch<true, false, LAYOUT_GENERIC>(pot->ddev, nbatch, t0, t1, h0, a, s)
           : mw ? launch_batch<false, false, LAYOUT_MW>(pot->ddev, nbatch, t0, t1, h0, a, s)
                : launch_batch<false, false, LAYOUT_GENERIC>(pot->ddev, nbatch, t0, t1, h0, a, s);
}
