// gamd_device.hpp -- device-side potential table and FP64 force kernels.
//
// A potential is a small flat table (DevPot) passed BY VALUE as a kernel
// argument, so every wave reads it through scalar loads from the kernarg
// segment: the component type switch is wave-uniform (a scalar branch, no
// divergence) and the parameters sit in SGPRs.  This replaces gala's table of
// function pointers (potential/potential/src/cpotential.h:10-36) with no
// indirect calls.  Parameter-only subexpressions (G*m, b*b, ...) are folded
// on the host into d[]; the products are the same IEEE values gala forms on
// every call (builtin_potentials.cpp:221-233 etc.), so nothing changes but
// the work.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gala_amd.h"

namespace gamd {

constexpr int MAXC = GAMD_MAX_COMPONENTS;

struct DevComp {
    int32_t type;
    int32_t shift_rotate;
    int32_t nmax, lmax;
    double p[8];          // raw parameters, p[0] = G (gala's pars layout)
    double d[4];          // host-folded constants, see fold_component()
    double q0[3];
    double R[9];
    const double *coef;   // SCF: device, per (l, m) block: (S_nlm, T_nlm) for n = 0..nmax
};

struct DevPot {
    int32_t n;
    int32_t has_scf;
    int32_t all_unshifted;
    int32_t layout;       // LAYOUT_* below: selects a straight-line evaluator
    DevComp c[MAXC];
};

// Component layouts with a dedicated straight-line evaluator.  With a runtime
// loop over components each component is its own basic block and the four
// dependent sqrt/divide/log chains run one after the other; written out
// straight, the scheduler interleaves them, which is what a lone lane in the
// tail of a persistent launch needs.
// MW: Hernquist, Hernquist, MN, NFW, unshifted.  SCF106: one unshifted SCF
// component with nmax = 10, lmax = 6 (the expansion inlined and unrolled).
enum { LAYOUT_GENERIC = 0, LAYOUT_MW = 1, LAYOUT_SCF106 = 2 };

// ---------------------------------------------------------------------------
// Natural log with a short dependency chain, for finite x > 0.  The library
// log costs ~470 cycles of latency for one lane on gfx950 (a divide ~76, an
// FMA ~6), and a lone lane in the tail of a persistent launch is latency-
// bound.  Method: x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m =
// 2 atanh(s) = 2 s sum_k s^(2k)/(2k+1), s = (m-1)/(m+1), |s| <= 0.1716, so
// twelve terms reach 1e-20; the series is summed by Estrin's scheme (depth
// ~6 FMAs) and e ln2 is added with ln2 split in two (the high part has 21
// trailing zero bits, so e * ln2_hi is exact).  Error ~1-2 ulp.
// argument reduction: x = 2^e (1 + f) with 1 + f in [sqrt(1/2), sqrt(2))
__device__ __forceinline__ void log_reduce(double x, double &f, int &e) {
    double m = frexp(x, &e);  // m in [0.5, 1)
    if (m < 0.7071067811865476) {
        m = m + m;
        e -= 1;
    }
    f = m - 1.0;
}

// ln(2^e (1 + f)) from s = f / (2 + f)
__device__ __forceinline__ double log_series(double s, int e) {
    const double z = s * s;
    const double z2 = z * z, z4 = z2 * z2, z8 = z4 * z4;
    const double p01 = fma(z, 0.3333333333333333, 1.0);
    const double p23 = fma(z, 0.14285714285714285, 0.2);
    const double p45 = fma(z, 0.09090909090909091, 0.1111111111111111);
    const double p67 = fma(z, 0.06666666666666667, 0.07692307692307693);
    const double p89 = fma(z, 0.05263157894736842, 0.058823529411764705);
    const double pab = fma(z, 0.043478260869565216, 0.047619047619047616);
    const double q03 = fma(z2, p23, p01);
    const double q47 = fma(z2, p67, p45);
    const double q8b = fma(z2, pab, p89);
    const double P = fma(z8, q8b, fma(z4, q47, q03));
    const double ed = (double)e;
    const double lm = (s + s) * P;
    return fma(ed, 0.6931471803691238, fma(ed, 1.9082149292705877e-10, lm));
}

__device__ __forceinline__ double log_fast(double x) {
    double f;
    int e;
    log_reduce(x, f, e);
    return log_series(f / (2.0 + f), e);
}

// 1/sqrt(x) for finite x > 0: the hardware estimate (v_rsq_f64) refined by
// two Newton steps y <- y + y (1/2 - (x/2) y^2) in FMA form (~1 ulp; a
// correctly rounded sqrt is ~110 cycles of latency on gfx950, this ~60).
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    double r = fma(-hx * y, y, 0.5);
    y = fma(y, r, y);
    r = fma(-hx * y, y, 0.5);
    return fma(y, r, y);
}

// 1/d for finite nonzero d: v_rcp_f64 refined by two Newton steps (~1 ulp).
__device__ __forceinline__ double rcp_nr(double d) {
    double y = __builtin_amdgcn_rcp(d);
    double e = fma(-d, y, 1.0);
    y = fma(y, e, y);
    e = fma(-d, y, 1.0);
    return fma(y, e, y);
}

// ---------------------------------------------------------------------------
// analytic components: accumulate the gradient (gala's *_gradient_single
// semantics, grad += ...).  r is |q| when the caller already has it (a
// composite shares one sqrt across its unshifted spherical components).

// Hernquist, builtin_potentials.cpp:221-233.  d = [G m, c]
__device__ __forceinline__ void hernquist_grad(const DevComp &c, double x, double y, double z,
                                               double r, double &gx, double &gy, double &gz) {
    const double rc = r + c.d[1];
    const double fac = c.d[0] / (rc * rc * r);
    gx += fac * x;
    gy += fac * y;
    gz += fac * z;
}

// Plummer, builtin_potentials.cpp:302-314.  d = [G m, b^2]
__device__ __forceinline__ void plummer_grad(const DevComp &c, double x, double y, double z,
                                             double r2, double &gx, double &gy, double &gz) {
    const double R2b = r2 + c.d[1];
    const double fac = c.d[0] / sqrt(R2b) / R2b;
    gx += fac * x;
    gy += fac * y;
    gz += fac * z;
}

// Miyamoto-Nagai, builtin_potentials.cpp:1298-1312.  d = [G m, a, b^2]
// pow(D, -1.5) is formed as 1/(D sqrt(D)) (a sqrt and a divide instead of a
// log/exp pair; within 2 ulp of gala's pow).
__device__ __forceinline__ void mn_grad(const DevComp &c, double x, double y, double z,
                                        double &gx, double &gy, double &gz) {
    const double sqrtz = sqrt(z * z + c.d[2]);
    const double zd = c.d[1] + sqrtz;
    const double D = x * x + y * y + zd * zd;
    const double fac = c.d[0] / (D * sqrt(D));
    gx += fac * x;
    gy += fac * y;
    gz += fac * z * (1.0 + c.d[1] / sqrtz);
}

// Spherical NFW, builtin_potentials.cpp:835-850.  d = [G m / r_s, r_s, r_s^2]
// NaN at r = 0 exactly as there (0/0).
__device__ __forceinline__ void nfw_grad(const DevComp &c, double x, double y, double z, double r,
                                         double &gx, double &gy, double &gz) {
    const double u = r / c.d[1];
    const double fac = c.d[0] / (u * u * u) / c.d[2] * (log(1.0 + u) - u / (1.0 + u));
    gx += fac * x;
    gy += fac * y;
    gz += fac * z;
}

__device__ __forceinline__ double comp_value_analytic(const DevComp &c, double x, double y,
                                                      double z) {
    const double *p = c.p;
    switch (c.type) {
    case GAMD_HERNQUIST: {  // :211-219
        const double r = sqrt(x * x + y * y + z * z);
        return -p[0] * p[1] / (r + p[2]);
    }
    case GAMD_PLUMMER: {  // :292-300
        const double r2 = x * x + y * y + z * z;
        return -p[0] * p[1] / sqrt(r2 + p[2] * p[2]);
    }
    case GAMD_MIYAMOTONAGAI: {  // :1287-1296
        const double zd = p[2] + sqrt(z * z + p[3] * p[3]);
        return -p[0] * p[1] / sqrt(x * x + y * y + zd * zd);
    }
    case GAMD_NFW: {  // :819-833
        const double v_h2 = -p[0] * p[1] / p[2];
        const double u = sqrt(x * x + y * y + z * z) / p[2];
        if (u == 0.0) return v_h2;
        return v_h2 * log(1.0 + u) / u;
    }
    default:
        return 0.0;
    }
}

__device__ __forceinline__ double comp_density_analytic(const DevComp &c, double x, double y,
                                                        double z) {
    const double *p = c.p;
    const double PI = 3.141592653589793;
    switch (c.type) {
    case GAMD_HERNQUIST: {  // :235-244
        const double r = sqrt(x * x + y * y + z * z);
        const double rho0 = p[1] / (2.0 * PI * p[2] * p[2] * p[2]);
        const double w = 1.0 + r / p[2];
        return rho0 / ((r / p[2]) * (w * w * w));
    }
    case GAMD_PLUMMER: {  // :316-324
        const double r2 = x * x + y * y + z * z;
        const double w = 1.0 + r2 / (p[2] * p[2]);
        return 3.0 * p[1] / (4.0 * PI * p[2] * p[2] * p[2]) / (w * w * sqrt(w));
    }
    case GAMD_MIYAMOTONAGAI: {  // :1314-1331
        const double M = p[1], a = p[2], b = p[3];
        const double R2 = x * x + y * y;
        const double sz = sqrt(z * z + b * b);
        const double numer = (b * b * M / (4.0 * PI)) * (a * R2 + (a + 3.0 * sz) * (a + sz) * (a + sz));
        const double D = R2 + (a + sz) * (a + sz);
        const double denom = D * D * sqrt(D) * sz * sz * sz;
        return numer / denom;
    }
    case GAMD_NFW: {  // :852-863
        const double r = sqrt(x * x + y * y + z * z);
        const double rho0 = p[1] / (4.0 * PI * p[2] * p[2] * p[2]);
        const double w = 1.0 + r / p[2];
        return rho0 / ((r / p[2]) * (w * w));
    }
    default:
        return 0.0;
    }
}

// Hessians (row-major 3x3, accumulated): the second derivatives of the same
// potentials, H = f(r) I + (f'(r)/r) q q^T for the spherical ones.
__device__ __forceinline__ void comp_hessian_analytic(const DevComp &c, double x, double y,
                                                      double z, double *H) {
    const double *p = c.p;
    const double r2 = x * x + y * y + z * z, r = sqrt(r2);
    double f = 0.0, g = 0.0;
    switch (c.type) {
    case GAMD_HERNQUIST: {
        const double GM = p[0] * p[1], rc = r + p[2];
        f = GM / (r * rc * rc);
        g = -GM * (3.0 * r + p[2]) / (r2 * r * rc * rc * rc);
        break;
    }
    case GAMD_PLUMMER: {
        const double GM = p[0] * p[1], D = r2 + p[2] * p[2];
        const double Dm32 = 1.0 / (D * sqrt(D));
        f = GM * Dm32;
        g = -3.0 * GM * Dm32 / D;
        break;
    }
    case GAMD_NFW: {
        const double GM = p[0] * p[1], rs = p[2], u = r / rs;
        f = GM / (r2 * r) * (log(1.0 + u) - u / (1.0 + u));
        g = (-3.0 * f / r + GM / (r2 * rs * rs * (1.0 + u) * (1.0 + u))) / r;
        break;
    }
    case GAMD_MIYAMOTONAGAI: {
        const double GM = p[0] * p[1], b = p[3];
        const double s = sqrt(z * z + b * b), zeta = p[2] + s;
        const double D = x * x + y * y + zeta * zeta;
        const double Dm32 = 1.0 / (D * sqrt(D)), Dm52 = Dm32 / D;
        const double w = zeta * z / s;
        H[0] += GM * (Dm32 - 3.0 * x * x * Dm52);
        H[4] += GM * (Dm32 - 3.0 * y * y * Dm52);
        const double hxy = -3.0 * GM * x * y * Dm52;
        const double hxz = -3.0 * GM * x * w * Dm52;
        const double hyz = -3.0 * GM * y * w * Dm52;
        H[1] += hxy;
        H[3] += hxy;
        H[2] += hxz;
        H[6] += hxz;
        H[5] += hyz;
        H[7] += hyz;
        H[8] += GM * ((z * z / (s * s) + zeta * b * b / (s * s * s)) * Dm32 - 3.0 * w * w * Dm52);
        return;
    }
    default:
        return;
    }
    const double q[3] = {x, y, z};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) H[3 * i + j] += (i == j ? f : 0.0) + g * q[i] * q[j];
}

// ---------------------------------------------------------------------------
// SCF basis expansion (Hernquist-Ostriker, Lowing+2011 real coefficients):
// the quantity gala computes term by term in bfe.cpp:114-189 with GSL calls
// per (n,l,m) (bfe_helper.cpp:24-90), reorganised for a GPU lane:
//   * cos(m phi), sin(m phi) by the angle-addition recurrence from x/R, y/R;
//   * normalised Y_lm(cos theta) and dY_lm/dtheta by the standard stable
//     recurrences (one pass over l, m);
//   * per l, Gegenbauer C_n^{(2l+3/2)}(xi) and C_{n-1}^{(2l+5/2)}(xi) by their
//     three-term recurrences, giving phi_nl(s) and dphi_nl/ds for all n;
//   * the n-sums A_lm = sum_n phi_nl S_nlm (and T, and the s-derivatives)
//     are FMA chains over coefficients read with wave-uniform (scalar) loads.
// coef layout: for l = 0..lmax, m = 0..l, n = 0..nmax: (S_nlm, T_nlm).
// Returns Phi / (G M / r_s) scaled; accumulates G M/r_s^2 * grad into g.

// 1/n for the Gegenbauer recurrences (n <= GAMD_SCF_MAX_NMAX + 1): a
// wave-uniform table read instead of two FP64 divisions per (l, n) term
static __constant__ double scf_inv_n[GAMD_SCF_MAX_NMAX + 2] = {0.0, 1.0 / 1.0, 1.0 / 2.0, 1.0 / 3.0, 1.0 / 4.0, 1.0 / 5.0, 1.0 / 6.0, 1.0 / 7.0, 1.0 / 8.0, 1.0 / 9.0, 1.0 / 10.0, 1.0 / 11.0, 1.0 / 12.0, 1.0 / 13.0, 1.0 / 14.0, 1.0 / 15.0, 1.0 / 16.0, 1.0 / 17.0, 1.0 / 18.0, 1.0 / 19.0, 1.0 / 20.0, 1.0 / 21.0};

// Position of (S_nlm, T_nlm) in the device coefficient table: per l, per n,
// the l + 1 pairs of m = 0..l contiguous, so one n step of the expansion reads
// one contiguous run of 2 (l + 1) doubles (wide scalar loads).
__host__ __device__ __forceinline__ int scf_pack_index(int l, int m, int n, int nmax) {
    return (nmax + 1) * (l * (l + 1) / 2) + n * (l + 1) + m;
}

// The Gegenbauer recurrence factors of scf_eval_t as a table over (l, n):
// for a = 2l + 3/2 and b = a + 1, [0] = n + a - 1, [1] = n + 2a - 2,
// [2] = n + b - 1, [3] = n + 2b - 2 (the same doubles the loop would form from
// its counter), so the loop reads them instead of converting and adding.
struct ScfRecTable {
    double f[GAMD_SCF_MAX_LMAX + 1][GAMD_SCF_MAX_NMAX + 2][4];
};
constexpr ScfRecTable scf_rec_table() {
    ScfRecTable t{};
    for (int l = 0; l <= GAMD_SCF_MAX_LMAX; l++) {
        const double a = 2.0 * l + 1.5, b = a + 1.0;
        for (int n = 0; n <= GAMD_SCF_MAX_NMAX + 1; n++) {
            t.f[l][n][0] = n + a - 1.0;
            t.f[l][n][1] = n + 2.0 * a - 2.0;
            t.f[l][n][2] = n + b - 1.0;
            t.f[l][n][3] = n + 2.0 * b - 2.0;
        }
    }
    return t;
}
static __constant__ ScfRecTable scf_rec = scf_rec_table();

template <int NMAX, int LMAX>
__device__ __forceinline__ void scf_eval_t(const double *__restrict__ cf, double G, double M,
                                           double rs, int nmax_rt, int lmax_rt, double x,
                                           double y, double z, bool want_grad, double &phi_out,
                                           double &gx, double &gy, double &gz) {
    const int nmax = NMAX >= 0 ? NMAX : nmax_rt;
    const int lmax = LMAX >= 0 ? LMAX : lmax_rt;
    // the coefficients are read-only for the whole launch and every lane reads
    // the same one: through the constant address space they become scalar
    // (SMEM) loads instead of a vector load and wait per term
    const __attribute__((address_space(4))) double *cfc =
        (const __attribute__((address_space(4))) double *)cf;
    const double FOURPI_SQRT = 3.544907701811031;
    const double R2 = x * x + y * y;
    const double r2 = R2 + z * z;
    const double r = sqrt(r2);
    const double Rc = sqrt(R2);
    const double s = r / rs;
    const double X = z / r;          // cos(theta)
    const double sinth = Rc / r;     // sin(theta)
    double cphi = 1.0, sphi = 0.0;   // atan2(0, 0) = 0 on the axis, as gala
    if (Rc > 0.0) {
        cphi = x / Rc;
        sphi = y / Rc;
    }
    const double ops = 1.0 + s;
    const double inv_ops = 1.0 / ops;
    const double inv_s = 1.0 / s;
    const double xi = (s - 1.0) * inv_ops;
    const double inv_ops2 = inv_ops * inv_ops;

    double Phi = 0.0, dr = 0.0, dth = 0.0, dph = 0.0;

    // running radial prefactor pre_l = -sqrt(4 pi) s^l (1+s)^(-2l-1)
    double pre = -FOURPI_SQRT / ops;
    // normalised Y_mm seeds: Y_00 = 1/sqrt(4 pi)
    double ymm = 1.0 / FOURPI_SQRT;
    // cos(m phi), sin(m phi) for m = 0..lmax kept in registers
    double cm[LMAX >= 0 ? LMAX + 1 : GAMD_SCF_MAX_LMAX + 1];
    double sm[LMAX >= 0 ? LMAX + 1 : GAMD_SCF_MAX_LMAX + 1];
    cm[0] = 1.0;
    sm[0] = 0.0;
#pragma unroll
    for (int m = 1; m <= (LMAX >= 0 ? LMAX : GAMD_SCF_MAX_LMAX); m++) {
        if (m <= lmax) {
            cm[m] = cm[m - 1] * cphi - sm[m - 1] * sphi;
            sm[m] = sm[m - 1] * cphi + cm[m - 1] * sphi;
        }
    }
    // Y_{l-1,m} and Y_{l-2,m} for every m, updated as l advances
    double y1[LMAX >= 0 ? LMAX + 1 : GAMD_SCF_MAX_LMAX + 1];
    double y2[LMAX >= 0 ? LMAX + 1 : GAMD_SCF_MAX_LMAX + 1];
#pragma unroll
    for (int m = 0; m <= (LMAX >= 0 ? LMAX : GAMD_SCF_MAX_LMAX); m++) {
        y1[m] = 0.0;
        y2[m] = 0.0;
    }

#pragma unroll
    for (int l = 0; l <= (LMAX >= 0 ? LMAX : GAMD_SCF_MAX_LMAX); l++) {
        if (l > lmax) break;
        // radial functions for this l: phi_nl = pre * C_n^{a}(xi), a = 2l + 3/2
        // dphi_nl/ds = pre * [ (l/s - (2l+1)/(1+s)) C_n^{a} + 2a C_{n-1}^{a+1} 2/(1+s)^2 ]
        const double a = 2.0 * l + 1.5;
        const double dlog = (l == 0 ? 0.0 : l * inv_s) - (2.0 * l + 1.0) * inv_ops;
        const double chain = 4.0 * a * inv_ops2;
        // Y_lm for all m <= l at this l
        double ylm[LMAX >= 0 ? LMAX + 1 : GAMD_SCF_MAX_LMAX + 1];
        // n-sums per m
        double As[LMAX >= 0 ? LMAX + 1 : GAMD_SCF_MAX_LMAX + 1];
        double Bs[LMAX >= 0 ? LMAX + 1 : GAMD_SCF_MAX_LMAX + 1];
        double dAs[LMAX >= 0 ? LMAX + 1 : GAMD_SCF_MAX_LMAX + 1];
        double dBs[LMAX >= 0 ? LMAX + 1 : GAMD_SCF_MAX_LMAX + 1];
#pragma unroll
        for (int m = 0; m <= (LMAX >= 0 ? LMAX : GAMD_SCF_MAX_LMAX); m++) {
            As[m] = 0.0;
            Bs[m] = 0.0;
            dAs[m] = 0.0;
            dBs[m] = 0.0;
        }
        // Gegenbauer recurrences in n
        double c0 = 1.0, c1 = 0.0;  // C_n^a, C_{n-1}^a
        double e0 = 0.0, e1 = 0.0;  // C_{n-1}^{a+1}, C_{n-2}^{a+1}
        // the fixed-size evaluator unrolls n by 10 (measured best for nmax 10;
        // fully unrolled, the loop spills); the runtime-size one keeps a loop
#pragma unroll(NMAX >= 0 ? 10 : 1)
        for (int n = 0; n <= nmax; n++) {
            double Cn, En;  // C_n^a(xi), C_{n-1}^{a+1}(xi)
            if (n == 0) {
                Cn = 1.0;
                En = 0.0;
            } else if (n == 1) {
                Cn = 2.0 * a * xi;
                En = 1.0;
            } else {
                const double *rn = scf_rec.f[l][n], *rk = scf_rec.f[l][n - 1];
                Cn = (2.0 * xi * rn[0] * c0 - rn[1] * c1) * scf_inv_n[n];
                const double b = a + 1.0;
                const int k = n - 1;  // index of E
                En = (k == 1) ? 2.0 * b * xi
                              : (2.0 * xi * rk[2] * e0 - rk[3] * e1) * scf_inv_n[k];
            }
            c1 = c0;
            c0 = Cn;
            e1 = e0;
            e0 = En;
            const double phinl = pre * Cn;
            const double dphinl = pre * (dlog * Cn + chain * En);
#pragma unroll
            for (int m = 0; m <= (LMAX >= 0 ? LMAX : GAMD_SCF_MAX_LMAX); m++) {
                if (m <= l) {
                    const double S = cfc[2 * scf_pack_index(l, m, n, nmax)];
                    const double T = cfc[2 * scf_pack_index(l, m, n, nmax) + 1];
                    As[m] = fma(phinl, S, As[m]);
                    Bs[m] = fma(phinl, T, Bs[m]);
                    dAs[m] = fma(dphinl, S, dAs[m]);
                    dBs[m] = fma(dphinl, T, dBs[m]);
                }
            }
        }
        // Y_lm(X) for this l (normalised, Condon-Shortley phase)
#pragma unroll
        for (int m = 0; m <= (LMAX >= 0 ? LMAX : GAMD_SCF_MAX_LMAX); m++) {
            if (m <= l) {
                double v;
                if (m == l) {
                    v = ymm;
                } else if (m == l - 1) {
                    v = sqrt(2.0 * m + 3.0) * X * y1[m];
                } else {
                    const double l2 = (double)l * l, m2 = (double)m * m;
                    const double al = sqrt((4.0 * l2 - 1.0) / (l2 - m2));
                    const double bl = sqrt(((l - 1.0) * (l - 1.0) - m2) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
                    v = al * (X * y1[m] - bl * y2[m]);
                }
                ylm[m] = v;
            }
        }
        // accumulate
#pragma unroll
        for (int m = 0; m <= (LMAX >= 0 ? LMAX : GAMD_SCF_MAX_LMAX); m++) {
            if (m <= l) {
                const double ac = As[m] * cm[m] + Bs[m] * sm[m];
                Phi = fma(ylm[m], ac, Phi);
                if (want_grad) {
                    const double dac = dAs[m] * cm[m] + dBs[m] * sm[m];
                    dr = fma(ylm[m], dac, dr);
                    // dY_lm/dtheta * sin(theta) = l X Y_lm - sqrt((2l+1)/(2l-1) (l-m)(l+m)) Y_{l-1,m}
                    double dy = l * X * ylm[m];
                    if (m < l)
                        dy -= sqrt((2.0 * l + 1.0) / (2.0 * l - 1.0) * (double)(l - m) * (l + m)) * y1[m];
                    dth = fma(dy, ac, dth);
                    dph = fma((double)m * ylm[m], Bs[m] * cm[m] - As[m] * sm[m], dph);
                }
            }
        }
        // shift Y history, next Y_mm seed, next radial prefactor
#pragma unroll
        for (int m = 0; m <= (LMAX >= 0 ? LMAX : GAMD_SCF_MAX_LMAX); m++) {
            if (m <= l) {
                y2[m] = y1[m];
                y1[m] = ylm[m];
            }
        }
        ymm = -sqrt((2.0 * l + 3.0) / (2.0 * l + 2.0)) * sinth * ymm;
        pre = pre * s * inv_ops2;
    }
    const double GMr = G * M / rs;
    phi_out = Phi * GMr;
    if (want_grad) {
        // (d/dr, (1/r) d/dtheta, 1/(r sin theta) d/dphi) in units of r_s, then Cartesian
        const double g_r = dr;
        const double g_t = dth / (s * sinth);
        const double g_p = dph / (s * sinth);
        const double fac = G * M / (rs * rs);
        gx += (sinth * cphi * g_r + X * cphi * g_t - sphi * g_p) * fac;
        gy += (sinth * sphi * g_r + X * sphi * g_t + cphi * g_p) * fac;
        gz += (X * g_r - sinth * g_t) * fac;
    }
}

// One out-of-line copy per kernel: the SCF sum is long, and inlining it at
// each of DOP853's stage evaluations would multiply code size (and compile
// time) with no gain, since the call overhead is small next to the sum.
// Returns (Phi, dPhi/dx, dPhi/dy, dPhi/dz).
__device__ __noinline__ double4 scf_call(const double *__restrict__ cf, double G, double M,
                                         double rs, int nmax, int lmax, double x, double y,
                                         double z, int want_grad) {
    double phi = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    if (nmax == 10 && lmax == 6)
        scf_eval_t<10, 6>(cf, G, M, rs, 10, 6, x, y, z, want_grad != 0, phi, gx, gy, gz);
    else
        scf_eval_t<-1, -1>(cf, G, M, rs, nmax, lmax, x, y, z, want_grad != 0, phi, gx, gy, gz);
    return make_double4(phi, gx, gy, gz);
}

__device__ __forceinline__ void scf_eval(const DevComp &c, double x, double y, double z,
                                         bool want_grad, double &phi, double &gx, double &gy,
                                         double &gz) {
    const double4 r = scf_call(c.coef, c.p[0], c.p[1], c.p[2], c.nmax, c.lmax, x, y, z,
                               want_grad ? 1 : 0);
    phi = r.x;
    gx += r.y;
    gy += r.z;
    gz += r.w;
}

// SCF density (bfe.cpp:20-59): rho_nl = K_nl/(2 pi) sqrt(4 pi) s^l / (s (1+s)^(2l+3)) C_n(xi)
__device__ inline double scf_density(const DevComp &c, double x, double y, double z) {
    const int nmax = c.nmax, lmax = c.lmax;
    const double FOURPI_SQRT = 3.544907701811031, PI = 3.141592653589793;
    const double M = c.p[1], rs = c.p[2];
    const double R2 = x * x + y * y, r = sqrt(R2 + z * z), Rc = sqrt(R2);
    const double s = r / rs, X = z / r, sinth = Rc / r;
    double cphi = 1.0, sphi = 0.0;
    if (Rc > 0.0) {
        cphi = x / Rc;
        sphi = y / Rc;
    }
    const double xi = (s - 1.0) / (s + 1.0), ops = 1.0 + s;
    double cm[GAMD_SCF_MAX_LMAX + 1], sm[GAMD_SCF_MAX_LMAX + 1];
    double y1[GAMD_SCF_MAX_LMAX + 1], y2[GAMD_SCF_MAX_LMAX + 1];
    cm[0] = 1.0;
    sm[0] = 0.0;
    for (int m = 1; m <= lmax; m++) {
        cm[m] = cm[m - 1] * cphi - sm[m - 1] * sphi;
        sm[m] = sm[m - 1] * cphi + cm[m - 1] * sphi;
    }
    for (int m = 0; m <= lmax; m++) y1[m] = y2[m] = 0.0;
    double ymm = 1.0 / FOURPI_SQRT, rho = 0.0;
    double pre = FOURPI_SQRT / (2.0 * PI) / (s * ops * ops * ops);  // s^l/(s (1+s)^(2l+3)) at l=0
    for (int l = 0; l <= lmax; l++) {
        const double a = 2.0 * l + 1.5;
        for (int m = 0; m <= l; m++) {
            double v;
            if (m == l) v = ymm;
            else if (m == l - 1) v = sqrt(2.0 * m + 3.0) * X * y1[m];
            else {
                const double l2 = (double)l * l, m2 = (double)m * m;
                v = sqrt((4.0 * l2 - 1.0) / (l2 - m2)) *
                    (X * y1[m] - sqrt(((l - 1.0) * (l - 1.0) - m2) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0)) * y2[m]);
            }
            double c0 = 1.0, c1 = 0.0, A = 0.0, B = 0.0;
            for (int n = 0; n <= nmax; n++) {
                double Cn = n == 0 ? 1.0 : (n == 1 ? 2.0 * a * xi : (2.0 * xi * (n + a - 1.0) * c0 - (n + 2.0 * a - 2.0) * c1) / n);
                c1 = c0;
                c0 = Cn;
                const double Knl = 0.5 * n * (n + 4 * l + 3) + (l + 1) * (2 * l + 1);
                const double rnl = Knl * pre * Cn;
                const int idx = 2 * scf_pack_index(l, m, n, nmax);
                A += rnl * c.coef[idx];
                B += rnl * c.coef[idx + 1];
            }
            rho += v * (A * cm[m] + B * sm[m]);
            y2[m] = y1[m];
            y1[m] = v;
        }
        ymm = -sqrt((2.0 * l + 3.0) / (2.0 * l + 2.0)) * sinth * ymm;
        pre = pre * s / (ops * ops);
    }
    return rho * M / (rs * rs * rs);
}

// ---------------------------------------------------------------------------
// composite evaluation (cpotential.cpp:170-287).  Components run in table
// order; a shifted/rotated component is evaluated at R (q - q0) and its
// gradient rotated back with R^T.

template <bool WITH_SCF>
__device__ __forceinline__ void comp_grad(const DevComp &c, double x, double y, double z,
                                          double &gx, double &gy, double &gz) {
    switch (c.type) {
    case GAMD_HERNQUIST:
        hernquist_grad(c, x, y, z, sqrt(x * x + y * y + z * z), gx, gy, gz);
        break;
    case GAMD_PLUMMER:
        plummer_grad(c, x, y, z, x * x + y * y + z * z, gx, gy, gz);
        break;
    case GAMD_MIYAMOTONAGAI:
        mn_grad(c, x, y, z, gx, gy, gz);
        break;
    case GAMD_NFW:
        nfw_grad(c, x, y, z, sqrt(x * x + y * y + z * z), gx, gy, gz);
        break;
    case GAMD_SCF:
        if (WITH_SCF) {
            double phi;
            scf_eval(c, x, y, z, true, phi, gx, gy, gz);
        }
        break;
    default:
        break;
    }
}

__device__ __forceinline__ void shift_rotate(const DevComp &c, double x, double y, double z,
                                             double &xt, double &yt, double &zt) {
    const double tx = x - c.q0[0], ty = y - c.q0[1], tz = z - c.q0[2];
    xt = c.R[0] * tx + c.R[1] * ty + c.R[2] * tz;
    yt = c.R[3] * tx + c.R[4] * ty + c.R[5] * tz;
    zt = c.R[6] * tx + c.R[7] * ty + c.R[8] * tz;
}

// grad Phi at (x, y, z): overwritten (c_gradient zeroes first).
template <bool WITH_SCF>
__device__ __forceinline__ void pot_grad(const DevPot &P, double x, double y, double z,
                                         double &gx, double &gy, double &gz) {
    gx = 0.0;
    gy = 0.0;
    gz = 0.0;
    if (P.all_unshifted) {
        // one |q| shared by every unshifted spherical component
        const double r2 = x * x + y * y + z * z;
        const double r = sqrt(r2);
        for (int i = 0; i < P.n; i++) {
            const DevComp &c = P.c[i];
            switch (c.type) {
            case GAMD_HERNQUIST: hernquist_grad(c, x, y, z, r, gx, gy, gz); break;
            case GAMD_PLUMMER: plummer_grad(c, x, y, z, r2, gx, gy, gz); break;
            case GAMD_MIYAMOTONAGAI: mn_grad(c, x, y, z, gx, gy, gz); break;
            case GAMD_NFW: nfw_grad(c, x, y, z, r, gx, gy, gz); break;
            case GAMD_SCF:
                if (WITH_SCF) {
                    double phi;
                    scf_eval(c, x, y, z, true, phi, gx, gy, gz);
                }
                break;
            default: break;
            }
        }
        return;
    }
    for (int i = 0; i < P.n; i++) {
        const DevComp &c = P.c[i];
        if (!c.shift_rotate) {
            comp_grad<WITH_SCF>(c, x, y, z, gx, gy, gz);
        } else {
            double xt, yt, zt, tx = 0.0, ty = 0.0, tz = 0.0;
            shift_rotate(c, x, y, z, xt, yt, zt);
            comp_grad<WITH_SCF>(c, xt, yt, zt, tx, ty, tz);
            gx += c.R[0] * tx + c.R[3] * ty + c.R[6] * tz;
            gy += c.R[1] * tx + c.R[4] * ty + c.R[7] * tz;
            gz += c.R[2] * tx + c.R[5] * ty + c.R[8] * tz;
        }
    }
}

// Gradient of a body's potential PP at q, centred on the body's position xp:
// c_gradient of the potential with every component's origin set to xp
// (c_nbody_acceleration / c_nbody_gradient_symplectic force do_shift_rotate
// = 1, cpotential.cpp:389-442): q' = R (q - xp), gradient rotated back with
// R^T, summed over components into a zeroed f2 that the caller adds.
template <bool WITH_SCF>
__device__ __forceinline__ void body_grad(const DevPot &PP, double x, double y, double z,
                                          double px, double py, double pz, double &fx,
                                          double &fy, double &fz) {
    fx = 0.0;
    fy = 0.0;
    fz = 0.0;
    for (int i = 0; i < PP.n; i++) {
        const DevComp &c = PP.c[i];
        const double tx = x - px, ty = y - py, tz = z - pz;
        const double xt = c.R[0] * tx + c.R[1] * ty + c.R[2] * tz;
        const double yt = c.R[3] * tx + c.R[4] * ty + c.R[5] * tz;
        const double zt = c.R[6] * tx + c.R[7] * ty + c.R[8] * tz;
        double ux = 0.0, uy = 0.0, uz = 0.0;
        comp_grad<WITH_SCF>(c, xt, yt, zt, ux, uy, uz);
        fx += c.R[0] * ux + c.R[3] * uy + c.R[6] * uz;
        fy += c.R[1] * ux + c.R[4] * uy + c.R[7] * uz;
        fz += c.R[2] * ux + c.R[5] * uy + c.R[8] * uz;
    }
}

// The Milky-Way layout (nucleus and bulge Hernquist, Miyamoto-Nagai disk, NFW
// halo, gala's MilkyWayPotential order): the same formulas as the per-
// component functions above, one shared |q|, contributions accumulated in
// table order.
__device__ __forceinline__ void mw_grad(const DevPot &P, double x, double y, double z,
                                        double &gx, double &gy, double &gz) {
    const DevComp &h1 = P.c[0], &h2 = P.c[1], &mn = P.c[2], &nf = P.c[3];
    // parameters used inside the branches below, read before them: loads in
    // a conditional block are not hoisted out of the caller's loop, and each
    // one would stall the evaluation on the scalar cache
    const double gm1 = h1.d[0], gm2 = h2.d[0], gmn = nf.d[0] * nf.d[1];
    const double R2 = x * x + y * y;
    const double r2 = R2 + z * z;
    // |q|, 1/|q|^3, sqrt(z^2 + b^2), its inverse and D^(-3/2) from three
    // reciprocal square roots instead of three correctly rounded roots and
    // their divides
    const double ir = rsqrt_nr(r2);
    const double r = r2 * ir;
    const double ir3 = ir * ir * ir;
    const double s2 = z * z + mn.d[2];
    const double isz = rsqrt_nr(s2);
    const double sqrtz = s2 * isz;
    const double zd = mn.d[1] + sqrtz;
    const double D = R2 + zd * zd;
    const double iD = rsqrt_nr(D);
    const double f3 = mn.d[0] * (iD * iD * iD);
    const double f3z = 1.0 + mn.d[1] * isz;
    // NFW in the equivalent form G m / r^3 (log(1+u) - u/(1+u)), u = r / r_s
    // (v_h2 / u^3 / r_s^2 = G m / r^3)
    const double u = r * nf.d[3];  // d[3] = 1 / r_s
    const double opu = 1.0 + u;
    double lf;
    int le;
    log_reduce(opu, lf, le);
    const double tf = 2.0 + lf;
    // The remaining denominators -- the two Hernquist ones, 1 + u and the
    // log's 2 + f -- share one reciprocal (batched inversion: invert the
    // product, peel the factors off with multiplies).
    const double rc1 = r + h1.d[1];
    const double rc2 = r + h2.d[1];
    const double d1 = rc1 * rc1 * r, d2 = rc2 * rc2 * r;
    const double p12 = d1 * d2;
    const double p3 = p12 * opu;
    const double p = p3 * tf;
    const bool batched = p > 1e-280 && p < 1e280;
    double f1, f2, f4, i4 = 0.0, slog;
    if (batched) {
        const double inv = rcp_nr(p);
        slog = lf * (inv * p3);
        const double inv3 = inv * tf;
        const double inv2 = inv3 * opu;
        i4 = inv3 * p12;
        f1 = gm1 * (inv2 * d2);
        f2 = gm2 * (inv2 * d1);
    } else {  // extreme radii (or r = 0): the per-component forms
        slog = lf / tf;
        f1 = gm1 / d1;
        f2 = gm2 / d2;
    }
    const double lg = log_series(slog, le);
    if (batched)
        f4 = gmn * (lg - u * i4) * ir3;
    else
        f4 = nf.d[0] / (u * u * u) / nf.d[2] * (lg - u / opu);
    gx = 0.0;
    gy = 0.0;
    gz = 0.0;
    gx += f1 * x;
    gy += f1 * y;
    gz += f1 * z;
    gx += f2 * x;
    gy += f2 * y;
    gz += f2 * z;
    gx += f3 * x;
    gy += f3 * y;
    gz += f3 * z * f3z;
    gx += f4 * x;
    gy += f4 * y;
    gz += f4 * z;
}

// grad Phi with the layout chosen at compile time (LAYOUT_GENERIC reads the
// table's runtime layout and falls back to the general loop)
template <bool WITH_SCF, int LAYOUT>
__device__ __forceinline__ void pot_grad_l(const DevPot &P, double x, double y, double z,
                                           double &gx, double &gy, double &gz) {
    if (LAYOUT == LAYOUT_MW) {
        mw_grad(P, x, y, z, gx, gy, gz);
    } else if (LAYOUT == LAYOUT_SCF106) {
        const DevComp &c = P.c[0];
        double phi;
        gx = 0.0;
        gy = 0.0;
        gz = 0.0;
        scf_eval_t<10, 6>(c.coef, c.p[0], c.p[1], c.p[2], 10, 6, x, y, z, true, phi, gx, gy, gz);
    } else {
        pot_grad<WITH_SCF>(P, x, y, z, gx, gy, gz);
    }
}

template <bool WITH_SCF>
__device__ __forceinline__ double comp_value(const DevComp &c, double x, double y, double z) {
    if (WITH_SCF && c.type == GAMD_SCF) {
        double phi, gx = 0, gy = 0, gz = 0;
        scf_eval(c, x, y, z, false, phi, gx, gy, gz);
        return phi;
    }
    return comp_value_analytic(c, x, y, z);
}

template <bool WITH_SCF>
__device__ __forceinline__ double pot_value(const DevPot &P, double x, double y, double z) {
    double v = 0.0;
    for (int i = 0; i < P.n; i++) {
        const DevComp &c = P.c[i];
        if (!c.shift_rotate) {
            v = v + comp_value<WITH_SCF>(c, x, y, z);
        } else {
            double xt, yt, zt;
            shift_rotate(c, x, y, z, xt, yt, zt);
            v = v + comp_value<WITH_SCF>(c, xt, yt, zt);
        }
    }
    return v;
}

// c_d2_dr2, cpotential.cpp:346-371
template <bool WITH_SCF>
__device__ __forceinline__ double pot_d2_dr2(const DevPot &P, double x, double y, double z) {
    const double h = 1e-2;
    const double r = sqrt(x * x + y * y + z * z);
    double d2 = pot_value<WITH_SCF>(P, x + h * x / r, y + h * y / r, z + h * z / r);
    d2 = d2 - 2.0 * pot_value<WITH_SCF>(P, x, y, z);
    d2 = d2 + pot_value<WITH_SCF>(P, x - h * x / r, y - h * y / r, z - h * z / r);
    return d2 / (h * h);
}

}  // namespace gamd
