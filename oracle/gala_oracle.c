/*
 * gala_oracle.c -- CPU restatement of gala's orbit-integration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gala_amd/) links, loads
 * or calls this file.  It is used by tests/ as the checker of the HIP path,
 * by __graft_entry__.smoke() as the checker of one small call, and by
 * bench.py's cpu_baseline leg as the all-core CPU reference path.
 *
 * Every function restates the reference algorithm it names (paths relative to
 * the gala source tree, src/gala/...):
 *   potentials     potential/potential/builtin/builtin_potentials.cpp
 *                  (Hernquist :211-286, Plummer :292-360, SphericalNFW
 *                  :819-920, Miyamoto-Nagai :1287-1379)
 *   composite      potential/potential/src/cpotential.cpp:80-189, 214-287,
 *                  290-314, 346-371, 418-442
 *   leapfrog       integrate/cyintegrators/leapfrog.pyx:24-121, 126-257
 *   Ruth4          integrate/cyintegrators/ruth4.pyx:24-106
 *   DOP853         integrate/cyintegrators/dopri/dop853.cpp:18-650, 653-866,
 *                  869-904; dop853.pyx:90-250; hamiltonian/src/chamiltonian.cpp:38-57
 *   N-body         cpotential.cpp:389-442, dop853.cpp:990-1006,
 *                  dynamics/nbody/nbody.pyx:30-115
 *   mock stream    dynamics/mockstream/mockstream.pyx:176-303, 442-620,
 *                  df.pyx:61-106, 363-456, _coord.pyx:23-48
 *   SCF            potential/scf/src/bfe.cpp:20-259, bfe_helper.cpp:14-90
 *                  (GSL special functions replaced by textbook recurrences)
 *
 * Build: oracle/Makefile (gcc, C99, OpenMP).  Two variants are built from
 * this one source: liboracle.so (-O2, IEEE arithmetic, the checker) and
 * liboracle_fast.so (-Ofast, the reference's own flags, setup.py:459-460,
 * used only for timing the CPU baseline).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "dop853_coef.h"
#include "gala_oracle.h"

/* ------------------------------------------------------------------------ */
/* SCF special functions (replacing GSL, bfe_helper.cpp:5-9)                 */

#define ORC_SQRT_FOURPI 3.544907701811031

/* Gegenbauer C_n^{(a)}(x) by the three-term recurrence
 * n C_n = 2 x (n + a - 1) C_{n-1} - (n + 2a - 2) C_{n-2}. */
static double gegenbauer(int n, double a, double x) {
    if (n == 0) return 1.0;
    double c0 = 1.0, c1 = 2.0 * a * x;
    for (int k = 2; k <= n; k++) {
        double c2 = (2.0 * x * (k + a - 1.0) * c1 - (k + 2.0 * a - 2.0) * c0) / k;
        c0 = c1;
        c1 = c2;
    }
    return c1;
}

/* Associated Legendre P_l^m(x) with the Condon-Shortley phase (the GSL
 * gsl_sf_legendre_Plm convention, also scipy lpmv's). */
static double legendre_plm(int l, int m, double x) {
    if (m > l) return 0.0;
    double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double pmm = 1.0, fact = 1.0;
    for (int i = 1; i <= m; i++) {
        pmm *= -fact * somx2;
        fact += 2.0;
    }
    if (l == m) return pmm;
    double pmmp1 = x * (2.0 * m + 1.0) * pmm;
    if (l == m + 1) return pmmp1;
    double pll = 0.0;
    for (int ll = m + 2; ll <= l; ll++) {
        pll = ((2.0 * ll - 1.0) * x * pmmp1 - (ll + m - 1.0) * pmm) / (ll - m);
        pmm = pmmp1;
        pmmp1 = pll;
    }
    return pll;
}

static double factorial_d(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; i++) f *= i;
    return f;
}

/* sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_l^m(x): gsl_sf_legendre_sphPlm */
static double legendre_sphplm(int l, int m, double x) {
    double norm = sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * factorial_d(l - m) / factorial_d(l + m));
    return norm * legendre_plm(l, m, x);
}

/* bfe_helper.cpp:24-26 */
static double scf_phi_nl(double s, int n, int l) {
    return -ORC_SQRT_FOURPI * pow(s, l) * pow(1.0 + s, -2 * l - 1) *
           gegenbauer(n, 2.0 * l + 1.5, (s - 1.0) / (s + 1.0));
}

/* bfe_helper.cpp:14-19 */
static double scf_rho_nl(double s, int n, int l) {
    double Knl = 0.5 * n * (n + 4 * l + 3) + (l + 1) * (2 * l + 1);
    double RR = Knl / (2.0 * M_PI) * pow(s, l) / (s * pow(1.0 + s, 2 * l + 3)) *
                gegenbauer(n, 2.0 * l + 1.5, (s - 1.0) / (s + 1.0));
    return ORC_SQRT_FOURPI * RR;
}

/* bfe_helper.cpp:31-90: spherical gradient of one (n,l,m) basis term */
static void scf_sph_grad_nlm(double s, double X, int n, int l, int m, double *sg) {
    double sintheta = sqrt(1.0 - X * X);
    double Phi_nl = scf_phi_nl(s, n, l);
    double Ylm = legendre_sphplm(l, m, X);
    double Plm = (m <= l) ? legendre_plm(l, m, X) : 0.0;
    double dPhinl_dr;
    if (n == 0) {
        dPhinl_dr = ORC_SQRT_FOURPI * pow(s, -1 + l) * pow(1.0 + s, -3 - 2 * l) * (1.0 + s) *
                    (l * (s - 1.0) + s);
    } else {
        double xi = (s - 1.0) / (s + 1.0);
        dPhinl_dr = ORC_SQRT_FOURPI * pow(s, -1 + l) * pow(1.0 + s, -3 - 2 * l) *
                    (-2.0 * (3 + 4 * l) * s * gegenbauer(n - 1, 2.5 + 2 * l, xi) +
                     (1.0 + s) * (l * (s - 1.0) + s) * gegenbauer(n, 1.5 + 2 * l, xi));
    }
    dPhinl_dr *= Ylm;
    double dYlm_dtheta;
    if (l == 0) {
        dYlm_dtheta = 0.0;
    } else {
        double Pl1m = (m <= l - 1) ? legendre_plm(l - 1, m, X) : 0.0;
        double A = sqrt(2.0 * l + 1.0) / ORC_SQRT_FOURPI *
                   sqrt(factorial_d(l - m) / factorial_d(l + m));
        dYlm_dtheta = A / sintheta * (l * X * Plm - (l + m) * Pl1m);
    }
    sg[0] = dPhinl_dr;
    sg[1] = dYlm_dtheta * Phi_nl / s;
    sg[2] = (m == 0 ? 0.0 : (double)m) * Ylm * Phi_nl;
}

static int scf_idx(int n, int l, int m, int lmax) { return m + (lmax + 1) * (l + (lmax + 1) * n); }

/* bfe.cpp:61-112 (one point) */
static double scf_value1(const orc_comp *c, const double *q) {
    double G = c->pars[0], M = c->pars[1], r_s = c->pars[2];
    int nmax = c->nmax, lmax = c->lmax;
    double r = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    double s = r / r_s, X = q[2] / r, phi = atan2(q[1], q[0]);
    double val = 0.0;
    for (int n = 0; n <= nmax; n++)
        for (int l = 0; l <= lmax; l++)
            for (int m = 0; m <= l; m++) {
                int i = scf_idx(n, l, m, lmax);
                if (c->S[i] == 0.0 && c->T[i] == 0.0) continue;
                val += scf_phi_nl(s, n, l) * legendre_sphplm(l, m, X) *
                       (c->S[i] * cos(m * phi) + c->T[i] * sin(m * phi));
            }
    return val * G * M / r_s;
}

/* bfe.cpp:20-59 (one point) */
static double scf_density1(const orc_comp *c, const double *q) {
    double M = c->pars[1], r_s = c->pars[2];
    int nmax = c->nmax, lmax = c->lmax;
    double r = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    double s = r / r_s, X = q[2] / r, phi = atan2(q[1], q[0]);
    double val = 0.0;
    for (int n = 0; n <= nmax; n++)
        for (int l = 0; l <= lmax; l++)
            for (int m = 0; m <= l; m++) {
                int i = scf_idx(n, l, m, lmax);
                if (c->S[i] == 0.0 && c->T[i] == 0.0) continue;
                val += scf_rho_nl(s, n, l) * legendre_sphplm(l, m, X) *
                       (c->S[i] * cos(m * phi) + c->T[i] * sin(m * phi));
            }
    return val * M / (r_s * r_s * r_s);
}

/* bfe.cpp:114-189 (one point; accumulates into g) */
static void scf_gradient1(const orc_comp *c, const double *q, double *g) {
    double G = c->pars[0], M = c->pars[1], r_s = c->pars[2];
    int nmax = c->nmax, lmax = c->lmax;
    double r = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    double s = r / r_s, X = q[2] / r, phi = atan2(q[1], q[0]);
    double sintheta = sqrt(1.0 - X * X), cosphi = cos(phi), sinphi = sin(phi);
    double acc[3] = {0.0, 0.0, 0.0}, sg[3];
    for (int n = 0; n <= nmax; n++)
        for (int l = 0; l <= lmax; l++)
            for (int m = 0; m <= l; m++) {
                int i = scf_idx(n, l, m, lmax);
                if (c->S[i] == 0.0 && c->T[i] == 0.0) continue;
                double cm = cos(m * phi), sm = sin(m * phi);
                double tmp = c->S[i] * cm + c->T[i] * sm;
                scf_sph_grad_nlm(s, X, n, l, m, sg);
                acc[0] += sg[0] * tmp;
                acc[1] += sg[1] * tmp;
                acc[2] += sg[2] * (c->T[i] * cm - c->S[i] * sm) / (s * sintheta);
            }
    double fac = G * M / (r_s * r_s);
    g[0] += (sintheta * cosphi * acc[0] + X * cosphi * acc[1] - sinphi * acc[2]) * fac;
    g[1] += (sintheta * sinphi * acc[0] + X * sinphi * acc[1] + cosphi * acc[2]) * fac;
    g[2] += (X * acc[0] - sintheta * acc[1]) * fac;
}

/* ------------------------------------------------------------------------ */
/* Analytic components (builtin_potentials.cpp)                              */

static double comp_value(const orc_comp *c, const double *q) {
    const double *p = c->pars;
    switch (c->type) {
    case ORC_HERNQUIST: { /* :211-219 */
        double r = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        return -p[0] * p[1] / (r + p[2]);
    }
    case ORC_PLUMMER: { /* :292-300 */
        double r2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
        return -p[0] * p[1] / sqrt(r2 + p[2] * p[2]);
    }
    case ORC_MIYAMOTONAGAI: { /* :1287-1296 */
        double zd = p[2] + sqrt(q[2] * q[2] + p[3] * p[3]);
        return -p[0] * p[1] / sqrt(q[0] * q[0] + q[1] * q[1] + zd * zd);
    }
    case ORC_NFW: { /* :819-833 */
        double v_h2 = -p[0] * p[1] / p[2];
        double u = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) / p[2];
        if (u == 0.0) return v_h2;
        return v_h2 * log(1.0 + u) / u;
    }
    case ORC_SCF:
        return scf_value1(c, q);
    default:
        return 0.0;
    }
}

static double comp_density(const orc_comp *c, const double *q) {
    const double *p = c->pars;
    switch (c->type) {
    case ORC_HERNQUIST: { /* :235-244 */
        double r = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        double rho0 = p[1] / (2.0 * M_PI * p[2] * p[2] * p[2]);
        return rho0 / ((r / p[2]) * pow(1.0 + r / p[2], 3));
    }
    case ORC_PLUMMER: { /* :316-324 */
        double r2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
        return 3.0 * p[1] / (4.0 * M_PI * p[2] * p[2] * p[2]) * pow(1.0 + r2 / (p[2] * p[2]), -2.5);
    }
    case ORC_MIYAMOTONAGAI: { /* :1314-1331 */
        double M = p[1], a = p[2], b = p[3];
        double R2 = q[0] * q[0] + q[1] * q[1];
        double sqrt_zb = sqrt(q[2] * q[2] + b * b);
        double numer = (b * b * M / (4.0 * M_PI)) * (a * R2 + (a + 3.0 * sqrt_zb) * (a + sqrt_zb) * (a + sqrt_zb));
        double denom = pow(R2 + (a + sqrt_zb) * (a + sqrt_zb), 2.5) * sqrt_zb * sqrt_zb * sqrt_zb;
        return numer / denom;
    }
    case ORC_NFW: { /* :852-863 */
        double v_h2 = p[0] * p[1] / p[2];
        double r = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        double rho0 = v_h2 / (4.0 * M_PI * p[0] * p[2] * p[2]);
        return rho0 / ((r / p[2]) * pow(1.0 + r / p[2], 2));
    }
    case ORC_SCF:
        return scf_density1(c, q);
    default:
        return 0.0;
    }
}

/* accumulates into g, builtin_potentials.cpp *_gradient_single */
static void comp_gradient(const orc_comp *c, const double *q, double *g) {
    const double *p = c->pars;
    switch (c->type) {
    case ORC_HERNQUIST: { /* :221-233 */
        double r = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        double fac = p[0] * p[1] / ((r + p[2]) * (r + p[2]) * r);
        g[0] += fac * q[0];
        g[1] += fac * q[1];
        g[2] += fac * q[2];
        break;
    }
    case ORC_PLUMMER: { /* :302-314 */
        double R2b = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + p[2] * p[2];
        double fac = p[0] * p[1] / sqrt(R2b) / R2b;
        g[0] += fac * q[0];
        g[1] += fac * q[1];
        g[2] += fac * q[2];
        break;
    }
    case ORC_MIYAMOTONAGAI: { /* :1298-1312 */
        double sqrtz = sqrt(q[2] * q[2] + p[3] * p[3]);
        double zd = p[2] + sqrtz;
        double fac = p[0] * p[1] * pow(q[0] * q[0] + q[1] * q[1] + zd * zd, -1.5);
        g[0] += fac * q[0];
        g[1] += fac * q[1];
        g[2] += fac * q[2] * (1.0 + p[2] / sqrtz);
        break;
    }
    case ORC_NFW: { /* :835-850 */
        double v_h2 = p[0] * p[1] / p[2];
        double u = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) / p[2];
        double fac = v_h2 / (u * u * u) / (p[2] * p[2]) * (log(1.0 + u) - u / (1.0 + u));
        g[0] += fac * q[0];
        g[1] += fac * q[1];
        g[2] += fac * q[2];
        break;
    }
    case ORC_SCF:
        scf_gradient1(c, q, g);
        break;
    default:
        break;
    }
}

/* Closed-form Hessians, accumulated (row-major 3x3).  The reference writes
 * sympy-generated expressions (builtin_potentials.cpp:246-286, 326-360,
 * 881-920, 1333-1379); these are the same second derivatives derived by
 * hand from the potentials, H = f(r) I + f'(r) q q^T / r for the spherical
 * ones.  SCF: the reference installs no Hessian (contributes nothing). */
static void comp_hessian(const orc_comp *c, const double *q, double *H) {
    const double *p = c->pars;
    double x = q[0], y = q[1], z = q[2];
    double r2 = x * x + y * y + z * z, r = sqrt(r2);
    double f, fp_over_r; /* H_ij = f delta_ij + fp_over_r x_i x_j */
    switch (c->type) {
    case ORC_HERNQUIST: {
        double GM = p[0] * p[1], rc = r + p[2];
        f = GM / (r * rc * rc);
        fp_over_r = -GM * (3.0 * r + p[2]) / (r2 * r * rc * rc * rc);
        break;
    }
    case ORC_PLUMMER: {
        double GM = p[0] * p[1], D = r2 + p[2] * p[2];
        double Dm32 = 1.0 / (D * sqrt(D));
        f = GM * Dm32;
        fp_over_r = -3.0 * GM * Dm32 / D;
        break;
    }
    case ORC_NFW: {
        double GM = p[0] * p[1], rs = p[2], u = r / rs;
        double bracket = log(1.0 + u) - u / (1.0 + u);
        f = GM / (r2 * r) * bracket;
        fp_over_r = (-3.0 * f / r + GM / (r2 * rs * rs * (1.0 + u) * (1.0 + u))) / r;
        break;
    }
    case ORC_MIYAMOTONAGAI: {
        double GM = p[0] * p[1], a = p[2], b = p[3];
        double s = sqrt(z * z + b * b), zeta = a + s;
        double D = x * x + y * y + zeta * zeta;
        double Dm32 = 1.0 / (D * sqrt(D)), Dm52 = Dm32 / D;
        double w = zeta * z / s; /* d(zeta^2/2)/dz */
        H[0] += GM * (Dm32 - 3.0 * x * x * Dm52);
        H[4] += GM * (Dm32 - 3.0 * y * y * Dm52);
        H[1] += -3.0 * GM * x * y * Dm52;
        H[3] += -3.0 * GM * x * y * Dm52;
        H[2] += -3.0 * GM * x * w * Dm52;
        H[6] += -3.0 * GM * x * w * Dm52;
        H[5] += -3.0 * GM * y * w * Dm52;
        H[7] += -3.0 * GM * y * w * Dm52;
        H[8] += GM * ((z * z / (s * s) + zeta * b * b / (s * s * s)) * Dm32 - 3.0 * w * w * Dm52);
        return;
    }
    default:
        return;
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H[3 * i + j] += (i == j ? f : 0.0) + fp_over_r * q[i] * q[j];
}

/* ------------------------------------------------------------------------ */
/* Composite (cpotential.cpp)                                               */

/* apply_shift_rotate with transpose=0 into a zeroed output, :80-106,137-149 */
static void shift_rotate(const orc_comp *c, const double *q, double *qt) {
    double tmp[3] = {q[0] - c->q0[0], q[1] - c->q0[1], q[2] - c->q0[2]};
    const double *R = c->R;
    qt[0] = 0.0 + R[0] * tmp[0] + R[1] * tmp[1] + R[2] * tmp[2];
    qt[1] = 0.0 + R[3] * tmp[0] + R[4] * tmp[1] + R[5] * tmp[2];
    qt[2] = 0.0 + R[6] * tmp[0] + R[7] * tmp[1] + R[8] * tmp[2];
}

double orc_value(const orc_pot *P, double t, const double *q) {
    (void)t;
    double v = 0.0, qt[3];
    for (int i = 0; i < P->n; i++) {
        const orc_comp *c = &P->c[i];
        if (!c->shift_rotate) {
            v = v + comp_value(c, q);
        } else {
            shift_rotate(c, q, qt);
            v = v + comp_value(c, qt);
        }
    }
    return v;
}

double orc_density(const orc_pot *P, double t, const double *q) {
    (void)t;
    double v = 0.0, qt[3];
    for (int i = 0; i < P->n; i++) {
        const orc_comp *c = &P->c[i];
        if (!c->shift_rotate) {
            v = v + comp_density(c, q);
        } else {
            shift_rotate(c, q, qt);
            v = v + comp_density(c, qt);
        }
    }
    return v;
}

/* c_hessian, :290-314 (rotation of the Hessian is not applied, as there) */
void orc_hessian(const orc_pot *P, double t, const double *q, double *H) {
    (void)t;
    double qt[3];
    for (int k = 0; k < 9; k++) H[k] = 0.0;
    for (int i = 0; i < P->n; i++) {
        const orc_comp *c = &P->c[i];
        if (!c->shift_rotate) {
            comp_hessian(c, q, H);
        } else {
            shift_rotate(c, q, qt);
            comp_hessian(c, qt, H);
        }
    }
}

/* c_gradient for one point (AoS q[3], g[3]); zeroes g first (:238-240) and
 * accumulates components in order, rotating shifted ones back (:242-280). */
static void grad1(const orc_pot *P, const double *q, double *g) {
    g[0] = g[1] = g[2] = 0.0;
    for (int i = 0; i < P->n; i++) {
        const orc_comp *c = &P->c[i];
        if (!c->shift_rotate) {
            comp_gradient(c, q, g);
        } else {
            double qt[3], tg[3] = {0.0, 0.0, 0.0};
            shift_rotate(c, q, qt);
            comp_gradient(c, qt, tg);
            const double *R = c->R; /* apply_rotate_T transpose=1, :108-134 */
            double gx = R[0] * tg[0] + R[3] * tg[1] + R[6] * tg[2];
            double gy = R[1] * tg[0] + R[4] * tg[1] + R[7] * tg[2];
            double gz = R[2] * tg[0] + R[5] * tg[1] + R[8] * tg[2];
            g[0] += gx;
            g[1] += gy;
            g[2] += gz;
        }
    }
}

/* c_gradient over SoA (3,N) arrays, :214-287 */
void orc_gradient(const orc_pot *P, size_t N, double t, const double *q, double *grad) {
    (void)t;
    for (size_t i = 0; i < N; i++) {
        double qi[3] = {q[i], q[N + i], q[2 * N + i]}, g[3];
        grad1(P, qi, g);
        grad[i] = g[0];
        grad[N + i] = g[1];
        grad[2 * N + i] = g[2];
    }
}

/* c_d2_dr2, :346-371 (central second difference along r, h = 1e-2) */
double orc_d2_dr2(const orc_pot *P, double t, const double *q) {
    double eps[3], h = 1e-2;
    double r = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    for (int j = 0; j < 3; j++) eps[j] = q[j] + h * q[j] / r;
    double d2 = orc_value(P, t, eps);
    d2 = d2 - 2.0 * orc_value(P, t, q);
    for (int j = 0; j < 3; j++) eps[j] = q[j] - h * q[j] / r;
    d2 = d2 + orc_value(P, t, eps);
    return d2 / (h * h);
}

/* ------------------------------------------------------------------------ */
/* Leapfrog (leapfrog.pyx:24-121)                                           */

/* w0 (6,N) SoA; t (ntimes); out (6,ntimes,N) if save_all else (6,N).
 * dt = t[1]-t[0] for every step (:78). */
void orc_leapfrog(const orc_pot *P, const double *w0, size_t N, const double *t, size_t ntimes,
                  int save_all, double *out) {
    double dt = t[1] - t[0];
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        size_t i = (size_t)ii;
        double x[3], v[3], vh[3], g[3];
        for (int k = 0; k < 3; k++) {
            x[k] = w0[k * N + i];
            v[k] = w0[(3 + k) * N + i];
        }
        if (save_all)
            for (int k = 0; k < 6; k++) out[(k * ntimes + 0) * N + i] = w0[k * N + i];
        /* c_init_velocity, :24-32 */
        grad1(P, x, g);
        for (int k = 0; k < 3; k++) vh[k] = v[k] - g[k] * dt / 2.0;
        for (size_t j = 1; j < ntimes; j++) {
            /* c_leapfrog_step, :35-51 */
            for (int k = 0; k < 3; k++) x[k] = x[k] + vh[k] * dt;
            grad1(P, x, g);
            for (int k = 0; k < 3; k++) {
                v[k] = vh[k] - g[k] * dt / 2.0;
                vh[k] = vh[k] - g[k] * dt;
            }
            if (save_all)
                for (int k = 0; k < 3; k++) {
                    out[(k * ntimes + j) * N + i] = x[k];
                    out[((3 + k) * ntimes + j) * N + i] = v[k];
                }
        }
        if (!save_all)
            for (int k = 0; k < 3; k++) {
                out[k * N + i] = x[k];
                out[(3 + k) * N + i] = ntimes > 1 ? v[k] : w0[(3 + k) * N + i];
            }
    }
}

/* Ruth4 (ruth4.pyx:24-106): per step, four substeps j of kick by d_j then
 * drift by c_j, each kick with the gradient at the current position (the
 * first, d_0 = 0, leaves v unchanged and is still evaluated, as c_ruth4_step
 * does).  Coefficients from 2^(1/3) (ruth4.pyx:67-80). */
void orc_ruth4(const orc_pot *P, const double *w0, size_t N, const double *t, size_t ntimes,
               int save_all, double *out) {
    double dt = t[1] - t[0];
    double two_13 = pow(2.0, 1.0 / 3.0);
    double cs[4], ds[4];
    cs[0] = 1.0 / (2.0 * (2.0 - two_13));
    cs[1] = (1.0 - two_13) / (2.0 * (2.0 - two_13));
    cs[2] = (1.0 - two_13) / (2.0 * (2.0 - two_13));
    cs[3] = 1.0 / (2.0 * (2.0 - two_13));
    ds[0] = 0.0;
    ds[1] = 1.0 / (2.0 - two_13);
    ds[2] = -two_13 / (2.0 - two_13);
    ds[3] = 1.0 / (2.0 - two_13);
#pragma omp parallel for schedule(static)
    for (long long ii = 0; ii < (long long)N; ii++) {
        size_t i = (size_t)ii;
        double x[3], v[3], g[3];
        for (int k = 0; k < 3; k++) {
            x[k] = w0[k * N + i];
            v[k] = w0[(3 + k) * N + i];
        }
        if (save_all)
            for (int k = 0; k < 6; k++) out[(k * ntimes + 0) * N + i] = w0[k * N + i];
        for (size_t j = 1; j < ntimes; j++) {
            for (int s = 0; s < 4; s++) {
                grad1(P, x, g);
                for (int k = 0; k < 3; k++) {
                    v[k] = v[k] - ds[s] * g[k] * dt;
                    x[k] = x[k] + cs[s] * v[k] * dt;
                }
            }
            if (save_all)
                for (int k = 0; k < 3; k++) {
                    out[(k * ntimes + j) * N + i] = x[k];
                    out[((3 + k) * ntimes + j) * N + i] = v[k];
                }
        }
        if (!save_all)
            for (int k = 0; k < 3; k++) {
                out[k * N + i] = x[k];
                out[(3 + k) * N + i] = v[k];
            }
    }
}

/* ------------------------------------------------------------------------ */
/* DOP853 (dop853.cpp)                                                      */

/* One ODE system for the DOP853 core: n components, right-hand side f. */
typedef struct orc_sys {
    int n;
    void (*f)(const struct orc_sys *S, const double *y, double *dy);
    const orc_pot *P;    /* external potential */
    int nb;              /* orbits (Fwrapper_T) or bodies (Fwrapper_direct_nbody) */
    const orc_pot *const *bodies; /* n-body systems: potentials of bodies 0..nbody-1 */
    int nbody;                    /* (a NULL entry is a massless body) */
} orc_sys;

/* hamiltonian_gradient_T for the static frame (chamiltonian.cpp:38-57,
 * builtin_frames.cpp:18-24): f = (p, -grad Phi(q)) over a (6,nb) system. */
static void orc_fcn(const orc_pot *P, int nb, const double *y, double *f) {
    for (int i = 0; i < nb; i++) {
        double q[3] = {y[i], y[nb + i], y[2 * nb + i]}, g[3];
        grad1(P, q, g);
        f[i] = y[3 * nb + i];
        f[nb + i] = y[4 * nb + i];
        f[2 * nb + i] = y[5 * nb + i];
        f[3 * nb + i] = -g[0];
        f[4 * nb + i] = -g[1];
        f[5 * nb + i] = -g[2];
    }
}

static void sys_orbits(const orc_sys *S, const double *y, double *f) { orc_fcn(S->P, S->nb, y, f); }

static void nbody_grad(const orc_pot *ext, const orc_pot *const *bodies, int nbody, int self,
                       const double *q, const double *w, double *g);

/* Fwrapper_direct_nbody (dop853.cpp:990-1006) for bodies in (nb, 6) rows:
 * f = (p, -grad Phi_ext(q)) per body, then c_nbody_acceleration
 * (cpotential.cpp:389-415) subtracts, body by body in index order, the
 * gradient of every other body's potential centred on that body.
 * -(g_ext + f_0 + f_1 ...) equals (-g_ext) + (-f_0) + (-f_1) ... exactly. */
static void sys_nbody(const orc_sys *S, const double *y, double *f) {
    for (int i = 0; i < S->nb; i++) {
        const double *q = y + 6 * i;
        double g[3];
        nbody_grad(S->P, S->bodies, S->nbody, i, q, y, g);
        f[6 * i + 0] = q[3];
        f[6 * i + 1] = q[4];
        f[6 * i + 2] = q[5];
        f[6 * i + 3] = -g[0];
        f[6 * i + 4] = -g[1];
        f[6 * i + 5] = -g[2];
    }
}

static double sign_d(double a, double b) { return (b < 0.0) ? -fabs(a) : fabs(a); }
static double min_d(double a, double b) { return (a < b) ? a : b; }
static double max_d(double a, double b) { return (a > b) ? a : b; }

typedef struct {
    long nstep, naccpt, nrejct, nfcn;
} orc_dopstats;

/* hinit, dop853.cpp:18-86 (scalar tolerances, iord = 8): Hairer's initial
 * step guess from the norms of f0 and y, one explicit Euler step and the
 * resulting estimate of the second derivative.  f1 and yy1 are work arrays. */
static double orc_hinit(const orc_sys *S, double x, const double *y, double posneg,
                        const double *f0, double *f1, double *yy1, double hmax, double rtoli,
                        double atoli) {
    const int n = S->n;
    double dnf = 0.0, dny = 0.0;
    (void)x;
    for (int i = 0; i < n; i++) {
        double sk = atoli + rtoli * fabs(y[i]);
        double sqr = f0[i] / sk;
        dnf += sqr * sqr;
        sqr = y[i] / sk;
        dny += sqr * sqr;
    }
    double h = ((dnf <= 1.0e-10) || (dny <= 1.0e-10)) ? 1.0e-6 : sqrt(dny / dnf) * 0.01;
    h = min_d(h, hmax);
    h = sign_d(h, posneg);
    for (int i = 0; i < n; i++) yy1[i] = y[i] + h * f0[i];
    S->f(S, yy1, f1);
    double der2 = 0.0;
    for (int i = 0; i < n; i++) {
        double sk = atoli + rtoli * fabs(y[i]);
        double sqr = (f1[i] - f0[i]) / sk;
        der2 += sqr * sqr;
    }
    der2 = sqrt(der2) / h;
    double der12 = max_d(fabs(der2), sqrt(dnf));
    double h1 = (der12 <= 1.0e-15) ? max_d(1.0e-6, fabs(h) * 1.0e-3)
                                   : pow(0.01 / der12, 1.0 / 8.0);
    h = min_d(100.0 * fabs(h), min_d(h1, hmax));
    return sign_d(h, posneg);
}

/* dopcor, :89-650, specialised to scalar tolerances, no solout, beta=0,
 * meth=1, with gala's dense-output fill at out_t (:584-612).  y (n=6 nb)
 * is advanced in place.  Returns the idid code. */
static int orc_dopcor(const orc_sys *S, double x, double *y, double xend, double hmax,
                      double h, double rtoli, double atoli, long nmax, long nstiff, double uround,
                      const double *out_t, int n_out, double *out_y, double *work,
                      orc_dopstats *st) {
    const int n = S->n;
    const double safe = 0.9, fac1 = 0.333, fac2 = 6.0, beta = 0.0;
    double *k1 = work, *k2 = k1 + n, *k3 = k2 + n, *k4 = k3 + n, *k5 = k4 + n, *k6 = k5 + n,
           *k7 = k6 + n, *k8 = k7 + n, *k9 = k8 + n, *k10 = k9 + n, *yy1 = k10 + n,
           *rc = yy1 + n; /* rcont1..8: rc + j*n */
    const double(*A)[16] = ORC_A;
    const double *C = ORC_C, *B = ORC_B, *E5 = ORC_E5, *BHH = ORC_BHH;
    const double(*D)[16] = ORC_D;
    int dense = (out_y != NULL);
    int output_idx = 0;
    double facold = 1.0e-4, expo1 = 1.0 / 8.0 - beta * 0.2;
    double facc1 = 1.0 / fac1, facc2 = 1.0 / fac2;
    double posneg = sign_d(1.0, xend - x);
    int last = 0, reject = 0, iasti = 0, nonsti = 0;
    double hlamb = 0.0, xold = x, hout = h;
    st->nstep = st->naccpt = st->nrejct = st->nfcn = 0;

    S->f(S, y, k1);
    hmax = fabs(hmax);
    if (h == 0.0) h = orc_hinit(S, x, y, posneg, k1, k2, k3, hmax, rtoli, atoli); /* :315-317 */
    st->nfcn += 2; /* :319 (counts the hinit evaluation slot too) */

    while (1) {
        if (st->nstep > nmax) return -2;                    /* :340-347 */
        if (0.1 * fabs(h) <= fabs(x) * uround) return -3;   /* :349-354 */
        if ((x + 1.01 * h - xend) * posneg > 0.0) {         /* :356-359 */
            h = xend - x;
            last = 1;
        }
        xold = x;
        hout = h;
        st->nstep++;

        /* the twelve stages, :369-408 */
        for (int i = 0; i < n; i++) yy1[i] = y[i] + h * A[1][0] * k1[i];
        S->f(S, yy1, k2);
        for (int i = 0; i < n; i++) yy1[i] = y[i] + h * (A[2][0] * k1[i] + A[2][1] * k2[i]);
        S->f(S, yy1, k3);
        for (int i = 0; i < n; i++) yy1[i] = y[i] + h * (A[3][0] * k1[i] + A[3][2] * k3[i]);
        S->f(S, yy1, k4);
        for (int i = 0; i < n; i++)
            yy1[i] = y[i] + h * (A[4][0] * k1[i] + A[4][2] * k3[i] + A[4][3] * k4[i]);
        S->f(S, yy1, k5);
        for (int i = 0; i < n; i++)
            yy1[i] = y[i] + h * (A[5][0] * k1[i] + A[5][3] * k4[i] + A[5][4] * k5[i]);
        S->f(S, yy1, k6);
        for (int i = 0; i < n; i++)
            yy1[i] = y[i] + h * (A[6][0] * k1[i] + A[6][3] * k4[i] + A[6][4] * k5[i] + A[6][5] * k6[i]);
        S->f(S, yy1, k7);
        for (int i = 0; i < n; i++)
            yy1[i] = y[i] + h * (A[7][0] * k1[i] + A[7][3] * k4[i] + A[7][4] * k5[i] +
                                 A[7][5] * k6[i] + A[7][6] * k7[i]);
        S->f(S, yy1, k8);
        for (int i = 0; i < n; i++)
            yy1[i] = y[i] + h * (A[8][0] * k1[i] + A[8][3] * k4[i] + A[8][4] * k5[i] +
                                 A[8][5] * k6[i] + A[8][6] * k7[i] + A[8][7] * k8[i]);
        S->f(S, yy1, k9);
        for (int i = 0; i < n; i++)
            yy1[i] = y[i] + h * (A[9][0] * k1[i] + A[9][3] * k4[i] + A[9][4] * k5[i] +
                                 A[9][5] * k6[i] + A[9][6] * k7[i] + A[9][7] * k8[i] +
                                 A[9][8] * k9[i]);
        S->f(S, yy1, k10);
        for (int i = 0; i < n; i++)
            yy1[i] = y[i] + h * (A[10][0] * k1[i] + A[10][3] * k4[i] + A[10][4] * k5[i] +
                                 A[10][5] * k6[i] + A[10][6] * k7[i] + A[10][7] * k8[i] +
                                 A[10][8] * k9[i] + A[10][9] * k10[i]);
        S->f(S, yy1, k2);
        double xph = x + h;
        for (int i = 0; i < n; i++)
            yy1[i] = y[i] + h * (A[11][0] * k1[i] + A[11][3] * k4[i] + A[11][4] * k5[i] +
                                 A[11][5] * k6[i] + A[11][6] * k7[i] + A[11][7] * k8[i] +
                                 A[11][8] * k9[i] + A[11][9] * k10[i] + A[11][10] * k2[i]);
        S->f(S, yy1, k3);
        st->nfcn += 11;
        for (int i = 0; i < n; i++) {
            k4[i] = B[0] * k1[i] + B[5] * k6[i] + B[6] * k7[i] + B[7] * k8[i] + B[8] * k9[i] +
                    B[9] * k10[i] + B[10] * k2[i] + B[11] * k3[i];
            k5[i] = y[i] + h * k4[i];
        }

        /* error estimation, :416-444 */
        double err = 0.0, err2 = 0.0;
        for (int i = 0; i < n; i++) {
            double sk = atoli + rtoli * max_d(fabs(y[i]), fabs(k5[i]));
            double erri = k4[i] - BHH[0] * k1[i] - BHH[8] * k9[i] - BHH[11] * k3[i];
            double sqr = erri / sk;
            err2 += sqr * sqr;
            erri = E5[0] * k1[i] + E5[5] * k6[i] + E5[6] * k7[i] + E5[7] * k8[i] + E5[8] * k9[i] +
                   E5[9] * k10[i] + E5[10] * k2[i] + E5[11] * k3[i];
            sqr = erri / sk;
            err += sqr * sqr;
        }
        double deno = err + 0.01 * err2;
        if (deno <= 0.0) deno = 1.0;
        err = fabs(h) * err * sqrt(1.0 / (deno * (double)n));

        /* computation of hnew, :446-452 */
        double fac11 = pow(err, expo1);
        double fac = fac11 / pow(facold, beta);
        fac = max_d(facc2, min_d(facc1, fac / safe));
        double hnew = h / fac;

        if (err <= 1.0) {
            /* step accepted, :454-637 */
            facold = max_d(err, 1.0e-4);
            st->naccpt++;
            S->f(S, k5, k4);
            st->nfcn++;

            /* stiffness detection, :463-486 (the early return fires on the
             * first detection, as written there) */
            if (!(st->naccpt % nstiff) || (iasti > 0)) {
                double stnum = 0.0, stden = 0.0;
                for (int i = 0; i < n; i++) {
                    double sqr = k4[i] - k3[i];
                    stnum += sqr * sqr;
                    sqr = k5[i] - yy1[i];
                    stden += sqr * sqr;
                }
                if (stden > 0.0) hlamb = h * sqrt(stnum / stden);
                if (hlamb > 6.1) {
                    nonsti = 0;
                    iasti++;
                    return -4;
                } else {
                    nonsti++;
                    if (nonsti == 6) iasti = 0;
                }
            }

            if (dense) {
                /* dense output coefficients, :492-582 */
                double *r1 = rc, *r2 = rc + n, *r3 = rc + 2 * n, *r4 = rc + 3 * n, *r5 = rc + 4 * n,
                       *r6 = rc + 5 * n, *r7 = rc + 6 * n, *r8 = rc + 7 * n;
                double *rr[4] = {r5, r6, r7, r8};
                for (int i = 0; i < n; i++) {
                    r1[i] = y[i];
                    double ydiff = k5[i] - y[i];
                    r2[i] = ydiff;
                    double bspl = h * k1[i] - ydiff;
                    r3[i] = bspl;
                    r4[i] = ydiff - h * k4[i] - bspl;
                    for (int j = 0; j < 4; j++)
                        rr[j][i] = D[j][0] * k1[i] + D[j][5] * k6[i] + D[j][6] * k7[i] +
                                   D[j][7] * k8[i] + D[j][8] * k9[i] + D[j][9] * k10[i] +
                                   D[j][10] * k2[i] + D[j][11] * k3[i];
                }
                for (int i = 0; i < n; i++)
                    yy1[i] = y[i] + h * (A[13][0] * k1[i] + A[13][6] * k7[i] + A[13][7] * k8[i] +
                                         A[13][8] * k9[i] + A[13][9] * k10[i] + A[13][10] * k2[i] +
                                         A[13][11] * k3[i] + A[13][12] * k4[i]);
                S->f(S, yy1, k10);
                for (int i = 0; i < n; i++)
                    yy1[i] = y[i] + h * (A[14][0] * k1[i] + A[14][5] * k6[i] + A[14][6] * k7[i] +
                                         A[14][7] * k8[i] + A[14][10] * k2[i] + A[14][11] * k3[i] +
                                         A[14][12] * k4[i] + A[14][13] * k10[i]);
                S->f(S, yy1, k2);
                for (int i = 0; i < n; i++)
                    yy1[i] = y[i] + h * (A[15][0] * k1[i] + A[15][5] * k6[i] + A[15][6] * k7[i] +
                                         A[15][7] * k8[i] + A[15][8] * k9[i] + A[15][12] * k4[i] +
                                         A[15][13] * k10[i] + A[15][14] * k2[i]);
                S->f(S, yy1, k3);
                st->nfcn += 3;
                for (int i = 0; i < n; i++)
                    for (int j = 0; j < 4; j++)
                        rr[j][i] = h * (rr[j][i] + D[j][12] * k4[i] + D[j][13] * k10[i] +
                                        D[j][14] * k2[i] + D[j][15] * k3[i]);

                /* fill requested output times inside [xold, xold+hout], :584-612 */
                double x0 = xold, x1 = xold + hout;
                while (output_idx < n_out) {
                    double t_out = out_t[output_idx];
                    if ((x0 <= t_out && t_out <= x1) || (x1 <= t_out && t_out <= x0)) {
                        double s = (t_out - xold) / hout, s1 = 1.0 - s;
                        for (int i = 0; i < n; i++)
                            out_y[(size_t)output_idx * n + i] =
                                r1[i] + s * (r2[i] + s1 * (r3[i] + s * (r4[i] + s1 * (r5[i] +
                                        s * (r6[i] + s1 * (r7[i] + s * r8[i]))))));
                        output_idx++;
                    } else {
                        break;
                    }
                }
            }

            memcpy(k1, k4, n * sizeof(double));
            memcpy(y, k5, n * sizeof(double));
            x = xph;
            if (last) return 1;
            if (fabs(hnew) > hmax) hnew = posneg * hmax;
            if (reject) hnew = posneg * min_d(fabs(hnew), fabs(h));
            reject = 0;
        } else {
            /* step rejected, :638-645 */
            hnew = h / min_d(facc1, fac11 / safe);
            reject = 1;
            if (st->naccpt >= 1) st->nrejct++;
            last = 0;
        }
        h = hnew;
    }
}

/* dop853_integrate_hamiltonian (dop853.pyx:196-250) with dop853_helper
 * (:90-193): orbits chunked nbatch at a time into one (6,nb) system with one
 * shared step.  w0 (6,N); out (6,ntimes,N) if save_all else (6,N).
 * status[N] gets each batch's idid; stats (N,4) = nstep,naccpt,nrejct,nfcn of
 * the batch the orbit belongs to (nullable). */
void orc_dop853(const orc_pot *P, const double *w0, size_t N, const double *t, size_t ntimes,
                double atol, double rtol, long nmax, long nstiff, double dt_max, int save_all,
                int nbatch, double *out, int *status, long *stats) {
    if (nmax == 0) nmax = 1000000;          /* :673-674 */
    if (nstiff == 0) nstiff = 1000;         /* :691-692 */
    else if (nstiff < 0) nstiff = nmax + 10; /* :693-694 */
    double hmax = dt_max == 0.0 ? t[ntimes - 1] - t[0] : dt_max; /* :787-788 */
    long nbl = (long)((N + nbatch - 1) / nbatch);
#pragma omp parallel for schedule(dynamic, 1)
    for (long b = 0; b < nbl; b++) {
        size_t i0 = (size_t)b * nbatch, i1 = i0 + nbatch < N ? i0 + nbatch : N;
        int nb = (int)(i1 - i0), n = 6 * nb;
        double *y = (double *)malloc(sizeof(double) * n);
        double *work = (double *)malloc(sizeof(double) * n * 19);
        double *oy = save_all ? (double *)malloc(sizeof(double) * n * ntimes) : NULL;
        for (int k = 0; k < 6; k++)
            for (int i = 0; i < nb; i++) y[k * nb + i] = w0[k * N + i0 + i];
        orc_dopstats st;
        orc_sys S = {n, sys_orbits, P, nb, NULL};
        int idid = orc_dopcor(&S, t[0], y, t[ntimes - 1], hmax, t[1] - t[0], rtol, atol, nmax,
                              nstiff, 2.220446049250313e-16, t, (int)ntimes, oy, work, &st);
        for (int i = 0; i < nb; i++) {
            if (status) status[i0 + i] = idid;
            if (stats) {
                stats[4 * (i0 + i) + 0] = st.nstep;
                stats[4 * (i0 + i) + 1] = st.naccpt;
                stats[4 * (i0 + i) + 2] = st.nrejct;
                stats[4 * (i0 + i) + 3] = st.nfcn;
            }
            for (int k = 0; k < 6; k++) {
                if (save_all) {
                    for (size_t j = 0; j < ntimes; j++)
                        out[(k * ntimes + j) * N + i0 + i] = oy[j * n + k * nb + i];
                } else {
                    out[k * N + i0 + i] = y[k * nb + i];
                }
            }
        }
        free(y);
        free(work);
        free(oy);
    }
}

/* Independent per-orbit integrations from t0[i] to t1[i] (one-orbit systems,
 * the reference with nbatch=1 and save_all=False), initial step h0[i].
 * w0 (6,N) -> out (6,N); status[N]; stats (N,4). */
void orc_dop853_orbits(const orc_pot *P, const double *w0, size_t N, const double *t0,
                       const double *t1, const double *h0, double atol, double rtol, long nmax,
                       long nstiff, double *out, int *status, long *stats) {
    if (nmax == 0) nmax = 1000000;
    if (nstiff == 0) nstiff = 1000;
    else if (nstiff < 0) nstiff = nmax + 10;
#pragma omp parallel for schedule(dynamic, 1)
    for (long long ii = 0; ii < (long long)N; ii++) {
        size_t i = (size_t)ii;
        double y[6], work[6 * 19];
        for (int k = 0; k < 6; k++) y[k] = w0[k * N + i];
        orc_dopstats st;
        orc_sys S = {6, sys_orbits, P, 1, NULL};
        int idid = orc_dopcor(&S, t0[i], y, t1[i], t1[i] - t0[i], h0[i], rtol, atol, nmax,
                              nstiff, 2.220446049250313e-16, NULL, 0, NULL, work, &st);
        for (int k = 0; k < 6; k++) out[k * N + i] = y[k];
        if (status) status[i] = idid;
        if (stats) {
            stats[4 * i + 0] = st.nstep;
            stats[4 * i + 1] = st.naccpt;
            stats[4 * i + 2] = st.nrejct;
            stats[4 * i + 3] = st.nfcn;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Mock stream (mockstream.pyx, df.pyx, leapfrog.pyx N-body helpers)       */

/* Gradient at q felt by body `self` (-1 for none) of an N-body system whose
 * bodies sit in rows of 6 in w: the external potential (c_gradient), then for
 * every body j < nbody other than self with a potential, that potential with
 * every component shifted to body j's position (do_shift_rotate = 1, q0 =
 * body j, its own R kept), added in body order (c_nbody_gradient_symplectic,
 * leapfrog.pyx:151-152, cpotential.cpp:418-442). */
static void nbody_grad(const orc_pot *ext, const orc_pot *const *bodies, int nbody, int self,
                       const double *q, const double *w, double *g) {
    grad1(ext, q, g);
    for (int j = 0; j < nbody; j++) {
        const orc_pot *bp = bodies ? bodies[j] : NULL;
        if (j == self || !bp || bp->n == 0) continue;
        orc_pot shifted = *bp;
        for (int i = 0; i < shifted.n; i++) {
            shifted.c[i].shift_rotate = 1;
            shifted.c[i].q0[0] = w[6 * j + 0];
            shifted.c[i].q0[1] = w[6 * j + 1];
            shifted.c[i].q0[2] = w[6 * j + 2];
        }
        double f2[3];
        grad1(&shifted, q, f2);
        g[0] += f2[0];
        g[1] += f2[1];
        g[2] += f2[2];
    }
}

/* leapfrog_integrate_nbody (leapfrog.pyx:161-257): n bodies in (n, 6) rows,
 * bodies 0..nbody-1 with the potentials `bodies` (NULL = massless), the rest
 * massless.  Half kicks for every body (c_init_velocity_nbody, :126-137), then
 * per step every body in index order drifts, takes its gradient against the
 * others' current positions and kicks (c_leapfrog_step_nbody, :140-158).
 * out (ntimes, n, 6) when save_all (row 0 = w0), else (n, 6). */
void orc_nbody_leapfrog(const orc_pot *ext, const orc_pot *const *bodies, int nbody,
                        const double *w0, int n, const double *t, size_t ntimes, int save_all,
                        double *out) {
    const double dt = t[1] - t[0]; /* :217 */
    double *w = (double *)malloc(sizeof(double) * 6 * (n > 0 ? n : 1));
    double *vh = (double *)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    memcpy(w, w0, sizeof(double) * 6 * n);
    if (save_all) memcpy(out, w0, sizeof(double) * 6 * n);
    for (int i = 0; i < n; i++) {
        double g[3];
        nbody_grad(ext, bodies, nbody, i, &w[6 * i], w, g);
        for (int k = 0; k < 3; k++) vh[3 * i + k] = w[6 * i + 3 + k] - g[k] * dt / 2.0;
    }
    for (size_t j = 1; j < ntimes; j++) {
        for (int i = 0; i < n; i++) {
            double g[3];
            for (int k = 0; k < 3; k++) w[6 * i + k] = w[6 * i + k] + vh[3 * i + k] * dt;
            nbody_grad(ext, bodies, nbody, i, &w[6 * i], w, g);
            for (int k = 0; k < 3; k++) {
                w[6 * i + 3 + k] = vh[3 * i + k] - g[k] * dt / 2.0;
                vh[3 * i + k] = vh[3 * i + k] - g[k] * dt;
            }
            if (save_all)
                for (int k = 0; k < 6; k++) out[(j * n + i) * 6 + k] = w[6 * i + k];
        }
    }
    if (!save_all) memcpy(out, w, sizeof(double) * 6 * n);
    free(w);
    free(vh);
}

/* ruth4_integrate_nbody (ruth4.pyx:138-261): per step, body by body in index
 * order, the four kick-then-drift substeps of c_ruth4_step_nbody
 * (:119-136), each kick from the gradient against the others' current
 * positions.  Layouts as orc_nbody_leapfrog. */
void orc_nbody_ruth4(const orc_pot *ext, const orc_pot *const *bodies, int nbody,
                     const double *w0, int n, const double *t, size_t ntimes, int save_all,
                     double *out) {
    const double dt = t[1] - t[0];
    const double two_13 = pow(2.0, 1.0 / 3.0);
    double cs[4], ds[4];
    cs[0] = 1.0 / (2.0 * (2.0 - two_13));
    cs[1] = (1.0 - two_13) / (2.0 * (2.0 - two_13));
    cs[2] = (1.0 - two_13) / (2.0 * (2.0 - two_13));
    cs[3] = 1.0 / (2.0 * (2.0 - two_13));
    ds[0] = 0.0;
    ds[1] = 1.0 / (2.0 - two_13);
    ds[2] = -two_13 / (2.0 - two_13);
    ds[3] = 1.0 / (2.0 - two_13);
    double *w = (double *)malloc(sizeof(double) * 6 * (n > 0 ? n : 1));
    memcpy(w, w0, sizeof(double) * 6 * n);
    if (save_all) memcpy(out, w0, sizeof(double) * 6 * n);
    for (size_t j = 1; j < ntimes; j++) {
        for (int i = 0; i < n; i++) {
            double *x = &w[6 * i], *v = &w[6 * i + 3];
            for (int s = 0; s < 4; s++) {
                double g[3];
                nbody_grad(ext, bodies, nbody, i, x, w, g);
                for (int k = 0; k < 3; k++) {
                    v[k] = v[k] - ds[s] * g[k] * dt;
                    x[k] = x[k] + cs[s] * v[k] * dt;
                }
            }
            if (save_all)
                for (int k = 0; k < 6; k++) out[(j * n + i) * 6 + k] = w[6 * i + k];
        }
    }
    if (!save_all) memcpy(out, w, sizeof(double) * 6 * n);
    free(w);
}

/* direct_nbody_dop853 (nbody.pyx:30-115) through dop853_helper
 * (dop853.pyx:90-193): the n bodies (n, 6) are one system with
 * Fwrapper_direct_nbody forces, integrated over t with h = t[1] - t[0],
 * uround = eps, hmax = dt_max (0: the whole span) and the stiffness test off
 * (nstiff = -1); dense output at t.  out (ntimes, n, 6) when save_all, else
 * (n, 6).  Returns the idid; stats[4] nullable. */
int orc_nbody_dop853(const orc_pot *ext, const orc_pot *const *bodies, int nbody,
                     const double *w0, int n, const double *t, size_t ntimes, double atol,
                     double rtol, long nmax, double dt_max, int save_all, double *out,
                     long *stats) {
    if (nmax == 0) nmax = 1000000;                               /* :673-674 */
    const long nstiff = nmax + 10;                               /* -1: :693-694 */
    double hmax = dt_max == 0.0 ? t[ntimes - 1] - t[0] : dt_max; /* :787-788 */
    const int m = 6 * n;
    double *y = (double *)malloc(sizeof(double) * m);
    double *work = (double *)malloc(sizeof(double) * m * 19);
    memcpy(y, w0, sizeof(double) * m);
    orc_sys S = {m, sys_nbody, ext, n, bodies, nbody};
    orc_dopstats st;
    int idid = orc_dopcor(&S, t[0], y, t[ntimes - 1], hmax, t[1] - t[0], rtol, atol, nmax,
                          nstiff, 2.220446049250313e-16, save_all ? t : NULL, (int)ntimes,
                          save_all ? out : NULL, work, &st);
    if (!save_all) memcpy(out, y, sizeof(double) * m);
    if (stats) {
        stats[0] = st.nstep;
        stats[1] = st.naccpt;
        stats[2] = st.nrejct;
        stats[3] = st.nfcn;
    }
    free(y);
    free(work);
    return idid;
}

/* mockstream_leapfrog (mockstream.pyx:442-620): nbody system bodies (body 0
 * the progenitor, then perturbers; potentials `bodies`, NULL = massless)
 * and massless stream particles.
 *   full_time (nfull): the forward time grid; body_w0 (nbody, 6): the bodies
 *   at full_time[0]; spawn_t/nstream (nrel): release times and counts;
 *   stream_w0 (ntot,6) AoS; out: body_final (nbody, 6), stream_final (ntot,6).
 * The bodies' trajectory comes from leapfrog_integrate_nbody over full_time
 * (:536-549); each release batch integrates the bodies again from their
 * states on that trajectory together with the batch's particles (:555-599). */
void orc_mockstream_leapfrog(const orc_pot *ext, const orc_pot *const *bodies, int nbody,
                             const double *full_time, size_t nfull, const double *body_w0,
                             size_t nrel, const double *spawn_t, const int *nstream,
                             const double *stream_w0, double tfinal, double *body_final,
                             double *stream_final) {
    double dt = full_time[1] - full_time[0];
    double *traj = (double *)malloc(sizeof(double) * 6 * nbody * nfull); /* (nfull, nbody, 6) */
    orc_nbody_leapfrog(ext, bodies, nbody, body_w0, nbody, full_time, nfull, 1, traj);
    size_t *offs = (size_t *)malloc(sizeof(size_t) * (nrel + 1));
    offs[0] = 0;
    for (size_t i = 0; i < nrel; i++) offs[i + 1] = offs[i] + (size_t)nstream[i];
    long last_rel = -1;
    for (size_t i = 0; i < nrel; i++)
        if (nstream[i] > 0) last_rel = (long)i;

#pragma omp parallel for schedule(dynamic, 1)
    for (long long ii = 0; ii < (long long)nrel; ii++) {
        size_t i = (size_t)ii;
        int ns = nstream[i];
        if (ns == 0) continue;
        int idx = (int)((spawn_t[i] - full_time[0]) / dt + 0.5); /* :546 */
        int nbod = nbody + ns;
        double *w = (double *)malloc(sizeof(double) * 6 * nbod);
        double *vh = (double *)malloc(sizeof(double) * 3 * nbod);
        for (int k = 0; k < 6 * nbody; k++) w[k] = traj[(size_t)idx * 6 * nbody + k];
        for (int j = 0; j < ns; j++)
            for (int k = 0; k < 6; k++)
                w[6 * (nbody + j) + k] = stream_w0[6 * (offs[i] + j) + k];
        /* c_init_velocity_nbody for each body (:571-578) */
        for (int j = 0; j < nbod; j++) {
            double g[3];
            nbody_grad(ext, bodies, nbody, j, &w[6 * j], w, g);
            for (int k = 0; k < 3; k++) vh[3 * j + k] = w[6 * j + 3 + k] - g[k] * dt / 2.0;
        }
        int n_steps = (int)((tfinal - spawn_t[i]) / dt + 0.5); /* :581 */
        for (int s = 0; s < n_steps; s++) {
            for (int j = 0; j < nbod; j++) { /* c_leapfrog_step_nbody, :140-158 */
                double g[3];
                for (int k = 0; k < 3; k++) w[6 * j + k] = w[6 * j + k] + vh[3 * j + k] * dt;
                nbody_grad(ext, bodies, nbody, j, &w[6 * j], w, g);
                for (int k = 0; k < 3; k++) {
                    w[6 * j + 3 + k] = vh[3 * j + k] - g[k] * dt / 2.0;
                    vh[3 * j + k] = vh[3 * j + k] - g[k] * dt;
                }
            }
        }
        for (int j = 0; j < ns; j++)
            for (int k = 0; k < 6; k++)
                stream_final[6 * (offs[i] + j) + k] = w[6 * (nbody + j) + k];
        if ((long)i == last_rel)
            for (int k = 0; k < 6 * nbody; k++) body_final[k] = w[k];
        free(w);
        free(vh);
    }
    free(traj);
    free(offs);
}

/* mockstream_dop853's release loop (mockstream.pyx:286-303): for each
 * release i with nstream[i] > 0, the nbody bodies start from body_w[i]
 * (nbody, 6): their dense-output states at the release time,
 * mockstream.pyx:274-282) and the release's particles (the next nstream[i]
 * rows of stream_w0, (ntot, 6)) follow; the (nbody + k, 6) system is
 * integrated by dop853_step (dop853.pyx:29-79) from stream_t1[i] to tfinal
 * with h = dt0 and the defaults dop853() picks for it (uround 2.3e-16, hmax
 * tfinal - t1, dop853.cpp:733-792) and nstiff = 1 as dop853_step passes.
 * out (ntot, 6); body_final (nbody, 6): the bodies after the last release's
 * system; status[nrel] (0 where nstream[i] = 0), stats (nrel, 4) nullable. */
void orc_mockstream_dop853(const orc_pot *ext, const orc_pot *const *bodies, int nbody,
                           const double *body_w, size_t nrel, const double *stream_t1,
                           const int *nstream, const double *stream_w0, double tfinal,
                           double dt0, double atol, double rtol, long nmax, double *out,
                           double *body_final, int *status, long *stats) {
    if (nmax == 0) nmax = 1000000; /* :673-674 */
    size_t *offs = (size_t *)malloc(sizeof(size_t) * (nrel + 1));
    offs[0] = 0;
    for (size_t i = 0; i < nrel; i++) offs[i + 1] = offs[i] + (size_t)nstream[i];
    long last_rel = -1;
    for (size_t i = 0; i < nrel; i++)
        if (nstream[i] > 0) last_rel = (long)i;
#pragma omp parallel for schedule(dynamic, 1)
    for (long long ii = 0; ii < (long long)nrel; ii++) {
        size_t i = (size_t)ii;
        int k = nstream[i];
        if (status) status[i] = 0;
        if (k <= 0) continue;
        int nb = nbody + k, n = 6 * nb;
        double *y = (double *)malloc(sizeof(double) * n);
        double *work = (double *)malloc(sizeof(double) * n * 19);
        for (int c = 0; c < 6 * nbody; c++) y[c] = body_w[6 * nbody * i + c];
        for (int j = 0; j < k; j++)
            for (int c = 0; c < 6; c++)
                y[6 * (nbody + j) + c] = stream_w0[6 * (offs[i] + j) + c];
        orc_sys S = {n, sys_nbody, ext, nb, bodies, nbody};
        orc_dopstats st;
        int idid = orc_dopcor(&S, stream_t1[i], y, tfinal, tfinal - stream_t1[i], dt0, rtol, atol,
                              nmax, 1, 2.3e-16, NULL, 0, NULL, work, &st);
        for (int j = 0; j < k; j++)
            for (int c = 0; c < 6; c++) out[6 * (offs[i] + j) + c] = y[6 * (nbody + j) + c];
        if ((long)i == last_rel)
            for (int c = 0; c < 6 * nbody; c++) body_final[c] = y[c];
        if (status) status[i] = idid;
        if (stats) {
            stats[4 * i + 0] = st.nstep;
            stats[4 * i + 1] = st.naccpt;
            stats[4 * i + 2] = st.nrejct;
            stats[4 * i + 3] = st.nfcn;
        }
        free(y);
        free(work);
    }
    free(offs);
}

/* FardalStreamDF._sample (df.pyx:363-456) with get_rj_vj_R (:61-92) and
 * transform_from_sat (:94-106).  The random draws are supplied: z holds the
 * standard normals in the order the reference draws them (per release time,
 * trailing then leading, per particle kx, z, vt, vz); the reference's
 * normal(loc, scale) is loc + scale * z.
 *   prog_x, prog_v (ntimes,3); prog_m (ntimes); nparticles (ntimes)
 *   out x, v (ntot,3), t1 (ntot).  Returns the number of particles. */
long orc_fardal_sample(const orc_pot *P, double G, const double *prog_x, const double *prog_v,
                       const double *prog_t, const double *prog_m, const int *nparticles,
                       size_t ntimes, int lead, int trail, int gala_modified, const double *z,
                       double *ox, double *ov, double *ot1) {
    double k_mean[6] = {0}, k_disp[6] = {0};
    k_mean[0] = 2.0;
    k_disp[0] = gala_modified ? 0.5 : 0.4;
    k_mean[2] = 0.0;
    k_disp[2] = 0.5;
    k_mean[4] = 0.3;
    k_disp[4] = gala_modified ? 0.5 : 0.4;
    k_mean[5] = 0.0;
    k_disp[5] = 0.5;
    long j = 0, zi = 0;
    for (size_t i = 0; i < ntimes; i++) {
        if (prog_m[i] == 0.0) continue;
        const double *px = &prog_x[3 * i], *pv = &prog_v[3 * i];
        /* get_rj_vj_R */
        double dist = sqrt(px[0] * px[0] + px[1] * px[1] + px[2] * px[2]);
        double L[3] = {px[1] * pv[2] - px[2] * pv[1], -px[0] * pv[2] + px[2] * pv[0],
                       px[0] * pv[1] - px[1] * pv[0]};
        double Lnorm = sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
        double R[9];
        for (int k = 0; k < 3; k++) {
            R[k] = px[k] / dist;
            R[6 + k] = L[k] / Lnorm;
        }
        double Om = Lnorm / (dist * dist);
        double d2r = orc_d2_dr2(P, prog_t[i], px);
        double rj = pow(G * prog_m[i] / (Om * Om - d2r), 1.0 / 3.0);
        double vj = Om * rj;
        R[3] = -(R[1] * R[8] - R[2] * R[7]);
        R[4] = -(-R[0] * R[8] + R[2] * R[6]);
        R[5] = -(R[0] * R[7] - R[1] * R[6]);
        for (int tail = 0; tail < 2; tail++) {
            int use = tail == 0 ? trail : lead;
            if (!use) continue;
            double sgn = tail == 0 ? 1.0 : -1.0;
            for (int k = 0; k < nparticles[i]; k++) {
                double tx[3] = {0, 0, 0}, tv[3] = {0, 0, 0};
                double kx = k_mean[0] + k_disp[0] * z[zi++];
                tx[0] = kx * (sgn * rj);
                tx[2] = (k_mean[2] + k_disp[2] * z[zi++]) * (sgn * rj);
                tv[1] = (k_mean[4] + k_disp[4] * z[zi++]) * (sgn * vj);
                if (gala_modified) tv[1] *= kx;
                tv[2] = (k_mean[5] + k_disp[5] * z[zi++]) * (sgn * vj);
                for (int a = 0; a < 3; a++) {
                    ox[3 * (j + k) + a] = R[a] * tx[0] + R[3 + a] * tx[1] + R[6 + a] * tx[2] + px[a];
                    ov[3 * (j + k) + a] = R[a] * tv[0] + R[3 + a] * tv[1] + R[6 + a] * tv[2] + pv[a];
                }
                ot1[j + k] = prog_t[i];
            }
            j += nparticles[i];
        }
    }
    return j;
}

/* Per-component entry points (one point, accumulating), used by the driver
 * that plugs these potentials into the reference's CPotential table. */
double orc_component_value(const orc_comp *c, const double *q) { return comp_value(c, q); }
double orc_component_density(const orc_comp *c, const double *q) { return comp_density(c, q); }
void orc_component_gradient(const orc_comp *c, const double *q, double *g) { comp_gradient(c, q, g); }
void orc_component_hessian(const orc_comp *c, const double *q, double *H) { comp_hessian(c, q, H); }

int orc_max_components(void) { return ORC_MAXC; }
size_t orc_sizeof_pot(void) { return sizeof(orc_pot); }
int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
