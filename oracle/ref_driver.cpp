/*
 * ref_driver.cpp -- drives gala's own compiled C++ (TEST INFRASTRUCTURE ONLY).
 *
 * Links, from the reference tree as it lies, potential/potential/src/cpotential.cpp,
 * potential/hamiltonian/src/chamiltonian.cpp, potential/frame/src/cframe.cpp,
 * potential/frame/builtin/builtin_frames.cpp and
 * integrate/cyintegrators/dopri/dop853.cpp (see oracle/Makefile, target "ref").
 * builtin_potentials.cpp cannot be built here (it includes the header that
 * gala's setup.py generates, setup.py:523-532), so the potential functions
 * installed in the reference CPotential table are the oracle's per-component
 * functions behind adapters with the reference function-pointer signatures
 * (potential/src/funcdefs.h:3-6).  What this pins: the reference's composite
 * logic (c_potential/c_gradient/c_hessian/c_d2_dr2 with shift/rotate) and its
 * DOP853 (dopcor + dense output + Fwrapper_T/hamiltonian_gradient_T), driven
 * the way dop853.pyx:90-250 drives it (that Cython file cannot be built here).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "potential/src/cpotential.h"
#include "frame/src/cframe.h"
#include "hamiltonian/src/chamiltonian.h"
#include "dopri/dop853.h"
#include "gala_oracle.h"

extern void static_frame_gradient(double t, double *pars, double *q, int n_dim, size_t N,
                                  double *grad, void *state);
extern double static_frame_hamiltonian(double t, double *pars, double *qp, int n_dim);

namespace {

// The reference passes only pars to a component function; the adapters find
// the component type from a compile-time tag and rebuild an orc_comp around
// the parameter pointer.
template <int TYPE>
orc_comp as_comp(const double *pars) {
    orc_comp c;
    memset(&c, 0, sizeof(c));
    c.type = TYPE;
    for (int k = 0; k < 8; k++) c.pars[k] = pars[k];
    return c;
}

template <int TYPE>
double ad_value(double t, double *pars, double *q, int n_dim, void *state) {
    orc_comp c = as_comp<TYPE>(pars);
    return orc_component_value(&c, q);
}

template <int TYPE>
double ad_density(double t, double *pars, double *q, int n_dim, void *state) {
    orc_comp c = as_comp<TYPE>(pars);
    return orc_component_density(&c, q);
}

template <int TYPE>
void ad_gradient(double t, double *pars, double *q, int n_dim, size_t N, double *grad,
                 void *state) {
    orc_comp c = as_comp<TYPE>(pars);
    for (size_t i = 0; i < N; i++) {
        double qi[3] = {q[i], q[N + i], q[2 * N + i]};
        double g[3] = {grad[i], grad[N + i], grad[2 * N + i]};
        orc_component_gradient(&c, qi, g);
        grad[i] = g[0];
        grad[N + i] = g[1];
        grad[2 * N + i] = g[2];
    }
}

template <int TYPE>
void ad_hessian(double t, double *pars, double *q, int n_dim, double *hess, void *state) {
    orc_comp c = as_comp<TYPE>(pars);
    orc_component_hessian(&c, q, hess);
}

struct RefPot {
    CPotential *cp;
    std::vector<double> store;
};

template <int TYPE>
void install(CPotential *cp, int i) {
    cp->value[i] = &ad_value<TYPE>;
    cp->density[i] = &ad_density<TYPE>;
    cp->gradient[i] = &ad_gradient<TYPE>;
    cp->hessian[i] = &ad_hessian<TYPE>;
}

// Mirrors CPotentialWrapper.init (cpotential.pyx:57-102) and the composite
// assembly (ccompositepotential.pyx:27-70): borrowed pars/q0/R per component.
RefPot *build(const orc_pot *P) {
    RefPot *rp = new RefPot;
    rp->store.resize((size_t)P->n * 20);
    CPotential *cp = allocate_cpotential(P->n);
    cp->n_dim = 3;
    cp->null = 0;
    for (int i = 0; i < P->n; i++) {
        const orc_comp &c = P->c[i];
        double *blk = &rp->store[(size_t)i * 20];
        memcpy(blk, c.pars, 8 * sizeof(double));
        memcpy(blk + 8, c.q0, 3 * sizeof(double));
        memcpy(blk + 11, c.R, 9 * sizeof(double));
        cp->parameters[i] = blk;
        cp->q0[i] = blk + 8;
        cp->R[i] = blk + 11;
        cp->n_params[i] = 8;
        cp->state[i] = NULL;
        cp->do_shift_rotate[i] = c.shift_rotate;
        switch (c.type) {
        case ORC_HERNQUIST: install<ORC_HERNQUIST>(cp, i); break;
        case ORC_PLUMMER: install<ORC_PLUMMER>(cp, i); break;
        case ORC_MIYAMOTONAGAI: install<ORC_MIYAMOTONAGAI>(cp, i); break;
        case ORC_NFW: install<ORC_NFW>(cp, i); break;
        default: install<ORC_NULL>(cp, i); break;
        }
    }
    rp->cp = cp;
    return rp;
}

void destroy(RefPot *rp) {
    free_cpotential(rp->cp);
    delete rp;
}

}  // namespace

extern "C" {

double ref_value(const orc_pot *P, double t, const double *q) {
    RefPot *rp = build(P);
    double v = c_potential(rp->cp, t, const_cast<double *>(q));
    destroy(rp);
    return v;
}

double ref_density(const orc_pot *P, double t, const double *q) {
    RefPot *rp = build(P);
    double v = c_density(rp->cp, t, const_cast<double *>(q));
    destroy(rp);
    return v;
}

void ref_gradient(const orc_pot *P, size_t N, double t, const double *q, double *grad) {
    RefPot *rp = build(P);
    c_gradient(rp->cp, N, t, const_cast<double *>(q), grad);
    destroy(rp);
}

void ref_hessian(const orc_pot *P, double t, const double *q, double *hess) {
    RefPot *rp = build(P);
    c_hessian(rp->cp, t, const_cast<double *>(q), hess);
    destroy(rp);
}

double ref_d2_dr2(const orc_pot *P, double t, const double *q) {
    RefPot *rp = build(P);
    double eps[3];
    double v = c_d2_dr2(rp->cp, t, const_cast<double *>(q), eps);
    destroy(rp);
    return v;
}

// dop853_integrate_hamiltonian (dop853.pyx:196-250) calling dop853_helper's
// dop853(...) invocation (dop853.pyx:157-182) per batch of nbatch orbits.
// w0 (6,N); out (6,ntimes,N) if save_all else (6,N); status[N].
void ref_dop853(const orc_pot *P, const double *w0, size_t N, const double *t, size_t ntimes,
                double atol, double rtol, long nmax, long nstiff, double dt_max, int save_all,
                int nbatch, double *out, int *status) {
    RefPot *rp = build(P);
    CFrameType cf;
    cf.energy = (energyfunc)static_frame_hamiltonian;
    cf.gradient = (gradientfunc)static_frame_gradient;
    cf.hessian = NULL;
    cf.n_params = 0;
    cf.parameters = NULL;
    for (size_t i0 = 0; i0 < N; i0 += nbatch) {
        size_t i1 = i0 + nbatch < N ? i0 + nbatch : N;
        unsigned nb = (unsigned)(i1 - i0), size = 6 * nb;
        std::vector<double> w(size), ow(save_all ? size * ntimes : 1);
        for (int k = 0; k < 6; k++)
            for (unsigned i = 0; i < nb; i++) w[k * nb + i] = w0[k * N + i0 + i];
        Dop853DenseState *state = save_all ? dop853_dense_state_alloc(size, size) : NULL;
        int res = dop853(size, (FcnEqDiff)Fwrapper_T, rp->cp, &cf, nb, 0, NULL, t[0], w.data(),
                         t[ntimes - 1], &rtol, &atol, 0, NULL, 0, NULL, 2.220446049250313e-16,
                         0.0, 0.0, 0.0, 0.0, dt_max, t[1] - t[0], nmax, 1, nstiff,
                         save_all ? size : 0, NULL, 0, state, const_cast<double *>(t),
                         (int)ntimes, save_all ? ow.data() : NULL);
        if (state) dop853_dense_state_free(state, size);
        for (unsigned i = 0; i < nb; i++) {
            status[i0 + i] = res;
            for (int k = 0; k < 6; k++) {
                if (save_all) {
                    for (size_t j = 0; j < ntimes; j++)
                        out[(k * ntimes + j) * N + i0 + i] = ow[j * size + k * nb + i];
                } else {
                    out[k * N + i0 + i] = w[k * nb + i];
                }
            }
        }
    }
    destroy(rp);
}

// The body potentials of an N-body system as the reference holds them: one
// CPotential per body, a null CPotential (null = 1, as NullWrapper) for a
// massless body (mockstream.pyx _setup_particle_potentials; DirectNBody's
// particle_potentials).  Bodies from nbody to n are massless.
struct RefBodies {
    std::vector<RefPot *> built;
    CPotential *nullp;
    std::vector<CPotential *> pots;
    RefBodies(const orc_pot *const *bodies, int nbody, int n) {
        nullp = allocate_cpotential(0);
        nullp->null = 1;
        nullp->n_dim = 3;
        pots.assign((size_t)(n > nbody ? n : nbody), nullp);
        for (int j = 0; j < nbody; j++) {
            if (bodies && bodies[j]) {
                RefPot *rp = build(bodies[j]);
                built.push_back(rp);
                pots[j] = rp->cp;
            }
        }
    }
    ~RefBodies() {
        for (RefPot *rp : built) destroy(rp);
        free_cpotential(nullp);
    }
};

CFrameType static_frame() {
    CFrameType cf;
    cf.energy = (energyfunc)static_frame_hamiltonian;
    cf.gradient = (gradientfunc)static_frame_gradient;
    cf.hessian = NULL;
    cf.n_params = 0;
    cf.parameters = NULL;
    return cf;
}

// mockstream_dop853's release loop (mockstream.pyx:286-303): one call of
// gala's dop853() per release through the arguments dop853_step passes
// (dop853.pyx:29-79): Fwrapper_direct_nbody, norbits = nbody + k, nbody,
// uround/safe/fac/beta/hmax 0 (defaults), h = dt0, meth 0, nstiff 1.
// body_w (nrel, nbody, 6); stream_w0/out (ntot, 6); body_final (nbody, 6);
// status[nrel].
void ref_mockstream_dop853(const orc_pot *ext, const orc_pot *const *bodies, int nbody,
                           const double *body_w, size_t nrel, const double *stream_t1,
                           const int *nstream, const double *stream_w0, double tfinal,
                           double dt0, double atol, double rtol, long nmax, double *out,
                           double *body_final, int *status) {
    RefPot *re = build(ext);
    CFrameType cf = static_frame();
    size_t off = 0;
    for (size_t i = 0; i < nrel; i++) {
        const int k = nstream[i];
        status[i] = 0;
        if (k <= 0) continue;
        const unsigned nb = (unsigned)(nbody + k), n = 6 * nb;
        // fresh body CPotentials per release: c_nbody_acceleration re-points
        // their origins at the bodies on every call
        RefBodies rb(bodies, nbody, (int)nb);
        std::vector<double> w(n);
        for (int c = 0; c < 6 * nbody; c++) w[c] = body_w[6 * nbody * i + c];
        for (int j = 0; j < k; j++)
            for (int c = 0; c < 6; c++) w[6 * (nbody + j) + c] = stream_w0[6 * (off + j) + c];
        double rt = rtol, at = atol;
        int res = dop853(n, (FcnEqDiff)Fwrapper_direct_nbody, re->cp, &cf, nb, (unsigned)nbody,
                         (void *)rb.pots.data(), stream_t1[i], w.data(), tfinal, &rt, &at, 0,
                         NULL, 0, NULL, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, dt0, nmax, 0, 1, 0, NULL, 0,
                         NULL, NULL, 0, NULL);
        status[i] = res;
        for (int j = 0; j < k; j++)
            for (int c = 0; c < 6; c++) out[6 * (off + j) + c] = w[6 * (nbody + j) + c];
        for (int c = 0; c < 6 * nbody; c++) body_final[c] = w[c];
        off += (size_t)k;
    }
    destroy(re);
}

// direct_nbody_dop853 (nbody.pyx:30-115) through dop853_helper's dop853()
// call (dop853.pyx:157-182): the n bodies (n, 6) as one system with
// Fwrapper_direct_nbody, h = t[1] - t[0], uround eps, hmax dt_max, nstiff -1,
// dense output at t.  out (ntimes, n, 6) when save_all, else (n, 6).
// Returns the idid.
int ref_nbody_dop853(const orc_pot *ext, const orc_pot *const *bodies, int nbody,
                     const double *w0, int n, const double *t, size_t ntimes, double atol,
                     double rtol, long nmax, double dt_max, int save_all, double *out) {
    RefPot *re = build(ext);
    CFrameType cf = static_frame();
    RefBodies rb(bodies, nbody, n);
    const unsigned size = 6 * (unsigned)n;
    std::vector<double> w(w0, w0 + size);
    Dop853DenseState *state = save_all ? dop853_dense_state_alloc(size, size) : NULL;
    int res = dop853(size, (FcnEqDiff)Fwrapper_direct_nbody, re->cp, &cf, (unsigned)n,
                     (unsigned)nbody, (void *)rb.pots.data(), t[0], w.data(), t[ntimes - 1],
                     &rtol, &atol, 0, NULL, 0, NULL, 2.220446049250313e-16, 0.0, 0.0, 0.0, 0.0,
                     dt_max, t[1] - t[0], nmax, 1, -1, save_all ? size : 0, NULL, 0, state,
                     const_cast<double *>(t), (int)ntimes, save_all ? out : NULL);
    if (state) dop853_dense_state_free(state, size);
    if (!save_all) memcpy(out, w.data(), sizeof(double) * size);
    destroy(re);
    return res;
}

// leapfrog_integrate_nbody (leapfrog.pyx:161-257): the Cython loops
// (c_init_velocity_nbody :126-137, c_leapfrog_step_nbody :140-158) restated
// around gala's compiled c_gradient and c_nbody_gradient_symplectic
// (cpotential.cpp:418-442).  out (ntimes, n, 6) when save_all, else (n, 6).
void ref_nbody_leapfrog(const orc_pot *ext, const orc_pot *const *bodies, int nbody,
                        const double *w0, int n, const double *t, size_t ntimes, int save_all,
                        double *out) {
    RefPot *re = build(ext);
    RefBodies rb(bodies, nbody, n);
    const double dt = t[1] - t[0];
    std::vector<double> w(w0, w0 + 6 * n), vh(3 * (size_t)n);
    if (save_all) memcpy(out, w0, sizeof(double) * 6 * n);
    double grad[3];
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) grad[k] = 0.0;
        c_gradient(re->cp, 1, t[0], &w[6 * i], grad);
        c_nbody_gradient_symplectic(rb.pots.data(), t[0], &w[6 * i], w.data(), nbody, i, 3,
                                    grad);
        for (int k = 0; k < 3; k++) vh[3 * i + k] = w[6 * i + 3 + k] - grad[k] * dt / 2.;
    }
    for (size_t j = 1; j < ntimes; j++) {
        for (int i = 0; i < n; i++) {
            for (int k = 0; k < 3; k++) grad[k] = 0.0;
            for (int k = 0; k < 3; k++) w[6 * i + k] = w[6 * i + k] + vh[3 * i + k] * dt;
            c_gradient(re->cp, 1, t[j], &w[6 * i], grad);
            c_nbody_gradient_symplectic(rb.pots.data(), t[j], &w[6 * i], w.data(), nbody, i, 3,
                                        grad);
            for (int k = 0; k < 3; k++) {
                w[6 * i + 3 + k] = vh[3 * i + k] - grad[k] * dt / 2.;
                vh[3 * i + k] = vh[3 * i + k] - grad[k] * dt;
            }
            if (save_all)
                for (int k = 0; k < 6; k++) out[(j * n + i) * 6 + k] = w[6 * i + k];
        }
    }
    if (!save_all) memcpy(out, w.data(), sizeof(double) * 6 * n);
    destroy(re);
}

// ruth4_integrate_nbody (ruth4.pyx:138-261): the Cython loops
// (c_ruth4_step_nbody, :119-136) restated around gala's compiled c_gradient
// and c_nbody_gradient_symplectic.  out (ntimes, n, 6) when save_all, else
// (n, 6).
void ref_nbody_ruth4(const orc_pot *ext, const orc_pot *const *bodies, int nbody,
                     const double *w0, int n, const double *t, size_t ntimes, int save_all,
                     double *out) {
    RefPot *re = build(ext);
    RefBodies rb(bodies, nbody, n);
    const double dt = t[1] - t[0];
    const double two_13 = pow(2., 1. / 3.);
    const double cs[4] = {1. / (2. * (2. - two_13)), (1. - two_13) / (2. * (2. - two_13)),
                          (1. - two_13) / (2. * (2. - two_13)), 1. / (2. * (2. - two_13))};
    const double ds[4] = {0., 1. / (2. - two_13), -two_13 / (2. - two_13), 1. / (2. - two_13)};
    std::vector<double> w(w0, w0 + 6 * n);
    if (save_all) memcpy(out, w0, sizeof(double) * 6 * n);
    double grad[3];
    for (size_t j = 1; j < ntimes; j++) {
        for (int i = 0; i < n; i++) {
            double *wi = &w[6 * i];
            for (int s = 0; s < 4; s++) {
                for (int k = 0; k < 3; k++) grad[k] = 0.;
                c_gradient(re->cp, 1, t[j], wi, grad);
                c_nbody_gradient_symplectic(rb.pots.data(), t[j], wi, w.data(), nbody, i, 3,
                                            grad);
                for (int k = 0; k < 3; k++) {
                    wi[3 + k] = wi[3 + k] - ds[s] * grad[k] * dt;
                    wi[k] = wi[k] + cs[s] * wi[3 + k] * dt;
                }
            }
            if (save_all)
                for (int k = 0; k < 6; k++) out[(j * n + i) * 6 + k] = w[6 * i + k];
        }
    }
    if (!save_all) memcpy(out, w.data(), sizeof(double) * 6 * n);
    destroy(re);
}

}  // extern "C"
