// gamd_potential.hip -- potential handles and batched evaluation kernels.
//
// C ABI: gamd_potential_create/destroy/value/density/gradient/hessian/
// d2_dr2, gamd_hamiltonian_energy (declared in include/gala_amd.h).
#include <math.h>
#include <string.h>

#include "gamd_host.hpp"

using namespace gamd;

namespace {

// Fold parameter-only products into d[] (the same IEEE products gala forms
// per call in builtin_potentials.cpp).
void fold_component(DevComp &c) {
    const double *p = c.p;
    for (int k = 0; k < 4; k++) c.d[k] = 0.0;
    switch (c.type) {
    case GAMD_HERNQUIST:  // fac = G m / ((r+c)^2 r)
        c.d[0] = p[0] * p[1];
        c.d[1] = p[2];
        break;
    case GAMD_PLUMMER:  // R2b = r^2 + b^2
        c.d[0] = p[0] * p[1];
        c.d[1] = p[2] * p[2];
        break;
    case GAMD_MIYAMOTONAGAI:  // sqrt(z^2 + b^2), a
        c.d[0] = p[0] * p[1];
        c.d[1] = p[2];
        c.d[2] = p[3] * p[3];
        break;
    case GAMD_NFW:  // v_h2 = G m / r_s; u = r / r_s; / r_s^2
        c.d[0] = p[0] * p[1] / p[2];
        c.d[1] = p[2];
        c.d[2] = p[2] * p[2];
        c.d[3] = 1.0 / p[2];
        break;
    default:
        break;
    }
}

template <bool SCF>
__global__ void k_value(const DevPot *__restrict__ Pp, const double *__restrict__ q, size_t N, double *__restrict__ out) {
    const DevPot &P = *Pp;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = pot_value<SCF>(P, q[i], q[N + i], q[2 * N + i]);
}

template <bool SCF>
__global__ void k_gradient(const DevPot *__restrict__ Pp, const double *__restrict__ q, size_t N,
                           double *__restrict__ g) {
    const DevPot &P = *Pp;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
         i += (size_t)gridDim.x * blockDim.x) {
        double gx, gy, gz;
        pot_grad<SCF>(P, q[i], q[N + i], q[2 * N + i], gx, gy, gz);
        g[i] = gx;
        g[N + i] = gy;
        g[2 * N + i] = gz;
    }
}

template <bool SCF>
__global__ void k_d2_dr2(const DevPot *__restrict__ Pp, const double *__restrict__ q, size_t N, double *__restrict__ out) {
    const DevPot &P = *Pp;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = pot_d2_dr2<SCF>(P, q[i], q[N + i], q[2 * N + i]);
}

template <bool SCF>
__global__ void k_energy(const DevPot *__restrict__ Pp, const double *__restrict__ w, size_t N, double *__restrict__ out) {
    const DevPot &P = *Pp;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
         i += (size_t)gridDim.x * blockDim.x) {
        const double px = w[3 * N + i], py = w[4 * N + i], pz = w[5 * N + i];
        // static_frame_hamiltonian (builtin_frames.cpp:8-16) then the
        // potential terms (chamiltonian.cpp:11-16)
        const double E = 0.5 * (px * px + py * py + pz * pz);
        out[i] = E + pot_value<SCF>(P, w[i], w[N + i], w[2 * N + i]);
    }
}

__global__ void k_density(const DevPot *__restrict__ Pp, const double *__restrict__ q, size_t N, double *__restrict__ out) {
    const DevPot &P = *Pp;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
         i += (size_t)gridDim.x * blockDim.x) {
        const double x = q[i], y = q[N + i], z = q[2 * N + i];
        double v = 0.0;
        for (int k = 0; k < P.n; k++) {
            const DevComp &c = P.c[k];
            double xt = x, yt = y, zt = z;
            if (c.shift_rotate) shift_rotate(c, x, y, z, xt, yt, zt);
            v = v + (c.type == GAMD_SCF ? scf_density(c, xt, yt, zt)
                                        : comp_density_analytic(c, xt, yt, zt));
        }
        out[i] = v;
    }
}

__global__ void k_hessian(const DevPot *__restrict__ Pp, const double *__restrict__ q, size_t N, double *__restrict__ h) {
    const DevPot &P = *Pp;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
         i += (size_t)gridDim.x * blockDim.x) {
        const double x = q[i], y = q[N + i], z = q[2 * N + i];
        double H[9];
#pragma unroll
        for (int k = 0; k < 9; k++) H[k] = 0.0;
        for (int k = 0; k < P.n; k++) {
            const DevComp &c = P.c[k];
            double xt = x, yt = y, zt = z;
            if (c.shift_rotate) shift_rotate(c, x, y, z, xt, yt, zt);
            comp_hessian_analytic(c, xt, yt, zt, H);
        }
#pragma unroll
        for (int k = 0; k < 9; k++) h[k * N + i] = H[k];
    }
}

constexpr unsigned EVAL_BLOCK = 256;

int check_eval(const gamd_potential *pot, const double *q, size_t N, const double *out) {
    if (!pot || (N > 0 && (!q || !out))) return GAMD_ERR_INVALID;
    return GAMD_OK;
}

}  // namespace

extern "C" {

const char *gamd_version(void) { return "gala_amd 0.1 (gfx950)"; }

int gamd_potential_create(const gamd_component *components, int ncomponents,
                          gamd_potential **out) {
    if (!out || ncomponents < 0 || ncomponents > GAMD_MAX_COMPONENTS ||
        (ncomponents > 0 && !components))
        return GAMD_ERR_INVALID;
    gamd_potential *pot = new gamd_potential();
    memset(&pot->dev, 0, sizeof(pot->dev));
    if (hipGetDevice(&pot->device) != hipSuccess) {
        delete pot;
        return GAMD_ERR_HIP;
    }
    pot->dev.n = ncomponents;
    pot->dev.all_unshifted = 1;
    for (int i = 0; i < ncomponents; i++) {
        const gamd_component &src = components[i];
        DevComp &c = pot->dev.c[i];
        if (src.type < GAMD_NULL || src.type > GAMD_SCF) {
            gamd_potential_destroy(pot);
            return GAMD_ERR_INVALID;
        }
        c.type = src.type;
        c.shift_rotate = src.shift_rotate ? 1 : 0;
        memcpy(c.p, src.pars, sizeof(c.p));
        memcpy(c.q0, src.q0, sizeof(c.q0));
        memcpy(c.R, src.R, sizeof(c.R));
        fold_component(c);
        if (c.shift_rotate) pot->dev.all_unshifted = 0;
        if (src.type == GAMD_SCF) {
            if (!src.S || !src.T || src.nmax < 0 || src.lmax < 0 ||
                src.nmax > GAMD_SCF_MAX_NMAX || src.lmax > GAMD_SCF_MAX_LMAX) {
                gamd_potential_destroy(pot);
                return GAMD_ERR_INVALID;
            }
            c.nmax = src.nmax;
            c.lmax = src.lmax;
            const int nmax = src.nmax, lmax = src.lmax;
            const size_t nlm = (size_t)(lmax + 1) * (lmax + 2) / 2;
            std::vector<double> packed(nlm * (nmax + 1) * 2);
            // gala's index i = m + (lmax+1)(l + (lmax+1) n) (bfe.cpp:160) ->
            // per l, per n, m = 0..l, (S, T) interleaved (scf_pack_index)
            for (int l = 0; l <= lmax; l++)
                for (int m = 0; m <= l; m++)
                    for (int n = 0; n <= nmax; n++) {
                        const size_t gi = (size_t)m + (lmax + 1) * (l + (size_t)(lmax + 1) * n);
                        const size_t pi = (size_t)scf_pack_index(l, m, n, nmax);
                        packed[2 * pi] = src.S[gi];
                        packed[2 * pi + 1] = src.T[gi];
                    }
            void *d = nullptr;
            if (hipMalloc(&d, packed.size() * sizeof(double)) != hipSuccess) {
                gamd_potential_destroy(pot);
                return GAMD_ERR_NOMEM;
            }
            pot->buffers.push_back(d);
            if (hipMemcpy(d, packed.data(), packed.size() * sizeof(double),
                          hipMemcpyHostToDevice) != hipSuccess) {
                gamd_potential_destroy(pot);
                return GAMD_ERR_HIP;
            }
            c.coef = static_cast<const double *>(d);
            pot->dev.has_scf = 1;
        }
    }
    pot->dev.layout = LAYOUT_GENERIC;
    if (ncomponents == 4 && pot->dev.all_unshifted && pot->dev.c[0].type == GAMD_HERNQUIST &&
        pot->dev.c[1].type == GAMD_HERNQUIST && pot->dev.c[2].type == GAMD_MIYAMOTONAGAI &&
        pot->dev.c[3].type == GAMD_NFW)
        pot->dev.layout = LAYOUT_MW;
    if (ncomponents == 1 && pot->dev.all_unshifted && pot->dev.c[0].type == GAMD_SCF &&
        pot->dev.c[0].nmax == 10 && pot->dev.c[0].lmax == 6)
        pot->dev.layout = LAYOUT_SCF106;
    // Kernels read the table through a restrict-qualified pointer: with a
    // wave-uniform component index these become scalar loads, where a by-value
    // kernel argument indexed by a runtime component number would be copied to
    // scratch memory by every lane.
    void *dtab = nullptr;
    if (hipMalloc(&dtab, sizeof(DevPot)) != hipSuccess) {
        gamd_potential_destroy(pot);
        return GAMD_ERR_NOMEM;
    }
    pot->buffers.push_back(dtab);
    if (hipMemcpy(dtab, &pot->dev, sizeof(DevPot), hipMemcpyHostToDevice) != hipSuccess) {
        gamd_potential_destroy(pot);
        return GAMD_ERR_HIP;
    }
    pot->ddev = static_cast<const DevPot *>(dtab);
    *out = pot;
    return GAMD_OK;
}

int gamd_potential_destroy(gamd_potential *pot) {
    if (!pot) return GAMD_OK;
    int rc = GAMD_OK;
    for (void *b : pot->buffers)
        if (hipFree(b) != hipSuccess) rc = GAMD_ERR_HIP;
    delete pot;
    return rc;
}

int gamd_potential_value(const gamd_potential *pot, double t, const double *q, size_t N,
                         double *out, void *stream) {
    int rc = check_eval(pot, q, N, out);
    if (rc || N == 0) return rc;
    const unsigned g = grid_for(N, EVAL_BLOCK);
    if (pot->dev.has_scf)
        hipLaunchKernelGGL(k_value<true>, dim3(g), dim3(EVAL_BLOCK), 0, as_stream(stream), pot->ddev, q, N, out);
    else
        hipLaunchKernelGGL(k_value<false>, dim3(g), dim3(EVAL_BLOCK), 0, as_stream(stream), pot->ddev, q, N, out);
    return launch_status();
}

int gamd_potential_gradient(const gamd_potential *pot, double t, const double *q, size_t N,
                            double *grad, void *stream) {
    int rc = check_eval(pot, q, N, grad);
    if (rc || N == 0) return rc;
    const unsigned g = grid_for(N, EVAL_BLOCK);
    if (pot->dev.has_scf)
        hipLaunchKernelGGL(k_gradient<true>, dim3(g), dim3(EVAL_BLOCK), 0, as_stream(stream), pot->ddev, q, N, grad);
    else
        hipLaunchKernelGGL(k_gradient<false>, dim3(g), dim3(EVAL_BLOCK), 0, as_stream(stream), pot->ddev, q, N, grad);
    return launch_status();
}

int gamd_potential_density(const gamd_potential *pot, double t, const double *q, size_t N,
                           double *out, void *stream) {
    int rc = check_eval(pot, q, N, out);
    if (rc || N == 0) return rc;
    hipLaunchKernelGGL(k_density, dim3(grid_for(N, EVAL_BLOCK)), dim3(EVAL_BLOCK), 0,
                       as_stream(stream), pot->ddev, q, N, out);
    return launch_status();
}

int gamd_potential_hessian(const gamd_potential *pot, double t, const double *q, size_t N,
                           double *hess, void *stream) {
    int rc = check_eval(pot, q, N, hess);
    if (rc || N == 0) return rc;
    hipLaunchKernelGGL(k_hessian, dim3(grid_for(N, EVAL_BLOCK)), dim3(EVAL_BLOCK), 0,
                       as_stream(stream), pot->ddev, q, N, hess);
    return launch_status();
}

int gamd_potential_d2_dr2(const gamd_potential *pot, double t, const double *q, size_t N,
                          double *out, void *stream) {
    int rc = check_eval(pot, q, N, out);
    if (rc || N == 0) return rc;
    const unsigned g = grid_for(N, EVAL_BLOCK);
    if (pot->dev.has_scf)
        hipLaunchKernelGGL(k_d2_dr2<true>, dim3(g), dim3(EVAL_BLOCK), 0, as_stream(stream), pot->ddev, q, N, out);
    else
        hipLaunchKernelGGL(k_d2_dr2<false>, dim3(g), dim3(EVAL_BLOCK), 0, as_stream(stream), pot->ddev, q, N, out);
    return launch_status();
}

int gamd_hamiltonian_energy(const gamd_potential *pot, double t, const double *w, size_t N,
                            double *out, void *stream) {
    int rc = check_eval(pot, w, N, out);
    if (rc || N == 0) return rc;
    const unsigned g = grid_for(N, EVAL_BLOCK);
    if (pot->dev.has_scf)
        hipLaunchKernelGGL(k_energy<true>, dim3(g), dim3(EVAL_BLOCK), 0, as_stream(stream), pot->ddev, w, N, out);
    else
        hipLaunchKernelGGL(k_energy<false>, dim3(g), dim3(EVAL_BLOCK), 0, as_stream(stream), pot->ddev, w, N, out);
    return launch_status();
}

}  // extern "C"
