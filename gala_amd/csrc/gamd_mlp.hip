// gamd_mlp.hip -- the learned SCF surrogate: a 4 -> 256 x 4 (SiLU) -> 4 MLP
// in BF16 on the matrix cores (FP32 accumulation, FP32 master weights), its
// training kernels (forward, backward, Adam) and the surrogate leapfrog with
// a per-orbit energy monitor and exact-SCF fallback.  No counterpart in gala:
// the exact potential it stands in for is the SCF expansion of
// potential/scf/src/bfe.cpp:20-259.
//
// Network, per point q (FP64 in, FP64 out):
//   features  f = (x/r, y/r, z/r, (ln r - mu)/sigma): direction and log
//             radius, so small radii are resolved as finely as large ones
//   hidden    h1 = silu(f W1 + b1), h_l = silu(h_{l-1} W_l + b_l), l = 2..4
//   output    o = h4 W5 + b5   (4 values)
//   physics   Phi_net(q) = Phi_b(r) (1 + c o0), with Phi_b = -G M / (r + r_s)
//             the expansion's Hernquist monopole (smooth, so Phi_net is);
//             the orbit force is -grad Phi_net, the exact gradient of this
//             potential (reverse mode through the network), so a surrogate
//             orbit conserves Phi_net + |v|^2/2 up to the integrator's error
//             and the energy monitor measures a conservative system.
//   o1..o3    auxiliary training targets, s_grad(r) o[1:4] ~ grad Phi (the
//             exact force scaled by a radial profile, piecewise linear in
//             ln r): they shape the hidden features during training; the
//             orbit force does not use them.
//
// Two evaluations of the same network:
//   * training (tile_forward, k_mlp_forward_train): BF16 activations and
//     weights, 128-point tiles;
//   * orbits and inference (tile_conservative): the hidden products in
//     three BF16 matrix-core products, hi*hi + hi*lo + lo*hi of the split
//     v = hi + lo (hi = bf16(v), lo = bf16(v - hi)), which carries ~16
//     significant bits: the BF16 network's rounding jitter (~1e-3 of o0)
//     would drift the monitored energy by more than 1e-5 within steps.
//
// Matrix-core mapping (gfx950, v_mfma_f32_16x16x32_bf16; operand layout
// A[i][k] in lane i + 16 (k/8), element k%8; B[k][j] in lane j + 16 (k/8);
// D[i][j] in lane j + 16 (i/4), element i%4):
//   * training: a workgroup (4 waves) owns a tile of 128 points; a layer's
//     activations live in LDS (128 x 256 BF16, rows padded to 264 to keep
//     the 16-byte fragment reads conflict-free); wave w computes output
//     units [64w, 64w+64) of all 128 rows: 8 x 4 MFMA tiles, 128 FP32
//     accumulators per lane;
//   * hidden-layer weights are pre-swizzled into fragment order (and, for
//     the reverse pass, their transposes), so each wave reads 1 KB
//     contiguous per (16-unit tile, K step) from L2.
#include <math.h>
#include <string.h>

#include "gamd_host.hpp"

using namespace gamd;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int H = 256;        // hidden width
constexpr int TILE = 128;     // points per workgroup
constexpr int LDSR = 264;     // padded LDS row (BF16 elements)
// flat FP32 parameter layout
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + 4 * H;
constexpr int OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H * H;
constexpr int OFF_W4 = OFF_B3 + H, OFF_B4 = OFF_W4 + H * H;
constexpr int OFF_W5 = OFF_B4 + H, OFF_B5 = OFF_W5 + H * 4;
constexpr int NPARAM = OFF_B5 + 4;  // 199,684
// BF16 copies of W2..W4, each block 3 H^2 (one H^2 per layer): fragment
// order (hi), row-major (hi, for the training backward), fragment order of
// the low parts, and fragment order of the transposes (hi, then lo) for the
// reverse pass of tile_conservative
constexpr int BF_SW = 0, BF_ROW = 3 * H * H, BF_SWLO = 6 * H * H, BF_TSW = 9 * H * H,
              BF_TSWLO = 12 * H * H;
constexpr int NBF = 15 * H * H;
constexpr int CP = 32;  // points per workgroup of tile_conservative

// logistic function from the hardware exp2 and reciprocal (v_exp_f32,
// v_rcp_f32, ~1 ulp each): the activations are rounded to BF16 (8 bits)
// right after, so an IEEE divide's extra accuracy is never seen.  Large
// negative z: exp2 overflows to +inf and the reciprocal gives 0, the limit.
__device__ __forceinline__ float sigmoid_fast(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid_fast(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigmoid_fast(z);
    return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ __bf16 tobf(float f) { return (__bf16)f; }
__device__ __forceinline__ float frombf(__bf16 b) { return (float)b; }

// ---------------------------------------------------------------------------
// features / physics scaling

__device__ __forceinline__ void features(const gamd_mlp_norm &nm, double x, double y, double z,
                                         float *f, double &r) {
    r = sqrt(x * x + y * y + z * z);
    const double ir = 1.0 / r;
    f[0] = (float)(x * ir);
    f[1] = (float)(y * ir);
    f[2] = (float)(z * ir);
    f[3] = (float)((log(r) - nm.mu) * nm.inv_sigma);
}

// the auxiliary heads' scale at radius r: linear interpolation in ln r
// between the profile knots, constant beyond the end knots (numpy.interp's
// convention)
__device__ __forceinline__ double out_scale_grad(const gamd_mlp_norm &nm, double r) {
    double u = (log(r) - nm.lnr0) * nm.inv_dlnr;
    if (!(u > 0.0)) u = 0.0;  // also maps NaN (r = 0 with a zero knot spacing) to the first knot
    if (u > (double)(GAMD_MLP_KNOTS - 1)) u = (double)(GAMD_MLP_KNOTS - 1);
    int k = (int)u;
    k = k > GAMD_MLP_KNOTS - 2 ? GAMD_MLP_KNOTS - 2 : k;
    const double f = u - (double)k;
    const double *sg = nm.scale;
    return sg[k] + f * (sg[k + 1] - sg[k]);
}

// Phi_net and its gradient at q = (x, y, z), |q| = r, from the network's o0
// and its derivatives g[c] = d o0 / d f_c (f the features): with
// f = (q/r, (ln r - mu)/sigma), d(q_i/r)/dq_k = (delta_ik - q_i q_k / r^2) / r
// and d f_3/dq_k = q_k / (r^2 sigma).
__device__ __forceinline__ void net_physics(const gamd_mlp_norm &nm, double x, double y, double z,
                                            double r, float o0, const float *g, double &phi,
                                            double &gx, double &gy, double &gz) {
    const double rb = r + nm.rs;
    const double pb = -nm.gm / rb;            // Phi_b
    const double dpb = nm.gm / (rb * rb);     // d Phi_b / dr
    const double ratio = 1.0 + nm.phi_c * (double)o0;
    phi = pb * ratio;
    const double ir = 1.0 / r;
    const double g0 = (double)g[0], g1 = (double)g[1], g2 = (double)g[2], g3 = (double)g[3];
    const double gq = (g0 * x + g1 * y + g2 * z) * ir * ir;
    const double gr = g3 * nm.inv_sigma * ir * ir;
    // d o0 / dq
    const double ox = (g0 - gq * x) * ir + gr * x;
    const double oy = (g1 - gq * y) * ir + gr * y;
    const double oz = (g2 - gq * z) * ir + gr * z;
    const double a = dpb * ratio * ir, b = pb * nm.phi_c;
    gx = a * x + b * ox;
    gy = a * y + b * oy;
    gz = a * z + b * oz;
}

// ---------------------------------------------------------------------------
// fused forward of one 128-point tile.  feat: LDS [TILE][4]; on return the
// raw outputs are in LDS out[TILE][4].  With TRAIN, pre-activations and
// activations of the four hidden layers go to global (BF16, [4][M][256]).
template <bool TRAIN>
__device__ __forceinline__ void tile_forward(const float *__restrict__ theta, const __bf16 *__restrict__ wsw,
                             const float (*feat)[4], __bf16 *buf, float (*out)[4],
                             int64_t row0, int64_t M, __bf16 *__restrict__ Zs,
                             __bf16 *__restrict__ Hs) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    // layer 1: 4 -> 256 in FP32; thread t owns unit j = t
    {
        const int j = t;
        const float w0 = theta[OFF_W1 + 0 * H + j], w1 = theta[OFF_W1 + 1 * H + j];
        const float w2 = theta[OFF_W1 + 2 * H + j], w3 = theta[OFF_W1 + 3 * H + j];
        const float bj = theta[OFF_B1 + j];
#pragma unroll 4
        for (int p = 0; p < TILE; p++) {
            const float z = bj + feat[p][0] * w0 + feat[p][1] * w1 + feat[p][2] * w2 + feat[p][3] * w3;
            const float h = silu(z);
            buf[p * LDSR + j] = tobf(h);
            if (TRAIN && row0 + p < M) {
                Zs[(0 * M + row0 + p) * H + j] = tobf(z);
                Hs[(0 * M + row0 + p) * H + j] = tobf(h);
            }
        }
    }
    __syncthreads();
    // layers 2..4 on the matrix cores
    // one activation buffer: every wave holds its 64 units of the whole
    // layer in accumulators, so the layer reads the buffer, waits for every
    // wave, then overwrites it (68 KB of LDS: two workgroups per CU, whose
    // matrix-core and SiLU phases overlap)
#pragma unroll 1
    for (int l = 0; l < 3; l++) {
        const float *bias = theta + (l == 0 ? OFF_B2 : l == 1 ? OFF_B3 : OFF_B4);
        const __bf16 *W = wsw + (size_t)l * H * H;
        // the wave's 64 units in one pass (128 FP32 accumulators per lane):
        // each activation fragment is read from LDS once per layer
        {
            constexpr int NT = 4;  // 16-unit tiles per wave
            f32x4 acc[8][NT];
#pragma unroll
            for (int mi = 0; mi < 8; mi++)
#pragma unroll
                for (int ni = 0; ni < NT; ni++) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
            for (int ks = 0; ks < 8; ks++) {
                bf16x8 b[NT];
#pragma unroll
                for (int ni = 0; ni < NT; ni++)
                    b[ni] = *reinterpret_cast<const bf16x8 *>(
                        W + (((4 * w + ni) * 8 + ks) * 64 + lane) * 8);
#pragma unroll
                for (int mi = 0; mi < 8; mi++) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(
                        buf + (16 * mi + (lane & 15)) * LDSR + 32 * ks + 8 * (lane >> 4));
                    // the transposed product, Z^T = W^T H^T: the weight fragment
                    // is the A operand and the activations the B operand, so a
                    // lane's four accumulators are four consecutive units of
                    // one point (one 8-byte store below instead of four 2-byte)
#pragma unroll
                    for (int ni = 0; ni < NT; ni++)
                        acc[mi][ni] =
                            __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[ni], a, acc[mi][ni], 0, 0, 0);
                }
            }
            __syncthreads();  // every wave has read the layer's input
#pragma unroll
            for (int mi = 0; mi < 8; mi++)
#pragma unroll
                for (int ni = 0; ni < NT; ni++) {
                    const int unit = 64 * w + 16 * ni + 4 * (lane >> 4);
                    const int row = 16 * mi + (lane & 15);
                    const f32x4 bc = *reinterpret_cast<const f32x4 *>(bias + unit);
                    __bf16 hv[4], zv[4];
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const float z = acc[mi][ni][e] + bc[e];
                        hv[e] = tobf(silu(z));
                        zv[e] = tobf(z);
                    }
                    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
                    *reinterpret_cast<bf16x4 *>(buf + row * LDSR + unit) =
                        bf16x4{hv[0], hv[1], hv[2], hv[3]};
                    if (TRAIN && row0 + row < M) {
                        *reinterpret_cast<bf16x4 *>(Zs + ((size_t)(l + 1) * M + row0 + row) * H + unit) =
                            bf16x4{zv[0], zv[1], zv[2], zv[3]};
                        *reinterpret_cast<bf16x4 *>(Hs + ((size_t)(l + 1) * M + row0 + row) * H + unit) =
                            bf16x4{hv[0], hv[1], hv[2], hv[3]};
                    }
                }
        }
        __syncthreads();
    }
    // layer 5: 256 -> 4 in FP32 from the BF16 activations; thread t: point
    // t/2, outputs 2(t%2), 2(t%2)+1
    {
        const int p = t >> 1, o0 = 2 * (t & 1);
        float s0 = theta[OFF_B5 + o0], s1 = theta[OFF_B5 + o0 + 1];
        const __bf16 *hrow = buf + p * LDSR;
#pragma unroll 8
        for (int j = 0; j < H; j++) {
            const float hj = frombf(hrow[j]);
            s0 += hj * theta[OFF_W5 + j * 4 + o0];
            s1 += hj * theta[OFF_W5 + j * 4 + o0 + 1];
        }
        out[p][o0] = s0;
        out[p][o0 + 1] = s1;
    }
    __syncthreads();
}

constexpr int LDS_BYTES = TILE * LDSR * 2;

// ---------------------------------------------------------------------------
// The precise evaluation of one CP-point tile: o0 and d o0 / d f (reverse
// mode), the hidden products in three BF16 matrix-core products of the
// hi/lo split (hi*hi + hi*lo + lo*hi), SiLU and its derivative in FP32, the
// first and last layers in FP32 on the VALU.
//   feat   LDS [CP][4] features;
//   bhi,   LDS [CP][LDSR] BF16 each: a layer's input (activations forward,
//   blo    error signals backward), split;
//   red    LDS [4][CP][5] per-wave partial sums;
//   o0,    LDS [CP] and [CP][4]: results.
//   dfo
// Wave w owns units [64w, 64w + 64) of every layer (forward and backward),
// so the SiLU derivatives it keeps in registers from the forward pass are
// exactly those its reverse-pass products need.  All 256 threads call it.

__device__ __forceinline__ void split_bf(float v, __bf16 &hi, __bf16 &lo) {
    hi = tobf(v);
    lo = tobf(v - frombf(hi));
}

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// acc[mi][ni] (+)= sum over the 256 inputs of the layer's weights (fragment
// order: hi at Whi, lo at Wlo) times the split input in bhi/blo
__device__ __forceinline__ void split_layer(const __bf16 *__restrict__ Whi,
                                            const __bf16 *__restrict__ Wlo, const __bf16 *bhi,
                                            const __bf16 *blo, f32x4 (&acc)[2][4]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int ni = 0; ni < 4; ni++) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int ks = 0; ks < 8; ks++) {
        bf16x8 wh[4], wl[4];
#pragma unroll
        for (int ni = 0; ni < 4; ni++) {
            const size_t o = (((size_t)(4 * w + ni) * 8 + ks) * 64 + lane) * 8;
            wh[ni] = *reinterpret_cast<const bf16x8 *>(Whi + o);
            wl[ni] = *reinterpret_cast<const bf16x8 *>(Wlo + o);
        }
#pragma unroll
        for (int mi = 0; mi < 2; mi++) {
            const int off = (16 * mi + (lane & 15)) * LDSR + 32 * ks + 8 * (lane >> 4);
            const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(bhi + off);
            const bf16x8 al = *reinterpret_cast<const bf16x8 *>(blo + off);
#pragma unroll
            for (int ni = 0; ni < 4; ni++) {
                acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[ni], ah, acc[mi][ni], 0, 0, 0);
                acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh[ni], al, acc[mi][ni], 0, 0, 0);
                acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl[ni], ah, acc[mi][ni], 0, 0, 0);
            }
        }
    }
}

__device__ __forceinline__ void tile_conservative(const float *__restrict__ theta,
                                                  const __bf16 *__restrict__ wbf,
                                                  const float (*feat)[4], __bf16 *bhi, __bf16 *blo,
                                                  float (*red)[CP][5], float *o0, float (*dfo)[4]) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    // layer 1 (FP32): thread t owns unit t
    {
        const int j = t;
        const float w0 = theta[OFF_W1 + 0 * H + j], w1 = theta[OFF_W1 + 1 * H + j];
        const float w2 = theta[OFF_W1 + 2 * H + j], w3 = theta[OFF_W1 + 3 * H + j];
        const float bj = theta[OFF_B1 + j];
#pragma unroll 4
        for (int p = 0; p < CP; p++) {
            const float z = bj + feat[p][0] * w0 + feat[p][1] * w1 + feat[p][2] * w2 + feat[p][3] * w3;
            split_bf(silu(z), bhi[p * LDSR + j], blo[p * LDSR + j]);
        }
    }
    __syncthreads();
    // SiLU derivatives of layers 2 and 3 at this wave's units, in the
    // accumulator layout: unit 64w + 16ni + 4(lane/16) + e, point 16mi + lane%16
    f32x4 ds[2][2][4];
    float po[2] = {0.f, 0.f};  // partial o0 of points lane%16 and 16 + lane%16
    f32x4 acc[2][4];
#pragma unroll 1
    for (int l = 0; l < 3; l++) {
        split_layer(wbf + BF_SW + (size_t)l * H * H, wbf + BF_SWLO + (size_t)l * H * H, bhi, blo,
                    acc);
        __syncthreads();  // every wave has read the layer's input
        const float *bias = theta + (l == 0 ? OFF_B2 : l == 1 ? OFF_B3 : OFF_B4);
#pragma unroll
        for (int mi = 0; mi < 2; mi++)
#pragma unroll
            for (int ni = 0; ni < 4; ni++) {
                const int unit = 64 * w + 16 * ni + 4 * (lane >> 4);
                const int row = 16 * mi + (lane & 15);
                const f32x4 bc = *reinterpret_cast<const f32x4 *>(bias + unit);
                __bf16 hh[4], hl[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const float z = acc[mi][ni][e] + bc[e];
                    const float sg = sigmoid_fast(z);
                    const float h = z * sg;
                    const float d = sg * (1.0f + z * (1.0f - sg));
                    if (l < 2) {
                        ds[l][mi][ni][e] = d;
                        split_bf(h, hh[e], hl[e]);
                    } else {
                        // layer 4: the output dot and the output's error signal
                        const float w5 = theta[OFF_W5 + (unit + e) * 4];
                        po[mi] += w5 * h;
                        split_bf(w5 * d, hh[e], hl[e]);
                    }
                }
                *reinterpret_cast<bf16x4 *>(bhi + row * LDSR + unit) = bf16x4{hh[0], hh[1], hh[2], hh[3]};
                *reinterpret_cast<bf16x4 *>(blo + row * LDSR + unit) = bf16x4{hl[0], hl[1], hl[2], hl[3]};
            }
        __syncthreads();
    }
    // o0: the four lanes of a point in this wave, then the four waves
#pragma unroll
    for (int mi = 0; mi < 2; mi++) {
        float v = po[mi];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        if (lane < 16) red[w][16 * mi + lane][4] = v;
    }
    // reverse pass: delta_{l-1} = (W_l^T delta_l) silu'(z_{l-1}), l = 4, 3, 2
    float pg[2][4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll 1
    for (int l = 2; l >= 0; l--) {
        split_layer(wbf + BF_TSW + (size_t)l * H * H, wbf + BF_TSWLO + (size_t)l * H * H, bhi, blo,
                    acc);
        __syncthreads();
#pragma unroll
        for (int mi = 0; mi < 2; mi++)
#pragma unroll
            for (int ni = 0; ni < 4; ni++) {
                const int unit = 64 * w + 16 * ni + 4 * (lane >> 4);
                const int row = 16 * mi + (lane & 15);
                __bf16 hh[4], hl[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    float d;
                    if (l > 0) {
                        d = acc[mi][ni][e] * ds[l - 1][mi][ni][e];
                        split_bf(d, hh[e], hl[e]);
                    } else {
                        // layer 1's derivative, recomputed from its inputs
                        const int u = unit + e;
                        const float z = theta[OFF_B1 + u] + feat[row][0] * theta[OFF_W1 + u] +
                                        feat[row][1] * theta[OFF_W1 + H + u] +
                                        feat[row][2] * theta[OFF_W1 + 2 * H + u] +
                                        feat[row][3] * theta[OFF_W1 + 3 * H + u];
                        d = acc[mi][ni][e] * dsilu(z);
#pragma unroll
                        for (int c = 0; c < 4; c++) pg[mi][c] += theta[OFF_W1 + c * H + u] * d;
                    }
                }
                if (l > 0) {
                    *reinterpret_cast<bf16x4 *>(bhi + row * LDSR + unit) = bf16x4{hh[0], hh[1], hh[2], hh[3]};
                    *reinterpret_cast<bf16x4 *>(blo + row * LDSR + unit) = bf16x4{hl[0], hl[1], hl[2], hl[3]};
                }
            }
        __syncthreads();
    }
#pragma unroll
    for (int mi = 0; mi < 2; mi++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            float v = pg[mi][c];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (lane < 16) red[w][16 * mi + lane][c] = v;
        }
    __syncthreads();
    if (t < CP) {
        o0[t] = theta[OFF_B5] + red[0][t][4] + red[1][t][4] + red[2][t][4] + red[3][t][4];
#pragma unroll
        for (int c = 0; c < 4; c++) dfo[t][c] = red[0][t][c] + red[1][t][c] + red[2][t][c] + red[3][t][c];
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// inference: q (3,M) FP64 -> Phi_net (M), grad Phi_net (3,M), one CP-point
// tile per workgroup (tile_conservative)
__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ theta,
                                                     const __bf16 *__restrict__ wbf,
                                                     gamd_mlp_norm nm,
                                                     const double *__restrict__ q, int64_t M,
                                                     double *__restrict__ phi,
                                                     double *__restrict__ grad) {
    __shared__ __align__(16) __bf16 bhi[CP * LDSR];
    __shared__ __align__(16) __bf16 blo[CP * LDSR];
    __shared__ float feat[CP][4];
    __shared__ float red[4][CP][5];
    __shared__ float o0[CP];
    __shared__ float dfo[CP][4];
    const int t = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * CP + t;
    double x = nm.pad_r, y = 0, z = 0, r = 1;  // padding rows: a harmless point
    if (t < CP) {
        if (i < M) {
            x = q[i];
            y = q[M + i];
            z = q[2 * M + i];
        }
        float f[4];
        features(nm, x, y, z, f, r);
#pragma unroll
        for (int c = 0; c < 4; c++) feat[t][c] = f[c];
    }
    __syncthreads();
    tile_conservative(theta, wbf, feat, bhi, blo, red, o0, dfo);
    if (t < CP && i < M) {
        double ph, gx, gy, gz;
        net_physics(nm, x, y, z, r, o0[t], dfo[t], ph, gx, gy, gz);
        phi[i] = ph;
        grad[i] = gx;
        grad[M + i] = gy;
        grad[2 * M + i] = gz;
    }
}

// training forward: feat (M,4) FP32 -> Zs, Hs (BF16 [4][M][256]), Y (M,4)
__global__ __launch_bounds__(256, 2) void k_mlp_forward_train(const float *__restrict__ theta,
                                                           const __bf16 *__restrict__ wsw,
                                                           const float *__restrict__ featg,
                                                           int64_t M, __bf16 *__restrict__ Zs,
                                                           __bf16 *__restrict__ Hs,
                                                           float *__restrict__ Y) {
    extern __shared__ __align__(16) unsigned char smem[];
    __bf16 *buf = reinterpret_cast<__bf16 *>(smem);
    __shared__ float feat[TILE][4];
    __shared__ float out[TILE][4];
    const int t = threadIdx.x;
    const int64_t row0 = (int64_t)blockIdx.x * TILE;
    if (t < TILE) {
        const int64_t i = row0 + t;
#pragma unroll
        for (int c = 0; c < 4; c++) feat[t][c] = i < M ? featg[i * 4 + c] : 0.0f;
    }
    __syncthreads();
    tile_forward<true>(theta, wsw, feat, buf, out, row0, M, Zs, Hs);
    if (t < TILE && row0 + t < M)
#pragma unroll
        for (int c = 0; c < 4; c++) Y[(row0 + t) * 4 + c] = out[t][c];
}

// ---------------------------------------------------------------------------
// generic BF16 GEMM on the matrix cores for the backward pass:
//   C[m][n] = sum_k A[m][k] B[k][n], A[m][k] = A[m*sam + k*sak], B[k][n] = B[k*sbk + n*sbn]
// EPI 0: C (FP32, ldc) = acc.  EPI 1: out (BF16, ldc) = acc * silu'(Zprev[m][n]).
// Tile 64 x 64 per workgroup (4 waves of 32 x 32); M, N multiples of 64, K of 32.
// KC: both operands contiguous along k (sak = sbk = 1), read as 16-byte vectors.
template <int EPI, bool KC = false>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const __bf16 *__restrict__ A,
                                              int64_t sam, int64_t sak,
                                              const __bf16 *__restrict__ B, int64_t sbk,
                                              int64_t sbn, float *__restrict__ C,
                                              __bf16 *__restrict__ Cb, int64_t ldc,
                                              const __bf16 *__restrict__ Zprev) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.x * 64 + 32 * (w >> 1);
    const int n0 = blockIdx.y * 64 + 32 * (w & 1);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int li = lane & 15, lk = 8 * (lane >> 4);
    // split-K over gridDim.z (each slice a multiple of 32; EPI 0 then adds)
    const int kslice = K / (int)gridDim.z;
    const int k_lo = (int)blockIdx.z * kslice, k_hi = k_lo + kslice;
    for (int k0 = k_lo; k0 < k_hi; k0 += 32) {
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int64_t row = m0 + 16 * i + li;
            if (KC) {
                a[i] = *reinterpret_cast<const bf16x8 *>(A + row * sam + k0 + lk);
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) a[i][e] = A[row * sam + (int64_t)(k0 + lk + e) * sak];
            }
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int64_t col = n0 + 16 * j + li;
            if (KC) {
                b[j] = *reinterpret_cast<const bf16x8 *>(B + k0 + lk + col * sbn);
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) b[j][e] = B[(int64_t)(k0 + lk + e) * sbk + col * sbn];
            }
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int64_t row = m0 + 16 * i + 4 * (lane >> 4) + e;
                const int64_t col = n0 + 16 * j + li;
                if (EPI == 0) {
                    if (gridDim.z > 1)
                        atomicAdd(&C[row * ldc + col], acc[i][j][e]);
                    else
                        C[row * ldc + col] = acc[i][j][e];
                } else
                    Cb[row * ldc + col] = tobf(acc[i][j][e] * dsilu(frombf(Zprev[row * ldc + col])));
            }
}

// Weight gradient of a hidden layer, C[m][n] += sum_p A[p][m] B[p][n] over
// this workgroup's slice of the P points (split over gridDim.z), with A and B
// row-major [P][256] BF16 (a layer's input activations and its pre-activation
// gradients) and C [256][256] FP32 (atomic adds across slices; the caller
// zeroes it).  The reduction runs along the points, which both operands store
// 512 bytes apart, so tiles of 64 points x 64 columns are read with coalesced
// 16-byte loads and written transposed into LDS: each matrix-core fragment (8
// consecutive points of one column) is then one 16-byte LDS read.  64 x 64
// outputs per workgroup, a 32 x 32 quarter per wave.  With gb, the
// workgroups of the first row tile also add the column sums of B (the bias
// gradient) into gb from the tiles they stage.
constexpr int WG_KC = 64;          // points per LDS stage
constexpr int WG_LDK = WG_KC + 8;  // padded LDS row in BF16 (144 bytes, 16-byte aligned)

__global__ __launch_bounds__(256) void k_wgrad(int P, const __bf16 *__restrict__ A,
                                               const __bf16 *__restrict__ B,
                                               float *__restrict__ C, float *__restrict__ gb) {
    __shared__ __align__(16) __bf16 As[64][WG_LDK];
    __shared__ __align__(16) __bf16 Bs[64][WG_LDK];
    __shared__ float red[4][64];
    const bool colsum = gb != nullptr && blockIdx.x == 0;  // uniform over the workgroup
    float cs = 0.f;  // column t % 64, points 16 (t / 64) .. + 15 of each stage
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int mt = blockIdx.x * 64, nt = blockIdx.y * 64;  // the workgroup's output tile
    const int mw = 32 * (w >> 1), nw = 32 * (w & 1);       // the wave's quarter of it
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int li = lane & 15, lk = 8 * (lane >> 4);
    const int pslice = P / (int)gridDim.z;
    const int p_lo = (int)blockIdx.z * pslice, p_hi = p_lo + pslice;
    // staging: thread t loads points r and r + 32 of the stage, columns c8 .. c8 + 7
    const int r = t >> 3, c8 = 8 * (t & 7);
    for (int p0 = p_lo; p0 < p_hi; p0 += WG_KC) {
        bf16x8 va[2], vb[2];
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
            const int64_t p = (int64_t)p0 + r + 32 * hh;
            va[hh] = *reinterpret_cast<const bf16x8 *>(A + p * H + mt + c8);
            vb[hh] = *reinterpret_cast<const bf16x8 *>(B + p * H + nt + c8);
        }
        __syncthreads();  // every wave has read the previous stage
#pragma unroll
        for (int hh = 0; hh < 2; hh++)
#pragma unroll
            for (int e = 0; e < 8; e++) {
                As[c8 + e][r + 32 * hh] = va[hh][e];
                Bs[c8 + e][r + 32 * hh] = vb[hh][e];
            }
        __syncthreads();
        if (colsum) {
            const bf16x8 *src = reinterpret_cast<const bf16x8 *>(&Bs[t & 63][16 * (t >> 6)]);
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const bf16x8 v = src[q];
#pragma unroll
                for (int e = 0; e < 8; e++) cs += frombf(v[e]);
            }
        }
#pragma unroll
        for (int ks = 0; ks < WG_KC; ks += 32) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++)
                a[i] = *reinterpret_cast<const bf16x8 *>(&As[mw + 16 * i + li][ks + lk]);
#pragma unroll
            for (int j = 0; j < 2; j++)
                b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[nw + 16 * j + li][ks + lk]);
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int row = mt + mw + 16 * i + 4 * (lane >> 4) + e;
                const int col = nt + nw + 16 * j + li;
                atomicAdd(&C[row * H + col], acc[i][j][e]);
            }
    if (colsum) {
        red[t >> 6][t & 63] = cs;
        __syncthreads();
        if (t < 64) atomicAdd(&gb[nt + t], red[0][t] + red[1][t] + red[2][t] + red[3][t]);
    }
}

// loss and dY = 2 (Y - T) / M; per-block partial loss sums
__global__ void k_loss(const float *__restrict__ Y, const float *__restrict__ T, int64_t M,
                       float *__restrict__ dY, float *__restrict__ partial) {
    __shared__ float red[256];
    float s = 0.f;
    const int64_t n = M * 4;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        dY[i] = 2.0f * d / (float)M;
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// output layer backward: gW5[j][o] = sum_m H4[m][j] dY[m][o]; gb5[o] = sum_m dY[m][o];
// dZ4[m][j] = (sum_o dY[m][o] W5[j][o]) silu'(Z4[m][j]).  One thread per (m, j)
// for dZ4; the reductions over m run per column block with atomics.
__global__ void k_out_backward_dz(const float *__restrict__ theta, const float *__restrict__ dY,
                                  const __bf16 *__restrict__ Z4, int64_t M,
                                  __bf16 *__restrict__ dZ4) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= M * H) return;
    const int64_t m = idx / H;
    const int j = (int)(idx % H);
    float s = 0.f;
#pragma unroll
    for (int o = 0; o < 4; o++) s += dY[m * 4 + o] * theta[OFF_W5 + j * 4 + o];
    dZ4[idx] = tobf(s * dsilu(frombf(Z4[idx])));
}

// column sums / small reductions: out[j*ncol_out + o] += sum_m X[m][j] Yv[m][o]
// with X BF16 (M x nx, row stride nx) and Yv FP32 (M x ny).  Grid: (nx, chunks).
__global__ void k_outer_reduce(const __bf16 *__restrict__ X, int nx, const float *__restrict__ Yv,
                               int ny, int64_t M, float *__restrict__ out) {
    const int j = blockIdx.x;
    const int64_t chunk = (M + gridDim.y - 1) / gridDim.y;
    const int64_t m_lo = blockIdx.y * chunk;
    const int64_t m_hi = m_lo + chunk < M ? m_lo + chunk : M;
    __shared__ float red[256][4];
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    for (int64_t m = m_lo + threadIdx.x; m < m_hi; m += blockDim.x) {
        const float xv = X ? frombf(X[m * nx + j]) : 1.0f;
        for (int o = 0; o < ny; o++) s[o] += xv * Yv[m * ny + o];
    }
    for (int o = 0; o < 4; o++) red[threadIdx.x][o] = s[o];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int o = 0; o < 4; o++) red[threadIdx.x][o] += red[threadIdx.x + st][o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int o = 0; o < ny; o++) atomicAdd(&out[j * ny + o], red[0][o]);
}

// Column reductions over the M rows of a BF16 (M x 256) matrix, one block per
// CS_ROWS rows: thread t reads 8 adjacent columns (one 16-byte load) of rows
// t / 32, t / 32 + 8, ...; the 8 row groups are summed in LDS and each block
// adds its 256 partial sums with atomics.
constexpr int CS_ROWS = 128;

// bias gradients: gb[j] += sum_m dZ[m][j]
__global__ __launch_bounds__(256) void k_colsum(const __bf16 *__restrict__ dZ, int64_t M,
                                                float *__restrict__ gb) {
    __shared__ float red[8][H];
    const int t = threadIdx.x, c8 = 8 * (t & 31), rg = t >> 5;
    const int64_t r0 = (int64_t)blockIdx.x * CS_ROWS;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int r = rg; r < CS_ROWS; r += 8) {
        const int64_t m = r0 + r;
        if (m >= M) break;
        const bf16x8 v = *reinterpret_cast<const bf16x8 *>(dZ + m * H + c8);
#pragma unroll
        for (int e = 0; e < 8; e++) s[e] += frombf(v[e]);
    }
#pragma unroll
    for (int e = 0; e < 8; e++) red[rg][c8 + e] = s[e];
    __syncthreads();
    float tot = 0.f;
#pragma unroll
    for (int g = 0; g < 8; g++) tot += red[g][t];
    atomicAdd(&gb[t], tot);
}

// Outer-product reduction of an (M x 4) FP32 and an (M x 256) BF16 matrix,
// g[c][j] += sum_m feat[m][c] dZ1[m][j]: the first layer's weight gradient
// (UNIT_MAJOR false, stored [4][256]) and the output layer's (feat = dY,
// dZ1 = the last hidden activations, UNIT_MAJOR true, stored [256][4]).
// Row-blocked like k_colsum: thread t reads 8 adjacent columns of every 8th
// row with one 16-byte load.
template <bool UNIT_MAJOR>
__global__ __launch_bounds__(256) void k_outer4(const float *__restrict__ feat,
                                                const __bf16 *__restrict__ dZ1, int64_t M,
                                                float *__restrict__ gW1) {
    __shared__ float red[8][4][H];
    const int t = threadIdx.x, c8 = 8 * (t & 31), rg = t >> 5;
    const int64_t r0 = (int64_t)blockIdx.x * CS_ROWS;
    float s[4][8];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int e = 0; e < 8; e++) s[c][e] = 0.f;
    for (int r = rg; r < CS_ROWS; r += 8) {
        const int64_t m = r0 + r;
        if (m >= M) break;
        const bf16x8 v = *reinterpret_cast<const bf16x8 *>(dZ1 + m * H + c8);
        const f32x4 f = *reinterpret_cast<const f32x4 *>(feat + m * 4);
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int e = 0; e < 8; e++) s[c][e] += f[c] * frombf(v[e]);
    }
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int e = 0; e < 8; e++) red[rg][c][c8 + e] = s[c][e];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; c++) {
        float tot = 0.f;
#pragma unroll
        for (int g = 0; g < 8; g++) tot += red[g][c][t];
        atomicAdd(&gW1[UNIT_MAJOR ? t * 4 + c : c * H + t], tot);
    }
}

// BF16 copies of W2..W4 (layout at BF_SW .. BF_TSWLO): fragment order of
// the high parts, row-major high parts, fragment order of the low parts,
// and fragment order of the transposes' high and low parts
__device__ __forceinline__ size_t frag_index(int k, int n) {
    // fragment order of V[k][n] (k input, n output): ((jt * 8 + ks) * 64 +
    // lane) * 8 + e with n = 16 jt + lane % 16, k = 32 ks + 8 (lane / 16) + e
    const int jt = n >> 4, ks = k >> 5, lane = (n & 15) + 16 * ((k & 31) >> 3), e = k & 7;
    return (((size_t)jt * 8 + ks) * 64 + lane) * 8 + e;
}

__global__ void k_sync_bf16(const float *__restrict__ theta, __bf16 *__restrict__ wbf) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= 3 * H * H) return;
    const int l = (int)(idx / (H * H));
    const int rem = (int)(idx % (H * H));
    const int k = rem / H, n = rem % H;  // W[k][n], row-major [in][out]
    const float v = theta[(l == 0 ? OFF_W2 : l == 1 ? OFF_W3 : OFF_W4) + rem];
    __bf16 hi, lo;
    split_bf(v, hi, lo);
    const size_t lb = (size_t)l * H * H;
    wbf[BF_SW + lb + frag_index(k, n)] = hi;
    wbf[BF_ROW + lb + rem] = hi;
    wbf[BF_SWLO + lb + frag_index(k, n)] = lo;
    wbf[BF_TSW + lb + frag_index(n, k)] = hi;    // W^T[n][k] = W[k][n]
    wbf[BF_TSWLO + lb + frag_index(n, k)] = lo;
}

// Adam (FP32 masters)
__global__ void k_adam(float *__restrict__ theta, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, int64_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    theta[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// training data: positions (3,M) FP64 with exact Phi and grad -> features
// (M,4) and targets (M,4): o0 the potential's departure from the monopole
// in units of c, o1..o3 the force over its radial scale
__global__ void k_make_targets(gamd_mlp_norm nm, const double *__restrict__ q,
                               const double *__restrict__ phi, const double *__restrict__ grad,
                               int64_t M, float *__restrict__ feat, float *__restrict__ target) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= M) return;
    const double x = q[i], y = q[M + i], z = q[2 * M + i];
    float f[4];
    double r;
    features(nm, x, y, z, f, r);
#pragma unroll
    for (int c = 0; c < 4; c++) feat[i * 4 + c] = f[c];
    const double sgrad = out_scale_grad(nm, r);
    // o0's target: Phi = Phi_b (1 + c o0) with Phi_b = -G M / (r + r_s)
    target[i * 4 + 0] = (float)((phi[i] / (-nm.gm / (r + nm.rs)) - 1.0) / nm.phi_c);
    target[i * 4 + 1] = (float)(grad[i] / sgrad);
    target[i * 4 + 2] = (float)(grad[M + i] / sgrad);
    target[i * 4 + 3] = (float)(grad[2 * M + i] / sgrad);
}

// ---------------------------------------------------------------------------
// surrogate leapfrog, one step per call (this kernel for the orbits on the
// network, then k_surrogate_fallback for those on the exact force).  Orbit i:
// w (6,N) holds x and the half-step velocity between launches (vsync the
// synchronised velocity after the last step); E0[i] the initial energy;
// fb[i] the step at which the orbit fell back to exact SCF (-1 while on the
// surrogate).  mode 0: initial half kick; mode 1: drift + kick.  The force
// is -grad Phi_net, the gradient of the network's own potential, so the
// monitored energy Phi_net + |v|^2/2 is that of a conservative system; once
// |E - E0| > tol |E0| the orbit switches to the exact SCF force for the rest
// of the run (and for the remainder of this step).
__global__ __launch_bounds__(256) void k_surrogate_step(
    const float *__restrict__ theta, const __bf16 *__restrict__ wbf, gamd_mlp_norm nm,
    double *__restrict__ wq, double *__restrict__ vsync, double *__restrict__ E0,
    int32_t *__restrict__ fb, double *__restrict__ wsw, int64_t N, double dt, int mode,
    int32_t step, double tol) {
    __shared__ __align__(16) __bf16 bhi[CP * LDSR];
    __shared__ __align__(16) __bf16 blo[CP * LDSR];
    __shared__ float feat[CP][4];
    __shared__ float red[4][CP][5];
    __shared__ float o0[CP];
    __shared__ float dfo[CP][4];
    const int t = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * CP + t;
    const bool own = t < CP && i < N;
    double x = nm.pad_r, y = 0, z = 0, hx = 0, hy = 0, hz = 0, r = 1;
    int32_t fallback = -1;
    if (own) {
        x = wq[i];
        y = wq[N + i];
        z = wq[2 * N + i];
        hx = wq[3 * N + i];
        hy = wq[4 * N + i];
        hz = wq[5 * N + i];
        fallback = fb[i];
        if (mode == 1) {
            x = x + hx * dt;
            y = y + hy * dt;
            z = z + hz * dt;
        }
    }
    if (t < CP) {
        float f[4];
        features(nm, x, y, z, f, r);
#pragma unroll
        for (int c = 0; c < 4; c++) feat[t][c] = f[c];
    }
    // the tile runs the network unless every orbit in it has fallen back
    // (uniform over the workgroup, so every thread reaches the same barriers)
    const int any_net = __syncthreads_or(own && fallback < 0);
    if (any_net) tile_conservative(theta, wbf, feat, bhi, blo, red, o0, dfo);
    if (!own || fallback >= 0) return;  // exact-force orbits: k_surrogate_fallback
    double phi, gx, gy, gz;
    net_physics(nm, x, y, z, r, o0[t], dfo[t], phi, gx, gy, gz);
    // energy monitor on the synchronised velocity this force gives
    const double vx = (mode == 0) ? hx : hx - gx * dt / 2.0;
    const double vy = (mode == 0) ? hy : hy - gy * dt / 2.0;
    const double vz = (mode == 0) ? hz : hz - gz * dt / 2.0;
    const double E = phi + 0.5 * (vx * vx + vy * vy + vz * vz);
    if (mode == 0) {
        E0[i] = E;
    } else if (fabs(E - E0[i]) > tol * fabs(E0[i])) {
        // from this step on the exact force: k_surrogate_fallback redoes it
        // from the state before the step, which is recorded when asked
        fb[i] = step;
        if (wsw) {
#pragma unroll
            for (int c = 0; c < 6; c++) wsw[c * N + i] = wq[c * N + i];
        }
        return;
    }
    if (mode == 0) {
        wq[3 * N + i] = hx - gx * dt / 2.0;
        wq[4 * N + i] = hy - gy * dt / 2.0;
        wq[5 * N + i] = hz - gz * dt / 2.0;
    } else {
        wq[i] = x;
        wq[N + i] = y;
        wq[2 * N + i] = z;
        vsync[i] = vx;
        vsync[N + i] = vy;
        vsync[2 * N + i] = vz;
        wq[3 * N + i] = hx - gx * dt;
        wq[4 * N + i] = hy - gy * dt;
        wq[5 * N + i] = hz - gz * dt;
    }
}

// The same leapfrog step for the orbits on the exact SCF force (fb >= 0,
// including those k_surrogate_step has just switched): the drift is redone
// from the stored state, then the force, kick and writes with the
// arithmetic and the SCF evaluator of gamd_leapfrog's loop (k_leapfrog for
// the same potential layout), so from its switch step on an orbit follows
// the exact leapfrog bit for bit.  A kernel of its own keeps the SCF
// evaluator's registers out of the network kernel.
template <int LAY>
__global__ __launch_bounds__(256) void k_surrogate_fallback(const DevPot *__restrict__ Pp,
                                                            double *__restrict__ wq,
                                                            double *__restrict__ vsync,
                                                            const int32_t *__restrict__ fb,
                                                            int64_t N, double dt, int mode) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= N || fb[i] < 0) return;
    double x = wq[i], y = wq[N + i], z = wq[2 * N + i];
    const double hx = wq[3 * N + i], hy = wq[4 * N + i], hz = wq[5 * N + i];
    if (mode == 1) {
        x = x + hx * dt;
        y = y + hy * dt;
        z = z + hz * dt;
    }
    double gx, gy, gz;
    pot_grad_l<true, LAY>(*Pp, x, y, z, gx, gy, gz);
    if (mode == 0) {
        wq[3 * N + i] = hx - gx * dt / 2.0;
        wq[4 * N + i] = hy - gy * dt / 2.0;
        wq[5 * N + i] = hz - gz * dt / 2.0;
    } else {
        wq[i] = x;
        wq[N + i] = y;
        wq[2 * N + i] = z;
        vsync[i] = hx - gx * dt / 2.0;
        vsync[N + i] = hy - gy * dt / 2.0;
        vsync[2 * N + i] = hz - gz * dt / 2.0;
        wq[3 * N + i] = hx - gx * dt;
        wq[4 * N + i] = hy - gy * dt;
        wq[5 * N + i] = hz - gz * dt;
    }
}

int check_norm(const gamd_mlp_norm *nm) {
    if (!nm || !nm->scale) return GAMD_ERR_INVALID;
    if (!(nm->inv_dlnr > 0.0) || !isfinite(nm->inv_dlnr) || !isfinite(nm->lnr0) ||
        !(nm->pad_r > 0.0) || !(nm->gm > 0.0) || !(nm->rs > 0.0) || !(nm->phi_c > 0.0) ||
        !isfinite(nm->gm) || !isfinite(nm->rs) || !isfinite(nm->phi_c))
        return GAMD_ERR_INVALID;
    return GAMD_OK;
}

int check_params(const float *theta, const uint16_t *wbf) {
    return (theta && wbf) ? GAMD_OK : GAMD_ERR_INVALID;
}

// the training forward uses 66 KB of dynamic LDS (one padded 128 x 256 BF16
// activation buffer), above the default 64 KB; raise its limit once per
// process (a function-local static: initialised once even when several host
// threads make their first call together)
int ensure_lds() {
    static const int rc = [] {
        const void *ks[1] = {(const void *)k_mlp_forward_train};
        for (const void *k : ks)
            if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, LDS_BYTES) !=
                hipSuccess)
                return (int)GAMD_ERR_HIP;
        return (int)GAMD_OK;
    }();
    return rc;
}

}  // namespace

extern "C" {

int64_t gamd_mlp_param_count(void) { return NPARAM; }
int64_t gamd_mlp_bf16_count(void) { return NBF; }

int64_t gamd_mlp_workspace_bytes(int64_t M) {
    // Zs, Hs: 2 x 4 x M x 256 BF16; Y, dY: 2 x M x 4 FP32; dZ ping-pong 2 x M x 256 BF16;
    // loss partials 1024 FP32
    return 2 * 4 * M * H * 2 + 2 * M * 4 * 4 + 2 * M * H * 2 + 1024 * 4 + 256;
}

int gamd_mlp_sync_bf16(const float *theta, uint16_t *wbf, void *stream) {
    if (check_params(theta, wbf)) return GAMD_ERR_INVALID;
    hipLaunchKernelGGL(k_sync_bf16, dim3(3 * H * H / 256), dim3(256), 0, as_stream(stream), theta,
                       reinterpret_cast<__bf16 *>(wbf));
    return launch_status();
}

int gamd_mlp_forward(const float *theta, const uint16_t *wbf, const gamd_mlp_norm *norm,
                     const double *q, size_t M, double *phi, double *grad, void *stream) {
    if (check_params(theta, wbf) || check_norm(norm) || (M && (!q || !phi || !grad)))
        return GAMD_ERR_INVALID;
    if (M == 0) return GAMD_OK;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((M + CP - 1) / CP)), dim3(256), 0,
                       as_stream(stream), theta, reinterpret_cast<const __bf16 *>(wbf), *norm, q,
                       (int64_t)M, phi, grad);
    return launch_status();
}

int gamd_mlp_make_targets(const gamd_mlp_norm *norm, const double *q, const double *phi,
                          const double *grad, size_t M, float *feat, float *target, void *stream) {
    if (check_norm(norm) || (M && (!q || !phi || !grad || !feat || !target)))
        return GAMD_ERR_INVALID;
    if (M == 0) return GAMD_OK;
    hipLaunchKernelGGL(k_make_targets, dim3(grid_for(M, 256)), dim3(256), 0, as_stream(stream),
                       *norm, q, phi, grad, (int64_t)M, feat, target);
    return launch_status();
}

int gamd_mlp_gradients(const float *theta, const uint16_t *wbf, const float *feat,
                       const float *target, size_t M, float *grad, float *loss, void *workspace,
                       size_t workspace_bytes, void *stream) {
    if (check_params(theta, wbf) || !feat || !target || !grad || !workspace) return GAMD_ERR_INVALID;
    if (M == 0 || M % 64) return GAMD_ERR_INVALID;
    if ((int64_t)workspace_bytes < gamd_mlp_workspace_bytes((int64_t)M)) return GAMD_ERR_INVALID;
    hipStream_t s = as_stream(stream);
    const int64_t Mi = (int64_t)M;
    unsigned char *ws = static_cast<unsigned char *>(workspace);
    __bf16 *Zs = reinterpret_cast<__bf16 *>(ws);
    __bf16 *Hs = Zs + 4 * Mi * H;
    float *Y = reinterpret_cast<float *>(Hs + 4 * Mi * H);
    float *dY = Y + Mi * 4;
    __bf16 *dZa = reinterpret_cast<__bf16 *>(dY + Mi * 4);
    __bf16 *dZb = dZa + Mi * H;
    float *partial = reinterpret_cast<float *>(dZb + Mi * H);
    const __bf16 *wsw = reinterpret_cast<const __bf16 *>(wbf);
    const __bf16 *wrow = wsw + 3 * H * H;
    if (ensure_lds()) return GAMD_ERR_HIP;
    GAMD_HIP_CHECK(hipMemsetAsync(grad, 0, NPARAM * sizeof(float), s));
    hipLaunchKernelGGL(k_mlp_forward_train, dim3((unsigned)((Mi + TILE - 1) / TILE)), dim3(256),
                       LDS_BYTES, s, theta, wsw, feat, Mi, Zs, Hs, Y);
    const unsigned lblocks = 512;
    hipLaunchKernelGGL(k_loss, dim3(lblocks), dim3(256), 0, s, Y, target, Mi, dY, partial);
    // output layer
    const unsigned cs_blocks = (unsigned)((Mi + CS_ROWS - 1) / CS_ROWS);
    hipLaunchKernelGGL(k_outer4<true>, dim3(cs_blocks), dim3(256), 0, s, dY, Hs + 3 * Mi * H, Mi,
                       grad + OFF_W5);
    hipLaunchKernelGGL(k_outer_reduce, dim3(1, 16), dim3(256), 0, s, (const __bf16 *)nullptr, 1, dY,
                       4, Mi, grad + OFF_B5);
    hipLaunchKernelGGL(k_out_backward_dz, dim3(grid_for(Mi * H, 256)), dim3(256), 0, s, theta, dY,
                       Zs + 3 * Mi * H, Mi, dZa);
    // hidden layers 4, 3, 2: dW_l = H_{l-1}^T dZ_l, db_l, dZ_{l-1} = (dZ_l W_l^T) silu'(Z_{l-1})
    __bf16 *dZ = dZa, *dZn = dZb;
    const int offW[3] = {OFF_W2, OFF_W3, OFF_W4}, offB[3] = {OFF_B2, OFF_B3, OFF_B4};
    for (int l = 2; l >= 0; l--) {  // l indexes W_{l+2}
        const __bf16 *Hin = Hs + (size_t)l * Mi * H;   // H_{l+1}
        // split the points over up to 16 slices of whole 64-point stages (grad is zeroed above)
        unsigned ksplit = 16;
        while (ksplit > 1 && Mi % (WG_KC * (int64_t)ksplit) != 0) ksplit /= 2;
        hipLaunchKernelGGL(k_wgrad, dim3(H / 64, H / 64, ksplit), dim3(256), 0, s, (int)Mi, Hin, dZ,
                           grad + offW[l], grad + offB[l]);
        hipLaunchKernelGGL((k_gemm<1, true>), dim3((unsigned)(Mi / 64), H / 64), dim3(256), 0, s, (int)Mi, H,
                           H, dZ, (int64_t)H, (int64_t)1, wrow + (size_t)l * H * H, (int64_t)1,
                           (int64_t)H, (float *)nullptr, dZn, (int64_t)H, Zs + (size_t)l * Mi * H);
        __bf16 *tmp = dZ;
        dZ = dZn;
        dZn = tmp;
    }
    // first layer
    hipLaunchKernelGGL(k_outer4<false>, dim3(cs_blocks), dim3(256), 0, s, feat, dZ, Mi,
                       grad + OFF_W1);
    hipLaunchKernelGGL(k_colsum, dim3(cs_blocks), dim3(256), 0, s, dZ, Mi, grad + OFF_B1);
    if (loss) {
        // sum of the per-block partials into loss[0]
        GAMD_HIP_CHECK(hipMemsetAsync(loss, 0, sizeof(float), s));
        hipLaunchKernelGGL(k_outer_reduce, dim3(1, 1), dim3(256), 0, s, (const __bf16 *)nullptr, 1,
                           partial, 1, (int64_t)lblocks, loss);
    }
    return launch_status();
}

int gamd_mlp_adam(float *theta, const float *grad, float *m, float *v, int64_t step, float lr,
                  float beta1, float beta2, float eps, uint16_t *wbf, void *stream) {
    if (!theta || !grad || !m || !v || !wbf || step < 1) return GAMD_ERR_INVALID;
    const float c1 = 1.0f / (1.0f - powf(beta1, (float)step));
    const float c2 = 1.0f / (1.0f - powf(beta2, (float)step));
    hipLaunchKernelGGL(k_adam, dim3(grid_for(NPARAM, 256)), dim3(256), 0, as_stream(stream), theta,
                       grad, m, v, (int64_t)NPARAM, lr, beta1, beta2, eps, c1, c2);
    int rc = launch_status();
    if (rc) return rc;
    return gamd_mlp_sync_bf16(theta, wbf, stream);
}

int gamd_surrogate_leapfrog_step(const float *theta, const uint16_t *wbf, const gamd_mlp_norm *norm,
                                 const gamd_potential *scf, double *w, double *vsync, double *E0,
                                 int32_t *fallback_step, double *w_switch, size_t N, double dt,
                                 int mode, int32_t step, double tol, void *stream) {
    if (check_params(theta, wbf) || check_norm(norm) || !scf || !w || !vsync || !E0 ||
        !fallback_step)
        return GAMD_ERR_INVALID;
    if (scf->dev.n != 1 || scf->dev.c[0].type != GAMD_SCF || scf->dev.c[0].shift_rotate)
        return GAMD_ERR_INVALID;
    if (mode != 0 && mode != 1) return GAMD_ERR_INVALID;
    if (N == 0) return GAMD_OK;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_surrogate_step, dim3((unsigned)((N + CP - 1) / CP)), dim3(256), 0, s,
                       theta, reinterpret_cast<const __bf16 *>(wbf), *norm, w, vsync, E0,
                       fallback_step, w_switch, (int64_t)N, dt, mode, step, tol);
    int rc = launch_status();
    if (rc) return rc;
    auto fk = scf->dev.layout == LAYOUT_SCF106 ? k_surrogate_fallback<LAYOUT_SCF106>
                                                : k_surrogate_fallback<LAYOUT_GENERIC>;
    hipLaunchKernelGGL(fk, dim3(grid_for(N, 256)), dim3(256), 0, s, scf->ddev, w, vsync,
                       fallback_step, (int64_t)N, dt, mode);
    return launch_status();
}

}  // extern "C"
