// libsrmg, learned coarse-level initial guess.
//
// A U-Net of 3^D convolutions maps the right-hand side f of the level-1
// (uniform, coarsest) problem L_1 u = f to an approximation of u. Activations
// and weights are BF16 and every product runs on the matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation). Forward pass, backward pass
// (data and weight gradients) and the Adam update are kernels of this file,
// driven from the host code below; nothing here calls a library GEMM.
//
// Layouts
//   activations  [batch][voxel][channel] BF16 (channels last, x fastest),
//                channels a power of two >= 8, so one lane loads the 8
//                channels of one stencil tap as one 16-byte vector;
//   parameters   FP32, per convolution W[cout][tap][cin] then b[cout], taps
//                ordered t = (dx+1) + 3 (dy+1) + 9 (dz+1);
//   packed       BF16 GEMM operands: forward Wf[cout (16-padded)][tap*cin
//                (32-padded)], data gradient Wb[cin][tap'*cout] with the taps
//                mirrored (the transposed convolution).
//
// Network (levels L, widths w_0..w_{L-1}): per level two ReLU convolutions
// (encoder; 2^D average pooling between levels); per decoder level a linear
// convolution at the coarse level, nearest upsampling added to the encoder
// output of the same level (skip connection) and one ReLU convolution; a last
// linear convolution to one FP32 channel. Input channels: f / max|f|, the
// normalized cell-centre coordinates, zeros. Output: u / (max|f| c_out).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/srmg.h"
#include "common.h"
#include "unet.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TPB = 256;
// convolutions over fewer voxels split their K loop over the waves of a work group
constexpr int64_t SPLITK_NVOX = 65536;
constexpr int MAXLEV = 6;
constexpr int FEAT = 8;  // input channels
constexpr int LOSS_WGS = 1024;

__device__ inline float bf2f(uint32_t b) { return __uint_as_float(b << 16); }
// round to nearest even; NaN stays NaN
__device__ inline uint32_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0xffffu) ? 0x40u : 0u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ inline bool bf_pos(uint32_t b) { return (b & 0xffffu) != 0 && !(b & 0x8000u); }

__device__ inline void unpack8(uint4 q, float *o) {
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        o[2 * i] = bf2f(w[i] & 0xffffu);
        o[2 * i + 1] = bf2f(w[i] >> 16);
    }
}
__device__ inline uint4 pack8(const float *o) {
    uint32_t w[4];
#pragma unroll
    for (int i = 0; i < 4; i++) w[i] = f2bf(o[2 * i]) | (f2bf(o[2 * i + 1]) << 16);
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// ---------------------------------------------------------------------------
// convolution as an implicit GEMM on the matrix cores
// ---------------------------------------------------------------------------

// out[v][co] = act(sum_{t, ci} W[co][t][ci] in[v + off(t)][ci] + bias[co]),
// zero outside the grid. GEMM view: D (cout x voxels) = W (cout x K) *
// Xcol (K x voxels), K = taps * cin. A wave owns 16 R output channels x 64
// voxels (R x 4 accumulators of 16 x 16); one 16x16x32 MFMA consumes, per
// lane, 8 consecutive k (one tap, 8 channels) of one voxel (B) and of one
// output channel (A). The same kernel is the data gradient when given the
// mirrored, transposed weights; `mask` then applies the ReLU derivative of
// the layer that produced the gradient's target (out *= mask > 0).
// OUT 0: BF16 output with cout_s stored channels; OUT 1: FP32, one channel.
// SK > 1 (grids with few voxels, where a launch of one wave per 64 voxels
// leaves the chip idle and each wave walks the whole K loop): the SK waves of
// a work group share one group of 64 voxels and take every SK-th 32-wide K
// step; waves 1..SK-1 hand their accumulators to wave 0 through LDS, which
// adds them in wave order before the epilogue.
template <int D, int R, int OUT, int SK = 1>
__global__ __launch_bounds__(TPB) void k_conv(const uint16_t *__restrict__ in, int cin_log2,
                                              const uint16_t *__restrict__ w, int kp, int ktot,
                                              const float *__restrict__ bias, int nbias, int relu,
                                              const uint16_t *__restrict__ mask, void *__restrict__ out, int cout_s,
                                              int n0, int n1, int n2, int64_t nvox) {
    static_assert(SK == 1 || SK == TPB / 64, "split K over the waves of a work group");
    const int lane = threadIdx.x & 63, col = lane & 15, g = lane >> 4;
    const int wv = threadIdx.x >> 6;
    const int64_t vbase = (SK > 1 ? (int64_t)blockIdx.x : (int64_t)blockIdx.x * (TPB / 64) + wv) * 64;
    const int64_t V = (int64_t)n0 * n1 * n2;
    int px[4], py[4], pz[4];
    int64_t pn[4];
    bool ok[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int64_t v = vbase + 16 * q + col;
        ok[q] = v < nvox;
        const int64_t nn = ok[q] ? v / V : 0, r = ok[q] ? v - nn * V : 0;
        pn[q] = nn * V;
        px[q] = (int)(r % n0);
        py[q] = (int)((r / n0) % n1);
        pz[q] = (int)(r / ((int64_t)n0 * n1));
    }
    f32x4 acc[R][4];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[r][q] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int cin = 1 << cin_log2;
    for (int kc = SK > 1 ? 32 * wv : 0; kc < kp; kc += 32 * SK) {
        const int kk = kc + 8 * g;
        bf16x8 a[R];
#pragma unroll
        for (int r = 0; r < R; r++)
            a[r] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(w + (size_t)(16 * r + col) * kp + kk));
        const int t = kk >> cin_log2, ci = kk & (cin - 1);
        const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = D == 3 ? t / 9 - 1 : 0;
        bf16x8 b[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int x = px[q] + dx, y = py[q] + dy, z = pz[q] + dz;
            uint4 val = make_uint4(0u, 0u, 0u, 0u);
            if (ok[q] && kk < ktot && x >= 0 && x < n0 && y >= 0 && y < n1 && z >= 0 && z < n2)
                val = *reinterpret_cast<const uint4 *>(in + (pn[q] + x + (int64_t)n0 * (y + (int64_t)n1 * z)) * cin +
                                                       ci);
            b[q] = __builtin_bit_cast(bf16x8, val);
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int q = 0; q < 4; q++) acc[r][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b[q], acc[r][q], 0, 0, 0);
    }
    if constexpr (SK > 1) {
        __shared__ f32x4 red[SK - 1][R][4][64];
        if (wv > 0) {
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int q = 0; q < 4; q++) red[wv - 1][r][q][lane] = acc[r][q];
        }
        __syncthreads();
        if (wv > 0) return;
#pragma unroll
        for (int k = 0; k < SK - 1; k++)
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int q = 0; q < 4; q++) acc[r][q] += red[k][r][q][lane];
    }
    // lane holds D[4 g + j][col]: output channels 16 r + 4 g + j of voxel col
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (!ok[q]) continue;
        const int64_t v = vbase + 16 * q + col;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int co0 = 16 * r + 4 * g;
            if (co0 >= cout_s) continue;
            float o[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                float val = acc[r][q][j];
                if (bias && co0 + j < nbias) val += bias[co0 + j];
                if (relu) val = fmaxf(val, 0.f);
                o[j] = val;
            }
            if constexpr (OUT == 0) {
                if (mask) {
                    const uint2 mm = *reinterpret_cast<const uint2 *>(mask + v * cout_s + co0);
                    if (!bf_pos(mm.x)) o[0] = 0.f;
                    if (!bf_pos(mm.x >> 16)) o[1] = 0.f;
                    if (!bf_pos(mm.y)) o[2] = 0.f;
                    if (!bf_pos(mm.y >> 16)) o[3] = 0.f;
                }
                uint2 pk;
                pk.x = f2bf(o[0]) | (f2bf(o[1]) << 16);
                pk.y = f2bf(o[2]) | (f2bf(o[3]) << 16);
                *reinterpret_cast<uint2 *>(static_cast<uint16_t *>(out) + v * cout_s + co0) = pk;
            } else {
                if (co0 == 0) static_cast<float *>(out)[v] = o[0];
            }
        }
    }
}

// Weight and bias gradient of one convolution over a chunk of voxels:
// partial[chunk][co][k] = sum_v dz[v][co] Xcol[v][k], k < ktot, and the bias
// column k = ktot (Xcol = 1). GEMM view with the voxels as the reduction
// dimension: a wave owns 16 output channels x 4 tiles of 16 k.
template <int D>
__global__ __launch_bounds__(TPB) void k_wgrad(const uint16_t *__restrict__ dz, int cg, int cout,
                                               const uint16_t *__restrict__ x, int cin_log2, int ktot, int kx,
                                               int n0, int n1, int n2, int64_t nvox, int64_t chunk,
                                               float *__restrict__ partial) {
    const int lane = threadIdx.x & 63, col = lane & 15, g = lane >> 4;
    const int kt0 = (blockIdx.x * (TPB / 64) + (threadIdx.x >> 6)) * 4;
    if (16 * kt0 >= kx) return;
    const int co = 16 * blockIdx.y + col;
    const int64_t v0 = (int64_t)blockIdx.z * chunk, v1 = min(nvox, v0 + chunk);
    const int64_t V = (int64_t)n0 * n1 * n2;
    const int cin = 1 << cin_log2;
    int kdx[4], kdy[4], kdz[4], kci[4], kmode[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int k = 16 * (kt0 + j) + col;
        const int t = k >> cin_log2;
        kci[j] = k & (cin - 1);
        kdx[j] = t % 3 - 1;
        kdy[j] = (t / 3) % 3 - 1;
        kdz[j] = D == 3 ? t / 9 - 1 : 0;
        kmode[j] = k < ktot ? 1 : (k == ktot ? 2 : 0);
    }
    f32x4 acc[4];
#pragma unroll
    for (int j = 0; j < 4; j++) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // rows of a multiple of 8 voxels: each lane's 8 consecutive voxels lie in
    // one row (chunks start at multiples of 32), so their coordinates are
    // found once per step
    const bool rows8 = (n0 & 7) == 0;
    for (int64_t vs = v0; vs < v1 && rows8; vs += 32) {
        const int64_t vb = vs + 8 * g;
        uint32_t av[8], bv[4][8];
        const bool okv = vb < v1;
        const int64_t nn = okv ? vb / V : 0, r = okv ? vb - nn * V : 0;
        const int rr = (int)r, px0 = rr % n0, py = (rr / n0) % n1, pz = rr / (n0 * n1);
#pragma unroll
        for (int i = 0; i < 8; i++) av[i] = (okv && co < cg) ? dz[(vb + i) * cg + co] : 0u;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int yy = py + kdy[j], zz = pz + kdz[j];
            const bool rowok = okv && kmode[j] == 1 && yy >= 0 && yy < n1 && zz >= 0 && zz < n2;
            const int64_t base = (nn * V + (int64_t)n0 * (yy + (int64_t)n1 * zz)) * cin + kci[j];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int xx = px0 + i + kdx[j];
                uint32_t val = 0u;
                if (okv && kmode[j] == 2) val = 0x3f80u;  // 1.0
                else if (rowok && xx >= 0 && xx < n0) val = x[base + (int64_t)xx * cin];
                bv[j][i] = val;
            }
        }
        const uint4 aq = make_uint4(av[0] | (av[1] << 16), av[2] | (av[3] << 16), av[4] | (av[5] << 16),
                                    av[6] | (av[7] << 16));
        const bf16x8 a = __builtin_bit_cast(bf16x8, aq);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint4 bq = make_uint4(bv[j][0] | (bv[j][1] << 16), bv[j][2] | (bv[j][3] << 16),
                                        bv[j][4] | (bv[j][5] << 16), bv[j][6] | (bv[j][7] << 16));
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8, bq), acc[j], 0, 0, 0);
        }
    }
    for (int64_t vs = v0; vs < v1 && !rows8; vs += 32) {
        uint32_t av[8], bv[4][8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int64_t vox = vs + 8 * g + i;
            const bool okv = vox < v1;
            av[i] = (okv && co < cg) ? dz[vox * cg + co] : 0u;
            const int64_t nn = okv ? vox / V : 0, r = okv ? vox - nn * V : 0;
            const int px = (int)(r % n0), py = (int)((r / n0) % n1), pz = (int)(r / ((int64_t)n0 * n1));
#pragma unroll
            for (int j = 0; j < 4; j++) {
                uint32_t val = 0u;
                if (okv) {
                    if (kmode[j] == 2) {
                        val = 0x3f80u;  // 1.0
                    } else if (kmode[j] == 1) {
                        const int xx = px + kdx[j], yy = py + kdy[j], zz = pz + kdz[j];
                        if (xx >= 0 && xx < n0 && yy >= 0 && yy < n1 && zz >= 0 && zz < n2)
                            val = x[(nn * V + xx + (int64_t)n0 * (yy + (int64_t)n1 * zz)) * cin + kci[j]];
                    }
                }
                bv[j][i] = val;
            }
        }
        const uint4 aq = make_uint4(av[0] | (av[1] << 16), av[2] | (av[3] << 16), av[4] | (av[5] << 16),
                                    av[6] | (av[7] << 16));
        const bf16x8 a = __builtin_bit_cast(bf16x8, aq);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint4 bq = make_uint4(bv[j][0] | (bv[j][1] << 16), bv[j][2] | (bv[j][3] << 16),
                                        bv[j][4] | (bv[j][5] << 16), bv[j][6] | (bv[j][7] << 16));
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8, bq), acc[j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const int coo = 16 * blockIdx.y + 4 * g + jj, k = 16 * (kt0 + j) + col;
            if (coo < cout && k < kx) partial[((int64_t)blockIdx.z * cout + coo) * kx + k] = acc[j][jj];
        }
}

// fixed-order sum of the chunk partials into the weight and bias gradients
__global__ __launch_bounds__(TPB) void k_wgrad_reduce(const float *__restrict__ partial, int nchunk, int cout, int kx,
                                                      int ktot, float *__restrict__ gw, float *__restrict__ gb) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= cout * kx) return;
    float s = 0.f;
    // loads of 16 chunks in flight, the sum in chunk order
#pragma unroll 16
    for (int c = 0; c < nchunk; c++) s += partial[(int64_t)c * cout * kx + i];
    const int co = i / kx, k = i - co * kx;
    if (k < ktot)
        gw[co * ktot + k] = s;
    else
        gb[co] = s;
}

// ---------------------------------------------------------------------------
// pooling, upsampling, skip connections (8 channels per thread)
// ---------------------------------------------------------------------------

struct Ext {
    int n0, n1, n2;
};

__device__ inline int64_t vindex(const Ext &e, int x, int y, int z) {
    return x + (int64_t)e.n0 * (y + (int64_t)e.n1 * z);
}

// coarse voxel -> its 2^D children
template <int D>
__global__ __launch_bounds__(TPB) void k_pool(const uint16_t *__restrict__ in, uint16_t *__restrict__ out, int c8,
                                              Ext ec, Ext ef, int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= total) return;
    const int cg = (int)(i % c8);
    const int64_t vc = i / c8, Vc = (int64_t)ec.n0 * ec.n1 * ec.n2, Vf = (int64_t)ef.n0 * ef.n1 * ef.n2;
    const int64_t n = vc / Vc, r = vc - n * Vc;
    const int xc = (int)(r % ec.n0), yc = (int)((r / ec.n0) % ec.n1), zc = (int)(r / ((int64_t)ec.n0 * ec.n1));
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < (1 << D); c++) {
        const int xf = 2 * xc + (c & 1), yf = 2 * yc + ((c >> 1) & 1), zf = D == 3 ? 2 * zc + (c >> 2) : 0;
        float v[8];
        unpack8(*reinterpret_cast<const uint4 *>(in + ((n * Vf + vindex(ef, xf, yf, zf)) * c8 + cg) * 8), v);
        for (int k = 0; k < 8; k++) acc[k] += v[k];
    }
    for (int k = 0; k < 8; k++) acc[k] *= 1.0f / (1 << D);
    *reinterpret_cast<uint4 *>(out + i * 8) = pack8(acc);
}

// fine voxel: out = coarse(parent) + skip
template <int D>
__global__ __launch_bounds__(TPB) void k_up_add(const uint16_t *__restrict__ coarse, const uint16_t *__restrict__ skip,
                                                uint16_t *__restrict__ out, int c8, Ext ec, Ext ef, int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= total) return;
    const int cg = (int)(i % c8);
    const int64_t vf = i / c8, Vc = (int64_t)ec.n0 * ec.n1 * ec.n2, Vf = (int64_t)ef.n0 * ef.n1 * ef.n2;
    const int64_t n = vf / Vf, r = vf - n * Vf;
    const int x = (int)(r % ef.n0), y = (int)((r / ef.n0) % ef.n1), z = (int)(r / ((int64_t)ef.n0 * ef.n1));
    float a[8], b[8];
    unpack8(*reinterpret_cast<const uint4 *>(coarse + ((n * Vc + vindex(ec, x >> 1, y >> 1, z >> 1)) * c8 + cg) * 8),
            a);
    unpack8(*reinterpret_cast<const uint4 *>(skip + i * 8), b);
    for (int k = 0; k < 8; k++) a[k] += b[k];
    *reinterpret_cast<uint4 *>(out + i * 8) = pack8(a);
}

// coarse voxel: d coarse = sum of the children's d fine (adjoint of nearest upsampling)
template <int D>
__global__ __launch_bounds__(TPB) void k_up_bwd(const uint16_t *__restrict__ dfine, uint16_t *__restrict__ dcoarse,
                                                int c8, Ext ec, Ext ef, int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= total) return;
    const int cg = (int)(i % c8);
    const int64_t vc = i / c8, Vc = (int64_t)ec.n0 * ec.n1 * ec.n2, Vf = (int64_t)ef.n0 * ef.n1 * ef.n2;
    const int64_t n = vc / Vc, r = vc - n * Vc;
    const int xc = (int)(r % ec.n0), yc = (int)((r / ec.n0) % ec.n1), zc = (int)(r / ((int64_t)ec.n0 * ec.n1));
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < (1 << D); c++) {
        const int xf = 2 * xc + (c & 1), yf = 2 * yc + ((c >> 1) & 1), zf = D == 3 ? 2 * zc + (c >> 2) : 0;
        float v[8];
        unpack8(*reinterpret_cast<const uint4 *>(dfine + ((n * Vf + vindex(ef, xf, yf, zf)) * c8 + cg) * 8), v);
        for (int k = 0; k < 8; k++) acc[k] += v[k];
    }
    *reinterpret_cast<uint4 *>(dcoarse + i * 8) = pack8(acc);
}

// fine voxel: d e = (d skip + d pooled(parent) / 2^D) * (e > 0)
template <int D>
__global__ __launch_bounds__(TPB) void k_pool_bwd(const uint16_t *__restrict__ dskip, const uint16_t *__restrict__ dp,
                                                  const uint16_t *__restrict__ e, uint16_t *__restrict__ de, int c8,
                                                  Ext ec, Ext ef, int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= total) return;
    const int cg = (int)(i % c8);
    const int64_t vf = i / c8, Vc = (int64_t)ec.n0 * ec.n1 * ec.n2, Vf = (int64_t)ef.n0 * ef.n1 * ef.n2;
    const int64_t n = vf / Vf, r = vf - n * Vf;
    const int x = (int)(r % ef.n0), y = (int)((r / ef.n0) % ef.n1), z = (int)(r / ((int64_t)ef.n0 * ef.n1));
    float a[8], b[8];
    unpack8(*reinterpret_cast<const uint4 *>(dskip + i * 8), a);
    unpack8(*reinterpret_cast<const uint4 *>(dp + ((n * Vc + vindex(ec, x >> 1, y >> 1, z >> 1)) * c8 + cg) * 8), b);
    const uint4 m = *reinterpret_cast<const uint4 *>(e + i * 8);
    const uint32_t mw[4] = {m.x, m.y, m.z, m.w};
    for (int k = 0; k < 8; k++) {
        const uint32_t bits = (mw[k >> 1] >> (16 * (k & 1))) & 0xffffu;
        a[k] = bf_pos(bits) ? a[k] + b[k] * (1.0f / (1 << D)) : 0.f;
    }
    *reinterpret_cast<uint4 *>(de + i * 8) = pack8(a);
}

// ---------------------------------------------------------------------------
// input normalization, loss, output, optimizer
// ---------------------------------------------------------------------------

// scale[n] = max |f[n]| (1 for an all-zero grid); one work group per grid
// maximum over the work group (all threads get it)
__device__ __forceinline__ double wg_max(double m) {
    __shared__ double red[TPB];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    return red[0];
}

// max |f| per grid in two stages: ABSMAX_PARTS work groups per grid reduce
// contiguous chunks, then one work group per grid reduces the chunk maxima
// (a maximum does not depend on the order: the same value as one pass)
constexpr int ABSMAX_PARTS = 64;
template <typename TF>
__global__ __launch_bounds__(TPB) void k_absmax_part(const TF *__restrict__ f, int64_t V,
                                                     double *__restrict__ part) {
    const int64_t chunk = (V + ABSMAX_PARTS - 1) / ABSMAX_PARTS;
    const TF *p = f + (int64_t)blockIdx.y * V;
    const int64_t i0 = (int64_t)blockIdx.x * chunk, i1 = min(V, i0 + chunk);
    double m = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += TPB) m = fmax(m, fabs((double)p[i]));
    m = wg_max(m);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.y * ABSMAX_PARTS + blockIdx.x] = m;
}

__global__ __launch_bounds__(TPB) void k_absmax(const double *__restrict__ part, double *__restrict__ scale) {
    const double m = wg_max(threadIdx.x < ABSMAX_PARTS ? part[(int64_t)blockIdx.x * ABSMAX_PARTS + threadIdx.x]
                                                       : 0.0);
    if (threadIdx.x == 0) scale[blockIdx.x] = m > 0.0 ? m : 1.0;
}

template <int D, typename TF>
__global__ __launch_bounds__(TPB) void k_features(const TF *__restrict__ f, const double *__restrict__ scale,
                                                  uint16_t *__restrict__ x, Ext e, int64_t nvox) {
    const int64_t v = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (v >= nvox) return;
    const int64_t V = (int64_t)e.n0 * e.n1 * e.n2, n = v / V, r = v - n * V;
    const int px = (int)(r % e.n0), py = (int)((r / e.n0) % e.n1), pz = (int)(r / ((int64_t)e.n0 * e.n1));
    float o[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    o[0] = (float)((double)f[v] / scale[n]);
    o[1] = 2.0f * ((float)px + 0.5f) / (float)e.n0 - 1.0f;
    o[2] = 2.0f * ((float)py + 0.5f) / (float)e.n1 - 1.0f;
    if (D == 3) o[3] = 2.0f * ((float)pz + 0.5f) / (float)e.n2 - 1.0f;
    *reinterpret_cast<uint4 *>(x + v * FEAT) = pack8(o);
}

// d = y - u / (scale c_out); dy[v][0] = 2 d / nvox (other channels 0);
// partial[wg] = sum d^2 over a fixed grid-stride set of voxels
template <typename TU>
__global__ __launch_bounds__(TPB) void k_mse(const float *__restrict__ y, const TU *__restrict__ u,
                                             const double *__restrict__ scale, double c_out, int64_t V, int64_t nvox,
                                             uint16_t *__restrict__ dy, double *__restrict__ partial) {
    __shared__ double red[TPB];
    double acc = 0.0;
    const float gs = (float)(2.0 / (double)nvox);
    for (int64_t v = (int64_t)blockIdx.x * TPB + threadIdx.x; v < nvox; v += (int64_t)gridDim.x * TPB) {
        const double t = (double)u[v] / (scale[v / V] * c_out);
        const double d = (double)y[v] - t;
        acc += d * d;
        *reinterpret_cast<uint4 *>(dy + v * 8) = make_uint4(f2bf(gs * (float)d), 0u, 0u, 0u);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(TPB) void k_loss_total(const double *__restrict__ partial, int n, double inv,
                                                    double *__restrict__ out) {
    __shared__ double red[TPB];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += TPB) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * inv;
}

template <typename TO>
__global__ __launch_bounds__(TPB) void k_output(const float *__restrict__ y, const double *__restrict__ scale,
                                                double c_out, int64_t V, int64_t nvox, TO *__restrict__ u) {
    const int64_t v = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (v >= nvox) return;
    u[v] = (TO)((double)y[v] * scale[v / V] * c_out);
}

__global__ __launch_bounds__(TPB) void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                                              float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Wf[row][k] (rows16 x kp) from W[cout][ktot]
__global__ __launch_bounds__(TPB) void k_pack_fwd(const float *__restrict__ W, int cout, int ktot, int rows16, int kp,
                                                  uint16_t *__restrict__ wf) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= (int64_t)rows16 * kp) return;
    const int r = (int)(i / kp), k = (int)(i - (int64_t)r * kp);
    wf[i] = (uint16_t)((r < cout && k < ktot) ? f2bf(W[(int64_t)r * ktot + k]) : 0u);
}

// Wb[ci][t' cg + co] = W[co][ntap - 1 - t'][ci]: the transposed convolution
__global__ __launch_bounds__(TPB) void k_pack_bwd(const float *__restrict__ W, int cout, int cin, int ntap, int cg,
                                                  int rows16, int kp, uint16_t *__restrict__ wb) {
    const int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= (int64_t)rows16 * kp) return;
    const int ci = (int)(i / kp), kb = (int)(i - (int64_t)ci * kp);
    const int t = kb / cg, co = kb - t * cg;
    uint32_t val = 0u;
    if (ci < cin && t < ntap && co < cout) val = f2bf(W[((int64_t)co * ntap + (ntap - 1 - t)) * cin + ci]);
    wb[i] = (uint16_t)val;
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + TPB - 1) / TPB); }
inline int ilog2(int x) {
    int l = 0;
    while ((1 << l) < x) l++;
    return l;
}
inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

struct ConvL {
    int level = 0;           // grid level of input and output (0 = full grid)
    int cin = 0, cout = 0;   // channels; cin a power of two >= 8
    int cout_s = 0;          // stored output channels (1: the FP32 output)
    int cg = 0;              // channels of the output gradient (>= 8)
    bool relu = false;
    int ktot = 0, kp = 0, rows = 0;    // forward GEMM
    int ktot_b = 0, kp_b = 0, rows_b = 0;  // data-gradient GEMM
    size_t woff = 0, boff = 0;
    DevBuf wf, wb;
};

struct Tensor {
    int level = 0, c = 0;
    DevBuf v, g;
};

}  // namespace

struct srmg_unet {
    int D = 3, L = 4, ntap = 27;
    Ext ext[MAXLEV];
    int64_t V[MAXLEV];
    std::vector<int> w;
    std::vector<ConvL> conv;
    std::vector<int> cA, cE, cC, cDd;
    int cY = -1;
    size_t nparam = 0;
    DevBuf p, gr, m, v;
    int64_t steps = 0;
    double c_out = 1.0;
    bool has_problem = false;
    int geometry = 0, disc = 0;
    double k2 = 0.0;
    int8_t bc[27] = {0};
    hipStream_t stream = nullptr;
    int cap = 0;
    Tensor X;
    std::vector<Tensor> A, E, P, C, S, Dd;
    DevBuf Y, gY, scale, absmax_part, lossp, loss, partial;
    bool packed = false;
};

namespace {

void alloc_tensor(Tensor &t, int level, int c, int64_t nvox, bool grad) {
    t.level = level;
    t.c = c;
    t.v.alloc((size_t)nvox * c * 2);
    if (grad) t.g.alloc((size_t)nvox * c * 2);
}

void ensure_batch(srmg_unet *n, int B) {
    if (B <= n->cap) return;
    const int L = n->L;
    auto nv = [&](int l) { return (int64_t)B * n->V[l]; };
    alloc_tensor(n->X, 0, FEAT, nv(0), false);
    n->A.resize(L);
    n->E.resize(L);
    n->P.resize(L);
    n->C.resize(L);
    n->S.resize(L);
    n->Dd.resize(L);
    for (int l = 0; l < L; l++) {
        alloc_tensor(n->A[l], l, n->w[l], nv(l), true);
        alloc_tensor(n->E[l], l, n->w[l], nv(l), true);
        if (l > 0) alloc_tensor(n->P[l], l, n->w[l - 1], nv(l), true);
        if (l < L - 1) {
            alloc_tensor(n->C[l], l + 1, n->w[l], nv(l + 1), true);
            alloc_tensor(n->S[l], l, n->w[l], nv(l), true);
            alloc_tensor(n->Dd[l], l, n->w[l], nv(l), true);
        }
    }
    n->Y.alloc((size_t)nv(0) * sizeof(float));
    n->gY.alloc((size_t)nv(0) * 8 * 2);
    n->scale.alloc((size_t)B * sizeof(double));
    n->absmax_part.alloc((size_t)B * ABSMAX_PARTS * sizeof(double));
    n->lossp.alloc(LOSS_WGS * sizeof(double));
    n->loss.alloc(sizeof(double));
    n->cap = B;
}

void pack(srmg_unet *n, hipStream_t st) {
    for (ConvL &cv : n->conv) {
        const float *W = n->p.as<float>() + cv.woff;
        hipLaunchKernelGGL(k_pack_fwd, dim3(nblocks((int64_t)cv.rows * 16 * cv.kp)), dim3(TPB), 0, st, W, cv.cout,
                           cv.ktot, cv.rows * 16, cv.kp, cv.wf.as<uint16_t>());
        hipLaunchKernelGGL(k_pack_bwd, dim3(nblocks((int64_t)cv.rows_b * 16 * cv.kp_b)), dim3(TPB), 0, st, W, cv.cout,
                           cv.cin, n->ntap, cv.cg, cv.rows_b * 16, cv.kp_b, cv.wb.as<uint16_t>());
    }
    n->packed = true;
}

template <int D, int OUT>
void conv_launch_r(int R, unsigned grid, hipStream_t st, const uint16_t *in, int cin_log2, const uint16_t *w, int kp,
                   int ktot, const float *bias, int nbias, int relu, const uint16_t *mask, void *out, int cout_s,
                   const Ext &e, int64_t nvox) {
    // few voxels: split K over the waves of a work group (one group of 64 voxels per work group)
    const bool split = nvox < SPLITK_NVOX;
    const unsigned gsk = (unsigned)((nvox + 63) / 64);
#define SRMG_CONV(RR)                                                                                              \
    do {                                                                                                           \
        if (split)                                                                                                 \
            hipLaunchKernelGGL((k_conv<D, RR, OUT, TPB / 64>), dim3(gsk), dim3(TPB), 0, st, in, cin_log2, w, kp,   \
                               ktot, bias, nbias, relu, mask, out, cout_s, e.n0, e.n1, e.n2, nvox);                \
        else                                                                                                       \
            hipLaunchKernelGGL((k_conv<D, RR, OUT>), dim3(grid), dim3(TPB), 0, st, in, cin_log2, w, kp, ktot,      \
                               bias, nbias, relu, mask, out, cout_s, e.n0, e.n1, e.n2, nvox);                      \
    } while (0)
    if (R == 1)
        SRMG_CONV(1);
    else if (R == 2)
        SRMG_CONV(2);
    else if (R == 4)
        SRMG_CONV(4);
    else
        fail("unsupported channel count");
#undef SRMG_CONV
}

template <int D>
void conv_fwd(srmg_unet *n, hipStream_t st, const ConvL &cv, const uint16_t *in, void *out, int B) {
    const int64_t nvox = (int64_t)B * n->V[cv.level];
    const unsigned grid = (unsigned)((nvox + TPB - 1) / TPB);
    const float *bias = n->p.as<float>() + cv.boff;
    if (cv.cout_s == 1)
        conv_launch_r<D, 1>(cv.rows, grid, st, in, ilog2(cv.cin), cv.wf.as<uint16_t>(), cv.kp, cv.ktot, bias, cv.cout,
                            0, nullptr, out, 1, n->ext[cv.level], nvox);
    else
        conv_launch_r<D, 0>(cv.rows, grid, st, in, ilog2(cv.cin), cv.wf.as<uint16_t>(), cv.kp, cv.ktot, bias, cv.cout,
                            cv.relu ? 1 : 0, nullptr, out, cv.cout_s, n->ext[cv.level], nvox);
}

// gradient w.r.t. the convolution's input from the gradient dz of its output
template <int D>
void conv_bwd_data(srmg_unet *n, hipStream_t st, const ConvL &cv, const uint16_t *dz, uint16_t *din,
                   const uint16_t *mask, int B) {
    const int64_t nvox = (int64_t)B * n->V[cv.level];
    const unsigned grid = (unsigned)((nvox + TPB - 1) / TPB);
    conv_launch_r<D, 0>(cv.rows_b, grid, st, dz, ilog2(cv.cg), cv.wb.as<uint16_t>(), cv.kp_b, cv.ktot_b, nullptr, 0,
                        0, mask, din, cv.cin, n->ext[cv.level], nvox);
}

template <int D>
void conv_bwd_weights(srmg_unet *n, hipStream_t st, const ConvL &cv, const uint16_t *dz, const uint16_t *x, int B) {
    const int64_t nvox = (int64_t)B * n->V[cv.level];
    const int kx = cv.ktot + 1;
    int64_t nchunk = std::min<int64_t>(512, (nvox + 255) / 256);
    int64_t chunk = (nvox + nchunk - 1) / nchunk;
    chunk = (chunk + 31) / 32 * 32;
    nchunk = (nvox + chunk - 1) / chunk;
    const size_t need = (size_t)nchunk * cv.cout * kx * sizeof(float);
    if (n->partial.bytes < need) n->partial.alloc(need);
    dim3 grid((unsigned)((kx + 255) / 256), (unsigned)((cv.cout + 15) / 16), (unsigned)nchunk);
    const Ext &e = n->ext[cv.level];
    hipLaunchKernelGGL((k_wgrad<D>), grid, dim3(TPB), 0, st, dz, cv.cg, cv.cout, x, ilog2(cv.cin), cv.ktot, kx, e.n0,
                       e.n1, e.n2, nvox, chunk, n->partial.as<float>());
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(nblocks((int64_t)cv.cout * kx)), dim3(TPB), 0, st,
                       n->partial.as<float>(), (int)nchunk, cv.cout, kx, cv.ktot, n->gr.as<float>() + cv.woff,
                       n->gr.as<float>() + cv.boff);
}

template <int D>
void forward(srmg_unet *n, hipStream_t st, int B) {
    const int L = n->L;
    for (int l = 0; l < L; l++) {
        const uint16_t *in;
        if (l == 0) {
            in = n->X.v.as<uint16_t>();
        } else {
            const int64_t total = (int64_t)B * n->V[l] * (n->w[l - 1] / 8);
            hipLaunchKernelGGL((k_pool<D>), dim3(nblocks(total)), dim3(TPB), 0, st, n->E[l - 1].v.as<uint16_t>(),
                               n->P[l].v.as<uint16_t>(), n->w[l - 1] / 8, n->ext[l], n->ext[l - 1], total);
            in = n->P[l].v.as<uint16_t>();
        }
        conv_fwd<D>(n, st, n->conv[n->cA[l]], in, n->A[l].v.p, B);
        conv_fwd<D>(n, st, n->conv[n->cE[l]], n->A[l].v.as<uint16_t>(), n->E[l].v.p, B);
    }
    for (int l = L - 2; l >= 0; l--) {
        const uint16_t *src = (l == L - 2) ? n->E[L - 1].v.as<uint16_t>() : n->Dd[l + 1].v.as<uint16_t>();
        conv_fwd<D>(n, st, n->conv[n->cC[l]], src, n->C[l].v.p, B);
        const int64_t total = (int64_t)B * n->V[l] * (n->w[l] / 8);
        hipLaunchKernelGGL((k_up_add<D>), dim3(nblocks(total)), dim3(TPB), 0, st, n->C[l].v.as<uint16_t>(),
                           n->E[l].v.as<uint16_t>(), n->S[l].v.as<uint16_t>(), n->w[l] / 8, n->ext[l + 1], n->ext[l],
                           total);
        conv_fwd<D>(n, st, n->conv[n->cDd[l]], n->S[l].v.as<uint16_t>(), n->Dd[l].v.p, B);
    }
    conv_fwd<D>(n, st, n->conv[n->cY], n->Dd[0].v.as<uint16_t>(), n->Y.p, B);
}

template <int D>
void backward(srmg_unet *n, hipStream_t st, int B) {
    const int L = n->L;
    auto g16 = [](Tensor &t) { return t.g.as<uint16_t>(); };
    auto v16 = [](Tensor &t) { return t.v.as<uint16_t>(); };
    {
        const ConvL &cv = n->conv[n->cY];
        conv_bwd_weights<D>(n, st, cv, n->gY.as<uint16_t>(), v16(n->Dd[0]), B);
        conv_bwd_data<D>(n, st, cv, n->gY.as<uint16_t>(), g16(n->Dd[0]), v16(n->Dd[0]), B);
    }
    for (int l = 0; l <= L - 2; l++) {
        const ConvL &cd = n->conv[n->cDd[l]];
        conv_bwd_weights<D>(n, st, cd, g16(n->Dd[l]), v16(n->S[l]), B);
        conv_bwd_data<D>(n, st, cd, g16(n->Dd[l]), g16(n->S[l]), nullptr, B);
        const int64_t total = (int64_t)B * n->V[l + 1] * (n->w[l] / 8);
        hipLaunchKernelGGL((k_up_bwd<D>), dim3(nblocks(total)), dim3(TPB), 0, st, g16(n->S[l]), g16(n->C[l]),
                           n->w[l] / 8, n->ext[l + 1], n->ext[l], total);
        const ConvL &cc = n->conv[n->cC[l]];
        Tensor &src = (l == L - 2) ? n->E[L - 1] : n->Dd[l + 1];
        conv_bwd_weights<D>(n, st, cc, g16(n->C[l]), v16(src), B);
        conv_bwd_data<D>(n, st, cc, g16(n->C[l]), g16(src), v16(src), B);
    }
    for (int l = L - 1; l >= 0; l--) {
        const ConvL &ce = n->conv[n->cE[l]];
        conv_bwd_weights<D>(n, st, ce, g16(n->E[l]), v16(n->A[l]), B);
        conv_bwd_data<D>(n, st, ce, g16(n->E[l]), g16(n->A[l]), v16(n->A[l]), B);
        const ConvL &ca = n->conv[n->cA[l]];
        const uint16_t *in = (l == 0) ? n->X.v.as<uint16_t>() : v16(n->P[l]);
        conv_bwd_weights<D>(n, st, ca, g16(n->A[l]), in, B);
        if (l > 0) {
            conv_bwd_data<D>(n, st, ca, g16(n->A[l]), g16(n->P[l]), nullptr, B);
            const int64_t total = (int64_t)B * n->V[l - 1] * (n->w[l - 1] / 8);
            hipLaunchKernelGGL((k_pool_bwd<D>), dim3(nblocks(total)), dim3(TPB), 0, st, g16(n->S[l - 1]),
                               g16(n->P[l]), v16(n->E[l - 1]), g16(n->E[l - 1]), n->w[l - 1] / 8, n->ext[l],
                               n->ext[l - 1], total);
        }
    }
}

template <typename TF>
void features(srmg_unet *n, hipStream_t st, const TF *f, int B) {
    hipLaunchKernelGGL((k_absmax_part<TF>), dim3(ABSMAX_PARTS, B), dim3(TPB), 0, st, f, n->V[0],
                       n->absmax_part.as<double>());
    hipLaunchKernelGGL(k_absmax, dim3(B), dim3(TPB), 0, st, n->absmax_part.as<double>(), n->scale.as<double>());
    const int64_t nvox = (int64_t)B * n->V[0];
    if (n->D == 3)
        hipLaunchKernelGGL((k_features<3, TF>), dim3(nblocks(nvox)), dim3(TPB), 0, st, f, n->scale.as<double>(),
                           n->X.v.as<uint16_t>(), n->ext[0], nvox);
    else
        hipLaunchKernelGGL((k_features<2, TF>), dim3(nblocks(nvox)), dim3(TPB), 0, st, f, n->scale.as<double>(),
                           n->X.v.as<uint16_t>(), n->ext[0], nvox);
}

void run_forward(srmg_unet *n, hipStream_t st, int B) {
    if (!n->packed) pack(n, st);
    if (n->D == 3)
        forward<3>(n, st, B);
    else
        forward<2>(n, st, B);
}

// loss and gradient of one batch into n->gr (and n->loss)
void loss_grad(srmg_unet *n, const float *f, const float *u, int B) {
    hipStream_t st = n->stream;
    ensure_batch(n, B);
    features<float>(n, st, f, B);
    run_forward(n, st, B);
    const int64_t nvox = (int64_t)B * n->V[0];
    hipLaunchKernelGGL(k_mse<float>, dim3(LOSS_WGS), dim3(TPB), 0, st, n->Y.as<float>(), u, n->scale.as<double>(),
                       n->c_out, n->V[0], nvox, n->gY.as<uint16_t>(), n->lossp.as<double>());
    hipLaunchKernelGGL(k_loss_total, dim3(1), dim3(TPB), 0, st, n->lossp.as<double>(), LOSS_WGS, 1.0 / (double)nvox,
                       n->loss.as<double>());
    if (n->D == 3)
        backward<3>(n, st, B);
    else
        backward<2>(n, st, B);
}

// splitmix64 + Box-Muller: reproducible He-normal initialization
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};

srmg_unet *checked(srmg_unet *n) {
    if (!n) fail("null network");
    return n;
}

void check_batch(const srmg_unet *n, int B) {
    if (B < 1) fail("batch must be >= 1");
    if ((int64_t)B * n->V[0] > (int64_t)1 << 31) fail("batch too large for one launch");
}

}  // namespace

namespace srmg_guess {

void extents(const srmg_unet *net, int *D, int n[3]) {
    *D = net->D;
    n[0] = net->ext[0].n0;
    n[1] = net->ext[0].n1;
    n[2] = net->ext[0].n2;
}

bool matches(const srmg_unet *net, int geometry, int disc, double k2, const int8_t bc[27]) {
    return net->has_problem && net->geometry == geometry && net->disc == disc && net->k2 == k2 &&
           std::memcmp(net->bc, bc, 27) == 0;
}

void infer(srmg_unet *n, const double *f, double *u, int ngrid, hipStream_t st) {
    check_batch(n, ngrid);
    ensure_batch(n, ngrid);
    features<double>(n, st, f, ngrid);
    run_forward(n, st, ngrid);
    const int64_t nvox = (int64_t)ngrid * n->V[0];
    hipLaunchKernelGGL(k_output<double>, dim3(nblocks(nvox)), dim3(TPB), 0, st, n->Y.as<float>(),
                       n->scale.as<double>(), n->c_out, n->V[0], nvox, u);
}

}  // namespace srmg_guess

extern "C" {

int srmg_unet_create(srmg_unet **out, int D, const int *n, int levels, const int *widths, uint64_t seed) {
    return guarded([&] {
        if (!out || !n || !widths) fail("null argument");
        if (D != 2 && D != 3) fail("D must be 2 or 3");
        if (levels < 2 || levels > MAXLEV) fail("levels must be in 2.." + std::to_string(MAXLEV));
        for (int l = 0; l < levels; l++) {
            const int w = widths[l];
            if (w < 8 || w > 64 || (w & (w - 1))) fail("widths must be powers of two in 8..64");
        }
        for (int d = 0; d < D; d++)
            if (n[d] < 1 || n[d] % (1 << (levels - 1))) fail("grid extents must be divisible by 2^(levels-1)");
        auto net = std::make_unique<srmg_unet>();
        net->D = D;
        net->L = levels;
        net->ntap = D == 3 ? 27 : 9;
        net->w.assign(widths, widths + levels);
        for (int l = 0; l < levels; l++) {
            net->ext[l] = Ext{n[0] >> l, n[1] >> l, D == 3 ? (n[2] >> l) : 1};
            net->V[l] = (int64_t)net->ext[l].n0 * net->ext[l].n1 * net->ext[l].n2;
        }
        auto add = [&](int level, int cin, int cout, bool relu) {
            ConvL cv;
            cv.level = level;
            cv.cin = cin;
            cv.cout = cout;
            cv.cout_s = cout == 1 ? 1 : cout;
            cv.cg = std::max(cout, 8);
            cv.relu = relu;
            cv.ktot = net->ntap * cin;
            cv.kp = round_up(cv.ktot, 32);
            cv.rows = (cout + 15) / 16;
            if (cv.rows == 3) cv.rows = 4;
            cv.ktot_b = net->ntap * cv.cg;
            cv.kp_b = round_up(cv.ktot_b, 32);
            cv.rows_b = (cin + 15) / 16;
            if (cv.rows_b == 3) cv.rows_b = 4;
            cv.woff = net->nparam;
            net->nparam += (size_t)cout * cv.ktot;
            cv.boff = net->nparam;
            net->nparam += cout;
            net->conv.push_back(std::move(cv));
            return (int)net->conv.size() - 1;
        };
        net->cA.resize(levels);
        net->cE.resize(levels);
        net->cC.assign(levels, -1);
        net->cDd.assign(levels, -1);
        for (int l = 0; l < levels; l++) {
            net->cA[l] = add(l, l == 0 ? FEAT : widths[l - 1], widths[l], true);
            net->cE[l] = add(l, widths[l], widths[l], true);
        }
        for (int l = levels - 2; l >= 0; l--) {
            net->cC[l] = add(l + 1, widths[l + 1], widths[l], false);
            net->cDd[l] = add(l, widths[l], widths[l], true);
        }
        net->cY = add(0, widths[0], 1, false);
        for (ConvL &cv : net->conv) {
            cv.wf.alloc((size_t)cv.rows * 16 * cv.kp * 2);
            cv.wb.alloc((size_t)cv.rows_b * 16 * cv.kp_b * 2);
        }
        std::vector<float> h(net->nparam, 0.f);
        Rng rng{seed};
        for (const ConvL &cv : net->conv) {
            const double sd = std::sqrt((cv.relu ? 2.0 : 1.0) / (double)cv.ktot);
            for (size_t i = 0; i < (size_t)cv.cout * cv.ktot; i++) h[cv.woff + i] = (float)(sd * rng.normal());
        }
        net->p.alloc(net->nparam * sizeof(float));
        net->gr.alloc(net->nparam * sizeof(float));
        net->m.alloc(net->nparam * sizeof(float));
        net->v.alloc(net->nparam * sizeof(float));
        HIPCHK(hipStreamCreate(&net->stream));
        HIPCHK(hipMemcpy(net->p.p, h.data(), net->nparam * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(net->m.p, 0, net->nparam * sizeof(float)));
        HIPCHK(hipMemset(net->v.p, 0, net->nparam * sizeof(float)));
        HIPCHK(hipMemset(net->gr.p, 0, net->nparam * sizeof(float)));
        *out = net.release();
    });
}

int srmg_unet_destroy(srmg_unet *n) {
    return guarded([&] {
        if (!n) return;
        if (n->stream) (void)hipStreamSynchronize(n->stream);
        hipStream_t s = n->stream;
        delete n;
        if (s) (void)hipStreamDestroy(s);
    });
}

int srmg_unet_param_count(const srmg_unet *n, int64_t *count) {
    return guarded([&] {
        if (!n || !count) fail("null argument");
        *count = (int64_t)n->nparam;
    });
}

int srmg_unet_get_params(srmg_unet *n, float *host) {
    return guarded([&] {
        checked(n);
        if (!host) fail("null argument");
        HIPCHK(hipStreamSynchronize(n->stream));
        HIPCHK(hipMemcpy(host, n->p.p, n->nparam * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int srmg_unet_set_params(srmg_unet *n, const float *host) {
    return guarded([&] {
        checked(n);
        if (!host) fail("null argument");
        HIPCHK(hipStreamSynchronize(n->stream));
        HIPCHK(hipMemcpy(n->p.p, host, n->nparam * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(n->m.p, 0, n->nparam * sizeof(float)));
        HIPCHK(hipMemset(n->v.p, 0, n->nparam * sizeof(float)));
        n->steps = 0;
        n->packed = false;
    });
}

int srmg_unet_set_output_scale(srmg_unet *n, double c_out) {
    return guarded([&] {
        checked(n);
        if (!(c_out > 0.0) || !std::isfinite(c_out)) fail("output scale must be positive and finite");
        n->c_out = c_out;
    });
}

int srmg_unet_get_output_scale(const srmg_unet *n, double *c_out) {
    return guarded([&] {
        if (!n || !c_out) fail("null argument");
        *c_out = n->c_out;
    });
}

int srmg_unet_set_problem(srmg_unet *n, const srmg_op *op) {
    return guarded([&] {
        checked(n);
        if (!op) fail("null operator");
        n->has_problem = true;
        n->geometry = op->geometry;
        n->disc = op->disc;
        n->k2 = op->disc ? op->k2 : 0.0;
        std::memcpy(n->bc, op->bc, 27);
    });
}

int srmg_unet_forward(srmg_unet *n, const float *f, float *u, int batch) {
    return guarded([&] {
        checked(n);
        if (!f || !u) fail("null argument");
        check_batch(n, batch);
        ensure_batch(n, batch);
        features<float>(n, n->stream, f, batch);
        run_forward(n, n->stream, batch);
        const int64_t nvox = (int64_t)batch * n->V[0];
        hipLaunchKernelGGL(k_output<float>, dim3(nblocks(nvox)), dim3(TPB), 0, n->stream, n->Y.as<float>(),
                           n->scale.as<double>(), n->c_out, n->V[0], nvox, u);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(n->stream));
    });
}

int srmg_unet_loss_grad(srmg_unet *n, const float *f, const float *u, int batch, double *loss, float *grad) {
    return guarded([&] {
        checked(n);
        if (!f || !u) fail("null argument");
        check_batch(n, batch);
        if (!n->packed) pack(n, n->stream);
        loss_grad(n, f, u, batch);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(n->stream));
        if (loss) HIPCHK(hipMemcpy(loss, n->loss.p, sizeof(double), hipMemcpyDeviceToHost));
        if (grad) HIPCHK(hipMemcpy(grad, n->gr.p, n->nparam * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int srmg_unet_train_step(srmg_unet *n, const float *f, const float *u, int batch, float lr, double *loss) {
    return guarded([&] {
        checked(n);
        if (!f || !u) fail("null argument");
        check_batch(n, batch);
        if (!(lr > 0.f)) fail("learning rate must be positive");
        if (!n->packed) pack(n, n->stream);
        loss_grad(n, f, u, batch);
        n->steps++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)n->steps), c2 = 1.f - std::pow(b2, (float)n->steps);
        hipLaunchKernelGGL(k_adam, dim3(nblocks((int64_t)n->nparam)), dim3(TPB), 0, n->stream, n->p.as<float>(),
                           n->gr.as<float>(), n->m.as<float>(), n->v.as<float>(), (int64_t)n->nparam, lr, b1, b2,
                           1e-8f, c1, c2);
        pack(n, n->stream);
        HIPCHK(hipGetLastError());
        if (loss) {
            HIPCHK(hipMemcpyAsync(loss, n->loss.p, sizeof(double), hipMemcpyDeviceToHost, n->stream));
            HIPCHK(hipStreamSynchronize(n->stream));
        }
    });
}

int srmg_unet_synchronize(srmg_unet *n) {
    return guarded([&] {
        checked(n);
        HIPCHK(hipStreamSynchronize(n->stream));
    });
}

}  // extern "C"
