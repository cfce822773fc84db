// HIP kernels of the multigrid path (gfx950). Every kernel is templated on the
// dimension D, the block side M and the element type T (double or float).
//
// Floating point: this translation unit is compiled with -ffp-contract=off so
// that products and sums round exactly as written; fma() appears only where the
// reference writes muladd (src/multigrid.jl:176-177). Stencil terms are summed
// in the order the reference's @stencil macro emits them (src/laplacian.jl:25-49),
// so the FP64 kernels reproduce the CPU restatement bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.h"

namespace srmg {

constexpr int WG = 256;  // work-group size of the elementwise/stencil kernels

template <typename T>
struct LevelArgs {
    int blk0;      // first device block of the level
    int nblk;      // blocks in the level
    T s;           // s_l = s / 4^(l-1)
    T cc;          // stencil centre = diagelm: -2D, or -2D - s_l k2 (Helmholtz)
    T nwc;         // -(omega / diagelm)
    T sg[6];       // ghost = sg[face] * mirror cell (GeneralHomogeneousBoundaryConditions)
    int rdim;      // radial dimension (cylindrical) or -1
    const T *wm;   // (2J-2)/(2J-1) by 1-based global radial index J at this level
    const T *wp;   // 2J/(2J-1)
};

template <typename T> __device__ __forceinline__ T fma_t(T a, T b, T c);
template <> __device__ __forceinline__ double fma_t<double>(double a, double b, double c) { return fma(a, b, c); }
template <> __device__ __forceinline__ float fma_t<float>(float a, float b, float c) { return fmaf(a, b, c); }

// Value of cell (c0,c1,c2) + dir e_dim of block blk, crossing the block face if
// needed: same-level neighbour cell, mirror (depth-1 ghost = sign * own cell,
// src/boundary.jl:45-47 with src/blockfield.jl:513-521), or the coarse-fine
// ghost buffer.
template <int D, int M, typename T>
__device__ __forceinline__ T nbv(const T *__restrict__ x, const T *__restrict__ ib,
                                 const int *__restrict__ nbr, int blk, int c0, int c1, int c2,
                                 int dim, int dir, const T *sg) {
    using LY = Lay<D, M>;
    int n0 = c0 + (dim == 0 ? dir : 0);
    int n1 = c1 + (dim == 1 ? dir : 0);
    int n2 = c2 + (dim == 2 ? dir : 0);
    int cd = dim == 0 ? n0 : (dim == 1 ? n1 : n2);
    if (cd >= 0 && cd < M) return x[(size_t)blk * LY::V + LY::sidx(n0, n1, n2)];
    int f = 2 * dim + (dir > 0 ? 1 : 0);
    int t = nbr[blk * 2 * D + f];
    if (t >= 0) {
        int w = dir > 0 ? 0 : M - 1;
        if (dim == 0) n0 = w; else if (dim == 1) n1 = w; else n2 = w;
        return x[(size_t)t * LY::V + LY::sidx(n0, n1, n2)];
    }
    if (t == NB_BOUNDARY) return sg[f] * x[(size_t)blk * LY::V + LY::sidx(c0, c1, c2)];
    return ib[(size_t)nb_slot(t) * LY::F + LY::faceidx(dim, c0, c1, c2)];
}

// applystencil(u, s, I, J, disc, geom, Val(:lhs)) for the order-2 operators
// (src/laplacian.jl:78-128, 278-303; src/helmholtz.jl:17-100).
template <int D, int M, typename T>
__device__ __forceinline__ T stencil(const T *__restrict__ x, const T *__restrict__ ib,
                                     const int *__restrict__ nbr, int blk, int c0, int c1, int c2,
                                     T own, const LevelArgs<T> &a, T wmr, T wpr) {
    T v = T(0);
#pragma unroll
    for (int d = D - 1; d >= 0; d--) {
        T w = (d == a.rdim) ? wmr : T(1);
        v += w * nbv<D, M, T>(x, ib, nbr, blk, c0, c1, c2, d, -1, a.sg);
    }
    v += a.cc * own;
#pragma unroll
    for (int d = 0; d < D; d++) {
        T w = (d == a.rdim) ? wpr : T(1);
        v += w * nbv<D, M, T>(x, ib, nbr, blk, c0, c1, c2, d, +1, a.sg);
    }
    return v;
}

template <int D, int M, typename T>
__device__ __forceinline__ void radial_weights(const LevelArgs<T> &a, const int *__restrict__ coord,
                                               int blk, int c0, int c1, T &wm, T &wp) {
    wm = T(1);
    wp = T(1);
    if (a.rdim >= 0) {
        int cr = a.rdim == 0 ? c0 : c1;
        int J = M * (coord[blk * D + a.rdim] - 1) + cr + 1;
        wm = a.wm[J];
        wp = a.wp[J];
    }
}

// One colour of gauss_seidel! (src/multigrid.jl:153-180) on every block of a
// level: pass 0 updates the cells with odd global index sum, pass 1 the rest.
// Same-colour cells never read each other, so the update is in place.
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_gs_pass(T *__restrict__ u, const T *__restrict__ b,
                                                const T *__restrict__ ib, const int *__restrict__ nbr,
                                                const int *__restrict__ coord, LevelArgs<T> a, int pass) {
    using LY = Lay<D, M>;
    int gt = blockIdx.x * blockDim.x + threadIdx.x;
    int lb = gt / LY::H;
    if (lb >= a.nblk) return;
    int blk = a.blk0 + lb;
    int q = pass * LY::H + (gt - lb * LY::H);
    int c0, c1, c2;
    LY::decode(q, c0, c1, c2);
    size_t o = (size_t)blk * LY::V + q;
    T own = u[o];
    T wm, wp;
    radial_weights<D, M, T>(a, coord, blk, c0, c1, wm, wp);
    T Lu = stencil<D, M, T>(u, ib, nbr, blk, c0, c1, c2, own, a, wm, wp);
    T v = fma_t<T>(a.s, b[o], Lu);
    u[o] = fma_t<T>(a.nwc, v, own);
}

// residual! over a level (src/multigrid.jl:309-323): r = -(b + L u / s)
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_residual(T *__restrict__ r, const T *__restrict__ u,
                                                 const T *__restrict__ b, const T *__restrict__ ib,
                                                 const int *__restrict__ nbr, const int *__restrict__ coord,
                                                 LevelArgs<T> a) {
    using LY = Lay<D, M>;
    int gt = blockIdx.x * blockDim.x + threadIdx.x;
    int lb = gt / LY::V;
    if (lb >= a.nblk) return;
    int blk = a.blk0 + lb;
    int q = gt - lb * LY::V;
    int c0, c1, c2;
    LY::decode(q, c0, c1, c2);
    size_t o = (size_t)blk * LY::V + q;
    T wm, wp;
    radial_weights<D, M, T>(a, coord, blk, c0, c1, wm, wp);
    T Lu = stencil<D, M, T>(u, ib, nbr, blk, c0, c1, c2, u[o], a, wm, wp);
    r[o] = -(b[o] + Lu / a.s);
}

// restrict! over the Child edges of a level (src/restrict.jl:5-37): each
// coarse cell of the parent subblock = mean of its 2^D children, summed with
// the first index fastest.
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_restrict(T *__restrict__ x, const int *__restrict__ parent,
                                                 const int *__restrict__ sub, int blk0, int nblk) {
    using LY = Lay<D, M>;
    constexpr int NS = LY::V / LY::NC;  // coarse cells per subblock
    int gt = blockIdx.x * blockDim.x + threadIdx.x;
    int lb = gt / NS;
    if (lb >= nblk) return;
    int fb = blk0 + lb;
    int j = gt - lb * NS;
    int j0 = j % LY::MH, j1 = (j / LY::MH) % LY::MH, j2 = (D == 3) ? j / (LY::MH * LY::MH) : 0;
    T f = T(0);
#pragma unroll
    for (int q = 0; q < LY::NC; q++) {
        int i0 = q & 1, i1 = (q >> 1) & 1, i2 = (q >> 2) & 1;
        f += x[(size_t)fb * LY::V + LY::sidx(2 * j0 + i0, 2 * j1 + i1, D == 3 ? 2 * j2 + i2 : 0)];
    }
    int pb = parent[fb];
    int sb = sub[fb];
    int X0 = (sb & 1) * LY::MH + j0, X1 = ((sb >> 1) & 1) * LY::MH + j1;
    int X2 = (D == 3) ? ((sb >> 2) & 1) * LY::MH + j2 : 0;
    x[(size_t)pb * LY::V + LY::sidx(X0, X1, X2)] = f / T(LY::NC);
}

// residual_postrestrict_level! (src/multigrid.jl:112-119, 331-344): FAS
// right-hand side at every coarse cell covered by a block of this level,
// b = -(r + L_{l-1} u / s_{l-1}); `a` describes the coarse level.
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_postrestrict(T *__restrict__ bout, const T *__restrict__ u,
                                                     const T *__restrict__ r, const T *__restrict__ ib,
                                                     const int *__restrict__ nbr, const int *__restrict__ coord,
                                                     const int *__restrict__ parent, const int *__restrict__ sub,
                                                     int fblk0, int nfine, LevelArgs<T> a) {
    using LY = Lay<D, M>;
    constexpr int NS = LY::V / LY::NC;
    int gt = blockIdx.x * blockDim.x + threadIdx.x;
    int lb = gt / NS;
    if (lb >= nfine) return;
    int fb = fblk0 + lb;
    int j = gt - lb * NS;
    int j0 = j % LY::MH, j1 = (j / LY::MH) % LY::MH, j2 = (D == 3) ? j / (LY::MH * LY::MH) : 0;
    int pb = parent[fb];
    int sb = sub[fb];
    int X0 = (sb & 1) * LY::MH + j0, X1 = ((sb >> 1) & 1) * LY::MH + j1;
    int X2 = (D == 3) ? ((sb >> 2) & 1) * LY::MH + j2 : 0;
    size_t o = (size_t)pb * LY::V + LY::sidx(X0, X1, X2);
    T wm, wp;
    radial_weights<D, M, T>(a, coord, pb, X0, X1, wm, wp);
    T Lu = stencil<D, M, T>(u, ib, nbr, pb, X0, X1, X2, u[o], a, wm, wp);
    bout[o] = -(r[o] + Lu / a.s);
}

// copyto! / diffto! over the valid cells of a block range (src/blockfield.jl:557-576)
template <typename T>
__global__ __launch_bounds__(WG) void k_copy(T *__restrict__ dst, const T *__restrict__ src, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <typename T>
__global__ __launch_bounds__(WG) void k_diff(T *__restrict__ dst, const T *__restrict__ a,
                                             const T *__restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = a[i] - b[i];
}

template <typename T>
__global__ __launch_bounds__(WG) void k_fill(T *__restrict__ dst, T v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

// Linear interpolation weights (src/interp.jl:7-11)
template <int D, typename T> __device__ __forceinline__ T w0() { return D == 2 ? T(0.5) : T(0.25); }
template <typename T> __device__ __forceinline__ T w1() { return T(0.25); }

// interp_add_level!(u, u1, child) (src/interp.jl:102-141, 185-193): every
// fine cell += linear interpolation of the parent's correction u1.
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_interp_add(T *__restrict__ u, const T *__restrict__ u1,
                                                   const T *__restrict__ ib1, const int *__restrict__ nbr,
                                                   const int *__restrict__ parent, const int *__restrict__ sub,
                                                   int blk0, int nblk, LevelArgs<T> a) {
    using LY = Lay<D, M>;
    int gt = blockIdx.x * blockDim.x + threadIdx.x;
    int lb = gt / LY::V;
    if (lb >= nblk) return;
    int fb = blk0 + lb;
    int q = gt - lb * LY::V;
    int x0, x1, x2;
    LY::decode(q, x0, x1, x2);
    int pb = parent[fb];
    int sb = sub[fb];
    int J0 = (sb & 1) * LY::MH + (x0 >> 1);
    int J1 = ((sb >> 1) & 1) * LY::MH + (x1 >> 1);
    int J2 = (D == 3) ? ((sb >> 2) & 1) * LY::MH + (x2 >> 1) : 0;
    T f = w0<D, T>() * u1[(size_t)pb * LY::V + LY::sidx(J0, J1, J2)];
    f += w1<T>() * nbv<D, M, T>(u1, ib1, nbr, pb, J0, J1, J2, 0, 2 * (x0 & 1) - 1, a.sg);
    f += w1<T>() * nbv<D, M, T>(u1, ib1, nbr, pb, J0, J1, J2, 1, 2 * (x1 & 1) - 1, a.sg);
    if (D == 3) f += w1<T>() * nbv<D, M, T>(u1, ib1, nbr, pb, J0, J1, J2, 2, 2 * (x2 & 1) - 1, a.sg);
    size_t o = (size_t)fb * LY::V + q;
    u[o] = f + u[o];
}

// fill_ghost_interp! (src/connectivity.jl:251-258, src/blockfield.jl:416-434)
// for the depth-1 ghost layer: slot k holds the ghost slab of one coarse-fine
// face. The coarse layer adjacent to the face feeds the fine ghost layer next
// to the fine block's valid cells, i.e. child offset 1 on a low face and 0 on a
// high face in the face dimension. `a` describes the coarse level (mirror signs).
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_fill_interp(T *__restrict__ ib, const T *__restrict__ x,
                                                    const int *__restrict__ nbr, const int4 *__restrict__ slots,
                                                    int slot0, int nslot, LevelArgs<T> a) {
    using LY = Lay<D, M>;
    int gt = blockIdx.x * blockDim.x + threadIdx.x;
    int ls = gt / LY::F;
    if (ls >= nslot) return;
    int k = slot0 + ls;
    int t = gt - ls * LY::F;
    int4 sl = slots[k];  // (fine, face, coarse, subblock code)
    int dim = sl.y >> 1, dir = (sl.y & 1) ? 1 : -1;
    int pb = sl.z, sb = sl.w;
    int c[3] = {0, 0, 0};
    // tangential coordinates in increasing dimension order
    int ta = t % M, tb = t / M;
    if (dim == 0) { c[1] = ta; c[2] = tb; }
    else if (dim == 1) { c[0] = ta; c[2] = tb; }
    else { c[0] = ta; c[1] = tb; }
    int J[3] = {0, 0, 0}, i[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; d++) {
        int sbd = (sb >> d) & 1;
        if (d == dim) {
            J[d] = sbd * LY::MH + (dir < 0 ? LY::MH - 1 : 0);
            i[d] = dir < 0 ? 1 : 0;
        } else {
            J[d] = sbd * LY::MH + (c[d] >> 1);
            i[d] = c[d] & 1;
        }
    }
    T f = w0<D, T>() * x[(size_t)pb * LY::V + LY::sidx(J[0], J[1], J[2])];
#pragma unroll
    for (int j = 0; j < D; j++)
        f += w1<T>() * nbv<D, M, T>(x, ib, nbr, pb, J[0], J[1], J[2], j, 2 * i[j] - 1, a.sg);
    ib[(size_t)k * LY::F + t] = f;
}

// Squared-residual sum over the leaf blocks of a level, one partial per work
// group (deterministic LDS tree); `wgt` = 2^(-D(l-1)). Equals what resnorm!
// obtains by restricting squared residuals down to level 1 (src/multigrid.jl:391-417).
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_resnorm(const T *__restrict__ u, const T *__restrict__ b,
                                                const T *__restrict__ ib, const int *__restrict__ nbr,
                                                const int *__restrict__ coord, const int *__restrict__ leaves,
                                                int nleaf, LevelArgs<T> a, double wgt,
                                                double *__restrict__ partial) {
    using LY = Lay<D, M>;
    __shared__ double red[WG];
    int gt = blockIdx.x * blockDim.x + threadIdx.x;
    int lb = gt / LY::V;
    double val = 0.0;
    if (lb < nleaf) {
        int blk = leaves[lb];
        int q = gt - lb * LY::V;
        int c0, c1, c2;
        LY::decode(q, c0, c1, c2);
        size_t o = (size_t)blk * LY::V + q;
        T wm, wp;
        radial_weights<D, M, T>(a, coord, blk, c0, c1, wm, wp);
        T Lu = stencil<D, M, T>(u, ib, nbr, blk, c0, c1, c2, u[o], a, wm, wp);
        double r = (double)(-(b[o] + Lu / a.s));
        val = r * r;
    }
    red[threadIdx.x] = val;
    __syncthreads();
#pragma unroll
    for (int w = WG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] * wgt;
}

// Sum n partials in a fixed order (single work group).
__global__ __launch_bounds__(WG) void k_sum(const double *__restrict__ partial, int n, double *__restrict__ out) {
    __shared__ double red[WG];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += WG) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = WG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// ---- exact level-1 solve (FP64) ----

// level-1 valid cells -> uniform grid (column-major, dimension 1 fastest), times c
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_grid_gather(double *__restrict__ g, const T *__restrict__ x,
                                                    const int *__restrict__ coord, int blk0, int nblk,
                                                    int N0, int N1, double c) {
    using LY = Lay<D, M>;
    int gt = blockIdx.x * blockDim.x + threadIdx.x;
    int lb = gt / LY::V;
    if (lb >= nblk) return;
    int blk = blk0 + lb;
    int q = gt - lb * LY::V;
    int c0, c1, c2;
    LY::decode(q, c0, c1, c2);
    size_t gi = (size_t)(M * (coord[blk * D] - 1) + c0) +
                (size_t)N0 * ((size_t)(M * (coord[blk * D + 1] - 1) + c1) +
                              (D == 3 ? (size_t)N1 * (size_t)(M * (coord[blk * D + 2] - 1) + c2) : 0));
    g[gi] = c * (double)x[(size_t)blk * LY::V + q];
}

template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_grid_scatter(T *__restrict__ x, const double *__restrict__ g,
                                                     const int *__restrict__ coord, int blk0, int nblk,
                                                     int N0, int N1) {
    using LY = Lay<D, M>;
    int gt = blockIdx.x * blockDim.x + threadIdx.x;
    int lb = gt / LY::V;
    if (lb >= nblk) return;
    int blk = blk0 + lb;
    int q = gt - lb * LY::V;
    int c0, c1, c2;
    LY::decode(q, c0, c1, c2);
    size_t gi = (size_t)(M * (coord[blk * D] - 1) + c0) +
                (size_t)N0 * ((size_t)(M * (coord[blk * D + 1] - 1) + c1) +
                              (D == 3 ? (size_t)N1 * (size_t)(M * (coord[blk * D + 2] - 1) + c2) : 0));
    x[(size_t)blk * LY::V + q] = (T)g[gi];
}

// out = Mx applied along one axis (extent n, stride `stride`) of a grid of `total` values
__global__ __launch_bounds__(WG) void k_axis(const double *__restrict__ Mx, int n, size_t stride, size_t total,
                                             const double *__restrict__ in, double *__restrict__ out) {
    size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total) return;
    size_t inner = gi % stride;
    size_t i = (gi / stride) % n;
    size_t outer = gi / (stride * n);
    const double *src = in + outer * stride * n + inner;
    const double *row = Mx + i * n;
    double acc = 0.0;
    for (int k = 0; k < n; k++) acc += row[k] * src[(size_t)k * stride];
    out[gi] = acc;
}

__global__ __launch_bounds__(WG) void k_eig_scale(double *__restrict__ g, const double *__restrict__ l0,
                                                  const double *__restrict__ l1, const double *__restrict__ l2,
                                                  int N0, int N1, int D, double shift, size_t total) {
    size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total) return;
    size_t r = gi;
    double ev = shift;
    ev += l0[r % N0]; r /= N0;
    ev += l1[r % N1]; r /= N1;
    if (D == 3) ev += l2[r];
    g[gi] /= ev;
}

// ---- mixed precision helpers ----

// FP64 residual of level l written as the FP32 right-hand side of the
// correction equation (b' = -r = b + L u / s), plus leaf squared sums.
template <int D, int M>
__global__ __launch_bounds__(WG) void k_defect(float *__restrict__ be, const double *__restrict__ u,
                                               const double *__restrict__ b, const double *__restrict__ ib,
                                               const int *__restrict__ nbr, const int *__restrict__ coord,
                                               LevelArgs<double> a) {
    using LY = Lay<D, M>;
    int gt = blockIdx.x * blockDim.x + threadIdx.x;
    int lb = gt / LY::V;
    if (lb >= a.nblk) return;
    int blk = a.blk0 + lb;
    int q = gt - lb * LY::V;
    int c0, c1, c2;
    LY::decode(q, c0, c1, c2);
    size_t o = (size_t)blk * LY::V + q;
    double wm, wp;
    radial_weights<D, M, double>(a, coord, blk, c0, c1, wm, wp);
    double Lu = stencil<D, M, double>(u, ib, nbr, blk, c0, c1, c2, u[o], a, wm, wp);
    be[o] = (float)(b[o] + Lu / a.s);
}

// u += e over all cells and interpolation-ghost slots (FP64 += FP32)
__global__ __launch_bounds__(WG) void k_add_correction(double *__restrict__ u, const float *__restrict__ e, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] += (double)e[i];
}

}  // namespace srmg
