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
    // several problems in one launch: (s, cc, nwc) of block blk are
    // pp[3 pix[blk] + 0..2]; nullptr when the context holds one problem
    const T *pp;
    const int *pix;
    // processing order of the level's sibling groups in the column sweeps:
    // work group slot i takes group gord[i]; nullptr: group i
    const int *gord;
};

template <typename T>
struct BP {
    T s, cc, nwc;
};

template <typename T>
__device__ __forceinline__ BP<T> bparams(const LevelArgs<T> &a, int blk) {
    if (a.pp) {
        const T *q = a.pp + 3 * a.pix[blk];
        return {q[0], q[1], q[2]};
    }
    return {a.s, a.cc, a.nwc};
}

template <typename T> __device__ __forceinline__ T fma_t(T a, T b, T c);
template <> __device__ __forceinline__ double fma_t<double>(double a, double b, double c) { return fma(a, b, c); }
template <> __device__ __forceinline__ float fma_t<float>(float a, float b, float c) { return fmaf(a, b, c); }

// f(integral_constant<int, i>) for i = 0 .. N-1, unrolled at compile time
template <int N, int I = 0, class Fn>
__device__ __forceinline__ void static_for(Fn &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>());
        static_for<N, I + 1>(f);
    }
}

// Work-group barrier that orders LDS accesses only: global loads still in
// flight stay in flight across it (__syncthreads would wait for them).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

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
                                     T own, const LevelArgs<T> &a, T cc, T wmr, T wpr) {
    T v = T(0);
#pragma unroll
    for (int d = D - 1; d >= 0; d--) {
        T w = (d == a.rdim) ? wmr : T(1);
        v += w * nbv<D, M, T>(x, ib, nbr, blk, c0, c1, c2, d, -1, a.sg);
    }
    v += cc * own;
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
    BP<T> bp = bparams(a, blk);
    T Lu = stencil<D, M, T>(u, ib, nbr, blk, c0, c1, c2, own, a, bp.cc, wm, wp);
    T v = fma_t<T>(bp.s, b[o], Lu);
    u[o] = fma_t<T>(bp.nwc, v, own);
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
    BP<T> bp = bparams(a, blk);
    T Lu = stencil<D, M, T>(u, ib, nbr, blk, c0, c1, c2, u[o], a, bp.cc, wm, wp);
    r[o] = -(b[o] + Lu / bp.s);
}

// Linear interpolation weights (src/interp.jl:7-11)
template <int D, typename T> __device__ __forceinline__ T w0() { return D == 2 ? T(0.5) : T(0.25); }
template <typename T> __device__ __forceinline__ T w1() { return T(0.25); }

template <typename T>
struct Bc27 {
    T v[27];
};

// ---- LaplacianDiscretization{D,4}: the 9- and 19-point operators ----
//
// Their stencils reach corner (2D) and edge (3D) ghosts, so the order-4 path
// keeps every block's depth-1 ghosts explicitly: a frame of (M + 2)^D values
// per block (the interior unused), filled as fill_ghost! fills the reference's
// ghost layers (src/connectivity.jl:155-258, with BoxStencil records): the
// copies from same-level neighbours, then the boundary mirrors (which read
// the frame where the mirrored cell is itself a ghost), then the coarse-fine
// interpolations. box[blk * 3^D + k] for the direction k = sum (dir_d + 1) 3^d
// is the same-level neighbour, -(1 + mask) where the direction leaves the
// domain in the dimensions of mask, or -16 - j for coarse-fine slot j
// (coarse block, subblock code).
template <int D, int M>
struct Frame {
    static constexpr int S = M + 2;
    static constexpr int P = D == 3 ? S * S * S : S * S;
    __host__ __device__ static constexpr int at(int c0, int c1, int c2) {
        return D == 3 ? ((c2 + 1) * S + (c1 + 1)) * S + (c0 + 1) : (c1 + 1) * S + (c0 + 1);
    }
};

// the value at cell c of block blk, c at most one cell outside the block
template <int D, int M, typename T>
__device__ __forceinline__ T boxv(const T *__restrict__ x, const T *__restrict__ g, int blk, int c0, int c1,
                                  int c2) {
    using LY = Lay<D, M>;
    const bool in = c0 >= 0 && c0 < M && c1 >= 0 && c1 < M && (D == 2 || (c2 >= 0 && c2 < M));
    return in ? x[(size_t)blk * LY::V + LY::sidx(c0, c1, c2)]
              : g[(size_t)blk * Frame<D, M>::P + Frame<D, M>::at(c0, c1, D == 3 ? c2 : 0)];
}

// the direction of a frame cell (-1, 0, 1 per dimension), its index k, and
// whether it is a ghost
template <int D, int M>
__device__ __forceinline__ bool frame_cell(int q, int *c, int *dir, int &k) {
    constexpr int S = M + 2;
    c[0] = q % S - 1;
    c[1] = (q / S) % S - 1;
    c[2] = D == 3 ? q / (S * S) - 1 : 0;
    bool ghost = false;
    k = 0;
    int m = 1;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        dir[d] = 0;
        if (d < D) {
            dir[d] = c[d] < 0 ? -1 : (c[d] >= M ? 1 : 0);
            ghost |= dir[d] != 0;
            k += (dir[d] + 1) * m;
            m *= 3;
        }
    }
    return ghost;
}

// The value fill_ghost_copy! then fill_ghost_bnd! give ghost cell c of blk (c
// at most one cell outside the block, modified). A ghost facing a same-level
// neighbour is that neighbour's cell. A ghost beyond the domain is a[bnd]
// times the cell mirrored in the dimensions that leave the domain; when other
// dimensions stay outside, the mirrored cell is a ghost of a direction inside
// the domain: a same-level neighbour's cell (read from the neighbour, the value
// the copy writes) or a coarse-fine ghost (read from the frame, as the
// interpolation that refills it comes after both fills). Coarse-fine ghosts
// themselves are read from the frame.
template <int D, int M, typename T>
__device__ __forceinline__ T ghost_value(const T *__restrict__ x, const T *__restrict__ g,
                                         const int *__restrict__ box, const Bc27<T> &bc, int blk, int *c) {
    using LY = Lay<D, M>;
    constexpr int P = Frame<D, M>::P, NB3 = D == 3 ? 27 : 9;
    int dir[3] = {0, 0, 0}, k = 0, m = 1;
#pragma unroll
    for (int d = 0; d < D; d++) {
        dir[d] = c[d] < 0 ? -1 : (c[d] >= M ? 1 : 0);
        k += (dir[d] + 1) * m;
        m *= 3;
    }
    const int e = box[blk * NB3 + k];
    if (e >= 0) {
#pragma unroll
        for (int d = 0; d < D; d++) c[d] -= dir[d] * M;
        return x[(size_t)e * LY::V + LY::sidx(c[0], c[1], c[2])];
    }
    if (e < -8) return g[(size_t)blk * P + Frame<D, M>::at(c[0], c[1], D == 3 ? c[2] : 0)];
    const int mask = -1 - e;
    int bi = 0, k2 = 0;
    m = 1;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const bool bd = (mask >> d) & 1;
        bi += ((bd ? dir[d] : 0) + 1) * m;
        if (bd) c[d] = dir[d] < 0 ? -1 - c[d] : 2 * M - 1 - c[d];
        k2 += ((bd ? 0 : dir[d]) + 1) * m;
        m *= 3;
    }
    T v;
    const int e2 = box[blk * NB3 + k2];
    if (k2 == NB3 / 2) {
        v = x[(size_t)blk * LY::V + LY::sidx(c[0], c[1], c[2])];
    } else if (e2 >= 0) {
#pragma unroll
        for (int d = 0; d < D; d++)
            if (!((mask >> d) & 1)) c[d] -= dir[d] * M;
        v = x[(size_t)e2 * LY::V + LY::sidx(c[0], c[1], c[2])];
    } else {
        v = g[(size_t)blk * P + Frame<D, M>::at(c[0], c[1], D == 3 ? c[2] : 0)];
    }
    return bc.v[bi] * v;
}

// frame cell q of block blk: same-level copy or boundary mirror (coarse-fine
// frame cells are k_frame_interp's)
template <int D, int M, typename T>
__device__ __forceinline__ void frame_copybnd_cell(T *__restrict__ g, const T *__restrict__ x,
                                                   const int *__restrict__ box, int blk, int q, const Bc27<T> &bc) {
    constexpr int P = Frame<D, M>::P, NB3 = D == 3 ? 27 : 9;
    int c[3], dir[3], k;
    if (!frame_cell<D, M>(q, c, dir, k)) return;
    if (box[blk * NB3 + k] < -8) return;  // coarse-fine: k_frame_interp
    g[(size_t)blk * P + q] = ghost_value<D, M, T>(x, g, box, bc, blk, c);
}

// fill_ghost_copy! then fill_ghost_bnd! in one pass (ghost_value); one pass
// needs no ordering between blocks and gives the bits of the two in order
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_frame_copybnd(T *__restrict__ g, const T *__restrict__ x,
                                                      const int *__restrict__ box, int blk0, int nblk, Bc27<T> bc) {
    constexpr int P = Frame<D, M>::P;
    // 64-bit: a level's nblk * P frame cells can pass 2^31 before its nblk * V cells do
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t lb = gt / P;
    if (lb >= nblk) return;
    frame_copybnd_cell<D, M, T>(g, x, box, blk0 + (int)lb, (int)(gt - lb * P), bc);
}

// reads of cells at most one cell outside a block for the order-4 stencils:
// from the frame as last filled (FrameRead), or live (LiveRead: the values a
// copy-then-boundary fill of the frame would give now)
template <int D, int M, typename T>
struct FrameRead {
    const T *__restrict__ x;
    const T *__restrict__ g;
    __device__ __forceinline__ T operator()(int blk, int c0, int c1, int c2) const {
        return boxv<D, M, T>(x, g, blk, c0, c1, c2);
    }
};
template <int D, int M, typename T>
struct LiveRead {
    const T *__restrict__ x;
    const T *__restrict__ g;
    const int *__restrict__ box;
    const Bc27<T> &bc;
    __device__ __forceinline__ T operator()(int blk, int c0, int c1, int c2) const {
        using LY = Lay<D, M>;
        if (c0 >= 0 && c0 < M && c1 >= 0 && c1 < M && (D == 2 || (c2 >= 0 && c2 < M)))
            return x[(size_t)blk * LY::V + LY::sidx(c0, c1, c2)];
        int c[3] = {c0, c1, D == 3 ? c2 : 0};
        return ghost_value<D, M, T>(x, g, box, bc, blk, c);
    }
};

// fill_ghost_interp! (src/connectivity.jl:251-258, src/interp.jl:23-60): a
// ghost facing a coarse block is the linear interpolation of the coarse cell
// under it with its axis neighbours (coarse ghosts from the coarse frame),
// the offsets towards this block in the dimensions of the direction
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_frame_interp(T *__restrict__ g, const T *__restrict__ x,
                                                     const int *__restrict__ box, const int2 *__restrict__ slots,
                                                     int blk0, int nblk) {
    using LY = Lay<D, M>;
    constexpr int P = Frame<D, M>::P, NB3 = D == 3 ? 27 : 9;
    // 64-bit: a level's nblk * P frame cells can pass 2^31 before its nblk * V cells do
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t lb = gt / P;
    if (lb >= nblk) return;
    const int q = (int)(gt - lb * P), blk = blk0 + (int)lb;
    int c[3], dir[3], k;
    if (!frame_cell<D, M>(q, c, dir, k)) return;
    const int e = box[blk * NB3 + k];
    if (e > -16) return;
    const int2 sl = slots[-16 - e];
    const int pb = sl.x, sb = sl.y;
    int J[3] = {0, 0, 0}, i[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; d++) {
        const int sbd = (sb >> d) & 1;
        if (dir[d] != 0) {
            J[d] = sbd * LY::MH + (dir[d] < 0 ? LY::MH - 1 : 0);
            i[d] = dir[d] < 0 ? 1 : 0;
        } else {
            J[d] = sbd * LY::MH + (c[d] >> 1);
            i[d] = c[d] & 1;
        }
    }
    T f = w0<D, T>() * x[(size_t)pb * LY::V + LY::sidx(J[0], J[1], J[2])];
#pragma unroll
    for (int j = 0; j < D; j++) {
        int K[3] = {J[0], J[1], J[2]};
        K[j] += 2 * i[j] - 1;
        f += w1<T>() * boxv<D, M, T>(x, g, pb, K[0], K[1], K[2]);
    }
    g[(size_t)blk * P + q] = f;
}

// interp_add_level! (src/interp.jl:102-115) reading the coarse ghosts from
// the frames
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_interp_add4(T *__restrict__ u, const T *__restrict__ u1,
                                                    const T *__restrict__ g1, const int *__restrict__ parent,
                                                    const int *__restrict__ sub, int blk0, int nblk) {
    using LY = Lay<D, M>;
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int lb = gt / LY::V;
    if (lb >= nblk) return;
    const int fb = blk0 + lb, q = gt - lb * LY::V;
    int x0, x1, x2;
    LY::decode(q, x0, x1, x2);
    const int pb = parent[fb], sb = sub[fb];
    const int J0 = (sb & 1) * LY::MH + (x0 >> 1);
    const int J1 = ((sb >> 1) & 1) * LY::MH + (x1 >> 1);
    const int J2 = (D == 3) ? ((sb >> 2) & 1) * LY::MH + (x2 >> 1) : 0;
    T f = w0<D, T>() * u1[(size_t)pb * LY::V + LY::sidx(J0, J1, J2)];
    f += w1<T>() * boxv<D, M, T>(u1, g1, pb, J0 + 2 * (x0 & 1) - 1, J1, J2);
    f += w1<T>() * boxv<D, M, T>(u1, g1, pb, J0, J1 + 2 * (x1 & 1) - 1, J2);
    if (D == 3) f += w1<T>() * boxv<D, M, T>(u1, g1, pb, J0, J1, J2 + 2 * (x2 & 1) - 1);
    const size_t o = (size_t)fb * LY::V + q;
    u[o] = f + u[o];
}

// applystencil (src/laplacian.jl:133-141, 305-320): the literal's nonzero
// terms in column-major order, then one division by 6
template <int D, int M, typename T, typename A>
__device__ __forceinline__ T stencil4(const A &at, int blk, int c0, int c1, int c2) {
    T v = T(0);
    constexpr int NK = D == 3 ? 1 : 0;
#pragma unroll
    for (int k = -NK; k <= NK; k++)
#pragma unroll
        for (int j = -1; j <= 1; j++)
#pragma unroll
            for (int i = -1; i <= 1; i++) {
                const int n = (i != 0) + (j != 0) + (k != 0);
                const int a = D == 2 ? (n == 0 ? -20 : (n == 1 ? 4 : 1))
                                     : (k == 0 ? (n == 0 ? -24 : (n == 1 ? 2 : 1)) : (n == 1 ? 2 : (n == 2 ? 1 : 0)));
                if (a == 0) continue;
                v += T(a) * at(blk, c0 + i, c1 + j, c2 + k);
            }
    return v / T(6);
}

// LaplacianDiscretization{2,4} x CylindricalGeometry{N+1} (src/laplacian.jl:
// 156-201) at 1-based global radial index i: the 3x3 literal's entries
// (integer ratios rounded once), visited column-major, divided by 6; and its
// diagelm
template <int M, typename T, typename A>
__device__ __forceinline__ T stencil4cyl(const A &at, int blk, int c0, int c1, int N, long long i) {
    const long long q = 2 * i - 1;
    const T lo = (T)(2 * (i - 1)) / (T)q, hi = (T)(2 * i) / (T)q;
    const T pm = (T)(4 * (-2 + 9 * i - 16 * i * i + 8 * i * i * i)) / (T)(q * q * q);
    const T pp = (T)(4 * (1 + i - 8 * i * i + 8 * i * i * i)) / (T)(q * q * q);
    const T cc = (T)(-16 * (1 - 5 * i + 5 * i * i)) / (T)((1 - 2 * i) * (1 - 2 * i));
    T w[3][3];
    if (N == 0) {
        w[0][0] = lo; w[0][1] = pm; w[0][2] = lo;
        w[1][0] = T(4); w[1][1] = cc; w[1][2] = T(4);
        w[2][0] = hi; w[2][1] = pp; w[2][2] = hi;
    } else {
        w[0][0] = lo; w[0][1] = T(4); w[0][2] = hi;
        w[1][0] = pm; w[1][1] = cc; w[1][2] = pp;
        w[2][0] = lo; w[2][1] = T(4); w[2][2] = hi;
    }
    T v = T(0);
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
        for (int a = 0; a < 3; a++) v += w[a][b] * at(blk, c0 + a - 1, c1 + b - 1, 0);
    return v / T(6);
}

template <typename T>
__device__ __forceinline__ T diag4cyl(long long i) {
    return (T)(-16 * (1 - 5 * i + 5 * i * i)) / (T)((1 - 2 * i) * (1 - 2 * i)) / T(6);
}

// the order-4 operator at cell (c0, c1, c2) of blk: Cartesian (rd < 0) or
// cylindrical with radial dimension rd (global radial index from coord)
template <int D, int M, typename T, typename A>
__device__ __forceinline__ T apply4a(const A &at, const int *__restrict__ coord, int blk, int c0, int c1, int c2,
                                     int rd, T &diag) {
    if constexpr (D == 2) {
        if (rd >= 0) {
            const long long i = (long long)(rd == 0 ? c0 : c1) + 1 + (long long)M * (coord[blk * D + rd] - 1);
            diag = diag4cyl<T>(i);
            return stencil4cyl<M, T>(at, blk, c0, c1, rd, i);
        }
    }
    diag = D == 2 ? T(-20.0 / 6.0) : T(-4);
    return stencil4<D, M, T>(at, blk, c0, c1, c2);
}
template <int D, int M, typename T>
__device__ __forceinline__ T apply4(const T *__restrict__ x, const T *__restrict__ g,
                                    const int *__restrict__ coord, int blk, int c0, int c1, int c2, int rd,
                                    T &diag) {
    return apply4a<D, M, T>(FrameRead<D, M, T>{x, g}, coord, blk, c0, c1, c2, rd, diag);
}

// the FP32 right-hand side of the order-4 correction equation of a mixed
// precision step: be = b + L4 u / s in FP64, rounded once
template <int D, int M>
__global__ __launch_bounds__(WG) void k_defect4(float *__restrict__ be, const double *__restrict__ u,
                                                const double *__restrict__ b, const double *__restrict__ g,
                                                const int *__restrict__ coord, int rd, int blk0, int nblk,
                                                double s) {
    using LY = Lay<D, M>;
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int lb = gt / LY::V;
    if (lb >= nblk) return;
    const int blk = blk0 + lb, q = gt - lb * LY::V;
    int c0, c1, c2;
    LY::decode(q, c0, c1, c2);
    const size_t o = (size_t)blk * LY::V + q;
    double c;
    const double Lu = apply4<D, M, double>(u, g, coord, blk, c0, c1, c2, rd, c);
    be[o] = (float)(b[o] + Lu / s);
}

// gauss_seidel!(..., ::FourColors{P1,P2}) (src/multigrid.jl:206-232) over a
// level: colour cells (x = P1, y = P2 mod 2 from the first valid cell; in 3D
// both shifted by the 1-based third index with G = 2 ghost layers, the
// reference's depth), u -= omega (Lu + s b) / diagelm; cells of one colour do
// not read each other
// LIVE: ghosts read live (LiveRead) instead of from the frame, which equals a
// pass that follows a copy-then-boundary fill: the cells one pass reads across
// a block face are of other colours, or, for a boundary mirror, the cell being
// updated itself (read before its write), so none changes during the pass
// colour cell q of block blk in the FourColors{P1,P2} pass
template <int D, int M, typename T, bool LIVE>
__device__ __forceinline__ void gs4_cell(T *__restrict__ u, const T *__restrict__ b, const T *__restrict__ g,
                                         const int *__restrict__ coord, int rd, int blk, int q, T s, T omega, int P1,
                                         int P2, const int *__restrict__ box, const Bc27<T> &bc) {
    using LY = Lay<D, M>;
    constexpr int H2 = M / 2;
    const int i0 = q % H2, i1 = (q / H2) % H2, z = D == 3 ? q / (H2 * H2) : 0;
    const int p1 = D == 3 ? (P1 + z + 3) & 1 : P1, p2 = D == 3 ? (P2 + z + 3) & 1 : P2;
    const int x0 = 2 * i0 + p1, x1 = 2 * i1 + p2;
    const size_t o = (size_t)blk * LY::V + LY::sidx(x0, x1, z);
    T c, Lu;
    if constexpr (LIVE)
        Lu = apply4a<D, M, T>(LiveRead<D, M, T>{u, g, box, bc}, coord, blk, x0, x1, z, rd, c);
    else
        Lu = apply4a<D, M, T>(FrameRead<D, M, T>{u, g}, coord, blk, x0, x1, z, rd, c);
    u[o] = u[o] - omega * (Lu + s * b[o]) / c;
}

template <int D, int M, typename T, bool LIVE>
__global__ __launch_bounds__(WG) void k_gs4_pass(T *__restrict__ u, const T *__restrict__ b,
                                                 const T *__restrict__ g, const int *__restrict__ coord, int rd,
                                                 int blk0, int nblk, T s, T omega, int P1, int P2,
                                                 const int *__restrict__ box, Bc27<T> bc) {
    constexpr int H2 = M / 2, NCC = D == 3 ? H2 * H2 * M : H2 * H2;
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int lb = gt / NCC;
    if (lb >= nblk) return;
    gs4_cell<D, M, T, LIVE>(u, b, g, coord, rd, blk0 + lb, gt - lb * NCC, s, omega, P1, P2, box, bc);
}


// residual! (src/multigrid.jl:309-323) with the order-4 operator
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_resid4(T *__restrict__ r, const T *__restrict__ u, const T *__restrict__ b,
                                               const T *__restrict__ g, const int *__restrict__ coord, int rd,
                                               int blk0, int nblk, T s) {
    using LY = Lay<D, M>;
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int lb = gt / LY::V;
    if (lb >= nblk) return;
    const int blk = blk0 + lb, q = gt - lb * LY::V;
    int c0, c1, c2;
    LY::decode(q, c0, c1, c2);
    const size_t o = (size_t)blk * LY::V + q;
    T c;
    const T Lu = apply4<D, M, T>(u, g, coord, blk, c0, c1, c2, rd, c);
    r[o] = -(b[o] + Lu / s);
}

// residual_postrestrict_level! (src/multigrid.jl:112-119, 331-344) with the
// order-4 operator on the coarse cells covered by the fine blocks
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_postrestrict4(T *__restrict__ bout, const T *__restrict__ u,
                                                      const T *__restrict__ r, const T *__restrict__ g,
                                                      const int *__restrict__ coord, int rd,
                                                      const int *__restrict__ parent, const int *__restrict__ sub,
                                                      int fblk0, int nfine, T s) {
    using LY = Lay<D, M>;
    constexpr int NS = LY::V / LY::NC;
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int lb = gt / NS;
    if (lb >= nfine) return;
    const int fb = fblk0 + lb, j = gt - lb * NS;
    const int j0 = j % LY::MH, j1 = (j / LY::MH) % LY::MH, j2 = (D == 3) ? j / (LY::MH * LY::MH) : 0;
    const int pb = parent[fb], sb = sub[fb];
    const int X0 = (sb & 1) * LY::MH + j0, X1 = ((sb >> 1) & 1) * LY::MH + j1;
    const int X2 = (D == 3) ? ((sb >> 2) & 1) * LY::MH + j2 : 0;
    const size_t o = (size_t)pb * LY::V + LY::sidx(X0, X1, X2);
    T c;
    const T Lu = apply4<D, M, T>(u, g, coord, pb, X0, X1, X2, rd, c);
    bout[o] = -(r[o] + Lu / s);
}

// rhs! (src/multigrid.jl:363-388) of the compact 4th-order Laplacian
// (LaplacianDiscretization{D,4}, src/laplacian.jl:142-149, 167-175, 187-195,
// 322-337): b1 = (1/12) [0 1 0; 1 8 1; 0 1 0] b in 2D (the radial weights of
// the cylindrical forms on the radial neighbours), the 7-point [1 .. 6 .. 1]
// in 3D; the @stencil order of terms is the order-2 stencil's with centre
// weight 12 - 2D, then one division by 12.
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_rhs4(T *__restrict__ b1, const T *__restrict__ b, const T *__restrict__ ib,
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
    T v = stencil<D, M, T>(b, ib, nbr, blk, c0, c1, c2, b[o], a, T(12 - 2 * D), wm, wp);
    b1[o] = v / T(12);
}

// restrict! over the Child edges of a level (src/restrict.jl:5-37): each
// coarse cell of the parent subblock = mean of its 2^D children, summed with
// the first index fastest. xf holds the fine level, xc the coarse one.
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_restrict(T *__restrict__ xc, const T *__restrict__ xf,
                                                 const int *__restrict__ parent, const int *__restrict__ sub,
                                                 int blk0, int nblk) {
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
        f += xf[(size_t)fb * LY::V + LY::sidx(2 * j0 + i0, 2 * j1 + i1, D == 3 ? 2 * j2 + i2 : 0)];
    }
    int pb = parent[fb];
    int sb = sub[fb];
    int X0 = (sb & 1) * LY::MH + j0, X1 = ((sb >> 1) & 1) * LY::MH + j1;
    int X2 = (D == 3) ? ((sb >> 2) & 1) * LY::MH + j2 : 0;
    xc[(size_t)pb * LY::V + LY::sidx(X0, X1, X2)] = f / T(LY::NC);
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
    BP<T> bp = bparams(a, pb);
    T Lu = stencil<D, M, T>(u, ib, nbr, pb, X0, X1, X2, u[o], a, bp.cc, wm, wp);
    bout[o] = -(r[o] + Lu / bp.s);
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

// r = a + b: a stream kernel of the solver's own size, for bandwidth calibration
template <typename T>
__global__ __launch_bounds__(WG) void k_triad(T *__restrict__ r, const T *__restrict__ a, const T *__restrict__ b,
                                              size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = a[i] + b[i];
}

// the same with 16-byte loads and nontemporal 16-byte stores per lane (n a
// multiple of 16 / sizeof(T)): the strongest plain stream form, the
// bandwidth ceiling of the calibration
template <typename T>
__global__ __launch_bounds__(WG) void k_triad16(T *__restrict__ r, const T *__restrict__ a, const T *__restrict__ b,
                                                size_t n) {
    constexpr int K = 16 / sizeof(T);
    typedef T V __attribute__((ext_vector_type(K)));
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * K;
    if (i >= n) return;
    const V x = *reinterpret_cast<const V *>(a + i), y = *reinterpret_cast<const V *>(b + i);
    __builtin_nontemporal_store(x + y, reinterpret_cast<V *>(r + i));
}

template <typename T>
__global__ __launch_bounds__(WG) void k_fill(T *__restrict__ dst, T v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}


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

// applydelta!'s field transfer (src/refinement.jl:339-370; interp! and
// restrict! of src/streamer.jl:191-221) from the old block layout (uo, its
// coarse-fine slabs ibo and neighbour table nbro) into the new one (un).
// Blocks that keep their (level, coordinate): moved to their new position;
// cp[k] = (old block, new block).
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_regrid_copy(T *__restrict__ un, const T *__restrict__ uo,
                                                    const int2 *__restrict__ cp, int n) {
    using LY = Lay<D, M>;
    const size_t gt = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = (int)(gt / LY::V);
    if (k >= n) return;
    const int q = (int)(gt - (size_t)k * LY::V);
    const int2 p = cp[k];
    un[(size_t)p.y * LY::V + q] = uo[(size_t)p.x * LY::V + q];
}

// New blocks: interp! (src/interp.jl:23-60, func = _first) from the parent's
// old values and its depth-1 ghosts, in k_interp_add's order of terms;
// ip[k] = (new block, old parent, subblock code).
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_regrid_interp(T *__restrict__ un, const T *__restrict__ uo,
                                                      const T *__restrict__ ibo, const int *__restrict__ nbro,
                                                      const int4 *__restrict__ ip, int n, LevelArgs<T> a) {
    using LY = Lay<D, M>;
    const size_t gt = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = (int)(gt / LY::V);
    if (k >= n) return;
    const int q = (int)(gt - (size_t)k * LY::V);
    int x0, x1, x2;
    LY::decode(q, x0, x1, x2);
    const int4 it = ip[k];
    const int pb = it.y, sb = it.z;
    const int J0 = (sb & 1) * LY::MH + (x0 >> 1);
    const int J1 = ((sb >> 1) & 1) * LY::MH + (x1 >> 1);
    const int J2 = (D == 3) ? ((sb >> 2) & 1) * LY::MH + (x2 >> 1) : 0;
    T f = w0<D, T>() * uo[(size_t)pb * LY::V + LY::sidx(J0, J1, J2)];
    f += w1<T>() * nbv<D, M, T>(uo, ibo, nbro, pb, J0, J1, J2, 0, 2 * (x0 & 1) - 1, a.sg);
    f += w1<T>() * nbv<D, M, T>(uo, ibo, nbro, pb, J0, J1, J2, 1, 2 * (x1 & 1) - 1, a.sg);
    if (D == 3) f += w1<T>() * nbv<D, M, T>(uo, ibo, nbro, pb, J0, J1, J2, 2, 2 * (x2 & 1) - 1, a.sg);
    un[(size_t)it.x * LY::V + q] = f;
}

// Removed blocks: restrict! (src/restrict.jl:5-26) into the parent's
// subblock, in k_restrict's order; rp[k] = (old block, new parent, subblock
// code).
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_regrid_restrict(T *__restrict__ un, const T *__restrict__ uo,
                                                        const int4 *__restrict__ rp, int n) {
    using LY = Lay<D, M>;
    constexpr int NS = LY::V / LY::NC;
    const size_t gt = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = (int)(gt / NS);
    if (k >= n) return;
    const int j = (int)(gt - (size_t)k * NS);
    const int j0 = j % LY::MH, j1 = (j / LY::MH) % LY::MH, j2 = (D == 3) ? j / (LY::MH * LY::MH) : 0;
    const int4 it = rp[k];
    T f = T(0);
#pragma unroll
    for (int c = 0; c < LY::NC; c++) {
        const int i0 = c & 1, i1 = (c >> 1) & 1, i2 = (c >> 2) & 1;
        f += uo[(size_t)it.x * LY::V + LY::sidx(2 * j0 + i0, 2 * j1 + i1, D == 3 ? 2 * j2 + i2 : 0)];
    }
    const int sb = it.z;
    const int X0 = (sb & 1) * LY::MH + j0, X1 = ((sb >> 1) & 1) * LY::MH + j1;
    const int X2 = (D == 3) ? ((sb >> 2) & 1) * LY::MH + j2 : 0;
    un[(size_t)it.y * LY::V + LY::sidx(X0, X1, X2)] = f / T(LY::NC);
}

// fill_ghost_interp! (src/connectivity.jl:251-258, src/blockfield.jl:416-434)
// for the depth-1 ghost layer: slot k holds the ghost slab of one coarse-fine
// face. The coarse layer adjacent to the face feeds the fine ghost layer next
// to the fine block's valid cells, i.e. child offset 1 on a low face and 0 on a
// high face in the face dimension. `a` describes the coarse level (mirror signs).
template <int D, int M, typename T>
__device__ __forceinline__ void fill_interp_one(int gt, T *__restrict__ ib, const T *__restrict__ x,
                                                const int *__restrict__ nbr, const int4 *__restrict__ slots,
                                                int slot0, int nslot, const LevelArgs<T> &a) {
    using LY = Lay<D, M>;
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

template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_fill_interp(T *__restrict__ ib, const T *__restrict__ x,
                                                    const int *__restrict__ nbr, const int4 *__restrict__ slots,
                                                    int slot0, int nslot, LevelArgs<T> a) {
    fill_interp_one<D, M, T>(blockIdx.x * blockDim.x + threadIdx.x, ib, x, nbr, slots, slot0, nslot, a);
}

template <typename T> struct Vec16;
template <> struct Vec16<float> { using V = float4; };
template <> struct Vec16<double> { using V = double2; };

// Two independent steps of the V-cycle ascent in one launch: the first
// nwg_diff work groups form dst = a1 - a2 over n values (k_diff; 16 bytes
// per lane, n a multiple of 16 / sizeof(T), the arrays 16-byte aligned; dst
// may be a2), the rest refresh nslot coarse-fine ghost slabs
// (k_fill_interp). The slabs read another level than the difference writes.
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_diff_fill(T *dst, const T *a1, const T *a2, size_t n, unsigned nwg_diff,
                                                  T *__restrict__ ib, const T *__restrict__ x,
                                                  const int *__restrict__ nbr, const int4 *__restrict__ slots,
                                                  int slot0, int nslot, LevelArgs<T> a) {
    if (blockIdx.x < nwg_diff) {
        using V = typename Vec16<T>::V;
        constexpr int K = sizeof(V) / sizeof(T);
        const size_t i = ((size_t)blockIdx.x * WG + threadIdx.x) * K;
        if (i < n) {
            V p = *reinterpret_cast<const V *>(a1 + i);
            const V q = *reinterpret_cast<const V *>(a2 + i);
            T *pv = reinterpret_cast<T *>(&p);
            const T *qv = reinterpret_cast<const T *>(&q);
#pragma unroll
            for (int k = 0; k < K; k++) pv[k] = pv[k] - qv[k];
            *reinterpret_cast<V *>(dst + i) = p;
        }
    } else {
        fill_interp_one<D, M, T>((blockIdx.x - nwg_diff) * WG + threadIdx.x, ib, x, nbr, slots, slot0, nslot, a);
    }
}

// acc + p[i0] + p[i0 + WG] + ... over i < end, added in that order, with
// four loads in flight at a time
__device__ __forceinline__ double strided_sum(const double *__restrict__ p, int i0, int end, double acc) {
    int i = i0;
    for (; i + 3 * WG < end; i += 4 * WG) {
        const double p0 = p[i], p1 = p[i + WG], p2 = p[i + 2 * WG], p3 = p[i + 3 * WG];
        acc += p0;
        acc += p1;
        acc += p2;
        acc += p3;
    }
    for (; i < end; i += WG) acc += p[i];
    return acc;
}

// First stage of a fixed-order sum: work group g sums the contiguous chunk
// [g * chunk, (g + 1) * chunk) of the n partials.
__global__ __launch_bounds__(WG) void k_sum_chunks(const double *__restrict__ partial, int n, int chunk,
                                                   double *__restrict__ out) {
    __shared__ double red[WG];
    int lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    const double acc = strided_sum(partial, lo + (int)threadIdx.x, hi, 0.0);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = WG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Work group g sums the ranges rng[g * nr .. (g + 1) * nr) of the partials
// (int2: first, count) in order into out[g]: the per-problem totals when one
// launch holds several problems.
__global__ __launch_bounds__(WG) void k_sum_ranges(const double *__restrict__ partial, const int2 *__restrict__ rng,
                                                   int nr, double *__restrict__ out) {
    __shared__ double red[WG];
    double acc = 0.0;
    for (int k = 0; k < nr; k++) {
        const int2 r = rng[blockIdx.x * nr + k];
        acc = strided_sum(partial, r.x + (int)threadIdx.x, r.x + r.y, acc);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = WG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
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

// ---- fused red-black sweep ----

// Work-group -> tile mapping that keeps consecutive tiles (Morton order) on one
// XCD: the hardware hands work group i to XCD i mod 8, so XCD x receives a
// contiguous range of tiles and neighbouring blocks meet in its L2.
__device__ __forceinline__ int xcd_remap(int i, int G) {
    constexpr int NX = 8;
    int x = i % NX, local = i / NX;
    int per = G / NX, rem = G % NX;
    return (x < rem) ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
}

// LDS tile of G^D blocks of M^D cells with a depth-1 halo: cells -1 .. G M.
// Rows hold R >= S entries (R > S pads rows against LDS bank conflicts).
// Helpers that work on one block of a tile take a pointer offset to that
// block's origin (the tile index is linear in the coordinates).
template <int D, int M, int G = 1, int R = G * M + 2>
struct Tile {
    static constexpr int S = G * M + 2;
    static_assert(R >= S, "row pitch below the row length");
    static constexpr int SZ = (D == 2) ? R * S : R * S * S;
    __device__ static __forceinline__ int at(int x, int y, int z) {
        return (x + 1) + R * ((y + 1) + (D == 3 ? S * (z + 1) : 0));
    }
    __device__ static __forceinline__ int at3(const int *c) { return at(c[0], c[1], D == 3 ? c[2] : 0); }
    __host__ __device__ static constexpr int stride(int d) { return d == 0 ? 1 : (d == 1 ? R : R * S); }
    // offset of block k of a G = 2 group (k's bit d = upper half along d)
    __device__ static __forceinline__ int origin(int k) {
        return M * ((k & 1) + R * (((k >> 1) & 1) + (D == 3 ? S * ((k >> 2) & 1) : 0)));
    }
};

// Per-block halo table used by the tiled kernels: the block's own 2D face
// entries, then for each face f the neighbour-table entries of the block across
// f on its 2(D-1) side faces (tangential dimensions in increasing order, low
// side first); unused when face f is not a same-level face.
template <int D>
struct Halo {
    static constexpr int ND = 2 * D;
    static constexpr int NS = 2 * (D - 1);
    static constexpr int N = ND + ND * NS;
};

// Stencil of the tile cell at p, all neighbours in the tile. RD is the radial
// dimension (cylindrical geometry) or -1; terms in the reference's order.
template <int D, int M, typename T, int RD, class TL = Tile<D, M>>
__device__ __forceinline__ T tstencil(const T *t, int p, T own, T cc, T wm, T wp) {
    T v = T(0);
#pragma unroll
    for (int d = D - 1; d >= 0; d--) {
        if (d == RD) v += wm * t[p - TL::stride(d)];
        else v += t[p - TL::stride(d)];
    }
    v += cc * own;
#pragma unroll
    for (int d = 0; d < D; d++) {
        if (d == RD) v += wp * t[p + TL::stride(d)];
        else v += t[p + TL::stride(d)];
    }
    return v;
}

template <int D, int M, typename T, int RD>
__device__ __forceinline__ void rweights(const LevelArgs<T> &a, const int *__restrict__ coord, int blk, int cr,
                                         T &wm, T &wp) {
    wm = T(1);
    wp = T(1);
    if (RD >= 0) {
        int J = M * (coord[blk * D + RD] - 1) + cr + 1;
        wm = a.wm[J];
        wp = a.wp[J];
    }
}

// Depth-1 halo cell number t of face F (a compile-time face) of block blk:
// its tile position, and its value (neighbour cell, mirror of the facing cell,
// or frozen slab).
template <int D, int M, int F, class TL = Tile<D, M>>
__device__ __forceinline__ int halo_at(int t) {
    constexpr int DIM = F >> 1, DIR = (F & 1) ? 1 : -1;
    int c[3] = {0, 0, 0};
    int ta = t % M, tb = t / M;
    if (DIM == 0) { c[1] = ta; c[2] = tb; }
    else if (DIM == 1) { c[0] = ta; c[2] = tb; }
    else { c[0] = ta; c[1] = tb; }
    c[DIM] = DIR > 0 ? M : -1;
    return TL::at(c[0], c[1], c[2]);
}

template <int D, int M, typename T, int F>
__device__ __forceinline__ T halo_value(int e, int blk, int t, const T *__restrict__ x, const T *__restrict__ ib,
                                        T sg) {
    using LY = Lay<D, M>;
    constexpr int DIM = F >> 1, DIR = (F & 1) ? 1 : -1;
    int c[3] = {0, 0, 0};
    int ta = t % M, tb = t / M;
    if (DIM == 0) { c[1] = ta; c[2] = tb; }
    else if (DIM == 1) { c[0] = ta; c[2] = tb; }
    else { c[0] = ta; c[1] = tb; }
    if (e >= 0) {
        c[DIM] = DIR > 0 ? 0 : M - 1;
        return x[(size_t)e * LY::V + LY::sidx(c[0], c[1], c[2])];
    } else if (e == NB_BOUNDARY) {
        c[DIM] = DIR > 0 ? M - 1 : 0;
        return sg * x[(size_t)blk * LY::V + LY::sidx(c[0], c[1], c[2])];
    }
    return ib[(size_t)nb_slot(e) * LY::F + t];
}

template <int D, int M, typename T, int F, class TL = Tile<D, M>>
__device__ __forceinline__ void halo_cell(T *tile, int e, int blk, int t, const T *__restrict__ x,
                                          const T *__restrict__ ib, T sg) {
    tile[halo_at<D, M, F, TL>(t)] = halo_value<D, M, T, F>(e, blk, t, x, ib, sg);
}

// face number f known only at run time (uniform per wave where the work
// layout allows it)
#define SRMG_FACE_SWITCH(f, EXPR)                \
    switch (f) {                                 \
        case 0: { constexpr int FF = 0; EXPR; } break; \
        case 1: { constexpr int FF = 1; EXPR; } break; \
        case 2: { constexpr int FF = 2; EXPR; } break; \
        case 3: { constexpr int FF = 3; EXPR; } break; \
        case 4: { constexpr int FF = (D == 3 ? 4 : 3); EXPR; } break; \
        default: { constexpr int FF = (D == 3 ? 5 : 3); EXPR; } break; \
    }

template <int D, int M, typename T, class TL = Tile<D, M>>
__device__ __forceinline__ void halo_switch(T *tl, int e, int f, int blk, int t, const T *__restrict__ x,
                                            const T *__restrict__ ib, const T *sg) {
    SRMG_FACE_SWITCH(f, (halo_cell<D, M, T, FF, TL>(tl, e, blk, t, x, ib, sg[FF])))
}

template <int D, int M, typename T>
__device__ __forceinline__ T halo_value_sw(int f, int e, int blk, int t, const T *__restrict__ x,
                                           const T *__restrict__ ib, const T *sg) {
    T v = T(0);
    SRMG_FACE_SWITCH(f, (v = halo_value<D, M, T, FF>(e, blk, t, x, ib, sg[FF])))
    return v;
}

template <int D, int M, class TL>
__device__ __forceinline__ int halo_at_sw(int f, int t) {
    int p = 0;
    SRMG_FACE_SWITCH(f, (p = halo_at<D, M, FF, TL>(t)))
    return p;
}

// Neighbour face-layer cells recomputed by the fused sweep: for the face
// normal to DIM on side dir, the neighbour's first-half-sweep cells of its
// layer facing this block, numbered tq in [0, F/2).
template <int D, int M, typename T, int DIM, class TL = Tile<D, M>>
struct RecFace {
    using LY = Lay<D, M>;
    using HT = Halo<D>;
    static constexpr int DA = DIM == 0 ? 1 : 0;
    static constexpr int DB = DIM == 2 ? 1 : 2;
    static constexpr int NS = 2 * (D - 1);

    __device__ static __forceinline__ void cell(int dir, int tq, int *cB) {
        constexpr int MH = LY::MH;
        int w = dir > 0 ? 0 : M - 1;
        int tb = (D == 3) ? tq / MH : 0;
        cB[0] = cB[1] = cB[2] = 0;
        cB[DIM] = w;
        cB[DA] = 2 * (tq % MH) + ((1 + w + tb + D) & 1);
        if (D == 3) cB[DB] = tb;
    }

    // operands outside the tile: one layer deeper in the neighbour, its
    // right-hand side, and its ghosts across its side faces. ZIN: the field
    // and its ghost slabs are zero (only the right-hand side is read).
    template <bool ZIN = false>
    __device__ static __forceinline__ void prefetch(int dir, int tq, int e, const int *gs, const T *__restrict__ x,
                                                    const T *__restrict__ b, const T *__restrict__ ib, T &rd, T &rb,
                                                    T *rg) {
        int cB[3];
        cell(dir, tq, cB);
        rb = b[(size_t)e * LY::V + LY::sidx(cB[0], cB[1], cB[2])];
        if constexpr (ZIN) {
            rd = T(0);
#pragma unroll
            for (int k = 0; k < NS; k++) rg[k] = T(0);
            return;
        }
        int c2[3] = {cB[0], cB[1], cB[2]};
        c2[DIM] = dir > 0 ? 1 : M - 2;
        rd = x[(size_t)e * LY::V + LY::sidx(c2[0], c2[1], c2[2])];
        // one guarded load per side face: the neighbour's cell (same level)
        // or the frozen slab (coarse-fine); nothing at a domain boundary or
        // off the edge of the face
#pragma unroll
        for (int k = 0; k < NS; k++) {
            const int dd = (k < 2) ? DA : DB;
            const int side = k & 1;
            const int g = gs[k];
            int c3[3] = {cB[0], cB[1], cB[2]};
            c3[dd] = side ? 0 : M - 1;
            const size_t o = g >= 0 ? (size_t)g * LY::V + LY::sidx(c3[0], c3[1], c3[2])
                                    : (size_t)max(nb_slot(g), 0) * LY::F + LY::faceidx(dd, cB[0], cB[1], cB[2]);
            const T *src = g >= 0 ? x : ib;
            rg[k] = T(0);
            if (cB[dd] == (side ? M - 1 : 0) && g != NB_BOUNDARY) rg[k] = src[o];
        }
    }

    // first-half-sweep value of the neighbour cell, written into this block's halo
    template <int RD>
    __device__ static __forceinline__ void compute(int dir, int tq, int e, const int *gs, T *tile,
                                                   const int *__restrict__ coord, const LevelArgs<T> &a,
                                                   const BP<T> &bp, T rd, T rb, const T *rg) {
        int cB[3];
        cell(dir, tq, cB);
        int ch[3] = {cB[0], cB[1], cB[2]};
        ch[DIM] = dir > 0 ? M : -1;
        const int ph = TL::at(ch[0], ch[1], ch[2]);
        const T xo = tile[ph];
        T wm, wp;
        rweights<D, M, T, RD>(a, coord, e, RD >= 0 ? cB[RD < 0 ? 0 : RD] : 0, wm, wp);
        T nbm[3], nbp[3];
        // across the shared face: this block's cell; one layer deeper: rd
        const int pa = TL::at(DIM == 0 ? (dir > 0 ? M - 1 : 0) : cB[0], DIM == 1 ? (dir > 0 ? M - 1 : 0) : cB[1],
                              DIM == 2 ? (dir > 0 ? M - 1 : 0) : cB[2]);
        if (dir > 0) { nbm[DIM] = tile[pa]; nbp[DIM] = rd; }
        else { nbm[DIM] = rd; nbp[DIM] = tile[pa]; }
        // along the face layer: in the halo; off the edge of the face: the
        // prefetched ghost, or the mirror at a domain boundary (the tile
        // position read for an off-edge cell is inside the tile and unused)
#pragma unroll
        for (int k = 0; k < NS; k++) {
            const int dd = (k < 2) ? DA : DB;
            const int side = k & 1;
            const int cc = cB[dd] + (side ? 1 : -1);
            const T vt = tile[ph + (side ? 1 : -1) * TL::stride(dd)];
            const T vg = (gs[k] == NB_BOUNDARY) ? a.sg[2 * dd + side] * xo : rg[k];
            const T v = (cc >= 0 && cc < M) ? vt : vg;
            if (side) nbp[dd] = v; else nbm[dd] = v;
        }
        T Lx = T(0);
#pragma unroll
        for (int dd = D - 1; dd >= 0; dd--) {
            if (dd == RD) Lx += wm * nbm[dd];
            else Lx += nbm[dd];
        }
        Lx += bp.cc * xo;
#pragma unroll
        for (int dd = 0; dd < D; dd++) {
            if (dd == RD) Lx += wp * nbp[dd];
            else Lx += nbp[dd];
        }
        tile[ph] = fma_t<T>(bp.nwc, fma_t<T>(bp.s, rb, Lx), xo);
    }
};

// One full gauss_seidel_iter! sweep (src/multigrid.jl:250-265: RedBlack{0} on
// all blocks, copy/boundary ghost fill, RedBlack{1}, fill) in one launch.
// Reads ui, writes uo (ping-pong, so neighbours always see the old values).
// Each work group stages BPW blocks with a depth-1 face halo in LDS. The
// second half-sweep needs the neighbours' face-layer cells after the first
// half-sweep; the work group recomputes them from the same inputs with the
// same operations as their owners, so the result is identical to the
// two-launch form. Boundary ghosts mirror the cell they face, and coarse-fine
// ghosts stay frozen during smoothing (src/multigrid.jl:256-263).
// Memory round trips before compute: (interior, b, halo table), then (halo
// cells, recompute operands). Faces are compile-time constants in unrolled
// loops, so work that is uniform per face carries no per-lane selects.
template <int D, int M, typename T, int BPW, int NT, int RD>
__global__ __launch_bounds__(NT) void k_gs_sweep(T *__restrict__ uo, const T *__restrict__ ui,
                                                 const T *__restrict__ b, const T *__restrict__ ib,
                                                 const int *__restrict__ htab, const int *__restrict__ coord,
                                                 LevelArgs<T> a) {
    using LY = Lay<D, M>;
    using TL = Tile<D, M>;
    using HT = Halo<D>;
    constexpr int F = LY::F, V = LY::V, H = LY::H, F2 = F / 2, ND = 2 * D, NS = 2 * (D - 1);
    static_assert((BPW * H) % NT == 0, "cells per colour must be a multiple of the work-group size");
    constexpr int CPT = BPW * H / NT;
    constexpr int NREC = BPW * ND * F2;
    constexpr int RPT = (NREC + NT - 1) / NT;
    constexpr int NHALO = BPW * ND * F;
    __shared__ T tile[BPW * TL::SZ];
    __shared__ int tab[BPW * HT::N];
    __shared__ BP<T> prm[BPW];
    const int lb0 = xcd_remap(blockIdx.x, gridDim.x) * BPW;
    const int tid = threadIdx.x;
    const int nloc = min(BPW, a.nblk - lb0);
    const int blk0 = a.blk0 + lb0;

    // round trip 1: interior cells, own right-hand side, halo tables, parameters
    T ux[CPT][2], bo[CPT], be[CPT];
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        int i = tid + k * NT, lb = i / H, r = i - lb * H;
        ux[k][0] = ux[k][1] = bo[k] = be[k] = T(0);
        if (lb < nloc) {
            size_t o = (size_t)(blk0 + lb) * V + r;
            ux[k][0] = ui[o];
            ux[k][1] = ui[o + H];
            bo[k] = b[o];
            be[k] = b[o + H];
        }
    }
    for (int i = tid; i < nloc * HT::N; i += NT) tab[i] = htab[(size_t)blk0 * HT::N + i];
    if (tid < nloc) prm[tid] = bparams(a, blk0 + tid);
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        int i = tid + k * NT, lb = i / H, r = i - lb * H;
        if (lb < nloc) {
            int x, y, z;
            LY::decode(r, x, y, z);
            tile[lb * TL::SZ + TL::at(x, y, z)] = ux[k][0];
            LY::decode(H + r, x, y, z);
            tile[lb * TL::SZ + TL::at(x, y, z)] = ux[k][1];
        }
    }
    __syncthreads();

    // round trip 2: halo cells (stored into the tile once every load of the
    // round trip is issued) ...
    constexpr int PH = (NHALO + NT - 1) / NT;
    T hv[PH];
#pragma unroll
    for (int j = 0; j < PH; j++) {
        const int i = tid + NT * j, lb = min(i / (ND * F), BPW - 1), rem = i - lb * (ND * F);
        const int f = min(rem / F, ND - 1), t = rem - f * F;
        hv[j] = T(0);
        if (i < NHALO && lb < nloc) hv[j] = halo_value_sw<D, M, T>(f, tab[lb * HT::N + f], blk0 + lb, t, ui, ib, a.sg);
    }
    // ... and the operands of the recomputed neighbour cells
    T rd[RPT], rb[RPT], rg[RPT][NS];
    int rj[RPT];
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        int j = tid + k * NT;
        rj[k] = -1;
        rd[k] = rb[k] = T(0);
        if (j >= NREC) continue;
        int lb = j / (ND * F2), rem = j - lb * (ND * F2);
        int f = rem / F2, tq = rem - f * F2;
        if (lb >= nloc) continue;
        const int *tb = tab + lb * HT::N, *gs = tb + HT::ND + f * HT::NS;
        int e = tb[f];
        if (e < 0) continue;
        rj[k] = j;
        int dim = f >> 1, dir = (f & 1) ? 1 : -1;
        if (dim == 0) RecFace<D, M, T, 0>::prefetch(dir, tq, e, gs, ui, b, ib, rd[k], rb[k], rg[k]);
        else if (dim == 1) RecFace<D, M, T, 1>::prefetch(dir, tq, e, gs, ui, b, ib, rd[k], rb[k], rg[k]);
        else RecFace<D, M, T, (D == 3 ? 2 : 1)>::prefetch(dir, tq, e, gs, ui, b, ib, rd[k], rb[k], rg[k]);
    }
#pragma unroll
    for (int j = 0; j < PH; j++) {
        const int i = tid + NT * j, lb = min(i / (ND * F), BPW - 1), rem = i - lb * (ND * F);
        const int f = min(rem / F, ND - 1), t = rem - f * F;
        if (i < NHALO && lb < nloc) tile[lb * TL::SZ + halo_at_sw<D, M, TL>(f, t)] = hv[j];
    }
    __syncthreads();

    // first half-sweep: cells with odd global index sum
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        int i = tid + k * NT, lb = i / H, r = i - lb * H;
        if (lb < nloc) {
            int x, y, z;
            LY::decode(r, x, y, z);
            int p = lb * TL::SZ + TL::at(x, y, z);
            T wm, wp;
            rweights<D, M, T, RD>(a, coord, blk0 + lb, RD == 0 ? x : y, wm, wp);
            const BP<T> bp = prm[lb];
            T Lu = tstencil<D, M, T, RD>(tile, p, ux[k][0], bp.cc, wm, wp);
            ux[k][0] = fma_t<T>(bp.nwc, fma_t<T>(bp.s, bo[k], Lu), ux[k][0]);
            tile[p] = ux[k][0];
        }
    }
    // the same half-sweep on the neighbours' face-layer cells
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        int j = rj[k];
        if (j < 0) continue;
        int lb = j / (ND * F2), rem = j - lb * (ND * F2);
        int f = rem / F2, tq = rem - f * F2;
        const int *tb = tab + lb * HT::N, *gs = tb + HT::ND + f * HT::NS;
        T *tl = tile + lb * TL::SZ;
        int e = tb[f];
        int dim = f >> 1, dir = (f & 1) ? 1 : -1;
        const BP<T> bp = prm[lb];
        if (dim == 0)
            RecFace<D, M, T, 0>::template compute<RD>(dir, tq, e, gs, tl, coord, a, bp, rd[k], rb[k], rg[k]);
        else if (dim == 1)
            RecFace<D, M, T, 1>::template compute<RD>(dir, tq, e, gs, tl, coord, a, bp, rd[k], rb[k], rg[k]);
        else
            RecFace<D, M, T, (D == 3 ? 2 : 1)>::template compute<RD>(dir, tq, e, gs, tl, coord, a, bp, rd[k], rb[k],
                                                                     rg[k]);
    }
    __syncthreads();

    // second half-sweep: the remaining cells
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        int i = tid + k * NT, lb = i / H, r = i - lb * H;
        if (lb < nloc) {
            int x, y, z;
            LY::decode(H + r, x, y, z);
            int p = lb * TL::SZ + TL::at(x, y, z);
            T wm, wp;
            rweights<D, M, T, RD>(a, coord, blk0 + lb, RD == 0 ? x : y, wm, wp);
            const BP<T> bp = prm[lb];
            T Lu = tstencil<D, M, T, RD>(tile, p, ux[k][1], bp.cc, wm, wp);
            size_t o = (size_t)(blk0 + lb) * V + r;
            uo[o] = ux[k][0];
            uo[o + H] = fma_t<T>(bp.nwc, fma_t<T>(bp.s, be[k], Lu), ux[k][1]);
        }
    }
}

// k_gs_sweep on sibling groups: the 2^D children of one parent (consecutive
// in the Morton order of a level, always created together) form one
// (2M)^D tile. The faces between siblings need no halo and no recompute; the
// group's exterior faces are handled per block exactly as in k_gs_sweep (own
// halo-table entries, neighbour recompute from the same inputs), so the result
// is bitwise the same. One group per work group.
template <int D, int M, typename T, int NT, int RD>
__global__ __launch_bounds__(NT) void k_gs_sweep_group(T *__restrict__ uo, const T *__restrict__ ui,
                                                       const T *__restrict__ b, const T *__restrict__ ib,
                                                       const int *__restrict__ htab, const int *__restrict__ coord,
                                                       LevelArgs<T> a) {
    using LY = Lay<D, M>;
    using TL = Tile<D, M, 2>;
    using HT = Halo<D>;
    constexpr int NB = 1 << D;
    constexpr int F = LY::F, V = LY::V, H = LY::H, F2 = F / 2, NS = 2 * (D - 1);
    static_assert((NB * H) % NT == 0, "cells per colour must be a multiple of the work-group size");
    constexpr int CPT = NB * H / NT;
    constexpr int NREC = NB * D * F2;  // exterior faces: D per block
    constexpr int RPT = (NREC + NT - 1) / NT;
    constexpr int NHALO = NB * D * F;
    __shared__ T tile[TL::SZ];
    __shared__ int tab[NB * HT::N];
    __shared__ BP<T> prm[NB];
    const int blk0 = a.blk0 + xcd_remap(blockIdx.x, gridDim.x) * NB;
    const int tid = threadIdx.x;

    T ux[CPT][2], bo[CPT], be[CPT];
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        const int i = tid + k * NT, lb = i / H, r = i - lb * H;
        const size_t o = (size_t)(blk0 + lb) * V + r;
        ux[k][0] = ui[o];
        ux[k][1] = ui[o + H];
        bo[k] = b[o];
        be[k] = b[o + H];
    }
    for (int i = tid; i < NB * HT::N; i += NT) tab[i] = htab[(size_t)blk0 * HT::N + i];
    if (tid < NB) prm[tid] = bparams(a, blk0 + tid);
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        const int i = tid + k * NT, lb = i / H, r = i - lb * H;
        int x, y, z;
        LY::decode(r, x, y, z);
        tile[TL::origin(lb) + TL::at(x, y, z)] = ux[k][0];
        LY::decode(H + r, x, y, z);
        tile[TL::origin(lb) + TL::at(x, y, z)] = ux[k][1];
    }
    __syncthreads();

    // halo of the exterior faces: block lb, dimension d, face side = its bit
    // d (stored into the tile once the recompute operands' loads are issued)
    constexpr int PH = (NHALO + NT - 1) / NT;
    T hv[PH];
#pragma unroll
    for (int j = 0; j < PH; j++) {
        const int i = tid + NT * j, lb = min(i / (D * F), NB - 1), rem = i - lb * (D * F);
        const int d = min(rem / F, D - 1), t = rem - d * F;
        const int f = 2 * d + ((lb >> d) & 1);
        hv[j] = T(0);
        if (i < NHALO) hv[j] = halo_value_sw<D, M, T>(f, tab[lb * HT::N + f], blk0 + lb, t, ui, ib, a.sg);
    }
    T rd[RPT], rb[RPT], rg[RPT][NS];
    int rj[RPT];
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int j = tid + k * NT;
        rj[k] = -1;
        rd[k] = rb[k] = T(0);
        if (j >= NREC) continue;
        const int lb = j / (D * F2), rem = j - lb * (D * F2);
        const int d = rem / F2, tq = rem - d * F2;
        const int f = 2 * d + ((lb >> d) & 1);
        const int *tb = tab + lb * HT::N, *gs = tb + HT::ND + f * HT::NS;
        const int e = tb[f];
        if (e < 0) continue;
        rj[k] = j;
        const int dir = (f & 1) ? 1 : -1;
        if (d == 0) RecFace<D, M, T, 0, TL>::prefetch(dir, tq, e, gs, ui, b, ib, rd[k], rb[k], rg[k]);
        else if (d == 1) RecFace<D, M, T, 1, TL>::prefetch(dir, tq, e, gs, ui, b, ib, rd[k], rb[k], rg[k]);
        else RecFace<D, M, T, (D == 3 ? 2 : 1), TL>::prefetch(dir, tq, e, gs, ui, b, ib, rd[k], rb[k], rg[k]);
    }
#pragma unroll
    for (int j = 0; j < PH; j++) {
        const int i = tid + NT * j, lb = min(i / (D * F), NB - 1), rem = i - lb * (D * F);
        const int d = min(rem / F, D - 1), t = rem - d * F;
        const int f = 2 * d + ((lb >> d) & 1);
        if (i < NHALO) tile[TL::origin(lb) + halo_at_sw<D, M, TL>(f, t)] = hv[j];
    }
    __syncthreads();

    // first half-sweep
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        const int i = tid + k * NT, lb = i / H, r = i - lb * H;
        int x, y, z;
        LY::decode(r, x, y, z);
        const int p = TL::origin(lb) + TL::at(x, y, z);
        T wm, wp;
        rweights<D, M, T, RD>(a, coord, blk0 + lb, RD == 0 ? x : y, wm, wp);
        const BP<T> bp = prm[lb];
        T Lu = tstencil<D, M, T, RD, TL>(tile, p, ux[k][0], bp.cc, wm, wp);
        ux[k][0] = fma_t<T>(bp.nwc, fma_t<T>(bp.s, bo[k], Lu), ux[k][0]);
        tile[p] = ux[k][0];
    }
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int j = rj[k];
        if (j < 0) continue;
        const int lb = j / (D * F2), rem = j - lb * (D * F2);
        const int d = rem / F2, tq = rem - d * F2;
        const int f = 2 * d + ((lb >> d) & 1);
        const int *tb = tab + lb * HT::N, *gs = tb + HT::ND + f * HT::NS;
        T *tl = tile + TL::origin(lb);
        const int e = tb[f];
        const int dir = (f & 1) ? 1 : -1;
        const BP<T> bp = prm[lb];
        if (d == 0)
            RecFace<D, M, T, 0, TL>::template compute<RD>(dir, tq, e, gs, tl, coord, a, bp, rd[k], rb[k], rg[k]);
        else if (d == 1)
            RecFace<D, M, T, 1, TL>::template compute<RD>(dir, tq, e, gs, tl, coord, a, bp, rd[k], rb[k], rg[k]);
        else
            RecFace<D, M, T, (D == 3 ? 2 : 1), TL>::template compute<RD>(dir, tq, e, gs, tl, coord, a, bp, rd[k],
                                                                         rb[k], rg[k]);
    }
    __syncthreads();

    // second half-sweep
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        const int i = tid + k * NT, lb = i / H, r = i - lb * H;
        int x, y, z;
        LY::decode(H + r, x, y, z);
        const int p = TL::origin(lb) + TL::at(x, y, z);
        T wm, wp;
        rweights<D, M, T, RD>(a, coord, blk0 + lb, RD == 0 ? x : y, wm, wp);
        const BP<T> bp = prm[lb];
        T Lu = tstencil<D, M, T, RD, TL>(tile, p, ux[k][1], bp.cc, wm, wp);
        const size_t o = (size_t)(blk0 + lb) * V + r;
        uo[o] = ux[k][0];
        uo[o + H] = fma_t<T>(bp.nwc, fma_t<T>(bp.s, be[k], Lu), ux[k][1]);
    }
}

// k_gs_sweep_group for D = 3, M = 8 with a column mapping: 256 threads per
// sibling group (16^3 cells), each thread owns one (X, Y) column of the group
// tile and keeps its 16 cells and right-hand sides in registers, so that the
// Z neighbours of a cell come from registers and only the X and Y neighbours
// from the LDS tile. Wave w covers rows Y in [8 (w >> 1), 8 (w >> 1) + 8) and
// the columns with X + Y = w (mod 2): the red cells of all its lanes then sit
// at Z = P (mod 2), P = w & 1, a compile-time constant of the code path the
// wave runs. Tile rows are padded to 24 entries (no bank conflicts in the X
// and Y neighbour reads of a wave). Halo cells, neighbour recompute and the
// order of every sum as in k_gs_sweep_group, so the same bits. Round trips:
// the group's halo-table rows (issued before the interior loads and waited
// for alone), then the halo cells and recompute operands with the interior
// loads still in flight.
template <int D, int M, typename T, int RD>
struct ColSweep {
    static_assert(D == 3 && M == 8 && RD < 0, "column sweep is written for 8^3 Cartesian blocks");
    using LY = Lay<D, M>;
    using TL = Tile<D, M, 2, 24>;
    using HT = Halo<D>;
    static constexpr int NB = 8, V = LY::V, H = LY::H, NG = 2 * M;
    static constexpr int SY = TL::stride(1), SZ = TL::stride(2);
    struct Smem {
        T tile[TL::SZ];
        int tab[NB * HT::N];
        BP<T> prm[NB];
    };

    // offset of cell Z of a column from the column's base in its block (the
    // lower block for Z < 8, the upper one, 4 blocks on, for Z >= 8; storage
    // half 0 holds the red cells)
    template <int P, int Z>
    __device__ static constexpr int zoff() {
        return LY::MH * M * (Z & 7) + (((Z + P) & 1) ? H : 0);
    }

    template <int P>
    __device__ static __forceinline__ void run(Smem &sm, T *__restrict__ uo, const T *__restrict__ ui,
                                               const T *__restrict__ b, const T *__restrict__ ib,
                                               const int *__restrict__ coord, const LevelArgs<T> &a, int blk0,
                                               int w, int lane, int tv, const BP<T> &bpl) {
        const int tid = w * 64 + lane;
        const int Y = 8 * (w >> 1) + (lane >> 3), X = 2 * (lane & 7) + ((Y + P) & 1);
        const int x = X & 7, y = Y & 7, kxy = (X >> 3) | ((Y >> 3) << 1);
        const size_t g0 = (size_t)(blk0 + kxy) * V + (size_t)((x >> 1) + LY::MH * y);
        const T *__restrict__ ul = ui + g0, *__restrict__ uh = ul + 4 * V;
        const T *__restrict__ bl = b + g0, *__restrict__ bh = bl + 4 * V;
        T uc[NG], bc[NG];
        static_for<NG>([&](auto zc) {
            constexpr int Z = decltype(zc)::value;
            uc[Z] = (Z < 8 ? ul : uh)[zoff<P, Z>()];
            bc[Z] = (Z < 8 ? bl : bh)[zoff<P, Z>()];
        });
        if (tid < NB * HT::N) sm.tab[tid] = tv;
        if (tid < NB) sm.prm[tid] = bpl;
        lds_barrier();

        // halo cells of the six group faces: wave w takes the block face w of
        // each (tangential bits of the block from w), lane = cell t of the
        // face; a boundary face's mirror factor is applied at the tile write
        T hv[6], hm[6];
        static_for<6>([&](auto fc) {
            constexpr int F = decltype(fc)::value;
            constexpr int DIM = F >> 1, DA = DIM == 0 ? 1 : 0, DB = DIM == 2 ? 1 : 2;
            const int k = ((F & 1) << DIM) | ((w & 1) << DA) | ((w >> 1) << DB);
            const int e = __builtin_amdgcn_readfirstlane(sm.tab[k * HT::N + F]);
            hv[F] = halo_value<D, M, T, F>(e, blk0 + k, lane, ui, ib, T(1));
            hm[F] = e == NB_BOUNDARY ? a.sg[F] : T(1);
        });
        // recompute operands: for dimension DIM, waves 0-1 the low group face,
        // waves 2-3 the high one; half-wave = block face, tq = lane & 31
        const int rside = w >> 1, rbf = ((w & 1) << 1) | ((lane >> 5) & 1), rtq = lane & 31;
        T rd[3], rb[3], rg[3][4];
        int re[3];
        static_for<3>([&](auto dc) {
            constexpr int DIM = decltype(dc)::value;
            constexpr int DA = DIM == 0 ? 1 : 0, DB = DIM == 2 ? 1 : 2;
            const int k = (rside << DIM) | ((rbf & 1) << DA) | ((rbf >> 1) << DB), f = 2 * DIM + rside;
            const int *tb = sm.tab + k * HT::N;
            re[DIM] = tb[f];
            rd[DIM] = rb[DIM] = T(0);
            if (re[DIM] >= 0)
                RecFace<D, M, T, DIM, TL>::prefetch(rside ? 1 : -1, rtq, re[DIM], tb + HT::ND + f * HT::NS, ui, b, ib,
                                                    rd[DIM], rb[DIM], rg[DIM]);
        });
        // the tile
        const int pc = TL::at(X, Y, 0);
        static_for<NG>([&](auto zc) {
            constexpr int Z = decltype(zc)::value;
            sm.tile[pc + Z * SZ] = uc[Z];
        });
        static_for<6>([&](auto fc) {
            constexpr int F = decltype(fc)::value;
            constexpr int DIM = F >> 1, DA = DIM == 0 ? 1 : 0, DB = DIM == 2 ? 1 : 2;
            const int k = ((F & 1) << DIM) | ((w & 1) << DA) | ((w >> 1) << DB);
            sm.tile[TL::origin(k) + halo_at<D, M, F, TL>(lane)] = hm[F] * hv[F];
        });
        lds_barrier();

        // one cell update of the column (tstencil's order of terms)
        const BP<T> bq[2] = {sm.prm[kxy], sm.prm[kxy | 4]};
        T *const t = sm.tile;
        auto update = [&](auto zc) {
            constexpr int Z = decltype(zc)::value;
            const int p = pc + Z * SZ;
            const BP<T> &bp = bq[Z >> 3];
            const T zm = (Z == 0) ? t[pc - SZ] : uc[Z > 0 ? Z - 1 : 0];
            const T zp = (Z == NG - 1) ? t[pc + NG * SZ] : uc[Z < NG - 1 ? Z + 1 : 0];
            T v = T(0);
            v += zm;
            v += t[p - SY];
            v += t[p - 1];
            v += bp.cc * uc[Z];
            v += t[p + 1];
            v += t[p + SY];
            v += zp;
            uc[Z] = fma_t<T>(bp.nwc, fma_t<T>(bp.s, bc[Z], v), uc[Z]);
        };
        // first half-sweep: red cells
        static_for<NG / 2>([&](auto ic) {
            constexpr int Z = 2 * decltype(ic)::value + P;
            update(std::integral_constant<int, Z>());
            t[pc + Z * SZ] = uc[Z];
        });
        // the neighbours' face-layer cells
        static_for<3>([&](auto dc) {
            constexpr int DIM = decltype(dc)::value;
            constexpr int DA = DIM == 0 ? 1 : 0, DB = DIM == 2 ? 1 : 2;
            const int k = (rside << DIM) | ((rbf & 1) << DA) | ((rbf >> 1) << DB), f = 2 * DIM + rside;
            if (re[DIM] >= 0)
                RecFace<D, M, T, DIM, TL>::template compute<RD>(rside ? 1 : -1, rtq, re[DIM],
                                                                sm.tab + k * HT::N + HT::ND + f * HT::NS,
                                                                t + TL::origin(k), coord, a, sm.prm[k], rd[DIM],
                                                                rb[DIM], rg[DIM]);
        });
        lds_barrier();

        // second half-sweep: black cells; then the column goes out
        static_for<NG / 2>([&](auto ic) {
            constexpr int Z = 2 * decltype(ic)::value + 1 - P;
            update(std::integral_constant<int, Z>());
        });
        T *__restrict__ ol = uo + g0, *__restrict__ oh = ol + 4 * V;
        static_for<NG>([&](auto zc) {
            constexpr int Z = decltype(zc)::value;
            (Z < 8 ? ol : oh)[zoff<P, Z>()] = uc[Z];
        });
    }
};

template <int D, int M, typename T, int RD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 5 : 1))) void k_gs_sweep_col(
    T *__restrict__ uo, const T *__restrict__ ui, const T *__restrict__ b, const T *__restrict__ ib,
    const int *__restrict__ htab, const int *__restrict__ coord, LevelArgs<T> a) {
    using CS = ColSweep<D, M, T, RD>;
    using HT = Halo<D>;
    __shared__ typename CS::Smem sm;
    const int blk0 = a.blk0 + xcd_remap(blockIdx.x, gridDim.x) * CS::NB;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    // halo-table rows and block parameters of the group, first
    int tv = 0;
    if (tid < CS::NB * HT::N) tv = htab[(size_t)blk0 * HT::N + tid];
    BP<T> bpl{};
    if (tid < CS::NB) bpl = bparams(a, blk0 + tid);
    if (w & 1) CS::template run<1>(sm, uo, ui, b, ib, coord, a, blk0, w, lane, tv, bpl);
    else CS::template run<0>(sm, uo, ui, b, ib, coord, a, blk0, w, lane, tv, bpl);
}

// k_gs_sweep_col with each column split between two threads (512 threads per
// sibling group): wave w covers the columns of wave w & 3 of k_gs_sweep_col in
// the block layer w >> 2 (8 cells each), so a group's work is spread over
// twice as many waves; the Z neighbours across the layers come from the tile.
// For the small levels, whose few groups leave most of the chip idle. The
// same operations on the same values as k_gs_sweep_col (same bits).
template <int D, int M, typename T, int RD>
struct ColSweepZ {
    using CS = ColSweep<D, M, T, RD>;
    using LY = Lay<D, M>;
    using TL = typename CS::TL;
    using HT = Halo<D>;
    static constexpr int NB = 8, V = LY::V, H = LY::H;
    static constexpr int SY = TL::stride(1), SZ = TL::stride(2);
    using Smem = typename CS::Smem;

    template <int DIM, bool ZIN = false>
    __device__ static __forceinline__ void rec_pre(Smem &sm, int q, const T *__restrict__ ui, const T *__restrict__ b,
                                                   const T *__restrict__ ib, int &re, T &rd, T &rb, T *rg) {
        constexpr int DA = DIM == 0 ? 1 : 0, DB = DIM == 2 ? 1 : 2;
        const int rside = q >> 7, rbf = (q >> 5) & 3, rtq = q & 31;
        const int k = (rside << DIM) | ((rbf & 1) << DA) | ((rbf >> 1) << DB), f = 2 * DIM + rside;
        const int *tb = sm.tab + k * HT::N;
        re = tb[f];
        rd = rb = T(0);
        if (re >= 0)
            RecFace<D, M, T, DIM, TL>::template prefetch<ZIN>(rside ? 1 : -1, rtq, re, tb + HT::ND + f * HT::NS, ui,
                                                              b, ib, rd, rb, rg);
    }
    template <int DIM>
    __device__ static __forceinline__ void rec_cmp(Smem &sm, int q, const int *__restrict__ coord,
                                                   const LevelArgs<T> &a, int re, T rd, T rb, const T *rg) {
        constexpr int DA = DIM == 0 ? 1 : 0, DB = DIM == 2 ? 1 : 2;
        const int rside = q >> 7, rbf = (q >> 5) & 3, rtq = q & 31;
        const int k = (rside << DIM) | ((rbf & 1) << DA) | ((rbf >> 1) << DB), f = 2 * DIM + rside;
        if (re >= 0)
            RecFace<D, M, T, DIM, TL>::template compute<RD>(rside ? 1 : -1, rtq, re,
                                                            sm.tab + k * HT::N + HT::ND + f * HT::NS,
                                                            sm.tile + TL::origin(k), coord, a, sm.prm[k], rd, rb, rg);
    }

    // VAR bit 0: the column goes out with nontemporal stores; bit 1: the
    // block's own right-hand side is read with nontemporal loads (both keep
    // L2 for the neighbours' face cells, which other work groups re-read);
    // bit 2: the input field and its ghost slabs are zero on this level (the
    // first sweep of a correction solve): nothing of it is read, the sweep
    // gives the bits of the same sweep over zeros
    template <int P, int VAR>
    __device__ static __forceinline__ void run(Smem &sm, T *__restrict__ uo, const T *__restrict__ ui,
                                               const T *__restrict__ b, const T *__restrict__ ib,
                                               const int *__restrict__ coord, const LevelArgs<T> &a, int blk0,
                                               int w, int lane, int tv, const BP<T> &bpl) {
        const int tid = w * 64 + lane;
        const int wc = w & 3, wz = w >> 2;
        const int Y = 8 * (wc >> 1) + (lane >> 3), X = 2 * (lane & 7) + ((Y + P) & 1);
        const int x = X & 7, y = Y & 7, kb = ((X >> 3) | ((Y >> 3) << 1)) | (wz << 2);
        const size_t g0 = (size_t)(blk0 + kb) * V + (size_t)((x >> 1) + LY::MH * y);
        constexpr bool ZIN = (VAR & 4) != 0;
        T uc[M], bc[M];
        static_for<M>([&](auto zc) {
            constexpr int z = decltype(zc)::value;
            uc[z] = ZIN ? T(0) : ui[g0 + CS::template zoff<P, z>()];
            if constexpr (VAR & 2) bc[z] = __builtin_nontemporal_load(&b[g0 + CS::template zoff<P, z>()]);
            else bc[z] = b[g0 + CS::template zoff<P, z>()];
        });
        if (tid < NB * HT::N) sm.tab[tid] = tv;
        if (tid < NB) sm.prm[tid] = bpl;
        lds_barrier();

        // halo cells of the six group faces: wave w takes the (face, block
        // face) pairs 3 w .. 3 w + 2 of the 24 (face uniform per wave), lane =
        // cell t of the face; a boundary face's mirror factor is applied at
        // the tile write
        T hv[3], hm[3];
        int hF[3], hk[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int pr = 3 * w + j, F = pr >> 2, bf = pr & 3;
            const int DIM = F >> 1, DA = DIM == 0 ? 1 : 0, DB = DIM == 2 ? 1 : 2;
            const int k = ((F & 1) << DIM) | ((bf & 1) << DA) | ((bf >> 1) << DB);
            const int e = __builtin_amdgcn_readfirstlane(sm.tab[k * HT::N + F]);
            hF[j] = F;
            hk[j] = k;
            T v = T(0), m = T(1);
            if constexpr (ZIN) {
                SRMG_FACE_SWITCH(F, (m = e == NB_BOUNDARY ? a.sg[FF] : T(1)));
            } else {
                SRMG_FACE_SWITCH(F, (v = halo_value<D, M, T, FF>(e, blk0 + k, lane, ui, ib, T(1)),
                                     m = e == NB_BOUNDARY ? a.sg[FF] : T(1)));
            }
            hv[j] = v;
            hm[j] = m;
        }
        // recompute operands: waves 0-3 dimensions 0 and 2, waves 4-7 dimension 1
        int re[2];
        T rd[2], rb[2], rg[2][4];
        if (w < 4) {
            rec_pre<0, ZIN>(sm, tid, ui, b, ib, re[0], rd[0], rb[0], rg[0]);
            rec_pre<2, ZIN>(sm, tid, ui, b, ib, re[1], rd[1], rb[1], rg[1]);
        } else {
            rec_pre<1, ZIN>(sm, tid - 256, ui, b, ib, re[0], rd[0], rb[0], rg[0]);
        }
        // the tile
        const int pc = TL::at(X, Y, 8 * wz);
        static_for<M>([&](auto zc) {
            constexpr int z = decltype(zc)::value;
            sm.tile[pc + z * SZ] = uc[z];
        });
#pragma unroll
        for (int j = 0; j < 3; j++)
            sm.tile[TL::origin(hk[j]) + halo_at_sw<D, M, TL>(hF[j], lane)] = hm[j] * hv[j];
        lds_barrier();

        const BP<T> bp = sm.prm[kb];
        T *const t = sm.tile;
        auto update = [&](auto zc) {
            constexpr int z = decltype(zc)::value;
            const int p = pc + z * SZ;
            const T zm = (z == 0) ? t[p - SZ] : uc[z > 0 ? z - 1 : 0];
            const T zp = (z == M - 1) ? t[p + SZ] : uc[z < M - 1 ? z + 1 : 0];
            T v = T(0);
            v += zm;
            v += t[p - SY];
            v += t[p - 1];
            v += bp.cc * uc[z];
            v += t[p + 1];
            v += t[p + SY];
            v += zp;
            uc[z] = fma_t<T>(bp.nwc, fma_t<T>(bp.s, bc[z], v), uc[z]);
        };
        // first half-sweep: red cells (z = P mod 2 in the layer, which starts at an even Z)
        static_for<M / 2>([&](auto ic) {
            constexpr int z = 2 * decltype(ic)::value + P;
            update(std::integral_constant<int, z>());
            t[pc + z * SZ] = uc[z];
        });
        if (w < 4) {
            rec_cmp<0>(sm, tid, coord, a, re[0], rd[0], rb[0], rg[0]);
            rec_cmp<2>(sm, tid, coord, a, re[1], rd[1], rb[1], rg[1]);
        } else {
            rec_cmp<1>(sm, tid - 256, coord, a, re[0], rd[0], rb[0], rg[0]);
        }
        lds_barrier();
        static_for<M / 2>([&](auto ic) {
            constexpr int z = 2 * decltype(ic)::value + 1 - P;
            update(std::integral_constant<int, z>());
        });
        static_for<M>([&](auto zc) {
            constexpr int z = decltype(zc)::value;
            if constexpr (VAR & 1) __builtin_nontemporal_store(uc[z], &uo[g0 + CS::template zoff<P, z>()]);
            else uo[g0 + CS::template zoff<P, z>()] = uc[z];
        });
    }
};

template <int D, int M, typename T, int RD, int VAR = 0>
__global__ __launch_bounds__(512) void k_gs_sweep_colz(T *__restrict__ uo, const T *__restrict__ ui,
                                                       const T *__restrict__ b, const T *__restrict__ ib,
                                                       const int *__restrict__ htab, const int *__restrict__ coord,
                                                       LevelArgs<T> a) {
    using CZ = ColSweepZ<D, M, T, RD>;
    using HT = Halo<D>;
    __shared__ typename CZ::Smem sm;
    int gi = xcd_remap(blockIdx.x, gridDim.x);
    if (a.gord) gi = a.gord[gi];
    const int blk0 = a.blk0 + gi * CZ::NB;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    int tv = 0;
    if (tid < CZ::NB * HT::N) tv = htab[(size_t)blk0 * HT::N + tid];
    BP<T> bpl{};
    if (tid < CZ::NB) bpl = bparams(a, blk0 + tid);
    if (w & 1) CZ::template run<1, VAR>(sm, uo, ui, b, ib, coord, a, blk0, w, lane, tv, bpl);
    else CZ::template run<0, VAR>(sm, uo, ui, b, ib, coord, a, blk0, w, lane, tv, bpl);
}

// ---- tiled residual kernels ----

// Stage BPW blocks (valid cells + depth-1 face halo) of x in the LDS tile.
// list != nullptr selects the blocks through an index list (leaf blocks).
// Two barriers: after (interior, halo table) and after the halo cells
// (LDSB: barriers that order LDS only, for callers with loads in flight).
template <int D, int M, typename T, int BPW, int NT, bool LDSB = false>
__device__ __forceinline__ void load_tile(T *tile, int *tab, int *blks, const T *__restrict__ x,
                                          const T *__restrict__ ib, const int *__restrict__ htab,
                                          const int *__restrict__ list, int first, int nloc, const T *sg) {
    using LY = Lay<D, M>;
    using TL = Tile<D, M>;
    using HT = Halo<D>;
    constexpr int F = LY::F, V = LY::V, ND = 2 * D;
    // each phase issues all of its loads before the first tile store (a loop
    // that stores between its loads waits out each load in turn)
    constexpr int PI = (BPW * V + NT - 1) / NT, PH = (BPW * ND * F + NT - 1) / NT;
    const int tid = threadIdx.x;
    if (tid < BPW) blks[tid] = (tid < nloc) ? (list ? list[first + tid] : first + tid) : 0;
    if (LDSB) lds_barrier(); else __syncthreads();
    T xv[PI];
#pragma unroll
    for (int j = 0; j < PI; j++) {
        const int i = tid + NT * j, lb = min(i / V, BPW - 1);
        xv[j] = (i < nloc * V) ? x[(size_t)blks[lb] * V + (i - lb * V)] : T(0);
    }
    for (int i = tid; i < nloc * ND; i += NT) {
        int lb = i / ND;
        tab[i] = htab[(size_t)blks[lb] * HT::N + (i - lb * ND)];
    }
#pragma unroll
    for (int j = 0; j < PI; j++) {
        const int i = tid + NT * j, lb = min(i / V, BPW - 1);
        if (i < nloc * V) {
            int cx, cy, cz;
            LY::decode(i - lb * V, cx, cy, cz);
            tile[lb * TL::SZ + TL::at(cx, cy, cz)] = xv[j];
        }
    }
    if (LDSB) lds_barrier(); else __syncthreads();
    T hv[PH];
#pragma unroll
    for (int j = 0; j < PH; j++) {
        const int i = tid + NT * j, lb = min(i / (ND * F), BPW - 1), rem = i - lb * (ND * F);
        const int f = min(rem / F, ND - 1), t = rem - f * F;
        hv[j] = T(0);
        if (i < nloc * ND * F) hv[j] = halo_value_sw<D, M, T>(f, tab[lb * ND + f], blks[lb], t, x, ib, sg);
    }
#pragma unroll
    for (int j = 0; j < PH; j++) {
        const int i = tid + NT * j, lb = min(i / (ND * F), BPW - 1), rem = i - lb * (ND * F);
        const int f = min(rem / F, ND - 1), t = rem - f * F;
        if (i < nloc * ND * F) tile[lb * TL::SZ + halo_at_sw<D, M, TL>(f, t)] = hv[j];
    }
    if (LDSB) lds_barrier(); else __syncthreads();
}

// k_interp_add on a level of complete sibling groups: the 2^D children of a
// parent are one work group, which stages the parent (with its depth-1 face
// halo: neighbour cells, mirrors, coarse-fine slabs of level l-1) in LDS and
// interpolates every child cell from there. Same operations and order as
// k_interp_add, so the result is bitwise the same.
template <int D, int M, typename T, int NT>
__device__ __forceinline__ void interp_add_group_one(int g, T *__restrict__ u, const T *__restrict__ u1,
                                                     const T *__restrict__ ib1, const int *__restrict__ htab,
                                                     const int *__restrict__ parent, int blk0,
                                                     const LevelArgs<T> &ac) {
    using LY = Lay<D, M>;
    using TL = Tile<D, M>;
    constexpr int NB = 1 << D, V = LY::V, MH = LY::MH;
    static_assert((NB * V) % NT == 0, "cells of a group must be a multiple of the work-group size");
    __shared__ T tile[TL::SZ];
    __shared__ int tab[2 * D];
    __shared__ int blks[1];
    const int g0 = blk0 + g * NB;
    const int pb = parent[g0];
    // the children's cells next: their loads stay in flight while the parent is staged
    T uo[NB * V / NT];
#pragma unroll
    for (int k = 0; k < NB * V / NT; k++) uo[k] = u[(size_t)g0 * V + threadIdx.x + k * NT];
    load_tile<D, M, T, 1, NT, true>(tile, tab, blks, u1, ib1, htab, nullptr, pb, 1, ac.sg);
#pragma unroll
    for (int k = 0; k < NB * V / NT; k++) {
        const int i = threadIdx.x + k * NT, lb = i / V, q = i - lb * V;
        int x0, x1, x2;
        LY::decode(q, x0, x1, x2);
        const int J0 = (lb & 1) * MH + (x0 >> 1);
        const int J1 = ((lb >> 1) & 1) * MH + (x1 >> 1);
        const int J2 = (D == 3) ? ((lb >> 2) & 1) * MH + (x2 >> 1) : 0;
        const int p = TL::at(J0, J1, J2);
        T f = w0<D, T>() * tile[p];
        f += w1<T>() * tile[p + (2 * (x0 & 1) - 1)];
        f += w1<T>() * tile[p + (2 * (x1 & 1) - 1) * TL::stride(1)];
        if (D == 3) f += w1<T>() * tile[p + (2 * (x2 & 1) - 1) * TL::stride(2)];
        const size_t o = (size_t)(g0 + lb) * V + q;
        u[o] = f + uo[k];
    }
}

template <int D, int M, typename T, int NT>
__global__ __launch_bounds__(NT) void k_interp_add_group(T *__restrict__ u, const T *__restrict__ u1,
                                                         const T *__restrict__ ib1, const int *__restrict__ htab,
                                                         const int *__restrict__ parent, int blk0,
                                                         LevelArgs<T> ac) {
    interp_add_group_one<D, M, T, NT>(blockIdx.x, u, u1, ib1, htab, parent, blk0, ac);
}

// k_interp_add_group on the first ngroups work groups and k_fill_interp of the
// fine level's slabs (from the coarse level of u, already final) on the rest:
// the prolongation writes the fine cells of u, the slab fill writes u's
// fine-level slabs; neither reads what the other writes.
template <int D, int M, typename T, int NT>
__global__ __launch_bounds__(NT) void k_interp_fill_group(T *__restrict__ u, const T *__restrict__ u1,
                                                          const T *__restrict__ ib1, const int *__restrict__ htab,
                                                          const int *__restrict__ parent, int blk0, unsigned ngroups,
                                                          T *__restrict__ ibu, const T *__restrict__ xc,
                                                          const int *__restrict__ nbr, const int4 *__restrict__ slots,
                                                          int slot0, int nslot, LevelArgs<T> ac) {
    if (blockIdx.x < ngroups)
        interp_add_group_one<D, M, T, NT>(blockIdx.x, u, u1, ib1, htab, parent, blk0, ac);
    else
        fill_interp_one<D, M, T>((blockIdx.x - ngroups) * NT + threadIdx.x, ibu, xc, nbr, slots, slot0, nslot, ac);
}

enum { RES_WRITE = 0, RES_NORM = 1, RES_DEFECT = 2 };

// Residual of a level from LDS tiles (src/multigrid.jl:309-323).
//   RES_WRITE:  r = -(b + L u / s) on the blocks [a.blk0, a.blk0 + a.nblk)
//   RES_NORM:   leaf blocks list[0..n): partial[k] = wgt * sum of r^2 over block list[k]
//   RES_DEFECT: out (float) = b + L u / s = -r on the level's blocks, and
//               partial[k] = wgt * sum of r^2 over block k if it is a leaf (else 0)
// Each block's partial is an LDS tree sum in a fixed order.
template <int D, int M, typename T, typename TO, int BPW, int NT, int MODE, int RD>
__global__ __launch_bounds__(NT) void k_resid_tile(TO *__restrict__ out, const T *__restrict__ u,
                                                   const T *__restrict__ b, const T *__restrict__ ib,
                                                   const int *__restrict__ htab, const int *__restrict__ coord,
                                                   const int *__restrict__ list, const int *__restrict__ isleaf,
                                                   int n, LevelArgs<T> a, double wgt, double *__restrict__ partial) {
    using LY = Lay<D, M>;
    using TL = Tile<D, M>;
    constexpr int V = LY::V;
    __shared__ T tile[BPW * TL::SZ];
    __shared__ int tab[BPW * 2 * D];
    __shared__ int blks[BPW];
    __shared__ BP<T> prm[BPW];
    __shared__ double red[MODE == RES_WRITE ? 1 : BPW * NT];
    const int first = xcd_remap(blockIdx.x, gridDim.x) * BPW;
    const int nloc = min(BPW, n - first);
    const int *lst = (MODE == RES_NORM) ? list : nullptr;
    const int base = (MODE == RES_NORM) ? first : a.blk0 + first;
    load_tile<D, M, T, BPW, NT>(tile, tab, blks, u, ib, htab, lst, base, nloc, a.sg);
    if ((int)threadIdx.x < nloc) prm[threadIdx.x] = bparams(a, blks[threadIdx.x]);
    // the right-hand sides of the thread's cells, all in flight at once
    constexpr int PER = (BPW * V + NT - 1) / NT;
    T bv[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int i = threadIdx.x + NT * j, lb = min(i / V, BPW - 1);
        bv[j] = (i < nloc * V) ? b[(size_t)blks[lb] * V + (i - lb * V)] : T(0);
    }
    __syncthreads();
    double acc[BPW];
#pragma unroll
    for (int lb = 0; lb < BPW; lb++) acc[lb] = 0.0;
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int i = threadIdx.x + NT * j;
        if (i >= nloc * V) break;
        int lb = i / V, q = i - lb * V;
        int blk = blks[lb];
        int x, y, z;
        LY::decode(q, x, y, z);
        int p = lb * TL::SZ + TL::at(x, y, z);
        T wm, wp;
        rweights<D, M, T, RD>(a, coord, blk, RD == 0 ? x : y, wm, wp);
        const BP<T> bp = prm[lb];
        T Lu = tstencil<D, M, T, RD>(tile, p, tile[p], bp.cc, wm, wp);
        size_t o = (size_t)blk * V + q;
        T rr = -(bv[j] + Lu / bp.s);
        if (MODE == RES_WRITE) out[o] = (TO)rr;
        if (MODE == RES_DEFECT) out[o] = (TO)(bv[j] + Lu / bp.s);
        if (MODE == RES_NORM || (MODE == RES_DEFECT && isleaf[blk])) {
#pragma unroll
            for (int k = 0; k < BPW; k++)
                if (k == lb) acc[k] += (double)rr * (double)rr;
        }
    }
    if (MODE != RES_WRITE) {
        // each block's halving tree over the NT threads, all blocks at once:
        // the levels of 64 and more across waves in LDS, the rest inside one
        // wave per block by lane shuffles (the same pairs in the same order)
        static_assert(NT % 64 == 0 && NT >= 128, "whole waves, at least two");
#pragma unroll
        for (int lb = 0; lb < BPW; lb++) red[lb * NT + threadIdx.x] = acc[lb];
        __syncthreads();
        for (int w = NT / 2; w >= 64; w >>= 1) {
            if ((int)threadIdx.x < w) {
#pragma unroll
                for (int lb = 0; lb < BPW; lb++) red[lb * NT + threadIdx.x] += red[lb * NT + threadIdx.x + w];
            }
            __syncthreads();
        }
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int lb = wave; lb < BPW; lb += NT / 64) {
            double v = red[lb * NT + lane];
#pragma unroll
            for (int w = 32; w > 0; w >>= 1) {
                const double o = __shfl_down(v, w, 64);
                if (lane < w) v += o;
            }
            if (lane == 0 && lb < nloc) partial[first + lb] = v * wgt;
        }
    }
}

// Stage sibling group g0 (2^D blocks from device block g0) of u into a
// Tile<D, M, 2> with the halo of the group's exterior faces, its halo table
// entries and block parameters; every load of a phase is issued before the
// phase's first LDS store. Ends with a barrier (which also waits for loads
// the caller issued before the call).
template <int D, int M, typename T, int NT>
__device__ __forceinline__ void stage_group(T *tile, int *tab, BP<T> *prm, const T *__restrict__ u,
                                            const T *__restrict__ ib, const int *__restrict__ htab,
                                            const LevelArgs<T> &a, int g0) {
    using LY = Lay<D, M>;
    using TL = Tile<D, M, 2>;
    using HT = Halo<D>;
    constexpr int NB = 1 << D, V = LY::V, F = LY::F;
    constexpr int PU = (NB * V + NT - 1) / NT, PH = (NB * D * F + NT - 1) / NT;
    const int tid = threadIdx.x;
    T uv[PU];
#pragma unroll
    for (int j = 0; j < PU; j++) {
        const int i = tid + NT * j;
        uv[j] = (i < NB * V) ? u[(size_t)g0 * V + i] : T(0);
    }
    for (int i = tid; i < NB * 2 * D; i += NT) {
        const int lb = i / (2 * D);
        tab[i] = htab[(size_t)(g0 + lb) * HT::N + (i - lb * 2 * D)];
    }
    if (tid < NB) prm[tid] = bparams(a, g0 + tid);
#pragma unroll
    for (int j = 0; j < PU; j++) {
        const int i = tid + NT * j, lb = min(i / V, NB - 1), q = i - lb * V;
        if (i < NB * V) {
            int x, y, z;
            LY::decode(q, x, y, z);
            tile[TL::origin(lb) + TL::at(x, y, z)] = uv[j];
        }
    }
    __syncthreads();
    T hv[PH];
#pragma unroll
    for (int j = 0; j < PH; j++) {
        const int i = tid + NT * j, lb = min(i / (D * F), NB - 1), rem = i - lb * (D * F);
        const int d = min(rem / F, D - 1), t = rem - d * F;
        const int f = 2 * d + ((lb >> d) & 1);
        hv[j] = T(0);
        if (i < NB * D * F) hv[j] = halo_value_sw<D, M, T>(f, tab[lb * 2 * D + f], g0 + lb, t, u, ib, a.sg);
    }
#pragma unroll
    for (int j = 0; j < PH; j++) {
        const int i = tid + NT * j, lb = min(i / (D * F), NB - 1), rem = i - lb * (D * F);
        const int d = min(rem / F, D - 1), t = rem - d * F;
        const int f = 2 * d + ((lb >> d) & 1);
        if (i < NB * D * F) tile[TL::origin(lb) + halo_at_sw<D, M, TL>(f, t)] = hv[j];
    }
    __syncthreads();
}

// k_resid_tile (RES_WRITE, RES_DEFECT; RES_DEFECT with out == nullptr forms
// only the partials, on a level of leaves in block order those of
// k_resid_tile<RES_NORM>) on a level of complete sibling groups:
// one group of 2^D blocks per work group of NT threads, staged as one
// (2M)^D tile with the halo of the group's exterior faces only. Teams of RT
// threads take one block at a time and visit its cells in the order of
// k_resid_tile<.., 1, RT, ..> (cells t, t + RT, ...; partials as the same
// RT-wide tree sum), so outputs and partials are bitwise the same.
template <int D, int M, typename T, typename TO, int RT, int MODE, int RD, int NT = 1024>
__global__ __launch_bounds__(NT) void k_resid_group(TO *__restrict__ out, const T *__restrict__ u,
                                                      const T *__restrict__ b, const T *__restrict__ ib,
                                                      const int *__restrict__ htab, const int *__restrict__ coord,
                                                      const int *__restrict__ isleaf, LevelArgs<T> a, double wgt,
                                                      double *__restrict__ partial) {
    using LY = Lay<D, M>;
    using TL = Tile<D, M, 2>;
    using HT = Halo<D>;
    constexpr int NB = 1 << D, V = LY::V, F = LY::F;
    constexpr int TEAMS = NT / RT;
    static_assert(NB % TEAMS == 0 && RT >= 64, "team layout");
    __shared__ T tile[TL::SZ];
    __shared__ int tab[NB * 2 * D];
    __shared__ BP<T> prm[NB];
    __shared__ double red[MODE == RES_DEFECT ? NT : 1];
    constexpr int NR = NB / TEAMS, QR = (V + RT - 1) / RT;
    const int gl = xcd_remap(blockIdx.x, gridDim.x);
    const int g0 = a.blk0 + gl * NB;
    const int tid = threadIdx.x;
    const int team = tid / RT, t = tid - team * RT;
    // the right-hand sides of the thread's cells, in flight during the staging
    T bv[NR][QR];
#pragma unroll
    for (int round = 0; round < NR; round++)
#pragma unroll
        for (int h = 0; h < QR; h++) {
            const int q = t + RT * h;
            bv[round][h] = q < V ? b[(size_t)(g0 + team + round * TEAMS) * V + q] : T(0);
        }
    stage_group<D, M, T, NT>(tile, tab, prm, u, ib, htab, a, g0);
#pragma unroll
    for (int round = 0; round < NR; round++) {
        const int lb = team + round * TEAMS, blk = g0 + lb;
        const BP<T> bp = prm[lb];
        const bool leaf = (MODE == RES_DEFECT) && isleaf[blk];
        double acc = 0.0;
#pragma unroll
        for (int h = 0; h < QR; h++) {
            const int q = t + RT * h;
            if (q >= V) break;
            int x, y, z;
            LY::decode(q, x, y, z);
            const int p = TL::origin(lb) + TL::at(x, y, z);
            T wm, wp;
            rweights<D, M, T, RD>(a, coord, blk, RD == 0 ? x : y, wm, wp);
            T Lu = tstencil<D, M, T, RD, TL>(tile, p, tile[p], bp.cc, wm, wp);
            const size_t o = (size_t)blk * V + q;
            T rr = -(bv[round][h] + Lu / bp.s);
            if (MODE == RES_WRITE) out[o] = (TO)rr;
            if (MODE == RES_DEFECT) {
                if (out != nullptr) out[o] = (TO)(bv[round][h] + Lu / bp.s);
                if (leaf) acc += (double)rr * (double)rr;
            }
        }
        if (MODE == RES_DEFECT) {
            // the team's halving tree: levels of 64 and more in LDS, the
            // rest inside the team's first wave by lane shuffles (same pairs)
            red[tid] = acc;
            __syncthreads();
            for (int w = RT / 2; w >= 64; w >>= 1) {
                if (t < w) red[tid] += red[tid + w];
                __syncthreads();
            }
            if (t < 64) {
                double v = red[tid];
#pragma unroll
                for (int w = 32; w > 0; w >>= 1) {
                    const double o = __shfl_down(v, w, 64);
                    if (t < w) v += o;
                }
                if (t == 0) partial[(blk - a.blk0)] = v * wgt;
            }
            __syncthreads();
        }
    }
}

// The V-cycle descent's residual and restrictions on a level of complete
// sibling groups, in one launch: residual of the group (as k_resid_group,
// kept in LDS instead of written), then on the parent's cells the restricted
// residual rc and the restricted u (into uc and, if u1c is not null, u1c),
// each the sum of the 2^D children in k_restrict's order divided by 2^D. Work
// groups past ngroups copy u -> u1 on the coarse level's leaf blocks (lists
// leaves[0..nleafc)), completing copyto!(u1, u) on that level. Same values
// and operations as residual + restrict(u) + copy + restrict(r).
template <int D, int M, typename T, int RT, int RD, int NT = 1024>
__global__ __launch_bounds__(NT) void k_resid_restrict(const T *__restrict__ u, const T *__restrict__ b,
                                                         const T *__restrict__ ib, const int *__restrict__ htab,
                                                         const int *__restrict__ coord,
                                                         const int *__restrict__ parent, LevelArgs<T> a,
                                                         unsigned ngroups, T *__restrict__ rc, T *__restrict__ uc,
                                                         T *__restrict__ u1c, const T *__restrict__ ucin,
                                                         const int *__restrict__ leaves, int nleafc) {
    using LY = Lay<D, M>;
    using TL = Tile<D, M, 2>;
    using HT = Halo<D>;
    constexpr int NB = 1 << D, V = LY::V, F = LY::F, MH = LY::MH;
    constexpr int TEAMS = NT / RT;
    static_assert(NB % TEAMS == 0 && RT >= 64, "team layout");
    const int tid = threadIdx.x;
    if (blockIdx.x >= ngroups) {
        // u1 = u on the coarse level's leaf blocks
        const int64_t i = (int64_t)(blockIdx.x - ngroups) * NT + tid;
        if (i < (int64_t)nleafc * V) {
            const size_t o = (size_t)leaves[i / V] * V + (size_t)(i % V);
            u1c[o] = ucin[o];
        }
        return;
    }
    __shared__ T tile[TL::SZ];
    __shared__ T rt[NB * V];
    __shared__ int tab[NB * 2 * D];
    __shared__ BP<T> prm[NB];
    constexpr int NR = NB / TEAMS, QR = (V + RT - 1) / RT;
    const int gl = xcd_remap(blockIdx.x, ngroups);
    const int g0 = a.blk0 + gl * NB;
    const int team = tid / RT, t = tid - team * RT;
    // the right-hand sides of the thread's cells, in flight during the staging
    T bv[NR][QR];
#pragma unroll
    for (int round = 0; round < NR; round++)
#pragma unroll
        for (int h = 0; h < QR; h++) {
            const int q = t + RT * h;
            bv[round][h] = q < V ? b[(size_t)(g0 + team + round * TEAMS) * V + q] : T(0);
        }
    stage_group<D, M, T, NT>(tile, tab, prm, u, ib, htab, a, g0);
#pragma unroll
    for (int round = 0; round < NR; round++) {
        const int lb = team + round * TEAMS, blk = g0 + lb;
        const BP<T> bp = prm[lb];
#pragma unroll
        for (int h = 0; h < QR; h++) {
            const int q = t + RT * h;
            if (q >= V) break;
            int x, y, z;
            LY::decode(q, x, y, z);
            const int p = TL::origin(lb) + TL::at(x, y, z);
            T wm, wp;
            rweights<D, M, T, RD>(a, coord, blk, RD == 0 ? x : y, wm, wp);
            T Lu = tstencil<D, M, T, RD, TL>(tile, p, tile[p], bp.cc, wm, wp);
            rt[lb * V + q] = -(bv[round][h] + Lu / bp.s);
        }
    }
    __syncthreads();
    // the parent's cells: coarse cell X of sub-block k = (X >= MH per dim)
    const int pb = parent[g0];
    for (int i = tid; i < V; i += NT) {
        int X0, X1, X2;
        LY::decode(i, X0, X1, X2);
        const int k = (X0 >= MH ? 1 : 0) | (X1 >= MH ? 2 : 0) | (D == 3 && X2 >= MH ? 4 : 0);
        const int j0 = X0 - (k & 1) * MH, j1 = X1 - ((k >> 1) & 1) * MH, j2 = (D == 3) ? X2 - ((k >> 2) & 1) * MH : 0;
        T fr = T(0), fu = T(0);
#pragma unroll
        for (int q = 0; q < LY::NC; q++) {
            const int i0 = q & 1, i1 = (q >> 1) & 1, i2 = (q >> 2) & 1;
            const int cx = 2 * j0 + i0, cy = 2 * j1 + i1, cz = D == 3 ? 2 * j2 + i2 : 0;
            fr += rt[k * V + LY::sidx(cx, cy, cz)];
            fu += tile[TL::origin(k) + TL::at(cx, cy, cz)];
        }
        const size_t o = (size_t)pb * V + i;
        rc[o] = fr / T(LY::NC);
        const T uv = fu / T(LY::NC);
        uc[o] = uv;
        if (u1c) u1c[o] = uv;
    }
}

// k_resid_restrict for D = 3, M = 8 with the column mapping of k_gs_sweep_col:
// 256 threads per sibling group, thread = (X, Y) column of the group with its
// 16 cells and right-hand sides in registers (Z neighbours from registers).
// The residual stays in registers while the parent's u is restricted from
// the u tile; then it takes u's place in the tile, from which the parent's
// residual is restricted (both in k_restrict's order). Same operations as
// k_resid_restrict, so the same bits. Work groups past ngroups copy u to u1
// on the coarse level's leaf blocks (256 cells each).
// Column staging shared by the column residual kernels (3D, 8^3, one sibling
// group of 16^3 cells per 256-thread work group): thread (w, lane) owns the
// (X, Y) column X + Y = w (mod 2) of rows 8 (w >> 1) .. 8 (w >> 1) + 7 (as in
// k_gs_sweep_col); its 16 cells of u and b go to registers, u with the group's
// depth-1 face halo to the LDS tile. tab holds the group's own face entries
// (2D per block, loaded by the caller into tv), prm the block parameters.
template <int D, int M, typename T, int RD>
struct ColStage {
    static_assert(D == 3 && M == 8 && RD < 0, "column kernels are written for 8^3 Cartesian blocks");
    using LY = Lay<D, M>;
    using TL = Tile<D, M, 2>;
    static constexpr int NB = 8, V = LY::V, NG = 2 * M, ND = 2 * D;
    static constexpr int SY = TL::stride(1), SZ = TL::stride(2);

    struct Col {
        int X, Y, x, y, kxy, pc;
        size_t gb;  // the column's base offset in its lower block
    };

    template <int P>
    __device__ static __forceinline__ Col column(int g0, int w, int lane) {
        Col c;
        c.Y = 8 * (w >> 1) + (lane >> 3);
        c.X = 2 * (lane & 7) + ((c.Y + P) & 1);
        c.x = c.X & 7;
        c.y = c.Y & 7;
        c.kxy = (c.X >> 3) | ((c.Y >> 3) << 1);
        c.gb = (size_t)(g0 + c.kxy) * V + (size_t)((c.x >> 1) + LY::MH * c.y);
        c.pc = TL::at(c.X, c.Y, 0);
        return c;
    }

    template <int P>
    __device__ static __forceinline__ void stage(T *tile, int *tab, BP<T> *prm, const Col &c, const T *__restrict__ u,
                                                 const T *__restrict__ b, const T *__restrict__ ib,
                                                 const LevelArgs<T> &a, int g0, int w, int lane, int tv,
                                                 const BP<T> &bpl, T *ux, T *bx) {
        using CS = ColSweep<D, M, T, RD>;
        const int tid = w * 64 + lane;
        const T *__restrict__ ul = u + c.gb, *__restrict__ uh = ul + 4 * V;
        const T *__restrict__ bl = b + c.gb, *__restrict__ bh = bl + 4 * V;
        static_for<NG>([&](auto zc) {
            constexpr int Z = decltype(zc)::value;
            ux[Z] = (Z < 8 ? ul : uh)[CS::template zoff<P, Z>()];
            bx[Z] = (Z < 8 ? bl : bh)[CS::template zoff<P, Z>()];
        });
        if (tid < NB * ND) tab[tid] = tv;
        if (tid < NB) prm[tid] = bpl;
        lds_barrier();
        // halo of the six group faces (wave w: block face w of each; lane: cell)
        T hv[6], hm[6];
        static_for<6>([&](auto fc) {
            constexpr int F = decltype(fc)::value;
            constexpr int DIM = F >> 1, DA = DIM == 0 ? 1 : 0, DB = DIM == 2 ? 1 : 2;
            const int k = ((F & 1) << DIM) | ((w & 1) << DA) | ((w >> 1) << DB);
            const int e = __builtin_amdgcn_readfirstlane(tab[k * ND + F]);
            hv[F] = halo_value<D, M, T, F>(e, g0 + k, lane, u, ib, T(1));
            hm[F] = e == NB_BOUNDARY ? a.sg[F] : T(1);
        });
        static_for<NG>([&](auto zc) {
            constexpr int Z = decltype(zc)::value;
            tile[c.pc + Z * SZ] = ux[Z];
        });
        static_for<6>([&](auto fc) {
            constexpr int F = decltype(fc)::value;
            constexpr int DIM = F >> 1, DA = DIM == 0 ? 1 : 0, DB = DIM == 2 ? 1 : 2;
            const int k = ((F & 1) << DIM) | ((w & 1) << DA) | ((w >> 1) << DB);
            tile[TL::origin(k) + halo_at<D, M, F, TL>(lane)] = hm[F] * hv[F];
        });
        lds_barrier();
    }

    // L u of cell Z of the column (tstencil's order of terms)
    template <int Z>
    __device__ static __forceinline__ T lu(const T *t, const Col &c, const T *ux, T cc) {
        const int p = c.pc + Z * SZ;
        const T zm = (Z == 0) ? t[c.pc - SZ] : ux[Z > 0 ? Z - 1 : 0];
        const T zp = (Z == NG - 1) ? t[c.pc + NG * SZ] : ux[Z < NG - 1 ? Z + 1 : 0];
        T v = T(0);
        v += zm;
        v += t[p - SY];
        v += t[p - 1];
        v += cc * ux[Z];
        v += t[p + 1];
        v += t[p + SY];
        v += zp;
        return v;
    }
};

template <int D, int M, typename T, int RD>
struct ColResid {
    using CT = ColStage<D, M, T, RD>;
    using LY = Lay<D, M>;
    using TL = Tile<D, M, 2>;
    static constexpr int NB = 8, V = LY::V, NG = 2 * M, ND = 2 * D;
    struct Smem {
        T tile[TL::SZ];  // u with its halo; then the residual at the group's cells
        int tab[NB * ND];
        BP<T> prm[NB];
    };

    template <int P>
    __device__ static __forceinline__ void run(Smem &sm, const T *__restrict__ u, const T *__restrict__ b,
                                               const T *__restrict__ ib, const int *__restrict__ parent,
                                               const LevelArgs<T> &a, int g0, int w, int lane, int tv,
                                               const BP<T> &bpl, T *__restrict__ rc, T *__restrict__ uc,
                                               T *__restrict__ u1c) {
        const int tid = w * 64 + lane;
        const typename CT::Col c = CT::template column<P>(g0, w, lane);
        T ux[NG], bx[NG];
        CT::template stage<P>(sm.tile, sm.tab, sm.prm, c, u, b, ib, a, g0, w, lane, tv, bpl, ux, bx);
        // residual r = -(b + L u / s), kept in registers
        const BP<T> bq[2] = {sm.prm[c.kxy], sm.prm[c.kxy | 4]};
        T rr[NG];
        static_for<NG>([&](auto zc) {
            constexpr int Z = decltype(zc)::value;
            const BP<T> &bp = bq[Z >> 3];
            const T v = CT::template lu<Z>(sm.tile, c, ux, bp.cc);
            rr[Z] = -(bx[Z] + v / bp.s);
        });
        // the parent's cells, in storage order: coarse cell (X0, X1, X2) sums
        // group cells (2 X0 + i0, 2 X1 + i1, 2 X2 + i2), q = i0 + 2 i1 + 4 i2;
        // u's sums first, then the residual takes u's place in the tile
        constexpr int PC = V / 256;
        T fu[PC];
#pragma unroll
        for (int h = 0; h < PC; h++) {
            int X0, X1, X2;
            LY::decode(tid + 256 * h, X0, X1, X2);
            fu[h] = T(0);
#pragma unroll
            for (int q = 0; q < LY::NC; q++)
                fu[h] += sm.tile[TL::at(2 * X0 + (q & 1), 2 * X1 + ((q >> 1) & 1), 2 * X2 + ((q >> 2) & 1))];
        }
        lds_barrier();  // every read of u in the tile is done
        static_for<NG>([&](auto zc) {
            constexpr int Z = decltype(zc)::value;
            sm.tile[TL::at(c.X, c.Y, Z)] = rr[Z];
        });
        lds_barrier();
        const int pb = parent[g0];
#pragma unroll
        for (int h = 0; h < PC; h++) {
            const int i = tid + 256 * h;
            int X0, X1, X2;
            LY::decode(i, X0, X1, X2);
            T fr = T(0);
#pragma unroll
            for (int q = 0; q < LY::NC; q++)
                fr += sm.tile[TL::at(2 * X0 + (q & 1), 2 * X1 + ((q >> 1) & 1), 2 * X2 + ((q >> 2) & 1))];
            const size_t o = (size_t)pb * V + i;
            rc[o] = fr / T(LY::NC);
            const T uv = fu[h] / T(LY::NC);
            uc[o] = uv;
            if (u1c) u1c[o] = uv;
        }
    }
};

// k_resid_group<RES_DEFECT> (FP64 in, FP32 out) with the column mapping: out =
// b + L u / s on every cell of the group, and per block the partial
// wgt * sum of r^2 over its cells if it is a leaf (0 otherwise), summed as
// k_resid_group sums it (thread t of 256: cells q = t, t + 256 in that order,
// then the halving tree over t), so the partials are the same bits. With
// out == nullptr only the partials are formed: on a level whose blocks are all
// leaves, in level order, these are the partials of k_resid_tile<RES_NORM>
// (the same per-block order).
template <int D, int M, int RD>
struct ColDefect {
    using T = double;
    using CT = ColStage<D, M, T, RD>;
    using LY = Lay<D, M>;
    using TL = Tile<D, M, 2>;
    static constexpr int NB = 8, V = LY::V, NG = 2 * M, ND = 2 * D;
    static_assert(TL::SZ >= NB * V, "the squares reuse the tile");
    struct Smem {
        T tile[TL::SZ];  // u with its halo; then the squares sq[k V + q] of block k, cell q
        int tab[NB * ND];
        BP<T> prm[NB];
        int leaf[NB];
    };

    template <int P>
    __device__ static __forceinline__ void run(Smem &sm, float *__restrict__ out, const T *__restrict__ u,
                                               const T *__restrict__ b, const T *__restrict__ ib,
                                               const LevelArgs<T> &a, int g0, int w, int lane, int tv,
                                               const BP<T> &bpl, double wgt, double *__restrict__ partial) {
        using CS = ColSweep<D, M, T, RD>;
        const int tid = w * 64 + lane;
        const typename CT::Col c = CT::template column<P>(g0, w, lane);
        T ux[NG], bx[NG];
        CT::template stage<P>(sm.tile, sm.tab, sm.prm, c, u, b, ib, a, g0, w, lane, tv, bpl, ux, bx);
        const BP<T> bq[2] = {sm.prm[c.kxy], sm.prm[c.kxy | 4]};
        const bool lf[2] = {sm.leaf[c.kxy] != 0, sm.leaf[c.kxy | 4] != 0};
        float *__restrict__ ol = out + c.gb, *__restrict__ oh = ol + 4 * V;
        double sq[NG];
        static_for<NG>([&](auto zc) {
            constexpr int Z = decltype(zc)::value;
            const BP<T> &bp = bq[Z >> 3];
            const T v = CT::template lu<Z>(sm.tile, c, ux, bp.cc);
            if (out != nullptr) (Z < 8 ? ol : oh)[CS::template zoff<P, Z>()] = (float)(bx[Z] + v / bp.s);
            const T rr = -(bx[Z] + v / bp.s);
            sq[Z] = lf[Z >> 3] ? (double)rr * (double)rr : 0.0;
        });
        lds_barrier();  // every tile read is done: the squares take its place
        static_for<NG>([&](auto zc) {
            constexpr int Z = decltype(zc)::value;
            const int k = c.kxy | ((Z >> 3) << 2);
            sm.tile[k * V + LY::sidx(c.x, c.y, Z & 7)] = sq[Z];
        });
        lds_barrier();
        // per block: acc_t = (0 + sq[t]) + sq[t + 256], then the halving tree
        for (int k = 0; k < NB; k++) {
            double acc = 0.0;
            acc += sm.tile[k * V + tid];
            acc += sm.tile[k * V + tid + 256];
            sm.tile[k * V + tid] = acc;
        }
        lds_barrier();
        // the tree's levels 128 and 64 across waves in LDS; 32 .. 1 inside
        // wave w for blocks w and w + 4 by lane shuffles (the same pairs)
        for (int wd = 128; wd >= 64; wd >>= 1) {
            for (int i = tid; i < NB * wd; i += 256) {
                const int k = i / wd, t = i - k * wd;
                sm.tile[k * V + t] += sm.tile[k * V + t + wd];
            }
            lds_barrier();
        }
#pragma unroll
        for (int j = 0; j < NB / 4; j++) {
            const int k = w + 4 * j;
            double v = sm.tile[k * V + lane];
#pragma unroll
            for (int wd = 32; wd > 0; wd >>= 1) {
                const double o = __shfl_down(v, wd, 64);
                if (lane < wd) v += o;
            }
            if (lane == 0) partial[g0 + k - a.blk0] = v * wgt;
        }
    }
};

template <int D, int M, int RD>
__global__ __launch_bounds__(256) void k_defect_col(float *__restrict__ out, const double *__restrict__ u,
                                                    const double *__restrict__ b, const double *__restrict__ ib,
                                                    const int *__restrict__ htab, const int *__restrict__ isleaf,
                                                    LevelArgs<double> a, double wgt, double *__restrict__ partial) {
    using CD = ColDefect<D, M, RD>;
    using HT = Halo<D>;
    constexpr int ND = CD::ND;
    __shared__ typename CD::Smem sm;
    const int tid = threadIdx.x;
    const int g0 = a.blk0 + xcd_remap(blockIdx.x, gridDim.x) * CD::NB;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    int tv = 0;
    if (tid < CD::NB * ND) {
        const int lb = tid / ND;
        tv = htab[(size_t)(g0 + lb) * HT::N + (tid - lb * ND)];
    }
    BP<double> bpl{};
    if (tid < CD::NB) {
        bpl = bparams(a, g0 + tid);
        sm.leaf[tid] = isleaf[g0 + tid];
    }
    if (w & 1) CD::template run<1>(sm, out, u, b, ib, a, g0, w, lane, tv, bpl, wgt, partial);
    else CD::template run<0>(sm, out, u, b, ib, a, g0, w, lane, tv, bpl, wgt, partial);
}

template <int D, int M, typename T, int RD>
__global__ __launch_bounds__(256) void k_resid_restrict_col(const T *__restrict__ u, const T *__restrict__ b,
                                                            const T *__restrict__ ib, const int *__restrict__ htab,
                                                            const int *__restrict__ parent, LevelArgs<T> a,
                                                            unsigned ngroups, T *__restrict__ rc, T *__restrict__ uc,
                                                            T *__restrict__ u1c, const T *__restrict__ ucin,
                                                            const int *__restrict__ leaves, int nleafc) {
    using CR = ColResid<D, M, T, RD>;
    using HT = Halo<D>;
    constexpr int V = CR::V, ND = CR::ND;
    const int tid = threadIdx.x;
    if (blockIdx.x >= ngroups) {
        const int64_t i = (int64_t)(blockIdx.x - ngroups) * 256 + tid;
        if (i < (int64_t)nleafc * V) {
            const size_t o = (size_t)leaves[i / V] * V + (size_t)(i % V);
            u1c[o] = ucin[o];
        }
        return;
    }
    __shared__ typename CR::Smem sm;
    const int g0 = a.blk0 + xcd_remap(blockIdx.x, ngroups) * CR::NB;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    int tv = 0;
    if (tid < CR::NB * ND) {
        const int lb = tid / ND;
        tv = htab[(size_t)(g0 + lb) * HT::N + (tid - lb * ND)];
    }
    BP<T> bpl{};
    if (tid < CR::NB) bpl = bparams(a, g0 + tid);
    if (w & 1) CR::template run<1>(sm, u, b, ib, parent, a, g0, w, lane, tv, bpl, rc, uc, u1c);
    else CR::template run<0>(sm, u, b, ib, parent, a, g0, w, lane, tv, bpl, rc, uc, u1c);
}

// k_postrestrict on a level of complete sibling groups: the parent of group g
// is fully covered, so one work group stages the parent's u (level l-1, with
// its depth-1 halo) in LDS and forms b = -(r + L u / s) on all its cells.
// Same operations as k_postrestrict (the tile stencil equals the gather one).
template <int D, int M, typename T, int RD>
__global__ __launch_bounds__(256) void k_postrestrict_group(T *__restrict__ bout, const T *__restrict__ u,
                                                            const T *__restrict__ r, const T *__restrict__ ib,
                                                            const int *__restrict__ htab,
                                                            const int *__restrict__ coord,
                                                            const int *__restrict__ parent, int fblk0,
                                                            LevelArgs<T> ac) {
    using LY = Lay<D, M>;
    using TL = Tile<D, M>;
    constexpr int NB = 1 << D, V = LY::V;
    constexpr int PER = (V + 255) / 256;
    __shared__ T tile[TL::SZ];
    __shared__ int tab[2 * D];
    __shared__ int blks[1];
    // the parent's r first: its loads stay in flight while the tile is staged
    const int pb = parent[fblk0 + (int)blockIdx.x * NB];
    T rv[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int q = threadIdx.x + 256 * j;
        rv[j] = q < V ? r[(size_t)pb * V + q] : T(0);
    }
    load_tile<D, M, T, 1, 256, true>(tile, tab, blks, u, ib, htab, nullptr, pb, 1, ac.sg);
    const BP<T> bp = bparams(ac, pb);
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int q = threadIdx.x + 256 * j;
        if (q >= V) break;
        int x, y, z;
        LY::decode(q, x, y, z);
        const int p = TL::at(x, y, z);
        T wm, wp;
        rweights<D, M, T, RD>(ac, coord, pb, RD == 0 ? x : y, wm, wp);
        T Lu = tstencil<D, M, T, RD>(tile, p, tile[p], bp.cc, wm, wp);
        bout[(size_t)pb * V + q] = -(rv[j] + Lu / bp.s);
    }
}

// ---- electric field ----

template <typename T>
struct EFieldArgs {
    int blk0;    // first device block of the level
    T hl;        // cell size of the level: h / 2^(l-1)
    T e0[3];     // constant external field
    T sg[6];     // boundary ghost signs of u
    // free boundary (fill_ghost_free!, src/freebc.jl:218-238), 2D only: the
    // ghosts on the +fdim face of the outermost blocks (coordinate fkd along
    // fdim) gain 2 mean(fbb over fp entries); fbb == nullptr: none
    const T *fbb;
    const int *coord;
    int fdim, fkd, fp;
};

// electric_field! (src/multigrid.jl:421-449) with a constant external field
// (src/multigrid.jl:451) on one block per work group: u with its depth-1 halo
// in LDS (the ghosts of the post-solve fills, src/streamer.jl:340-349);
// E_d on the M+1 faces of every line along d, (u[c - e_d] - u[c]) / h_l + e0_d
// (face i is the lower face of cell i), stored as ef[blk][d][i + (M+1) o]
// with o the other coordinates in increasing dimension order; |E| at the cell
// centres, sqrt(sum_d (E_d(lower) + E_d(upper))^2 / 4), in the block layout.
template <int D, int M, typename T>
__global__ __launch_bounds__(256) void k_efield(T *__restrict__ eabs, T *__restrict__ ef, const T *__restrict__ u,
                                                const T *__restrict__ ib, const int *__restrict__ htab,
                                                EFieldArgs<T> ea) {
    using LY = Lay<D, M>;
    using TL = Tile<D, M>;
    constexpr int V = LY::V, NF = (M + 1) * (D == 2 ? M : M * M);
    __shared__ T tile[TL::SZ];
    __shared__ int tab[2 * D];
    __shared__ int blks[1];
    const int blk = ea.blk0 + blockIdx.x;
    load_tile<D, M, T, 1, 256>(tile, tab, blks, u, ib, htab, nullptr, blk, 1, ea.sg);
    if constexpr (D == 2) {
        if (ea.fbb != nullptr && ea.coord[blk * D + ea.fdim] == ea.fkd) {
            const int od = 1 - ea.fdim;
            const int i0 = (ea.coord[blk * D + od] - 1) * M;
            for (int i = threadIdx.x; i < M; i += 256) {
                T sum = T(0);
                for (int j = ea.fp * (i0 + i); j < ea.fp * (i0 + i + 1); j++) sum += ea.fbb[j];
                const T bmean = sum / T(ea.fp);
                int c[3] = {0, 0, 0};
                c[ea.fdim] = M;
                c[od] = i;
                tile[TL::at(c[0], c[1], 0)] += T(2) * bmean;
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int d = 0; d < D; d++) {
        const int da = d == 0 ? 1 : 0, db = d == 2 ? 1 : 2;
        for (int i = threadIdx.x; i < NF; i += 256) {
            const int fi = i % (M + 1), o = i / (M + 1);
            int c[3] = {0, 0, 0};
            c[d] = fi;
            c[da] = o % M;
            if (D == 3) c[db] = o / M;
            const int p = TL::at(c[0], c[1], c[2]);
            ef[((size_t)blk * D + d) * NF + i] = (tile[p - TL::stride(d)] - tile[p]) / ea.hl + ea.e0[d];
        }
    }
    for (int q = threadIdx.x; q < V; q += 256) {
        int x, y, z;
        LY::decode(q, x, y, z);
        const int p = TL::at(x, y, z);
        T f2 = T(0);
#pragma unroll
        for (int d = 0; d < D; d++) {
            const T lo = (tile[p - TL::stride(d)] - tile[p]) / ea.hl + ea.e0[d];
            const T hi = (tile[p] - tile[p + TL::stride(d)]) / ea.hl + ea.e0[d];
            const T t = lo + hi;
            f2 += t * t / T(4);
        }
        eabs[(size_t)blk * V + q] = sqrt(f2);
    }
}

// ---- free boundary conditions (src/freebc.jl, 2D cylindrical) ----

// getouter! (src/freebc.jl:112-142) as a gather: u1[j] = u[src[j]] / hl[j],
// where src names the outermost cell of the finest level covering entry j
__global__ __launch_bounds__(WG) void k_fb_gather(double *__restrict__ u1, const double *__restrict__ u,
                                                  const int64_t *__restrict__ src, const double *__restrict__ hl,
                                                  int64_t n) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) u1[j] = u[src[j]] / hl[j];
}

// y = S x for an n x n sine matrix given transposed (ST[j n + k] = S[k][j]),
// summed in j order; then y_k = fm[k] * y_k (fm != nullptr: am = f .* u1,
// src/freebc.jl:103) or y_k = y_k h / 2^(level-1) / 2 (src/freebc.jl:106)
__global__ __launch_bounds__(WG) void k_fb_dst(const double *__restrict__ ST, const double *__restrict__ x,
                                               double *__restrict__ y, int64_t n, const double *__restrict__ fm,
                                               double h, double lev) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double acc = 0.0;
    for (int64_t j = 0; j < n; j++) acc += ST[j * n + k] * x[j];
    y[k] = fm ? fm[k] * acc : acc * h / lev / 2.0;
}

// setouter! (src/freebc.jl:149-215): cell cells[i] receives, in order, the
// contributions r in [rs[i], rs[i+1]): q += coef[r] (sum of b over
// [bs[r], bs[r] + bl[r]) / bd[r]) / sl[r]
__global__ __launch_bounds__(WG) void k_fb_setouter(double *__restrict__ q, const int64_t *__restrict__ cells,
                                                    const int *__restrict__ rs, const double *__restrict__ coef,
                                                    const int64_t *__restrict__ bs, const int *__restrict__ bl,
                                                    const double *__restrict__ bd, const double *__restrict__ sl,
                                                    const double *__restrict__ b, int ncells) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncells) return;
    double v = q[cells[i]];
    for (int r = rs[i]; r < rs[i + 1]; r++) {
        double sum = 0.0;
        for (int64_t j = bs[r]; j < bs[r] + bl[r]; j++) sum += b[j];
        v += coef[r] * (sum / bd[r]) / sl[r];
    }
    q[cells[i]] = v;
}

// resnorm partials of u = 0 (the residual is -b): the same blocks, visiting
// order and tree sums as k_resid_tile<RES_NORM>, so the partials are bitwise
// those of a residual norm of a zero field, without reading one. be (may be
// null) receives (float) b on the visited blocks.
template <int D, int M, typename T, int BPW, int NT>
__global__ __launch_bounds__(NT) void k_bnorm(const T *__restrict__ b, const int *__restrict__ list, int n,
                                              double wgt, double *__restrict__ partial, float *__restrict__ be) {
    constexpr int V = Lay<D, M>::V;
    __shared__ int blks[BPW];
    __shared__ double red[BPW * NT];
    const int first = xcd_remap(blockIdx.x, gridDim.x) * BPW;
    const int nloc = min(BPW, n - first);
    if ((int)threadIdx.x < BPW) blks[threadIdx.x] = ((int)threadIdx.x < nloc) ? list[first + threadIdx.x] : 0;
    __syncthreads();
    // every load of the thread's cells before the first store and sum
    constexpr int PER = (BPW * V + NT - 1) / NT;
    double acc[BPW];
#pragma unroll
    for (int lb = 0; lb < BPW; lb++) acc[lb] = 0.0;
    T bv[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int i = threadIdx.x + NT * j, lb = min(i / V, BPW - 1);
        bv[j] = (i < nloc * V) ? b[(size_t)blks[lb] * V + (i - lb * V)] : T(0);
    }
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int i = threadIdx.x + NT * j, lb = min(i / V, BPW - 1);
        if (i >= nloc * V) break;
        if (be) be[(size_t)blks[lb] * V + (i - lb * V)] = (float)bv[j];
        const T rr = -bv[j];
#pragma unroll
        for (int k = 0; k < BPW; k++)
            if (k == lb) acc[k] += (double)rr * (double)rr;
    }
    // the halving tree of every block at once (the same pairs as one block
    // at a time; blocks past nloc are summed and dropped): the levels of 64
    // and more across waves in LDS, the rest inside one wave per block by
    // lane shuffles
    static_assert(NT % 64 == 0 && NT >= 128, "whole waves, at least two");
#pragma unroll
    for (int lb = 0; lb < BPW; lb++) red[lb * NT + threadIdx.x] = acc[lb];
    __syncthreads();
    for (int w = NT / 2; w >= 64; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int lb = 0; lb < BPW; lb++) red[lb * NT + threadIdx.x] += red[lb * NT + threadIdx.x + w];
        }
        __syncthreads();
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int lb = wave; lb < BPW; lb += NT / 64) {
        double v = red[lb * NT + lane];
#pragma unroll
        for (int w = 32; w > 0; w >>= 1) {
            const double o = __shfl_down(v, w, 64);
            if (lane < w) v += o;
        }
        if (lane == 0 && lb < nloc) partial[first + lb] = v * wgt;
    }
}

// ---- exact level-1 solve (FP64) ----

// level-1 valid cells -> uniform grid (column-major, dimension 1 fastest), times c
template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_grid_gather(double *__restrict__ g, const T *__restrict__ x,
                                                    const int *__restrict__ coord, int blk0, int nblk,
                                                    int N0, int N1, int nper, size_t total,
                                                    const double *__restrict__ cs) {
    using LY = Lay<D, M>;
    int gt = blockIdx.x * blockDim.x + threadIdx.x;
    int lb = gt / LY::V;
    if (lb >= nblk) return;
    int blk = blk0 + lb;
    int q = gt - lb * LY::V;
    int c0, c1, c2;
    LY::decode(q, c0, c1, c2);
    int grid = lb / nper;
    size_t gi = (size_t)grid * total + (size_t)(M * (coord[blk * D] - 1) + c0) +
                (size_t)N0 * ((size_t)(M * (coord[blk * D + 1] - 1) + c1) +
                              (D == 3 ? (size_t)N1 * (size_t)(M * (coord[blk * D + 2] - 1) + c2) : 0));
    g[gi] = cs[grid] * (double)x[(size_t)blk * LY::V + q];
}

template <int D, int M, typename T>
__global__ __launch_bounds__(WG) void k_grid_scatter(T *__restrict__ x, const double *__restrict__ g,
                                                     const int *__restrict__ coord, int blk0, int nblk,
                                                     int N0, int N1, int nper, size_t total) {
    using LY = Lay<D, M>;
    int gt = blockIdx.x * blockDim.x + threadIdx.x;
    int lb = gt / LY::V;
    if (lb >= nblk) return;
    int blk = blk0 + lb;
    int q = gt - lb * LY::V;
    int c0, c1, c2;
    LY::decode(q, c0, c1, c2);
    int grid = lb / nper;
    size_t gi = (size_t)grid * total + (size_t)(M * (coord[blk * D] - 1) + c0) +
                (size_t)N0 * ((size_t)(M * (coord[blk * D + 1] - 1) + c1) +
                              (D == 3 ? (size_t)N1 * (size_t)(M * (coord[blk * D + 2] - 1) + c2) : 0));
    x[(size_t)blk * LY::V + q] = (T)g[gi];
}

// Y = A X along one axis of a column-major grid: A is n x n (row-major), the
// grid is viewed as [outer][n][inner] (inner = stride of the axis) and every
// (outer, inner) pair is a column of X. Register-tiled FP64 GEMM: a work
// group computes a 64 x TC tile of Y (rows i, columns) with 256 threads, each
// an RA x 4 sub-tile (RA = TC / 16), over k-chunks of 16 staged in LDS
// (zero-padded past n). Each output is one dot product summed in k order, so
// every TC gives the same bits; narrow tiles give small grids more work groups.
template <int TC>
__global__ __launch_bounds__(256) void k_axis_gemm(const double *__restrict__ A, int n, int inner, int ncol,
                                                   const double *__restrict__ in, double *__restrict__ out) {
    constexpr int TI = 64, TK = 16, RA = TC / 16, TCQ = TC / 4;
    __shared__ double As[TK][TI + 1];
    __shared__ double Xs[TK][TC + 1];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.y * TI, c0 = blockIdx.x * TC;
    const int ti = (tid / TCQ) * RA, tc = (tid % TCQ) * 4;
    double acc[RA][4];
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[a][q] = 0.0;
    for (int k0 = 0; k0 < n; k0 += TK) {
        // A tile: rows i0..i0+63, columns k0..k0+15
        for (int e = tid; e < TI * TK; e += 256) {
            int ii = e / TK, kk = e - ii * TK;
            int gi = i0 + ii;
            As[kk][ii] = (gi < n && k0 + kk < n) ? A[(size_t)gi * n + k0 + kk] : 0.0;
        }
        // X tile: rows k0..k0+15, columns c0..c0+63
        for (int e = tid; e < TK * TC; e += 256) {
            int kk, cc;
            if (inner == 1) { cc = e / TK; kk = e - cc * TK; }   // k contiguous per column
            else { kk = e / TC; cc = e - kk * TC; }              // columns contiguous
            int col = c0 + cc;
            double v = 0.0;
            if (col < ncol && k0 + kk < n) {
                size_t o = (size_t)(col / inner), j = (size_t)(col % inner);
                v = in[o * (size_t)n * inner + (size_t)(k0 + kk) * inner + j];
            }
            Xs[kk][cc] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; kk++) {
            double av[RA], xv[4];
#pragma unroll
            for (int a = 0; a < RA; a++) av[a] = As[kk][ti + a];
#pragma unroll
            for (int q = 0; q < 4; q++) xv[q] = Xs[kk][tc + q];
#pragma unroll
            for (int a = 0; a < RA; a++)
#pragma unroll
                for (int q = 0; q < 4; q++) acc[a][q] += av[a] * xv[q];
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < RA; a++) {
        int gi = i0 + ti + a;
        if (gi >= n) continue;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int col = c0 + tc + q;
            if (col >= ncol) continue;
            size_t o = (size_t)(col / inner), j = (size_t)(col % inner);
            out[o * (size_t)n * inner + (size_t)gi * inner + j] = acc[a][q];
        }
    }
}

// k_axis_gemm for n <= 64 in one load phase: all of A (n x n) and the n x TC
// column tile of X in LDS behind one barrier, then every thread sums its
// 4 outputs over k = 0 .. n-1 in order (the same bits as k_axis_gemm).
template <int TC>
__global__ __launch_bounds__(256) void k_axis_gemm_small(const double *__restrict__ A, int n, int inner, int ncol,
                                                         const double *__restrict__ in, double *__restrict__ out) {
    constexpr int NMAX = 64, RA = TC / 16, TCQ = TC / 4;
    __shared__ double As[NMAX][NMAX + 1];   // As[k][i] = A[i][k]
    __shared__ double Xs[NMAX][TC + 1];
    const int tid = threadIdx.x;
    const int c0 = blockIdx.x * TC;
    for (int e = tid; e < n * n; e += 256) {
        const int i = e / n, k = e - i * n;
        As[k][i] = A[e];
    }
    for (int e = tid; e < n * TC; e += 256) {
        int kk, cc;
        if (inner == 1) { cc = e / n; kk = e - cc * n; }   // k contiguous per column
        else { kk = e / TC; cc = e - kk * TC; }            // columns contiguous
        const int col = c0 + cc;
        double v = 0.0;
        if (col < ncol) {
            const size_t o = (size_t)(col / inner), j = (size_t)(col % inner);
            v = in[o * (size_t)n * inner + (size_t)kk * inner + j];
        }
        Xs[kk][cc] = v;
    }
    __syncthreads();
    const int ti = (tid / TCQ) * RA, tc = (tid % TCQ) * 4;
    double acc[RA][4];
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[a][q] = 0.0;
    for (int k = 0; k < n; k++) {
        double av[RA], xv[4];
#pragma unroll
        for (int a = 0; a < RA; a++) av[a] = As[k][ti + a];
#pragma unroll
        for (int q = 0; q < 4; q++) xv[q] = Xs[k][tc + q];
#pragma unroll
        for (int a = 0; a < RA; a++)
#pragma unroll
            for (int q = 0; q < 4; q++) acc[a][q] += av[a] * xv[q];
    }
#pragma unroll
    for (int a = 0; a < RA; a++) {
        const int gi = ti + a;
        if (gi >= n) continue;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int col = c0 + tc + q;
            if (col >= ncol) continue;
            const size_t o = (size_t)(col / inner), j = (size_t)(col % inner);
            out[o * (size_t)n * inner + (size_t)gi * inner + j] = acc[a][q];
        }
    }
}

// k_axis_gemm for n <= 256 without LDS; each output is the k-ordered sum of
// products of k_axis_gemm (same bits).
// Axis with stride 1 (a column is n contiguous values): thread i computes row
// i of RB consecutive columns (n threads rounded up to whole waves). Row k of
// At (= column k of A) is a coalesced load; X[col][k] is the same for every
// lane (a scalar load).
template <int RB>
__global__ __launch_bounds__(256) void k_axis_gemm_c(const double *__restrict__ At, int n, int ncol,
                                                    const double *__restrict__ in, double *__restrict__ out) {
    const int i = threadIdx.x;
    const int c0 = blockIdx.x * RB;
    if (i >= n) return;
    double acc[RB];
#pragma unroll
    for (int r = 0; r < RB; r++) acc[r] = 0.0;
#pragma unroll 8
    for (int k = 0; k < n; k++) {
        const double a = At[(size_t)k * n + i];
#pragma unroll
        for (int r = 0; r < RB; r++) {
            const int col = min(c0 + r, ncol - 1);
            acc[r] += a * in[(size_t)col * n + k];
        }
    }
#pragma unroll
    for (int r = 0; r < RB; r++)
        if (c0 + r < ncol) out[(size_t)(c0 + r) * n + i] = acc[r];
}

// Axis with a stride (inner) that is a multiple of 64: lane j is one of 64
// consecutive columns (the same outer index), and the work group computes RB
// rows of them. X[o][k][j] is a coalesced load; A[i][k] is the same for every
// lane (a scalar load).
template <int RB>
__global__ __launch_bounds__(64) void k_axis_gemm_r(const double *__restrict__ A, int n, int inner,
                                                    const double *__restrict__ in, double *__restrict__ out) {
    const int col = blockIdx.x * 64 + threadIdx.x;
    const int i0 = blockIdx.y * RB;
    const size_t o = (size_t)(col / inner), jj = (size_t)(col % inner);
    const double *__restrict__ x = in + o * (size_t)n * inner + jj;
    double acc[RB];
#pragma unroll
    for (int r = 0; r < RB; r++) acc[r] = 0.0;
#pragma unroll 8
    for (int k = 0; k < n; k++) {
        const double xv = x[(size_t)k * inner];
#pragma unroll
        for (int r = 0; r < RB; r++) {
            const int i = min(i0 + r, n - 1);
            acc[r] += A[(size_t)i * n + k] * xv;
        }
    }
    double *__restrict__ y = out + o * (size_t)n * inner + jj;
#pragma unroll
    for (int r = 0; r < RB; r++)
        if (i0 + r < n) y[(size_t)(i0 + r) * inner] = acc[r];
}

// k_axis_gemm for 64 < n <= 256 with the whole k range of a 16 x 16 output
// tile in LDS: the tile's 16 rows of A and 16 columns of X are loaded at once
// (every load in flight before the first product), then each thread sums its
// output over k in increasing order from 0.0 (the bits of k_axis_gemm). One
// load round trip per launch instead of one per k chunk or per few k.
constexpr int AXF = 16;  // output rows and columns per work group
__global__ __launch_bounds__(256) void k_axis_gemm_full(const double *__restrict__ A, int n, int inner, int ncol,
                                                        const double *__restrict__ in, double *__restrict__ out) {
    __shared__ double As[AXF * 257];  // As[r][k], row stride n + 1
    __shared__ double Xs[256 * (AXF + 1)];  // Xs[k][c]
    const int tid = threadIdx.x;
    const int i0 = blockIdx.y * AXF, c0 = blockIdx.x * AXF;
    const int sa = n + 1;
    double va[16], vx[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int e = tid + q * 256, r = e / n, k = e - r * n;
        va[q] = (r < AXF && i0 + r < n) ? A[(size_t)(i0 + r) * n + k] : 0.0;
    }
    // X tile: element (k, c) of column c0 + c; consecutive threads take
    // consecutive k for the stride-1 axis and consecutive columns otherwise
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int e = tid + q * 256;
        int k, c;
        if (inner == 1) { c = e / n; k = e - c * n; }
        else { k = e / AXF; c = e - k * AXF; }
        const int col = c0 + c;
        double v = 0.0;
        if (k < n && c < AXF && col < ncol) {
            const size_t o = (size_t)(col / inner), jj = (size_t)(col % inner);
            v = in[(o * (size_t)n + k) * inner + jj];
        }
        vx[q] = v;
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int e = tid + q * 256, r = e / n, k = e - r * n;
        if (r < AXF) As[r * sa + k] = va[q];
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int e = tid + q * 256;
        int k, c;
        if (inner == 1) { c = e / n; k = e - c * n; }
        else { k = e / AXF; c = e - k * AXF; }
        if (k < n && c < AXF) Xs[k * (AXF + 1) + c] = vx[q];
    }
    __syncthreads();
    const int r = tid / AXF, c = tid - r * AXF;
    const int i = i0 + r, col = c0 + c;
    double acc = 0.0;
#pragma unroll 8
    for (int k = 0; k < n; k++) acc += As[r * sa + k] * Xs[k * (AXF + 1) + c];
    if (i < n && col < ncol) {
        const size_t o = (size_t)(col / inner), jj = (size_t)(col % inner);
        out[(o * (size_t)n + i) * inner + jj] = acc;
    }
}

// divide by the eigenvalues of the Kronecker-sum operator; the grid holds
// `ngrid` stacked level-1 grids of `total` points, grid g with shift[g]
// (Helmholtz: -s k^2; Laplacian: 0)
__global__ __launch_bounds__(WG) void k_eig_scale(double *__restrict__ g, const double *__restrict__ l0,
                                                  const double *__restrict__ l1, const double *__restrict__ l2,
                                                  int N0, int N1, int D, const double *__restrict__ shift,
                                                  size_t total, int ngrid) {
    size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total * ngrid) return;
    size_t gg = gi / total;
    size_t r = gi - gg * total;
    double ev = shift[gg];
    ev += l0[r % N0]; r /= N0;
    ev += l1[r % N1]; r /= N1;
    if (D == 3) ev += l2[r];
    g[gi] /= ev;
}

// ---- level-1 exact solve of 3D grids in three launches ----
//
// For stacked 3D grids [g][z][y][x] (x fastest) whose extents are multiples of
// 8 and at most 64, the transforms of k_axis_gemm (each output one sum over k
// in increasing order, from 0.0, of A[i][k] x[k], no fused multiply-add) in
// the order forward x, y, z; the eigenvalue division of k_eig_scale; inverse
// z, y, x. Each launch keeps a plane or a set of columns in LDS, so the nine
// launches of the axis-by-axis form (gather, three transforms, scale, three
// transforms, scatter) become three with the same bits:
//   k_l1_fwd_xy  plane (g, z), IC output columns: gather, forward x, forward y
//   k_l1_z       CC columns (y, x) of grid g: forward z, divide, inverse z
//   k_l1_inv_yx  plane (g, z), JC output rows: inverse y, inverse x, scatter
// BLK: the input (output) is the level-1 block field, gathered through btab
// (btab[g * nper + (bz * nb1 + by) * nb0 + bx] is the block at block position
// (bx, by, bz) of grid g) and scaled by cs[g]; otherwise a plain grid.

constexpr int L1T = 256;  // threads per work group of the three kernels (four waves)
constexpr int L1S = 65;   // LDS row stride of a staged plane (doubles)

// out[r] = sum over k < K of a(r, k) b(r, k), in increasing k from 0.0 (no
// fused multiply-add), for the R outputs of a thread at once
template <int R, class Fa, class Fb>
__device__ __forceinline__ void l1_dots(double (&out)[R], int K, Fa fa, Fb fb) {
#pragma unroll
    for (int r = 0; r < R; r++) out[r] = 0.0;
#pragma unroll 8
    for (int k = 0; k < K; k++) {
#pragma unroll
        for (int r = 0; r < R; r++) out[r] += fa(r, k) * fb(r, k);
    }
}

// rows of an n x m matrix (m <= 64) staged row by row: wave w takes rows w,
// w + 4, ..., lane = column; every load is issued before the first store
template <class Ld, class St>
__device__ __forceinline__ void l1_rows(int n, int m, Ld ld, St st) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = w + 4 * r;
        v[r] = (row < n && lane < m) ? ld(row, lane) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = w + 4 * r;
        if (row < n && lane < m) st(row, lane, v[r]);
    }
}

template <int M, typename T, bool BLK, int IC>
__global__ __launch_bounds__(L1T) void k_l1_fwd_xy(double *__restrict__ out, const T *__restrict__ src,
                                                   const double *__restrict__ gin, const int *__restrict__ btab,
                                                   const double *__restrict__ cs, const double *__restrict__ A0t,
                                                   const double *__restrict__ A1t, int N0, int N1, int N2) {
    using LY = Lay<3, M>;
    constexpr int RO = 64 * IC / L1T;  // outputs per thread
    constexpr int KS = L1T / IC;       // rows of a0 loaded per pass
    __shared__ double xs[64 * L1S];    // the plane X[y][x]
    __shared__ double a1[64 * 64];     // a1[k][j] = A1[j][k]
    __shared__ double a0[64 * IC];     // a0[k][il] = A0[i0 + il][k]
    __shared__ double y1[64 * IC];     // Y1[y][il]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int z = blockIdx.x % N2, g = blockIdx.x / N2, i0 = blockIdx.y * IC;
    const int il = tid % IC, yb = tid / IC;
    double a0r[RO];
#pragma unroll
    for (int r = 0; r < RO; r++) {
        const int k = yb + KS * r;
        a0r[r] = k < N0 ? A0t[(size_t)k * N0 + i0 + il] : 0.0;
    }
    if constexpr (BLK) {
        const int nb0 = N0 / M, nb1 = N1 / M, nper = nb0 * nb1 * (N2 / M);
        const int *tb = btab + (size_t)g * nper + (z / M) * nb0 * nb1;
        const double c = cs[g];
        int blk[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int y = w + 4 * r;
            blk[r] = (y < N1 && lane < N0) ? tb[(y / M) * nb0 + lane / M] : 0;
        }
        double v[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int y = w + 4 * r;
            v[r] = (y < N1 && lane < N0)
                       ? c * (double)src[(size_t)blk[r] * LY::V + LY::sidx(lane % M, y % M, z % M)]
                       : 0.0;
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int y = w + 4 * r;
            if (y < N1 && lane < N0) xs[y * L1S + lane] = v[r];
        }
    } else {
        const double *p = gin + ((size_t)g * N2 + z) * N1 * N0;
        l1_rows(N1, N0, [&](int y, int x) { return p[(size_t)y * N0 + x]; },
                [&](int y, int x, double v) { xs[y * L1S + x] = v; });
    }
    l1_rows(N1, N1, [&](int k, int j) { return A1t[(size_t)k * N1 + j]; },
            [&](int k, int j, double v) { a1[k * 64 + j] = v; });
#pragma unroll
    for (int r = 0; r < RO; r++) {
        const int k = yb + KS * r;
        if (k < N0) a0[k * IC + il] = a0r[r];
    }
    __syncthreads();
    // Y1[y][il] = sum_k A0[i0 + il][k] X[y][k], y = yb + KS r
    {
        double acc[RO];
        l1_dots(acc, N0, [&](int r, int k) { return a0[k * IC + il]; },
                [&](int r, int k) { return xs[(yb + KS * r) * L1S + k]; });
#pragma unroll
        for (int r = 0; r < RO; r++) y1[(yb + KS * r) * IC + il] = acc[r];
    }
    __syncthreads();
    // Y2[j][il] = sum_k A1[j][k] Y1[k][il], j = yb + KS r
    double acc[RO];
    l1_dots(acc, N1, [&](int r, int k) { return a1[k * 64 + yb + KS * r]; },
            [&](int r, int k) { return y1[k * IC + il]; });
    double *po = out + ((size_t)g * N2 + z) * N1 * N0 + i0 + il;
#pragma unroll
    for (int r = 0; r < RO; r++) {
        const int j = yb + KS * r;
        if (j < N1) po[(size_t)j * N0] = acc[r];
    }
}

// forward z, division by shift[g] + l0[x] + l1[y] + l2[z] (k_eig_scale's sum
// order), inverse z on CC consecutive columns c = y N0 + x of grid g
template <int CC>
__global__ __launch_bounds__(L1T) void k_l1_z(double *__restrict__ out, const double *__restrict__ in,
                                              const double *__restrict__ A2t, const double *__restrict__ B2t,
                                              const double *__restrict__ l0, const double *__restrict__ l1,
                                              const double *__restrict__ l2, const double *__restrict__ shift,
                                              int N0, int N1, int N2) {
    constexpr int RO = 64 * CC / L1T;  // outputs per thread
    constexpr int ZS = L1T / CC;       // z rows per pass
    __shared__ double am[64 * 64];     // am[k][zo] = A2[zo][k]
    __shared__ double bm[64 * 64];     // bm[k][zo] = B2[zo][k]
    __shared__ double ts[64 * CC];     // the columns, Y[z][cc]
    __shared__ double ws[64 * CC];     // the divided forward transform
    const int tid = threadIdx.x;
    const int ncb = N0 * N1 / CC;
    const int g = blockIdx.x / ncb, c0 = (blockIdx.x - g * ncb) * CC;
    const size_t plane = (size_t)N0 * N1;
    const int cc = tid % CC, zb = tid / CC;
    const double *pin = in + (size_t)g * N2 * plane + c0 + cc;
    double tv[RO];
#pragma unroll
    for (int r = 0; r < RO; r++) {
        const int zz = zb + ZS * r;
        tv[r] = zz < N2 ? pin[(size_t)zz * plane] : 0.0;
    }
    const int c = c0 + cc, y = c / N0, x = c - y * N0;
    double ev0 = shift[g];
    ev0 += l0[x];
    ev0 += l1[y];
    double ev[RO];
#pragma unroll
    for (int r = 0; r < RO; r++) {
        const int zo = zb + ZS * r;
        ev[r] = ev0 + (zo < N2 ? l2[zo] : 0.0);
    }
    l1_rows(N2, N2, [&](int k, int zo) { return A2t[(size_t)k * N2 + zo]; },
            [&](int k, int zo, double v) { am[k * 64 + zo] = v; });
    l1_rows(N2, N2, [&](int k, int zo) { return B2t[(size_t)k * N2 + zo]; },
            [&](int k, int zo, double v) { bm[k * 64 + zo] = v; });
#pragma unroll
    for (int r = 0; r < RO; r++) {
        const int zz = zb + ZS * r;
        if (zz < N2) ts[zz * CC + cc] = tv[r];
    }
    __syncthreads();
    {
        double acc[RO];
        l1_dots(acc, N2, [&](int r, int k) { return am[k * 64 + zb + ZS * r]; },
                [&](int r, int k) { return ts[k * CC + cc]; });
#pragma unroll
        for (int r = 0; r < RO; r++) ws[(zb + ZS * r) * CC + cc] = acc[r] / ev[r];
    }
    __syncthreads();
    double acc[RO];
    l1_dots(acc, N2, [&](int r, int k) { return bm[k * 64 + zb + ZS * r]; },
            [&](int r, int k) { return ws[k * CC + cc]; });
    double *po = out + (size_t)g * N2 * plane + c0 + cc;
#pragma unroll
    for (int r = 0; r < RO; r++) {
        const int zo = zb + ZS * r;
        if (zo < N2) po[(size_t)zo * plane] = acc[r];
    }
}

template <int M, typename T, bool BLK, int JC>
__global__ __launch_bounds__(L1T) void k_l1_inv_yx(T *__restrict__ dst, double *__restrict__ gout,
                                                   const double *__restrict__ in, const int *__restrict__ btab,
                                                   const double *__restrict__ B1t, const double *__restrict__ B0t,
                                                   int N0, int N1, int N2) {
    using LY = Lay<3, M>;
    constexpr int RO = JC / 4;       // outputs per thread: rows jl = w + 4 r, column = lane
    constexpr int KS = L1T / JC;     // rows of b1 loaded per pass
    __shared__ double vs[64 * L1S];  // the plane V[y][x]
    __shared__ double b0[64 * 64];   // b0[k][i] = B0[i][k]
    __shared__ double b1[64 * JC];   // b1[k][jl] = B1[j0 + jl][k]
    __shared__ double ts[JC * L1S];  // T[jl][x]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int z = blockIdx.x % N2, g = blockIdx.x / N2, j0 = blockIdx.y * JC;
    const int jq = tid % JC, kb = tid / JC;
    double b1r[64 / KS];
#pragma unroll
    for (int r = 0; r < 64 / KS; r++) {
        const int k = kb + KS * r;
        b1r[r] = k < N1 ? B1t[(size_t)k * N1 + j0 + jq] : 0.0;
    }
    const double *pin = in + ((size_t)g * N2 + z) * N1 * N0;
    l1_rows(N1, N0, [&](int y, int x) { return pin[(size_t)y * N0 + x]; },
            [&](int y, int x, double v) { vs[y * L1S + x] = v; });
    l1_rows(N0, N0, [&](int k, int i) { return B0t[(size_t)k * N0 + i]; },
            [&](int k, int i, double v) { b0[k * 64 + i] = v; });
#pragma unroll
    for (int r = 0; r < 64 / KS; r++) {
        const int k = kb + KS * r;
        if (k < N1) b1[k * JC + jq] = b1r[r];
    }
    __syncthreads();
    // T[jl][x] = sum_k B1[j0 + jl][k] V[k][x], jl = w + 4 r, x = lane
    {
        double acc[RO];
        l1_dots(acc, N1, [&](int r, int k) { return b1[k * JC + w + 4 * r]; },
                [&](int r, int k) { return vs[k * L1S + lane]; });
#pragma unroll
        for (int r = 0; r < RO; r++) ts[(w + 4 * r) * L1S + lane] = acc[r];
    }
    __syncthreads();
    // U[jl][i] = sum_k B0[i][k] T[jl][k], i = lane
    double acc[RO];
    l1_dots(acc, N0, [&](int r, int k) { return b0[k * 64 + lane]; },
            [&](int r, int k) { return ts[(w + 4 * r) * L1S + k]; });
    if (lane >= N0) return;
    const int nb0 = N0 / M, nb1 = N1 / M, nper = nb0 * nb1 * (N2 / M);
#pragma unroll
    for (int r = 0; r < RO; r++) {
        const int j = j0 + w + 4 * r;
        if constexpr (BLK) {
            const int blk = btab[(size_t)g * nper + ((z / M) * nb1 + j / M) * nb0 + lane / M];
            dst[(size_t)blk * LY::V + LY::sidx(lane % M, j % M, z % M)] = (T)acc[r];
        } else {
            gout[(((size_t)g * N2 + z) * N1 + j) * N0 + lane] = acc[r];
        }
    }
}

// ---- mixed precision helpers ----

// u += e over n cells (FP64 += FP32)
__global__ __launch_bounds__(WG) void k_add_correction(double *__restrict__ u, const float *__restrict__ e, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] += (double)e[i];
}

// The end of a mixed-precision outer step on a grouped level, run from the
// finest level down: on every block of sibling group g, u += e if the block is
// a leaf (a covered block already holds the restriction of its children,
// written by the finer level's launch, and u += e there would be overwritten
// anyway) and e = 0; then the parent's cells get the restriction of the
// group's u (k_restrict's order) and e = 0. Work groups past ngroups do
// u += e, e = 0 on the leaf blocks leaves[0..nleafc) of the coarser level
// (used for the coarsest level, which has no parent). The values equal those
// of u += e everywhere followed by restrict! level by level. ZERO: the leaf
// blocks' u is known to be 0 (the first step of a solve from u = 0) and is
// not read: u = 0.0 + e. KEEPE: e is not cleared on the group's own blocks
// (the next correction's first sweep there reads none of it; the parents'
// e is cleared).
template <int D, int M, bool ZERO = false, bool KEEPE = false>
__global__ __launch_bounds__(256) void k_add_restrict(double *__restrict__ u, float *__restrict__ e,
                                                      const int *__restrict__ isleaf, const int *__restrict__ parent,
                                                      int blk0, unsigned ngroups, const int *__restrict__ leaves,
                                                      int nleafc) {
    using LY = Lay<D, M>;
    constexpr int NB = 1 << D, V = LY::V, MH = LY::MH;
    const int tid = threadIdx.x;
    if (blockIdx.x >= ngroups) {
        const int c0 = (int)(blockIdx.x - ngroups) * 256;
        for (int i = c0 + tid; i < c0 + 256 && i < nleafc * V; i += 256) {
            const size_t o = (size_t)leaves[i / V] * V + (size_t)(i % V);
            u[o] = (ZERO ? 0.0 : u[o]) + (double)e[o];
            e[o] = 0.0f;
        }
        return;
    }
    // every load of the thread's cells is issued before the first store
    // (a loop that stores between its loads waits out each load in turn)
    constexpr int PER = NB * V / 256;
    static_assert(NB * V % 256 == 0, "group cells must be a multiple of the work-group size");
    __shared__ double ut[NB * V];
    __shared__ int leaf[NB];
    const int g0 = blk0 + (int)blockIdx.x * NB;
    if (tid < NB) leaf[tid] = isleaf[g0 + tid];
    __syncthreads();
    float ev[PER];
    double uv[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int i = tid + 256 * j;
        const size_t o = (size_t)g0 * V + i;
        ev[j] = e[o];
        uv[j] = (ZERO && leaf[i / V]) ? 0.0 : u[o];
    }
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int i = tid + 256 * j;
        const size_t o = (size_t)g0 * V + i;
        double v = uv[j];
        if (leaf[i / V]) {
            v = (ZERO ? 0.0 : uv[j]) + (double)ev[j];
            u[o] = v;
        }
        if (!KEEPE) e[o] = 0.0f;
        ut[i] = v;
    }
    __syncthreads();
    const int pb = parent[g0];
    for (int i = tid; i < V; i += 256) {
        int X0, X1, X2;
        LY::decode(i, X0, X1, X2);
        const int k = (X0 >= MH ? 1 : 0) | (X1 >= MH ? 2 : 0) | (D == 3 && X2 >= MH ? 4 : 0);
        const int j0 = X0 - (k & 1) * MH, j1 = X1 - ((k >> 1) & 1) * MH, j2 = (D == 3) ? X2 - ((k >> 2) & 1) * MH : 0;
        double f = 0.0;
#pragma unroll
        for (int q = 0; q < LY::NC; q++) {
            const int i0 = q & 1, i1 = (q >> 1) & 1, i2 = (q >> 2) & 1;
            f += ut[k * V + LY::sidx(2 * j0 + i0, 2 * j1 + i1, D == 3 ? 2 * j2 + i2 : 0)];
        }
        const size_t o = (size_t)pb * V + i;
        u[o] = f / double(LY::NC);
        e[o] = 0.0f;
    }
}

}  // namespace srmg
