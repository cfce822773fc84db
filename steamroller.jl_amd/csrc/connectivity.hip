// connectivity!(conn, tree, stencil) (src/connectivity.jl:106-149) on the GPU.
//
// The reference scans, for every block of every level and every offset of the
// stencil (StarStencil for the order-2 operators, BoxStencil for order 4 and
// the refinement marks), whether the neighbouring position is outside the
// level's domain (a Boundary record), holds a block (Neighbor) or lies in the
// parent level (RefBoundary), and gives every block above level 1 its Child
// record. Here one thread handles one block: the block positions go into a
// hash table keyed by (level, Morton code of the coordinate), a first pass
// counts each block's records, exclusive scans give every block its offsets
// in the four lists (three launches: per-chunk totals, a scan of the totals,
// the chunks rescanned from their offsets), and a second pass writes the
// records. The lists come out
// in the reference's order: level by level, blocks in `pairs` order, offsets
// in stencil order. Block positions may not repeat within a level; a missing
// parent (proper nesting broken) is an error, as the reference asserts.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/srmg.h"
#include "common.h"

namespace srmg {
namespace conn {

constexpr uint64_t EMPTY = ~0ull;

// ---- exclusive prefix sums of K int arrays of n entries (array k at x + k n) ----

constexpr int SCAN_WG = 256, SCAN_ITEMS = 16, SCAN_CHUNK = SCAN_WG * SCAN_ITEMS;

// exclusive prefix of v over the work group (threads in order) and the total
__device__ __forceinline__ int wg_exclusive(int v, int *wave_tot, int &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) wave_tot[w] = inc;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < SCAN_WG / 64; k++) {
        const int t = wave_tot[k];
        if (k < w) before += t;
        total += t;
    }
    __syncthreads();
    return before + inc - v;
}

// per chunk of SCAN_CHUNK entries: its total (part[k nchunk + chunk])
__global__ __launch_bounds__(SCAN_WG) void k_scan_totals(const int *__restrict__ x, int64_t n, int nchunk,
                                                         int *__restrict__ part) {
    __shared__ int wt[SCAN_WG / 64];
    const int k = blockIdx.y, c = blockIdx.x;
    const int *xa = x + (size_t)k * n;
    const int64_t i0 = (int64_t)c * SCAN_CHUNK + (int64_t)threadIdx.x * SCAN_ITEMS;
    int v = 0;
    for (int j = 0; j < SCAN_ITEMS; j++)
        if (i0 + j < n) v += xa[i0 + j];
    int total;
    (void)wg_exclusive(v, wt, total);
    if (threadIdx.x == 0) part[(size_t)k * nchunk + c] = total;
}

// exclusive scan of each array's chunk totals, in place (one work group per array)
__global__ __launch_bounds__(SCAN_WG) void k_scan_chunk_offsets(int *__restrict__ part, int nchunk) {
    __shared__ int wt[SCAN_WG / 64];
    int *pa = part + (size_t)blockIdx.x * nchunk;
    int carry = 0;
    for (int base = 0; base < nchunk; base += SCAN_WG) {
        const int i = base + (int)threadIdx.x;
        const int v = i < nchunk ? pa[i] : 0;
        int total;
        const int ex = wg_exclusive(v, wt, total);
        if (i < nchunk) pa[i] = carry + ex;
        carry += total;
    }
}

// each chunk rescanned from its offset into y
__global__ __launch_bounds__(SCAN_WG) void k_scan_apply(const int *__restrict__ x, int64_t n, int nchunk,
                                                        const int *__restrict__ part, int *__restrict__ y) {
    __shared__ int wt[SCAN_WG / 64];
    const int k = blockIdx.y, c = blockIdx.x;
    const int *xa = x + (size_t)k * n;
    int *ya = y + (size_t)k * n;
    const int64_t i0 = (int64_t)c * SCAN_CHUNK + (int64_t)threadIdx.x * SCAN_ITEMS;
    int loc[SCAN_ITEMS];
    int v = 0;
    for (int j = 0; j < SCAN_ITEMS; j++) {
        loc[j] = i0 + j < n ? xa[i0 + j] : 0;
        v += loc[j];
    }
    int total;
    int run = part[(size_t)k * nchunk + c] + wg_exclusive(v, wt, total);
    for (int j = 0; j < SCAN_ITEMS; j++) {
        if (i0 + j < n) ya[i0 + j] = run;
        run += loc[j];
    }
}

// y = exclusive prefix sums of the K arrays of n entries in x; part holds
// K * scan_chunks(n) ints
static int scan_chunks(int64_t n) { return (int)std::max<int64_t>(1, (n + SCAN_CHUNK - 1) / SCAN_CHUNK); }

static void exclusive_scan(const int *x, int *y, int64_t n, int K, int *part, hipStream_t st) {
    const int nchunk = scan_chunks(n);
    hipLaunchKernelGGL(k_scan_totals, dim3((unsigned)nchunk, (unsigned)K), dim3(SCAN_WG), 0, st, x, n, nchunk, part);
    hipLaunchKernelGGL(k_scan_chunk_offsets, dim3((unsigned)K), dim3(SCAN_WG), 0, st, part, nchunk);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nchunk, (unsigned)K), dim3(SCAN_WG), 0, st, x, n, nchunk, part,
                       y);
}

struct ConnArgs {
    int D, nlevels, ns;
    int off[26][3];     // stencil offsets, in the reference's iteration order
    int64_t ext1[3];    // level-1 domain in blocks
    int logcap;         // hash table: 2^logcap slots
};

__device__ __forceinline__ uint64_t spread3d(uint64_t x) {
    x &= 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
__device__ __forceinline__ uint64_t spread2d(uint64_t x) {
    x &= 0xffffffffull;
    x = (x | x << 16) & 0x0000ffff0000ffffull;
    x = (x | x << 8) & 0x00ff00ff00ff00ffull;
    x = (x | x << 4) & 0x0f0f0f0f0f0f0f0full;
    x = (x | x << 2) & 0x3333333333333333ull;
    x = (x | x << 1) & 0x5555555555555555ull;
    return x;
}

// (level, 1-based coordinate) -> key; coordinates are checked on the host to
// fit (level < 32, coordinates below 2^19 in 3D and 2^26 in 2D)
__device__ __forceinline__ uint64_t ckey(int D, int l, const int64_t *c) {
    uint64_t m = D == 3 ? (spread3d(c[0] - 1) | spread3d(c[1] - 1) << 1 | spread3d(c[2] - 1) << 2)
                        : (spread2d(c[0] - 1) | spread2d(c[1] - 1) << 1);
    return (uint64_t)l << 58 | m;
}

__device__ __forceinline__ uint32_t slot_of(uint64_t key, int logcap) {
    return (uint32_t)((key * 0x9e3779b97f4a7c15ull) >> (64 - logcap));
}

__device__ __forceinline__ int level_of(const int64_t *pairs_off, int nlevels, int64_t i) {
    int l = 0;
    while (l + 1 < nlevels && i >= pairs_off[l + 1]) l++;
    return l;
}

// block id of (level, coordinate), or 0 (the table is never full)
__device__ __forceinline__ int64_t lookup(const uint64_t *keys, const int64_t *vals, int logcap, uint64_t key) {
    const uint32_t mask = (1u << logcap) - 1;
    for (uint32_t s = slot_of(key, logcap);; s = (s + 1) & mask) {
        const uint64_t k = keys[s];
        if (k == key) return vals[s];
        if (k == EMPTY) return 0;
    }
}

__global__ void k_conn_insert(uint64_t *keys, int64_t *vals, const int64_t *pairs, const int64_t *pairs_off,
                              int64_t n, ConnArgs a, int *err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int D = a.D;
    const int64_t *p = pairs + i * (D + 1);
    const int l = level_of(pairs_off, a.nlevels, i);
    const uint64_t key = ckey(D, l, p);
    const uint32_t mask = (1u << a.logcap) - 1;
    for (uint32_t s = slot_of(key, a.logcap);; s = (s + 1) & mask) {
        const unsigned long long prev = atomicCAS((unsigned long long *)&keys[s], (unsigned long long)EMPTY,
                                                  (unsigned long long)key);
        if (prev == EMPTY) {
            vals[s] = p[D];
            return;
        }
        if (prev == key) {
            atomicOr(err, 1);  // a position holds two blocks
            return;
        }
    }
}

// WRITE = false: count the records of block i; true: write them at the offsets
template <bool WRITE>
__global__ void k_conn_scan(const uint64_t *keys, const int64_t *vals, const int64_t *pairs,
                            const int64_t *pairs_off, int64_t n, ConnArgs a, int *cnt_b, int *cnt_n, int *cnt_r,
                            const int *off_b, const int *off_n, const int *off_r, int64_t *bnd, int64_t *nbr,
                            int64_t *ref, int64_t *chd, int *err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int D = a.D;
    const int64_t *p = pairs + i * (D + 1);
    const int64_t blk = p[D];
    const int l = level_of(pairs_off, a.nlevels, i);
    int64_t ext[3];
    for (int d = 0; d < D; d++) ext[d] = a.ext1[d] << l;
    int nb = 0, nn = 0, nr = 0;
    int64_t *wb = WRITE ? bnd + (int64_t)off_b[i] * (2 + 2 * D) : nullptr;
    int64_t *wn = WRITE ? nbr + (int64_t)off_n[i] * (2 + D) : nullptr;
    int64_t *wr = WRITE ? ref + (int64_t)off_r[i] * (2 + 2 * D) : nullptr;
    for (int k = 0; k < a.ns; k++) {
        int64_t q[3] = {1, 1, 1};
        bool inside = true;
        for (int d = 0; d < D; d++) {
            q[d] = p[d] + a.off[k][d];
            if (q[d] < 1 || q[d] > ext[d]) inside = false;
        }
        if (!inside) {
            if (WRITE) {
                int64_t *r = wb + (int64_t)nb * (2 + 2 * D);
                r[0] = blk;
                for (int d = 0; d < D; d++) {
                    r[1 + d] = a.off[k][d];
                    r[1 + D + d] = (q[d] >= 1 && q[d] <= ext[d]) ? 0 : a.off[k][d];
                }
                r[1 + 2 * D] = l + 1;
            }
            nb++;
            continue;
        }
        const int64_t other = lookup(keys, vals, a.logcap, ckey(D, l, q));
        if (other != 0) {
            if (WRITE) {
                int64_t *r = wn + (int64_t)nn * (2 + D);
                r[0] = blk;
                r[1] = other;
                for (int d = 0; d < D; d++) r[2 + d] = a.off[k][d];
            }
            nn++;
        } else if (l > 0) {
            int64_t pc[3] = {1, 1, 1};
            for (int d = 0; d < D; d++) pc[d] = (q[d] - 1) / 2 + 1;
            const int64_t pblk = lookup(keys, vals, a.logcap, ckey(D, l - 1, pc));
            if (pblk == 0) {
                atomicOr(err, 2);  // proper nesting broken
                continue;
            }
            if (WRITE) {
                int64_t *r = wr + (int64_t)nr * (2 + 2 * D);
                r[0] = blk;
                r[1] = pblk;
                for (int d = 0; d < D; d++) {
                    r[2 + d] = a.off[k][d];
                    r[2 + D + d] = (q[d] - 1) % 2 + 1;
                }
            }
            nr++;
        }
    }
    if (l > 0) {
        int64_t pc[3] = {1, 1, 1};
        for (int d = 0; d < D; d++) pc[d] = (p[d] - 1) / 2 + 1;
        const int64_t pblk = lookup(keys, vals, a.logcap, ckey(D, l - 1, pc));
        if (pblk == 0) atomicOr(err, 4);  // a block without a parent
        if (WRITE && pblk != 0) {
            // Child records: one per block above level 1, in pairs order
            int64_t *r = chd + (i - pairs_off[1]) * (2 + D);
            r[0] = blk;
            r[1] = pblk;
            for (int d = 0; d < D; d++) r[2 + d] = (p[d] - 1) % 2 + 1;
        }
    }
    if (!WRITE) {
        cnt_b[i] = nb;
        cnt_n[i] = nn;
        cnt_r[i] = nr;
    }
}

}  // namespace conn
}  // namespace srmg

using namespace srmg;
using namespace srmg::conn;

int srmg_connectivity(int D, int nlevels, const int64_t *rootsize, const int64_t *pairs, const int64_t *pairs_off,
                      int stencil, int64_t *neighbor, int64_t *neighbor_off, int64_t neighbor_cap, int64_t *boundary,
                      int64_t *boundary_off, int64_t boundary_cap, int64_t *refboundary, int64_t *refboundary_off,
                      int64_t refboundary_cap, int64_t *child, int64_t *child_off, int64_t child_cap) {
    return guarded([&] {
        if (D != 2 && D != 3) fail("D must be 2 or 3");
        if (nlevels < 1 || nlevels > 31) fail("nlevels must be 1..31");
        if (!rootsize || !pairs || !pairs_off || !neighbor || !neighbor_off || !boundary || !boundary_off ||
            !refboundary || !refboundary_off || !child || !child_off)
            fail("null argument");
        if (stencil != SRMG_STAR_STENCIL && stencil != SRMG_BOX_STENCIL) fail("bad stencil");
        if (pairs_off[0] != 0) fail("pairs_off[0] must be 0");
        for (int l = 0; l < nlevels; l++)
            if (pairs_off[l + 1] < pairs_off[l]) fail("pairs_off must not decrease");
        const int64_t n = pairs_off[nlevels];
        const int cbits = D == 3 ? 19 : 26;
        for (int d = 0; d < D; d++)
            if (rootsize[d] < 1 || ((rootsize[d] << (nlevels - 1)) >> cbits) != 0)
                fail("level domain too large for the position keys");
        for (int64_t i = 0; i < n; i++) {
            int l = 0;
            while (l + 1 < nlevels && i >= pairs_off[l + 1]) l++;
            for (int d = 0; d < D; d++)
                if (pairs[i * (D + 1) + d] < 1 || pairs[i * (D + 1) + d] > (rootsize[d] << l))
                    fail("block coordinate outside its level's domain");
        }
        ConnArgs a{};
        a.D = D;
        a.nlevels = nlevels;
        for (int d = 0; d < D; d++) a.ext1[d] = rootsize[d];
        // StarStencil (src/stencil.jl:16-33) or BoxStencil (-1..1 first index
        // fastest, the zero offset dropped; src/stencil.jl:36-56)
        if (stencil == SRMG_STAR_STENCIL) {
            for (int d = 0; d < D; d++)
                for (int v = -1; v <= 1; v += 2) {
                    a.off[a.ns][d] = v;
                    a.ns++;
                }
        } else {
            const int n3 = D == 3 ? 27 : 9;
            for (int k = 0; k < n3; k++) {
                int o[3] = {k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1};
                if (o[0] == 0 && o[1] == 0 && (D == 2 || o[2] == 0)) continue;
                for (int d = 0; d < D; d++) a.off[a.ns][d] = o[d];
                a.ns++;
            }
        }
        a.logcap = 4;
        while ((int64_t)1 << a.logcap < 2 * std::max<int64_t>(n, 1)) a.logcap++;
        const int64_t cap = (int64_t)1 << a.logcap;
        const int64_t nchild = n - pairs_off[1];
        if (nchild > child_cap) fail("child capacity too small: " + std::to_string(nchild) + " records");

        hipStream_t st;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{st};
        DevBuf dpairs, doff, dkeys, dvals, dcnt, dofs, derr, dtmp;
        dpairs.alloc((size_t)std::max<int64_t>(n, 1) * (D + 1) * 8);
        doff.alloc((size_t)(nlevels + 1) * 8);
        dkeys.alloc((size_t)cap * 8);
        dvals.alloc((size_t)cap * 8);
        dcnt.alloc((size_t)3 * (n + 1) * 4);
        dofs.alloc((size_t)3 * (n + 1) * 4);
        derr.alloc(4);
        HIPCHK(hipMemcpyAsync(dpairs.p, pairs, (size_t)n * (D + 1) * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(doff.p, pairs_off, (size_t)(nlevels + 1) * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(dkeys.p, 0xff, (size_t)cap * 8, st));
        HIPCHK(hipMemsetAsync(dcnt.p, 0, (size_t)3 * (n + 1) * 4, st));
        HIPCHK(hipMemsetAsync(derr.p, 0, 4, st));
        const unsigned grid = (unsigned)((std::max<int64_t>(n, 1) + 255) / 256);
        int *cb = dcnt.as<int>(), *cn = cb + (n + 1), *cr = cn + (n + 1);
        int *ob = dofs.as<int>(), *on = ob + (n + 1), *orr = on + (n + 1);
        if (n) {
            hipLaunchKernelGGL(k_conn_insert, dim3(grid), dim3(256), 0, st, dkeys.as<uint64_t>(),
                               dvals.as<int64_t>(), dpairs.as<int64_t>(), doff.as<int64_t>(), n, a, derr.as<int>());
            hipLaunchKernelGGL((k_conn_scan<false>), dim3(grid), dim3(256), 0, st, dkeys.as<uint64_t>(),
                               dvals.as<int64_t>(), dpairs.as<int64_t>(), doff.as<int64_t>(), n, a, cb, cn, cr,
                               nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, derr.as<int>());
        }
        // exclusive scans of the three count arrays over n + 1 entries: entry n is the total
        dtmp.alloc((size_t)3 * scan_chunks(n + 1) * 4);
        exclusive_scan(cb, ob, n + 1, 3, dtmp.as<int>(), st);
        // per-level offsets: the scans at the levels' first blocks
        std::vector<int> scans((size_t)3 * (n + 1));
        HIPCHK(hipMemcpyAsync(scans.data(), dofs.p, scans.size() * 4, hipMemcpyDeviceToHost, st));
        int err = 0;
        HIPCHK(hipMemcpyAsync(&err, derr.p, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        std::vector<int> lofs(3 * (nlevels + 1));
        for (int l = 0; l <= nlevels; l++)
            for (int k = 0; k < 3; k++) lofs[3 * l + k] = scans[(size_t)k * (n + 1) + pairs_off[l]];
        if (err & 1) fail("two blocks at the same position of a level");
        if (err & 2) fail("proper nesting broken: a neighbour position has no parent block");
        if (err & 4) fail("malformed tree: a block above level 1 has no parent");
        const int64_t tb = lofs[3 * nlevels], tn = lofs[3 * nlevels + 1], tr = lofs[3 * nlevels + 2];
        if (tb > boundary_cap) fail("boundary capacity too small: " + std::to_string(tb) + " records");
        if (tn > neighbor_cap) fail("neighbor capacity too small: " + std::to_string(tn) + " records");
        if (tr > refboundary_cap) fail("refboundary capacity too small: " + std::to_string(tr) + " records");
        DevBuf db, dn, dr, dc;
        db.alloc((size_t)std::max<int64_t>(tb, 1) * (2 + 2 * D) * 8);
        dn.alloc((size_t)std::max<int64_t>(tn, 1) * (2 + D) * 8);
        dr.alloc((size_t)std::max<int64_t>(tr, 1) * (2 + 2 * D) * 8);
        dc.alloc((size_t)std::max<int64_t>(nchild, 1) * (2 + D) * 8);
        if (n)
            hipLaunchKernelGGL((k_conn_scan<true>), dim3(grid), dim3(256), 0, st, dkeys.as<uint64_t>(),
                               dvals.as<int64_t>(), dpairs.as<int64_t>(), doff.as<int64_t>(), n, a, nullptr, nullptr,
                               nullptr, ob, on, orr, db.as<int64_t>(), dn.as<int64_t>(), dr.as<int64_t>(),
                               dc.as<int64_t>(), derr.as<int>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(boundary, db.p, (size_t)tb * (2 + 2 * D) * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(neighbor, dn.p, (size_t)tn * (2 + D) * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(refboundary, dr.p, (size_t)tr * (2 + 2 * D) * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(child, dc.p, (size_t)nchild * (2 + D) * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int l = 0; l <= nlevels; l++) {
            boundary_off[l] = lofs[3 * l];
            neighbor_off[l] = lofs[3 * l + 1];
            refboundary_off[l] = lofs[3 * l + 2];
            child_off[l] = l == 0 ? 0 : pairs_off[l] - pairs_off[1];
        }
        child_off[0] = 0;
    });
}
