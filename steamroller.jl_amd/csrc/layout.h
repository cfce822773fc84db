// Device block layout shared by the kernels and the host-side conversions.
//
// A level is a contiguous run of blocks (Morton order of block coordinates);
// a block stores only its M^D valid cells, split by colour: the first half
// holds the cells updated by the first Gauss-Seidel half-sweep (global index
// sum odd, the RedBlack{0} cells of src/multigrid.jl:161-167), the second half
// the others. Within a half, cells follow the natural order with x halved:
// position (x >> 1) + (M/2) (y + M z). Face ghosts are not stored: same-level
// and boundary ghosts are read from the neighbour's (or the block's own) valid
// cells, and coarse-fine ghosts live in per-face buffers of M^(D-1) values.
#pragma once

#include <cstdint>

#ifndef __HIPCC__
#define __host__
#define __device__
#define __forceinline__ inline
#endif

namespace srmg {

template <int D, int M>
struct Lay {
    static constexpr int V = (D == 2) ? M * M : M * M * M;  // valid cells per block
    static constexpr int H = V / 2;                          // cells per colour
    static constexpr int F = (D == 2) ? M : M * M;           // cells per face
    static constexpr int MH = M / 2;
    static constexpr int NC = 1 << D;                        // children per block

    // storage position of local 0-based cell (x, y, z)
    __host__ __device__ static __forceinline__ int sidx(int x, int y, int z) {
        int odd = (x + y + z + D) & 1;  // == parity of the global 1-based index sum
        return (odd ? 0 : H) + (x >> 1) + MH * (y + M * z);
    }

    // inverse of sidx
    __host__ __device__ static __forceinline__ void decode(int q, int &x, int &y, int &z) {
        int h = q >= H ? 1 : 0;
        int r = q - h * H;
        int hx = r % MH;
        int row = r / MH;
        y = row % M;
        z = row / M;
        int odd = h ? 0 : 1;
        x = 2 * hx + ((odd ^ (y + z + D)) & 1);
    }

    // position of the cell with tangential coordinates inside a face slab
    // perpendicular to dimension dim (remaining dimensions in increasing order)
    __host__ __device__ static __forceinline__ int faceidx(int dim, int x, int y, int z) {
        return dim == 0 ? (y + M * z) : (dim == 1 ? (x + M * z) : (x + M * y));
    }
};

// neighbour-table entries: >= 0 same-level block; -1 domain boundary;
// <= -2 coarse-fine face whose ghost slab is buffer slot (-2 - entry)
constexpr int NB_BOUNDARY = -1;
__host__ __device__ __forceinline__ int nb_slot(int entry) { return -2 - entry; }

}  // namespace srmg
