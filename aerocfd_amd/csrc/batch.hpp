// One launch for all blocks of a context.
//
// Every block's fields live in one allocation (the arena); a device table
// holds per block its ranges and the offsets of its fields from the arena
// base.  A batched launch concatenates the blocks' own grids along x: a map
// gives each flat workgroup its block, and the block's grid entry turns the
// flat index back into the (x, y, z) workgroup coordinates the per-block
// kernel would have had.  The arena base is a kernel argument, so field
// pointers formed from it compile to global (not flat) memory operations.
#pragma once
#include "common.hpp"

namespace acfd {

// offsets (doubles) of a block's fields from the arena base; -1 = absent
struct FieldOff {
    long q, pri, rhsi, rhsv, rhs;
    long dxidx, inv_j, amu, dt;
    long dudx, tao, vor;
    long aal, bbe, gga, dist, spacing, length;
    long nut, sa_rhs, q0;
    long w_flux, w_dqs, w_dq, w_diag, w_r, w_sa;
};

struct BlockRec {
    Ext e;
    FieldOff o;
};

// a block's part of a batched launch: first flat workgroup and its grid
struct KGrid {
    int wg0, gx, gy, gz;
};

// host-side handle of one batched grid (device arrays)
struct BatchGrid {
    KGrid *kg = nullptr;  // per block
    int *map = nullptr;   // per flat workgroup: block index
    int total = 0;        // workgroups
};

// XCD-aware order of a 1-D launch: workgroup w goes to XCD w % 8, each XCD
// with its own L2; this hands XCD x the x-th eighth of the workgroup range,
// so each XCD walks a contiguous slab of the block and a node's j and k
// neighbours are mostly read through the same L2.  A bijection for any
// grid size.
__device__ inline long xcd_slab(long w, long total)
{
    const long q = total / 8, r = total % 8, x = w % 8;
    return x * q + (x < r ? x : r) + w / 8;
}

// Block index and per-block workgroup coordinates of this workgroup (flat
// workgroup numbers in XCD slabs, so that each XCD walks contiguous parts
// of the blocks).
__device__ inline int batch_wg(const int *__restrict__ map, const KGrid *__restrict__ kg, int &bx, int &by,
                               int &bz)
{
    const int flat = (int)xcd_slab(blockIdx.x, gridDim.x);
    const int m = map[flat];
    const KGrid g = kg[m];
    int local = flat - g.wg0;
    bx = local % g.gx;
    local /= g.gx;
    by = local % g.gy;
    bz = local / g.gy;
    return m;
}

// box_node with explicit workgroup coordinates
// Linear tiling of a block: workgroup w (256 threads) takes the nodes
// w*256 .. w*256+255 of the block's storage order (i fastest), so a
// wavefront reads whole 128-byte lines whatever the row length (a 64-wide
// tile of a row that does not start on a line boundary touches five lines
// for four lines of data).  Nodes outside the box b are skipped.
__device__ inline bool lin_node_at(const Ext &e, const Box &b, long w, int &i, int &j, int &k, long &n)
{
    n = w * 256 + (long)(threadIdx.x + blockDim.x * threadIdx.y);
    if (n >= e.n) return false;
    // blocks hold fewer than 2^31 nodes: 32-bit division
    const unsigned plane = (unsigned)e.ni * (unsigned)e.nj, un = (unsigned)n;
    const unsigned kk = un / plane, r = un - kk * plane, jj = r / (unsigned)e.ni;
    i = e.is + (int)(r - jj * (unsigned)e.ni);
    j = e.js + (int)jj;
    k = e.ks + (int)kk;
    return i >= b.i0 && i <= b.i1 && j >= b.j0 && j <= b.j1 && k >= b.k0 && k <= b.k1;
}

inline dim3 lin_grid(const Ext &e) { return dim3((unsigned)((e.n + 255) / 256)); }

__device__ inline bool box_node_at(const Box &b, int bx, int by, int bz, int &i, int &j, int &k)
{
    i = b.i0 + (int)(bx * blockDim.x + threadIdx.x);
    j = b.j0 + (int)(by * blockDim.y + threadIdx.y);
    k = b.k0 + bz;
    return i <= b.i1 && j <= b.j1 && k <= b.k1;
}

}  // namespace acfd
