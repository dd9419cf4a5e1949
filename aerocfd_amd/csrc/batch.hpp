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

// Block index and per-block workgroup coordinates of this workgroup.
__device__ inline int batch_wg(const int *__restrict__ map, const KGrid *__restrict__ kg, int &bx, int &by,
                               int &bz)
{
    const int flat = (int)blockIdx.x;
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
__device__ inline bool box_node_at(const Box &b, int bx, int by, int bz, int &i, int &j, int &k)
{
    i = b.i0 + (int)(bx * blockDim.x + threadIdx.x);
    j = b.j0 + (int)(by * blockDim.y + threadIdx.y);
    k = b.k0 + bz;
    return i <= b.i1 && j <= b.j1 && k <= b.k1;
}

}  // namespace acfd
