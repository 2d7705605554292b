// gz_groups.h -- host side of row groups: the active rows split into groups that share one cell
// of one grid (the cell of the level of each row's radius in the search over many levels, the
// cell of the row's box in a batch), so that the cell-walk kernels (one wavefront per group,
// candidates staged through LDS) serve those searches as they serve the single grid.
//
// Two launches with a host read of the group count between them (the stencil table is sized
// by it):  group_rows() keys every row (rows outside the pass last), sorts the keys unless the
// rows already come in key order, and lists the groups; group_stencils() then writes each
// group's 27-cell stencil.
#pragma once
#include <hip/hip_runtime.h>
#include "gz_prim.h"

#include "gz_kernels.h"

namespace gz {

struct RowGroups {
    uint32_t *key = nullptr, *skey = nullptr, *gkey = nullptr;  // row keys, sorted keys, group keys
    int *val = nullptr, *row = nullptr;                          // row slots before and after the sort
    int *flag = nullptr, *scan = nullptr, *start = nullptr, *count = nullptr;
    int *lev = nullptr, *code = nullptr, *nst = nullptr;
    int4 *sten = nullptr;
    long rows_alloc = 0, groups_alloc = 0;  // capacities (rows + 1, groups)
    int ngroups = 0, nrows = 0;
    bool sorted = false;                    // row holds the permutation

    CellRows view(BoxTable bt) const
    {
        CellRows cr{};
        cr.ncells = ngroups;
        cr.nrows = nrows;
        cr.grow = sorted ? row : nullptr;
        cr.gstart = start;
        cr.glev = lev;
        cr.gcode = code;
        cr.gnst = nst;
        cr.gsten = sten;
        cr.bt = bt;
        return cr;
    }
};

// temporary storage the sort and the scan of nrows rows need
inline hipError_t group_tmp_bytes(int nrows, size_t &bytes)
{
    size_t a = 0, b = 0;
    hipError_t e = gz_sort_pairs(nullptr, a, (const uint32_t *)nullptr, (uint32_t *)nullptr, (const int *)nullptr,
                                                      (int *)nullptr, nrows, 0, 32);
    if (e != hipSuccess) return e;
    e = gz_exclusive_sum(nullptr, b, (const int *)nullptr, (int *)nullptr, nrows + 1);
    bytes = a > b ? a : b;
    return e;
}

// keys (cells of the rows' own levels, radii from hslot[r] when given, else from H; box cells
// when bt.b), the sort when sort is set (else the keys must not decrease along the rows), group
// heads, their scan and the group list; the count is left in G.count on the device
inline hipError_t group_rows(RowGroups &G, hipStream_t s, void *tmp, size_t tmp_bytes, int nrows, const int *act, const uint8_t *need,
                             const uint32_t *key, const double *hslot, const double *H, double box, double margin,
                             const WideGrid *W, BoxTable bt, bool sort)
{
    constexpr int B = 256;
    const unsigned nbk = (unsigned)((nrows + B) / B);
    G.nrows = nrows;
    G.sorted = sort;
    k_row_group_key<<<nbk, B, 0, s>>>(nrows, act, need, key, hslot, H, box, margin, W, bt, G.key, sort ? G.val : nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t *sk = G.key;
    if (sort) {
        size_t bytes = tmp_bytes;
        e = gz_sort_pairs(tmp, bytes, G.key, G.skey, G.val, G.row, nrows, 0, 32, s);
        if (e != hipSuccess) return e;
        sk = G.skey;
    }
    k_group_heads<<<nbk, B, 0, s>>>(nrows, sk, G.flag);
    size_t bytes = tmp_bytes;
    e = gz_exclusive_sum(tmp, bytes, G.flag, G.scan, nrows + 1, s);
    if (e != hipSuccess) return e;
    k_group_list<<<nbk, B, 0, s>>>(nrows, sk, G.flag, G.scan, G.start, G.gkey, G.count);
    return hipGetLastError();
}

// the stencils of the G.ngroups groups (the count read back by the caller)
inline hipError_t group_stencils(RowGroups &G, hipStream_t s, GridView g, const int *act, BoxTable bt)
{
    if (G.ngroups == 0) return hipSuccess;
    k_group_stencil<<<(unsigned)G.ngroups, 64, 0, s>>>(g, G.ngroups, G.gkey, G.start, G.sorted ? G.row : nullptr, act, bt, G.lev,
                                                       G.code, G.nst, G.sten);
    return hipGetLastError();
}

}  // namespace gz
