// gz_kernels_impl.h -- gfx950 kernels of the MFM hydro pipeline (templates; instantiated per
// arithmetic form in gz_kernels.hip (FP64, strict) and gz_kernels_f32.hip (FP32-mixed)).
//
// Layout: particles live on the device in structure-of-arrays form, ordered by a 30-bit
// Morton key of a fixed 1024^3 grid over the periodic box.  Any power-of-two search grid
// (cell side >= the search radius) is then a contiguous range of that order per cell, and
// the 27 neighbour-cell ranges of a cell, sorted by start, visit candidates in ascending
// device index: the canonical summation order, shared with the CPU oracle.
//
// Neighbour lists are ELLPACK: entry q of row r sits at ell[ell_at(ld, r, q)], column-major
// (the 64 rows of a wave read entry q from four 64-byte lines).  One lane per active particle;
// wave64.
#pragma once
#include <type_traits>
#include "gz_kernels.h"

namespace gz {

// ---------------------------------------------------------------------------------------
// keys, gathers, scatters
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t compact10(uint32_t v)
{
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030c30c3u;
    v = (v | (v >> 4)) & 0x0300f00fu;
    v = (v | (v >> 8)) & 0x030000ffu;
    v = (v | (v >> 16)) & 0x000003ffu;
    return v;
}

__device__ __forceinline__ int fine_coord(double x, double box)
{
    double w = x - box * floor(x / box);
    int c = (int)(w * (double)kFine / box);
    return c < 0 ? 0 : (c >= kFine ? kFine - 1 : c);
}

#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_morton(long n, const double *__restrict__ pos, double box, uint32_t *__restrict__ key, int *__restrict__ val)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t cx = fine_coord(pos[3 * i + 0], box), cy = fine_coord(pos[3 * i + 1], box), cz = fine_coord(pos[3 * i + 2], box);
    key[i] = (spread10(cx) << 2) | (spread10(cy) << 1) | spread10(cz);
    val[i] = (int)i;
}
#endif


#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_morton_soa(long n, const double *__restrict__ X, const double *__restrict__ Y, const double *__restrict__ Z,
                             double box, uint32_t *__restrict__ key, int *__restrict__ val)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t cx = fine_coord(X[i], box), cy = fine_coord(Y[i], box), cz = fine_coord(Z[i], box);
    key[i] = (spread10(cx) << 2) | (spread10(cy) << 1) | spread10(cz);
    val[i] = (int)i;
}

// reorder width columns of 8-byte values (SoA, column stride n): out[c*n + s] = in[c*n + order[s]]
__global__ void k_permute_cols(long n, int width, const int *__restrict__ order, const unsigned long long *__restrict__ in,
                               unsigned long long *__restrict__ out)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    const long u = order[s];
    for (int c = 0; c < width; c++) out[(size_t)c * n + s] = in[(size_t)c * n + u];
}

__global__ void k_permute_int(long n, const int *__restrict__ order, const int *__restrict__ in, int *__restrict__ out)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    out[s] = in[order[s]];
}

__global__ void k_permute_u8(long n, const int *__restrict__ order, const uint8_t *__restrict__ in, uint8_t *__restrict__ out)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    out[s] = in[order[s]];
}

// up to kPermCols 8-byte columns gathered by one order in one launch (every load of a lane
// issued before its stores: the columns are distinct buffers)
__global__ void __launch_bounds__(256) k_permute_many(long n, const int *__restrict__ order, PermCols P)
{
    const long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    const long u = order[s];
    unsigned long long v[kPermCols];
#pragma unroll
    for (int c = 0; c < kPermCols; c++)
        if (c < P.ncol) v[c] = P.in[c][u];
#pragma unroll
    for (int c = 0; c < kPermCols; c++)
        if (c < P.ncol) P.out[c][s] = v[c];
}

__global__ void k_gather_ll(long n, const int *__restrict__ perm, const long long *__restrict__ in, long long *__restrict__ out)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    out[s] = in[perm[s]];
}

__global__ void k_scatter_ll(long n, const int *__restrict__ perm, const long long *__restrict__ in, long long *__restrict__ out)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    out[perm[s]] = in[s];
}
#endif


// sorted field s <- user field perm[s] (width doubles per particle; interleaved in, SoA out
// when soa != 0 with column stride n)
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_gather(long n, int width, const int *__restrict__ perm, const double *__restrict__ in,
                         double *__restrict__ out, int soa)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    long u = perm[s];
    for (int c = 0; c < width; c++) {
        double v = in[(size_t)u * width + c];
        if (soa) out[(size_t)c * n + s] = v; else out[(size_t)s * width + c] = v;
    }
}
#endif


#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_gather_int(long n, const int *__restrict__ perm, const int *__restrict__ in, int *__restrict__ out)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    out[s] = in[perm[s]];
}
#endif


#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_gather_u8(long n, const int *__restrict__ perm, const uint8_t *__restrict__ in, uint8_t *__restrict__ out)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    out[s] = in[perm[s]];
}
#endif


// user field perm[s] <- sorted field s
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_scatter(long n, int width, const int *__restrict__ perm, const double *__restrict__ in,
                          double *__restrict__ out, int soa)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    long u = perm[s];
    for (int c = 0; c < width; c++) {
        double v = soa ? in[(size_t)c * n + s] : in[(size_t)s * width + c];
        out[(size_t)u * width + c] = v;
    }
}
#endif


#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_scatter_int(long n, const int *__restrict__ perm, const int *__restrict__ in, int *__restrict__ out)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    out[perm[s]] = in[s];
}
#endif


// active rows: act[r] = sorted index of the r-th active particle (ascending)
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_flag_active(long n, const uint8_t *__restrict__ active, const double *__restrict__ mass, int *__restrict__ flag)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    flag[s] = (active == nullptr || active[s]) ? 1 : 0;
    (void)mass;
}
#endif


#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_compact(long n, const int *__restrict__ flag, const int *__restrict__ scan, int *__restrict__ act)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    if (flag[s]) act[scan[s]] = (int)s;
}
#endif


// ---------------------------------------------------------------------------------------
// cell grid at level L (nc = 2^L): cells indexed by Morton code, [start, end) of the order
// ---------------------------------------------------------------------------------------
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_cell_clear(long ncell, int *__restrict__ cstart, int *__restrict__ cend, double *__restrict__ chmax)
{
    long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    cstart[c] = 0;
    cend[c] = 0;
    chmax[c] = 0;
}
#endif


#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_cell_bounds(long n, const uint32_t *__restrict__ key, int shift, int *__restrict__ cstart, int *__restrict__ cend)
{
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t c = key[s] >> shift;
    if (s == 0 || (key[s - 1] >> shift) != c) cstart[c] = (int)s;
    if (s == n - 1 || (key[s + 1] >> shift) != c) cend[c] = (int)(s + 1);
}
#endif


// largest kernel radius per cell (the symmetric search reaches max(h_i, h_j))
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_cell_hmax(long ncell, const int *__restrict__ cstart, const int *__restrict__ cend,
                            const double *__restrict__ h, double *__restrict__ chmax)
{
    long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    double m = 0;
    for (int s = cstart[c]; s < cend[c]; s++) m = fmax(m, h[s]);
    chmax[c] = m;
}
#endif


// the (deduplicated) 27 neighbour cells of each cell, non-empty ones only, sorted by start;
// each entry packs start, end and the offset code (dx+1)*9+(dy+1)*3+(dz+1)
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_cell_stencil(int level, const int *__restrict__ cstart, const int *__restrict__ cend,
                               int4 *__restrict__ sten, int *__restrict__ nsten)
{
    long ncell = 1L << (3 * level);
    long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int nc = 1 << level;
    int cx = compact10((uint32_t)c >> 2), cy = compact10((uint32_t)c >> 1), cz = compact10((uint32_t)c);
    int4 e[27];
    int m = 0;
    for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) {
                int a = (cx + dx + nc) % nc, b = (cy + dy + nc) % nc, d = (cz + dz + nc) % nc;
                uint32_t code = (spread10(a) << 2) | (spread10(b) << 1) | spread10(d);
                int st = cstart[code], en = cend[code];
                if (en <= st) continue;
                bool dup = false;
                for (int q = 0; q < m; q++) if (e[q].x == st) { dup = true; break; }
                if (dup) continue;
                e[m++] = make_int4(st, en, (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1), (int)code);
            }
    for (int a = 1; a < m; a++) {  // insertion sort by start
        int4 t = e[a];
        int b = a - 1;
        while (b >= 0 && e[b].x > t.x) { e[b + 1] = e[b]; b--; }
        e[b + 1] = t;
    }
    for (int q = 0; q < m; q++) sten[c * 27 + q] = e[q];
    nsten[c] = m;
}
#endif


// ---------------------------------------------------------------------------------------
// neighbour lists (gpu_ngb_list_build, src/mesh/gpu_neighbor_list.cc:1950, exact FP64 accept)
//   SYMMETRIC = false: rows r2 < (scale*h_i)^2, self included (density, scale = 1.3 buffer)
//   SYMMETRIC = true : rows r2 < max(h_i,h_j)^2, self excluded (gradients, force)
// rows: the slots listed in rows[0..nrows) (row r of the list is slot rows[k]); act[slot] =
// sorted index of that slot's particle.
// ---------------------------------------------------------------------------------------
template <bool SYMMETRIC>
__global__ void __launch_bounds__(256) k_build_ell(int nrows, const int *__restrict__ rows, const int *__restrict__ act,
                                                   const double *__restrict__ hrow, double scale, const double *__restrict__ X,
                                                   const double *__restrict__ Y, const double *__restrict__ Z, const double *__restrict__ H,
                                                   const uint32_t *__restrict__ key, GridView g, int *__restrict__ ell, long ld, int cap,
                                                   int *__restrict__ cnt, int *__restrict__ overflow)
{
    const long k = xcd_block() * blockDim.x + threadIdx.x;
    if (k >= nrows) return;
    const int r = rows ? rows[k] : (int)k;
    const int i = act[r];
    const double xi = X[i], yi = Y[i], zi = Z[i];
    const double hi = hrow ? hrow[r] * scale : H[i];
    const uint32_t home = key[i] >> g.shift;
    const int nc = 1 << g.level;
    const int cx = compact10(home >> 2), cy = compact10(home >> 1), cz = compact10(home);
    const bool prune = nc >= 3;
    const int m = g.nsten[home];
    int c = 0;
    for (int q = 0; q < m; q++) {
        const int4 e = g.sten[(size_t)home * 27 + q];
        if (prune) {
            double reach = hi;
            if (SYMMETRIC) reach = fmax(hi, g.chmax[e.w]);
            reach = reach * (1.0 + 1e-9) + 1e-6 * g.cell;
            const int dx = e.z / 9 - 1, dy = (e.z / 3) % 3 - 1, dz = e.z % 3 - 1;
            const double lx = (cx + dx) * g.cell, ly = (cy + dy) * g.cell, lz = (cz + dz) * g.cell;
            const double xw = xi - g.box * floor(xi / g.box), yw = yi - g.box * floor(yi / g.box), zw = zi - g.box * floor(zi / g.box);
            double ex = fmax(fmax(lx - xw, xw - (lx + g.cell)), 0.0);
            double ey = fmax(fmax(ly - yw, yw - (ly + g.cell)), 0.0);
            double ez = fmax(fmax(lz - zw, zw - (lz + g.cell)), 0.0);
            if (ex * ex + ey * ey + ez * ez > reach * reach) continue;
        }
        for (int j = e.x; j < e.y; j++) {
            if (SYMMETRIC && j == i) continue;
            if (accept_pair<SYMMETRIC>(xi, yi, zi, X[j], Y[j], Z[j], hi, H[j], g.box, g.half)) {
                if (c < cap) ell[ell_at(ld, r, c)] = j;
                c++;
            }
        }
    }
    cnt[r] = min(c, cap);  // readers stay inside the list; the overflow carries the full length
    if (c > cap) atomicMax(overflow, c);
}

// ---------------------------------------------------------------------------------------
// density: fused pair sums (DensitySpec::pair_kernel, src/hydro/density_loop.h:1141-1385),
// after_iter_body (:516-961) and, on convergence, density_finalize_post_runner
// (src/hydro/density_loop.cc:303-733)
// ---------------------------------------------------------------------------------------
__device__ inline void density_finalize(const DevParams &P, const DensityOut &o, int i, double h, const double *acc, const DensityNorm &dn,
                                        double mass, double u)
{
    o.h[i] = h;
    const double nn = dn.nn, normalizer = dn.normalizer;
    o.drk[i] = dn.dr;
    o.divv[i] = dn.dv;
    for (int k = 0; k < 3; k++) o.fw[(size_t)k * o.n + i] = acc[10 + k];
    o.t[0 * o.n + i] = dn.Ti[0] / normalizer;
    o.t[1 * o.n + i] = dn.Ti[4] / normalizer;
    o.t[2 * o.n + i] = dn.Ti[8] / normalizer;
    o.t[3 * o.n + i] = dn.Ti[1] / normalizer;
    o.t[4 * o.n + i] = dn.Ti[5] / normalizer;
    o.t[5 * o.n + i] = dn.Ti[2] / normalizer;
    o.cn[i] = dn.cn;
    o.fce[i] = dn.fce;
    double rho = acc[1];
    if (mass > 0) {
        if ((h > 0) && (nn > 0)) rho = mass * nn / (kVolNorm * pow(h, 3.0));
        else if (h <= 0) rho = 0;
        else rho = mass / (kVolNorm * pow(h, 3.0));
        if (P.fd4alpha > 0 && acc[16] != 0) {
            // the first-derivative partition-of-unity volume correction (src/hydro/density_loop.cc:622-633):
            // grad H = -GradH_numer / GradH_denom, the density divided by 1 + 4 alpha |grad H|^2
            const double inv_denom = -1.0 / acc[16];
            const double g0 = acc[13] * inv_denom, g1 = acc[14] * inv_denom, g2 = acc[15] * inv_denom;
            rho /= 1.0 + P.fd4alpha * (g0 * g0 + g1 * g1 + g2 * g2);
        }
        o.p[i] = (P.gamma - 1) * u * rho;
    }
    o.rho[i] = rho;
    o.nn[i] = (nn > 0) ? pow(nn, 1.0 / 3.0) : 0.0;
}

// ---------------------------------------------------------------------------------------
// per-cell active rows: the active slots of one grid cell are a contiguous range of act[]
// (act ascends in device order, cells are contiguous in it); one entry per occupied cell
// ---------------------------------------------------------------------------------------
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_cell_rows_mark(int nact, const int *__restrict__ act, const uint32_t *__restrict__ key, int shift,
                                 int *__restrict__ ca0, int *__restrict__ ca1, int *__restrict__ flag)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nact) return;
    const uint32_t c = key[act[r]] >> shift;
    const bool first = (r == 0) || ((key[act[r - 1]] >> shift) != c);
    const bool last = (r == nact - 1) || ((key[act[r + 1]] >> shift) != c);
    if (first) ca0[c] = r;
    if (last) ca1[c] = r + 1;
    flag[r] = first ? 1 : 0;
}
#endif


#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void k_cell_rows_list(int nact, const int *__restrict__ act, const uint32_t *__restrict__ key, int shift,
                                 const int *__restrict__ flag, const int *__restrict__ scan, int *__restrict__ cells)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nact) return;
    if (flag[r]) cells[scan[r]] = (int)(key[act[r]] >> shift);
}
#endif


__device__ __forceinline__ bool wave_any(bool v) { return __any(v ? 1 : 0) != 0; }

// ---------------------------------------------------------------------------------------
// leaf culling for the cell walks.  A wavefront's particles (one leaf) have a bounding box in
// coordinates relative to the home cell's origin; a candidate farther from that box than the
// reach of the pair test cannot be accepted by any lane and is dropped while its tile is
// staged, keeping the survivors in ascending order (stable compaction over the ballot mask).
// The test is conservative: the slack covers rounding of the relative coordinates.
// ---------------------------------------------------------------------------------------
struct LeafBox {
    double ox, oy, oz;          // home cell origin
    double lo[3], hi[3];        // leaf box, relative to the origin
    bool cull;                  // false on grids too coarse for relative coordinates
};

__device__ __forceinline__ double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ LeafBox leaf_box(const GridView &g, int c, bool on, double xi, double yi, double zi)
{
    LeafBox b;
    const int cx = compact10((uint32_t)c >> 2), cy = compact10((uint32_t)c >> 1), cz = compact10((uint32_t)c);
    b.ox = cx * g.cell; b.oy = cy * g.cell; b.oz = cz * g.cell;
    const double rx = wrap(xi - b.ox, g.box, g.half), ry = wrap(yi - b.oy, g.box, g.half), rz = wrap(zi - b.oz, g.box, g.half);
    b.lo[0] = wave_min(on ? rx : 1e300); b.hi[0] = wave_max(on ? rx : -1e300);
    b.lo[1] = wave_min(on ? ry : 1e300); b.hi[1] = wave_max(on ? ry : -1e300);
    b.lo[2] = wave_min(on ? rz : 1e300); b.hi[2] = wave_max(on ? rz : -1e300);
    b.cull = (1 << g.level) >= 4;
    return b;
}

__device__ __forceinline__ bool leaf_reach(const GridView &g, const LeafBox &b, double xj, double yj, double zj, double reach)
{
    if (!b.cull) return true;
    const double rx = wrap(xj - b.ox, g.box, g.half), ry = wrap(yj - b.oy, g.box, g.half), rz = wrap(zj - b.oz, g.box, g.half);
    const double ex = fmax(fmax(b.lo[0] - rx, rx - b.hi[0]), 0.0);
    const double ey = fmax(fmax(b.lo[1] - ry, ry - b.hi[1]), 0.0);
    const double ez = fmax(fmax(b.lo[2] - rz, rz - b.hi[2]), 0.0);
    const double rr = reach * (1.0 + 1e-9) + 1e-12 * g.box;
    return ex * ex + ey * ey + ez * ez <= rr * rr;
}

// position of this lane among the kept lanes below it (stable compaction)
__device__ __forceinline__ int kept_rank(bool keep, int &count)
{
    const unsigned long long m = __ballot(keep ? 1 : 0);
    count = __popcll(m);
    const unsigned lane = threadIdx.x & 63;
    return __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
}

// ---------------------------------------------------------------------------------------
// density as a cell walk: one wavefront per occupied cell, its active particles one per lane;
// the candidates of the cell's 27 neighbour ranges are staged through LDS 64 at a time (after
// leaf culling) and every lane tests them in ascending order (the oracle's summation order).
// need[r] marks the rows still iterating.  The 13 pair sums of DensitySpec::pair_kernel
// (src/hydro/density_loop.h:1141-1385) go to acc[13*r]; k_density_after closes the round.
// ---------------------------------------------------------------------------------------
template <int KERN, typename R>
__global__ void __launch_bounds__(64) k_density_cells(DevParams P, GridView g, CellRows cr, const int *__restrict__ act,
                                                      const uint8_t *__restrict__ need, const double *__restrict__ X,
                                                      const double *__restrict__ Y, const double *__restrict__ Z,
                                                      const double *__restrict__ VX, const double *__restrict__ VY,
                                                      const double *__restrict__ VZ, const double *__restrict__ M,
                                                      const double *__restrict__ hcur, double *__restrict__ accout)
{
    __shared__ double sx[64], sy[64], sz[64], svx[64], svy[64], svz[64], sm[64];
    const long cb = xcd_block();
    if (cb >= (cr.ncells_dev ? *cr.ncells_dev : cr.ncells)) return;
    int c, a0, a1, nst;
    const int4 *sten;
    if (cr.gsten) {
        c = cr.gcode[cb]; a0 = cr.gstart[cb]; a1 = cr.gstart[cb + 1]; nst = cr.gnst[cb]; sten = cr.gsten + (size_t)27 * cb;
        if (cr.bt.b) {
            resolve_box(cr.bt, act[cr.grow ? cr.grow[a0] : a0], P, g);
        } else {
            g.level = cr.glev[cb];
            g.cell = g.box / (double)(1 << g.level);
        }
    } else {
        c = cr.cells[cb]; a0 = cr.ca0[c]; a1 = cr.ca1[c]; nst = g.nsten[c]; sten = g.sten + (size_t)c * 27;
    }
    const int t = threadIdx.x;
    for (int base = a0; base < a1; base += 64) {
        const bool in = base + t < a1;
        const int r = in ? (cr.grow ? cr.grow[base + t] : base + t) : 0;
        const bool on = in && need[r];
        if (!wave_any(on)) continue;
        int i = 0;
        double xi = 0, yi = 0, zi = 0, h = 0;
        R vxi = 0, vyi = 0, vzi = 0;
        if (on) {
            i = act[r];
            xi = X[i]; yi = Y[i]; zi = Z[i];
            vxi = (R)VX[i]; vyi = (R)VY[i]; vzi = (R)VZ[i];
            h = hcur[r];
        }
        const LeafBox lb = leaf_box(g, c, on, xi, yi, zi);
        const double reach = wave_max(on ? h : 0.0);
        // FP32 form on grids of at least 8 cells a side: positions relative to the home cell's
        // origin (one FP64 difference and wrap per row and per staged candidate, rounded to
        // FP32; every partner then lies within half a box of the origin), so the pair
        // differences are FP32 subtractions
        const bool rel = std::is_same<R, float>::value && (1 << g.level) >= 8;
        const R rxi = (R)wrap(xi - lb.ox, g.box, g.half), ryi = (R)wrap(yi - lb.oy, g.box, g.half),
                rzi = (R)wrap(zi - lb.oz, g.box, g.half);
        const bool fd = P.fd4alpha > 0;
        R a[kDensSumsFD];
        for (int q = 0; q < kDensSumsFD; q++) a[q] = 0;
        const R hR = (R)(on ? h : 1.0);
        const R h2 = hR * hR;
        R hinv, hinv3, hinv4;
        kernel_hinv(hR, hinv, hinv3, hinv4);
        for (int q = 0; q < nst; q++) {
            const int4 e = sten[q];
            for (int ts = e.x; ts < e.y; ts += 64) {
                const int j = ts + t;
                bool keep = false;
                double xj = 0, yj = 0, zj = 0;
                if (j < e.y) {
                    xj = X[j]; yj = Y[j]; zj = Z[j];
                    keep = leaf_reach(g, lb, xj, yj, zj, reach);
                }
                int nt;
                const int pos = kept_rank(keep, nt);
                __syncthreads();
                if (keep) {
                    if (rel) {
                        sx[pos] = (R)wrap(xj - lb.ox, g.box, g.half);
                        sy[pos] = (R)wrap(yj - lb.oy, g.box, g.half);
                        sz[pos] = (R)wrap(zj - lb.oz, g.box, g.half);
                    } else {
                        sx[pos] = xj; sy[pos] = yj; sz[pos] = zj;
                    }
                    svx[pos] = VX[j]; svy[pos] = VY[j]; svz[pos] = VZ[j]; sm[pos] = M[j];
                }
                __syncthreads();
                if (!on) continue;
                for (int k = 0; k < nt; k++) {
                    const double mj = sm[k];
                    if (mj <= 0) continue;
                    R dp0, dp1, dp2;
                    if (rel) {
                        dp0 = rxi - (R)sx[k]; dp1 = ryi - (R)sy[k]; dp2 = rzi - (R)sz[k];
                    } else {
                        double d0 = xi - sx[k], d1 = yi - sy[k], d2 = zi - sz[k];
                        d1 = wrap(d1, P.box, P.half);
                        d0 = wrap(d0, P.box, P.half);
                        d2 = wrap(d2, P.box, P.half);
                        dp0 = (R)d0; dp1 = (R)d1; dp2 = (R)d2;
                    }
                    const R r2 = dp0 * dp0 + dp1 * dp1 + dp2 * dp2;
                    if (r2 >= h2) continue;
                    const R rr = sqrt(r2);
                    const R u = rr * hinv;
                    R wk = 0, dwk = 0;
                    kernel_main<KERN, 0>(u, hinv3, hinv4, wk, dwk);
                    const R mj_wk = (R)mj * wk;
                    a[0] += rounded(wk);
                    a[1] += rounded(mj_wk);
                    a[2] += rounded(-(R(3.0) * hinv * wk + u * dwk));
                    if (rr > 0) {
                        a[4] += rounded(wk * (dp0 * dp0));
                        a[5] += rounded(wk * (dp1 * dp1));
                        a[6] += rounded(wk * (dp2 * dp2));
                        a[7] += rounded(wk * (dp0 * dp1));
                        a[8] += rounded(wk * (dp1 * dp2));
                        a[9] += rounded(wk * (dp0 * dp2));
                        a[10] += rounded(wk * dp0);
                        a[11] += rounded(wk * dp1);
                        a[12] += rounded(wk * dp2);
                        const R dv0 = vxi - (R)svx[k], dv1 = vyi - (R)svy[k], dv2 = vzi - (R)svz[k];
                        a[3] -= rounded(dwk * (dp0 * dv0 + dp1 * dv1 + dp2 * dv2) / rr);
                        if (fd) {  // GradH_numer, GradH_denom (src/hydro/density_loop.h:1281-1286)
                            a[13] += rounded((dwk / rr) * dp0);
                            a[14] += rounded((dwk / rr) * dp1);
                            a[15] += rounded((dwk / rr) * dp2);
                            a[16] += rounded(u * dwk);
                        }
                    }
                }
            }
        }
        if (on) {
            // static bounds: the sums stay in registers
            for (int q = 0; q < kDensSums; q++) accout[(size_t)q * cr.nrows + r] = (double)a[q];
            if (fd)
                for (int q = kDensSums; q < kDensSumsFD; q++) accout[(size_t)q * cr.nrows + r] = (double)a[q];
        }
    }
}

// the maximum of v over the block into *dst (one conditional atomic per block)
__device__ __forceinline__ void block_max_to(double v, unsigned long long *dst)
{
    __shared__ double wv[16];
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) wv[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)((blockDim.x + 63) >> 6); k++) v = fmax(v, wv[k]);
        if (v > 0) atomic_max_bits(dst, v);
    }
}

// density()'s pre-iteration radius clamp (src/hydro/density_loop.cc:818-838) and the start of
// the iteration: radius of each row, empty bracket, round counters, the largest radius (the
// bits of a positive double order like the double)
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void __launch_bounds__(256) k_density_init(int na, const int *__restrict__ act, double *__restrict__ H, double hclamp,
                                                      double *__restrict__ hcur, double *__restrict__ left, double *__restrict__ right,
                                                      double *__restrict__ cn_cur, int *__restrict__ cn_init, int *__restrict__ iter,
                                                      uint8_t *__restrict__ need, int *__restrict__ counters,
                                                      unsigned long long *__restrict__ hmax, unsigned long long *__restrict__ hfinal,
                                                      BoxTable bt)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < 8) counters[r] = 0;
    if (r == 0) *hfinal = 0;
    double h = 0;
    if (r < na) {
        const int i = act[r];
        if (bt.b) hclamp = 0.99 * bt.b[bt.BOX[i]].max_h;
        h = H[i];
        if ((h < 0) || !isfinite(h) || (h > hclamp)) { h = hclamp; H[i] = h; }
        hcur[r] = h;
        left[r] = 0; right[r] = 0; cn_cur[r] = 0; cn_init[r] = 0; iter[r] = 0;
        need[r] = 1;
    }
    if (!hmax) return;
    block_max_to(h, hmax);
}
#endif

// the maximum of v over the wave into slot (one atomic when the wave holds one box; the bits of
// a positive double order like the double)
__device__ __forceinline__ void wave_max_to(double v, bool valid, int box, unsigned long long *single, const BoxTable &bt, int slot)
{
    if (!bt.b) {
        for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
        if ((threadIdx.x & 63) == 0 && v > 0) atomic_max_bits(single, v);
        return;
    }
    const unsigned long long vm = __ballot(valid);
    if (!vm) return;
    const int b0 = __shfl(box, __ffsll((long long)vm) - 1, 64);
    if (__all(!valid || box == b0)) {
        double w = valid ? v : 0.0;
        for (int o = 32; o > 0; o >>= 1) w = fmax(w, __shfl_xor(w, o, 64));
        if ((threadIdx.x & 63) == 0 && w > 0) atomic_max_bits(&bt.slot[(size_t)b0 * kBoxSlots + slot], w);
    } else if (valid && v > 0) {
        atomic_max_bits(&bt.slot[(size_t)box * kBoxSlots + slot], v);
    }
}

// the sum of the list lengths (the pairs a force pass evaluates): one atomic per block
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void __launch_bounds__(256) k_sum_counts(int nrows, const int *__restrict__ cnt, unsigned long long *__restrict__ out,
                                                    const int *__restrict__ gate)
{
    if (gate && *gate) return;
    __shared__ unsigned long long part[4];
    unsigned long long s = 0;
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < nrows; r += (long)gridDim.x * blockDim.x) s += (unsigned long long)cnt[r];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        s = part[0] + part[1] + part[2] + part[3];
        if (s) atomicAdd(out, s);
    }
}

#endif

// after_iter_body (src/hydro/density_loop.h:516-961) on the sums of the round and, on
// convergence, density_finalize_post_runner (src/hydro/density_loop.cc:303-733)
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void __launch_bounds__(256) k_density_after(DevParams P, int nrows, const int *__restrict__ act, uint8_t *__restrict__ need,
                                                       const double *__restrict__ accin, const double *__restrict__ M,
                                                       const double *__restrict__ U, double cell, DensityIter it, DensityOut out,
                                                       int *__restrict__ nnext, int *__restrict__ rebuild, int *__restrict__ err,
                                                       unsigned long long *__restrict__ hfinal, BoxTable bt, double hlim,
                                                       int *__restrict__ gate, unsigned int *__restrict__ done,
                                                       const int *__restrict__ rows, int nlist, int *__restrict__ next_rows)
{
    // every lane reaches the wave-wide counts and maxima at the end.  Lane t takes row t, or the
    // t-th row of the list of rows of this round (rows, nlist); next_rows (may be null) receives
    // the rows that iterate again, in any order (each row's sums do not depend on the others)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = rows ? (t < nlist ? rows[t] : nrows) : t;
    const bool row = r < nrows && need[r];
    double h_final = 0, h_grow = 0;
    bool cont = false;
    int box = 0;
    if (row) {
        const int i = act[r];
        if (bt.b) {
            GridView g{};
            resolve_box(bt, i, P, g);
            box = bt.BOX[i];
            cell = bt.b[box].rowlim;
        }
        double acc[kDensSumsFD];
        for (int q = 0; q < kDensSums; q++) acc[q] = accin[(size_t)q * nrows + r];
        if (P.fd4alpha > 0)
            for (int q = kDensSums; q < kDensSumsFD; q++) acc[q] = accin[(size_t)q * nrows + r];
        else
            for (int q = kDensSums; q < kDensSumsFD; q++) acc[q] = 0;
        const double h = it.hcur[r];
        IterState sc;
        sc.left = it.left[r]; sc.right = it.right[r]; sc.cn_cur = it.cn_cur[r]; sc.cn_init = it.cn_init[r];
        const int iter = it.iter[r];
        double nh = h;
        DensityNorm dn;
        density_norm(acc, h, dn);
        const int rc = density_after_iter(P, out.cn[i], iter, h, dn, sc, nh);
        it.iter[r] = iter + 1;
        if (rc <= 0 || iter + 1 >= kMaxIter) {
            if (rc < 0) atomicMax(err, 90001009);
            else if (rc > 0) atomicMax(err, 1155);
            density_finalize(P, out, i, h, acc, dn, M[i], U[i]);
            need[r] = 0;
            if (h > 0) h_final = h;
        } else {
            it.left[r] = sc.left; it.right[r] = sc.right; it.cn_cur[r] = sc.cn_cur; it.cn_init[r] = sc.cn_init;
            it.hcur[r] = nh;
            if (nh * (1.0 + 1e-6) > cell) h_grow = nh;
            cont = true;
        }
    }
    // rows still iterating and the grid flag: one atomic per block
    const unsigned long long cm = __ballot(cont), gm = __ballot(h_grow > 0);
    __shared__ int wcont[16], wgrow[16];
    if ((threadIdx.x & 63) == 0) { wcont[threadIdx.x >> 6] = (int)__popcll(cm); wgrow[threadIdx.x >> 6] = gm != 0; }
    if (bt.b) {
        wave_max_to(h_final, row, box, hfinal, bt, kSlotHfinal);
        if (gm) wave_max_to(h_grow, row, box, nullptr, bt, kSlotHgrow);
        __syncthreads();
    } else {
        block_max_to(h_final, hfinal);  // synchronises the block
    }
    __shared__ int wbase[16];
    if (threadIdx.x == 0) {
        int c = 0, g = 0;
        for (int k = 0; k < (int)((blockDim.x + 63) >> 6); k++) { wbase[k] = c; c += wcont[k]; g |= wgrow[k]; }
        if (next_rows) {
            const int base = c ? atomicAdd(nnext, c) : 0;
            for (int k = 0; k < (int)((blockDim.x + 63) >> 6); k++) wbase[k] += base;
        } else if (c) {
            atomicAdd(nnext, c);  // no result: no wait for it
        }
        if (g) atomicOr(rebuild, 1);
        if (gate) {
            // the last block to finish, when every block's counts and maxima are in, writes
            // whether the passes launched behind this round before its counters are read stand:
            // the round's stop code; else -1 while rows still iterate; else -2 when a final
            // radius exceeds hlim, the bound every radius was taken under when the symmetric
            // list's grid was sized; else 0.  The passes' kernels with effects that a second
            // run would not undo (wake-up flags, pair counts) read it and do nothing unless 0.
            __threadfence();
            if (atomicAdd(done, 1u) == gridDim.x - 1) {
                __threadfence();
                const int e = atomicOr(err, 0), nrem = atomicOr(nnext, 0);
                const double hf = __longlong_as_double((long long)atomicOr(hfinal, 0ull));
                *gate = e ? e : (nrem ? -1 : (hf > hlim ? -2 : 0));
                *done = 0u;
            }
        }
    }
    if (next_rows) {
        // the rows that iterate again, each wave's at the place its block reserved
        __syncthreads();
        if (cont) {
            const unsigned long long below = (1ull << (threadIdx.x & 63)) - 1ull;
            next_rows[wbase[threadIdx.x >> 6] + (int)__popcll(cm & below)] = r;
        }
    }
}
#endif



// ---------------------------------------------------------------------------------------
// records: the neighbour-side fields of a pass, packed per particle (device order)
// ---------------------------------------------------------------------------------------
// the gradient record of particle j, from the state (the record a pass packs, or builds where it
// reads it)
template <typename R>
__device__ __forceinline__ GradRec<R> make_grad_rec(long j, const GradSrc &s, const double *__restrict__ X, const double *__restrict__ Y,
                                                    const double *__restrict__ Z, const BoxTable &bt)
{
    GradRec<R> g;
    g.h = (R)s.H[j]; g.m = (R)s.M[j]; g.rho = s.RHO[j]; g.p = s.PR[j];
    g.v0 = s.VX[j]; g.v1 = s.VY[j]; g.v2 = s.VZ[j];
    if constexpr (std::is_same<R, float>::value) g.fp = fine_pos(X[j], Y[j], Z[j], bt.b ? bt.b[bt.BOX[j]].box : s.box);
    return g;
}

// the gradient record of particle j
template <typename R>
__device__ __forceinline__ void pack_grad_one(long j, const double *__restrict__ H, const double *__restrict__ M,
                                              const double *__restrict__ RHO, const double *__restrict__ PR,
                                              const double *__restrict__ VX, const double *__restrict__ VY,
                                              const double *__restrict__ VZ, GradRec<R> *__restrict__ rec,
                                              const double *__restrict__ X, const double *__restrict__ Y,
                                              const double *__restrict__ Z, double box, const BoxTable &bt)
{
    rec[j] = make_grad_rec<R>(j, GradSrc{H, M, RHO, PR, VX, VY, VZ, box}, X, Y, Z, bt);
}

// every record (passes over many rows)
template <typename R>
__global__ void k_pack_grad(long n, const double *__restrict__ H, const double *__restrict__ M,
                            const double *__restrict__ RHO, const double *__restrict__ PR, const double *__restrict__ VX,
                            const double *__restrict__ VY, const double *__restrict__ VZ, GradRec<R> *__restrict__ rec,
                            const double *__restrict__ X, const double *__restrict__ Y, const double *__restrict__ Z, double box,
                            BoxTable bt)
{
    const long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (j >= n) return;
    pack_grad_one<R>(j, H, M, RHO, PR, VX, VY, VZ, rec, X, Y, Z, box, bt);
}

// the force record of particle j (effective_soundspeed, src/declarations/cell_data.h:703-712)
template <typename R>
__device__ __forceinline__ Cell<R> force_rec_of(long j, double gamma, const ForceView &F, double box, const BoxTable &bt)
{
    const long n = F.n;
    Cell<R> c;
    c.h = F.H[j]; c.m = (R)F.M[j]; c.rho = (R)F.RHO[j]; c.p = (R)F.PR[j];
    c.cn = (R)F.CN[j]; c.fce = (R)F.FCE[j]; c.drk = (R)F.DRK[j];
    c.v0 = (R)F.VX[j]; c.v1 = (R)F.VY[j]; c.v2 = (R)F.VZ[j];
    c.gr0 = (R)F.GR[j]; c.gr1 = (R)F.GR[n + j]; c.gr2 = (R)F.GR[2 * n + j];
    c.gp0 = (R)F.GP[j]; c.gp1 = (R)F.GP[n + j]; c.gp2 = (R)F.GP[2 * n + j];
    for (int k = 0; k < 9; k++) c.gv[k] = (R)F.GV[(size_t)k * n + j];
    for (int k = 0; k < 6; k++) c.t[k] = (R)F.T[(size_t)k * n + j];
    const double rho = F.RHO[j];
    c.cs = (R)((rho > 0) ? sqrt(gamma * F.PR[j] / rho) : sqrt(gamma * (gamma - 1.) * F.U[j]));
    if constexpr (std::is_same<R, float>::value) c.fp = fine_pos(F.X[j], F.Y[j], F.Z[j], bt.b ? bt.b[bt.BOX[j]].box : box);
    return c;
}

template <typename R>
__device__ __forceinline__ void pack_force_one(long j, double gamma, const ForceView &F, Cell<R> *__restrict__ rec, double box,
                                               const BoxTable &bt)
{
    rec[j] = force_rec_of<R>(j, gamma, F, box, bt);
}

// the force pass's error and wake-up flags, the first pair without a Riemann solution and the
// count of rows left to the whole Riemann ladder
__device__ __forceinline__ void force_flags_clear(int *zero2)
{
    zero2[0] = 0; zero2[2] = 0; zero2[10] = -1; zero2[11] = -1; zero2[12] = 0;
}

template <typename R>
__global__ void k_pack_force(double gamma, ForceView F, Cell<R> *__restrict__ rec, int *__restrict__ zero2, double box, BoxTable bt)
{
    const long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (j == 0 && zero2) force_flags_clear(zero2);
    if (j >= F.n) return;
    pack_force_one<R>(j, gamma, F, rec, box, bt);
}

// Few rows: the records of the rows and of their list entries only, one wavefront per row
// (a record shared by several rows is written several times with the same bytes)
template <typename R>
__global__ void __launch_bounds__(64) k_pack_grad_list(int nrows, const int *__restrict__ act, const int *__restrict__ ell, long ld,
                                                       const int *__restrict__ cnt, const double *__restrict__ H,
                                                       const double *__restrict__ M, const double *__restrict__ RHO,
                                                       const double *__restrict__ PR, const double *__restrict__ VX,
                                                       const double *__restrict__ VY, const double *__restrict__ VZ,
                                                       GradRec<R> *__restrict__ rec, const double *__restrict__ X,
                                                       const double *__restrict__ Y, const double *__restrict__ Z, double box,
                                                       BoxTable bt)
{
    const long r = xcd_block();
    if (r >= nrows) return;
    const int c = cnt[r];
    for (int q = threadIdx.x; q <= c; q += 64) {
        const int j = q == 0 ? act[r] : ell[ell_at(ld, r, q - 1)];
        pack_grad_one<R>(j, H, M, RHO, PR, VX, VY, VZ, rec, X, Y, Z, box, bt);
    }
}

template <typename R>
__global__ void __launch_bounds__(64) k_pack_force_list(int nrows, const int *__restrict__ act, const int *__restrict__ ell, long ld,
                                                        const int *__restrict__ cnt, double gamma, ForceView F,
                                                        Cell<R> *__restrict__ rec, int *__restrict__ zero2, double box, BoxTable bt)
{
    const long r = xcd_block();
    if (r == 0 && threadIdx.x == 0 && zero2) force_flags_clear(zero2);
    if (r >= nrows) return;
    const int c = cnt[r];
    for (int q = threadIdx.x; q <= c; q += 64) {
        const int j = q == 0 ? act[r] : ell[ell_at(ld, r, q - 1)];
        pack_force_one<R>(j, gamma, F, rec, box, bt);
    }
}

// ---------------------------------------------------------------------------------------
// gradients: gradient_accumulate_neighbor (src/hydro/gradient_functions.h:305-665),
// construct_gradient (src/hydro/gradients_loop.cc:129-136) and the limiter block
// (:1077-1130), fused per active particle
// ---------------------------------------------------------------------------------------
// construct_gradient (NV_T^-1 times the sums) and the limiter block of one particle
template <typename R>
__device__ inline void gradients_tail(const GradRec<R> &I, bool enabled, const R (&G)[3][5], const R (&mx)[5], const R (&mn)[5],
                                      R maxd, const double *__restrict__ T, const double *__restrict__ CN, long n, int i,
                                      double *__restrict__ GR, double *__restrict__ GP, double *__restrict__ GV)
{
    const R hi = I.h;
    const R t0 = (R)T[0 * n + i], t1 = (R)T[1 * n + i], t2 = (R)T[2 * n + i], t3 = (R)T[3 * n + i], t4 = (R)T[4 * n + i], t5 = (R)T[5 * n + i];
    R o[5][3];
    for (int b = 0; b < 5; b++) {
        const R v0 = G[0][b], v1 = G[1][b], v2 = G[2][b];
        o[b][0] = t0 * v0 + t3 * v1 + t5 * v2;
        o[b][1] = t3 * v0 + t1 * v1 + t4 * v2;
        o[b][2] = t5 * v0 + t4 * v1 + t2 * v2;
    }
    if (enabled) {
        const R stol = R(0.1);
        const R dmaxv = dmax(hi, maxd);
        const double cni = CN[i];
        R alim = R(0.25);
        if (cni > 100) alim = (R)dmin(0.5, 0.25 + 0.25 * (cni - 100) / 100);
        slopelimiter<R>(o[0], mx[0], mn[0], alim, hi, R(0), 1, dmaxv, (R)I.rho);
        slopelimiter<R>(o[1], mx[1], mn[1], alim, hi, stol, 1, dmaxv, (R)I.p);
        for (int kk = 0; kk < 3; kk++) slopelimiter<R>(o[2 + kk], mx[2 + kk], mn[2 + kk], alim, hi, stol, 0, R(0), R(0));
    }
    for (int kk = 0; kk < 3; kk++) {
        GR[(size_t)kk * n + i] = (double)o[0][kk];
        GP[(size_t)kk * n + i] = (double)o[1][kk];
        for (int b = 0; b < 3; b++) GV[(size_t)(3 * b + kk) * n + i] = (double)o[2 + b][kk];
    }
}

// four waves per SIMD (the FP64 form then fits 128 registers: 2.18 -> 1.99 ms on a 2M-row pass)
template <int KERN, typename R, bool BOXES>
__global__ void __launch_bounds__(256, 4) k_gradients(DevParams P, int nrows, const int *__restrict__ act, const double *__restrict__ X,
                                                   const double *__restrict__ Y, const double *__restrict__ Z,
                                                   const double *__restrict__ H, const GradRec<R> *__restrict__ rec, const double *__restrict__ T,
                                                   const double *__restrict__ CN, long n, const int *__restrict__ ell, long ld,
                                                   const int *__restrict__ cnt, double *__restrict__ GR, double *__restrict__ GP,
                                                   double *__restrict__ GV, BoxTable bt)
{
    const int r = (int)(xcd_block() * blockDim.x + threadIdx.x);
    if (r >= nrows) return;
    const int i = act[r];
    if (BOXES) {
        GridView g{};
        resolve_box(bt, i, P, g);
    }
    R G[3][5];
    for (int a = 0; a < 3; a++) for (int b = 0; b < 5; b++) G[a][b] = 0;
    R mx[5], mn[5];
    for (int b = 0; b < 5; b++) { mx[b] = 0; mn[b] = 0; }
    R maxd = 0;
    const GradRec<R> I = rec[i];
    const bool enabled = (I.m > 0 && I.rho > 0);
    const R hi = I.h;
    if (enabled) {
        const double xi = X[i], yi = Y[i], zi = Z[i];
        // the row's radius in FP64 (the kernel's edge factor, pair_kernel)
        const double hid = std::is_same<R, double>::value ? (double)hi : H[i];
        const double h2i = hid * hid;
        const int c = cnt[r];
        for (int q = 0; q < c; q++) {
            const int j = ell[ell_at(ld, r, q)];
            const GradRec<R> J = rec[j];
            if (!(J.m > 0) || !(J.rho > 0) || !(J.h > 0)) continue;
            R dp[3];
            record_dp<R>(P, I, J, xi, yi, zi, X, Y, Z, j, dp[0], dp[1], dp[2]);
            const R r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
            if (r2 <= 0) continue;
            const double r2d = pair_r2d<R>(dp[0], dp[1], dp[2], r2);
            if ((r2d >= h2i) && (r2d >= (double)J.h * (double)J.h)) continue;
            const R rr = sqrt(r2);
            const double rd = std::is_same<R, double>::value ? 0.0 : sqrt(r2d);
            const bool inside = std::is_same<R, double>::value ? (rr < hi) : (rd < hid);
            R wk = 0, dwk = 0;
            pair_kernel<KERN, -1, R>(rr, rd, hid, wk, dwk);
            if (rr > maxd) maxd = rr;
            R d[5];
            d[0] = (R)(J.rho - I.rho);
            d[1] = (R)(J.p - I.p);
            d[2] = (R)(J.v0 - I.v0);
            d[3] = (R)(J.v1 - I.v1);
            d[4] = (R)(J.v2 - I.v2);
            for (int b = 0; b < 5; b++) {
                if (d[b] < mn[b]) mn[b] = d[b];
                else if (d[b] > mx[b]) mx[b] = d[b];
            }
            if (inside) {
                for (int kk = 0; kk < 3; kk++) {
                    const R w = -wk * dp[kk];
                    for (int b = 0; b < 5; b++) G[kk][b] += rounded(w * d[b]);
                }
            }
        }
    }
    gradients_tail<R>(I, enabled, G, mx, mn, maxd, T, CN, n, i, GR, GP, GV);
}

// Running sum of a conserved flux total.  FP32: Neumaier-compensated (the lost low bits of each
// addition are kept in c, so sums that cancel to a small net keep their accuracy); FP64: the
// plain sum in list order (the oracle's additions).
template <typename R>
struct CSum {
    R s = 0, c = 0;
    __device__ __forceinline__ void add(R x)
    {
        const R t = s + x;
        c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s + (double)c; }
};
template <>
struct CSum<double> {
    double s = 0;
    __device__ __forceinline__ void add(double x) { s += x; }
    __device__ __forceinline__ double value() const { return s; }
};

// x_i - x_j of a pair of records: FP32 form from the records' fine cells and offsets (no FP64
// position reads); FP64 form the wrapped FP64 difference of the positions
template <typename R, typename Rec>
__device__ __forceinline__ void record_dp(const DevParams &P, const Rec &I, const Rec &J, double xi, double yi, double zi,
                                          const double *__restrict__ X, const double *__restrict__ Y,
                                          const double *__restrict__ Z, int j, R &dp0, R &dp1, R &dp2)
{
    if constexpr (std::is_same<R, float>::value) {
        fine_dp(I.fp, J.fp, (float)(P.box * (1.0 / 1024.0)), dp0, dp1, dp2);
    } else {
        double d0 = xi - X[j], d1 = yi - Y[j], d2 = zi - Z[j];
        d1 = wrap(d1, P.box, P.half);
        d0 = wrap(d0, P.box, P.half);
        d2 = wrap(d2, P.box, P.half);
        dp0 = (R)d0; dp1 = (R)d1; dp2 = (R)d2;
    }
}

// ---------------------------------------------------------------------------------------
// force: the pair loop of hydro_force() (src/hydro/hydro_force_loop.cc:423-480), two-sided
// (each active particle accumulates its own side of every pair), then the writeback
// (:227-321) and hydro_final_operations_and_cleanup (src/hydro/hydro_toplevel.cc:133-250).
// ---------------------------------------------------------------------------------------
// the writeback (src/hydro/hydro_force_loop.cc:302-321) and hydro_final_operations_and_cleanup
// (src/hydro/hydro_toplevel.cc:133-250, non-comoving MFM) of one particle
__device__ inline void force_tail(const DevParams &P, const ForceView &F, int i, bool enabled, double sound_i, double A0, double A1,
                                  double A2, double dE, double msv, double *__restrict__ ACC, double *__restrict__ DTU,
                                  double *__restrict__ MSV)
{
    const long n = F.n;
    const double mi = F.M[i];
    // MaxSignalVel = max(accumulated, c_s) over MIN_REAL_NUMBER (hydro_force_loop.cc:318-321)
    double ms = kMinReal;
    double mxs = fmax(msv, sound_i);
    if (ms < mxs) ms = mxs;
    double e = dE;
    if (enabled) {
        const double inv1 = 1.0 / 1.0;
        const double vp0 = F.VX[i] * inv1, vp1 = F.VY[i] * inv1, vp2 = F.VZ[i] * inv1;
        e -= vp0 * A0 + vp1 * A1 + vp2 * A2;
        const double invm = 1.0 / mi;
        A0 *= invm; A1 *= invm; A2 *= invm;
        e /= mi;
        if (F.H[i] >= 0.99 * P.max_kernel_radius) e = 0;
    }
    ACC[i] = A0; ACC[n + i] = A1; ACC[2 * n + i] = A2;
    DTU[i] = e;
    MSV[i] = ms;
}

// the record of i a pair loop reads: the record itself, or its split over registers and LDS
template <typename R>
__device__ __forceinline__ Cell<R> make_irec(const Cell<R> &I, const ColdRec<R> &, std::false_type) { return I; }
template <typename R>
__device__ __forceinline__ SplitCell<R> make_irec(const Cell<R> &I, const ColdRec<R> &C, std::true_type)
{
    return SplitCell<R>{I.h, I.rho, I.cs, I.v0, I.v1, I.v2, C.m, C.p, C.cn, C.fce, C.drk, C.gr0, C.gr1, C.gr2, C.gp0, C.gp1, C.gp2,
                        C.gv, C.t};
}

// One row of the lane-per-row force kernels; false (COMMON only): a face of the row needs a rare
// rung of the Riemann ladder, and nothing of the row was written but its neighbours' wake-up
// flags (which the whole-ladder pass over the row sets again, to the same values)
template <int KERN, typename R, bool BOXES, bool STATS, bool COMMON>
__device__ __forceinline__ bool force_row(const DevParams &P0, const Gam<R> &G, const ForceView &F, const Cell<R> *__restrict__ rec,
                                          const ForceRows &a, const BoxTable &bt, int r)
{
    DevParams P = P0;
    const int i = a.act[r];
    if (BOXES) {
        GridView g{};
        resolve_box(bt, i, P, g);
    }
    // hydro_force_initial_operations_preloop (src/hydro/hydro_toplevel.cc:457-565): an active
    // particle's own flag is cleared; only inactive neighbours are flagged below
    if (a.wakeup) a.wakeup[i] = 0;
    const double mi = F.M[i], rhoi = F.RHO[i];
    double A0 = 0, A1 = 0, A2 = 0, dE = 0, msv = 0;
    // effective_soundspeed of i, src/declarations/cell_data.h:703-712 (floor of MaxSignalVel)
    double sound_i = (rhoi > 0) ? sqrt(P.gamma * F.PR[i] / rhoi) : sqrt(P.gamma * (P.gamma - 1.) * F.U[i]);
    const bool enabled = (mi > 0 && rhoi > 0);
    if (enabled) {
        const Cell<R> I0 = rec[i];
        // the FP64 common pass keeps most of i's record in LDS (SplitCell); the other kernels
        // hold it in registers
        constexpr bool split = COMMON && std::is_same<R, double>::value;
        __shared__ ColdRec<R> cold[split ? 256 : 1];
        ColdRec<R> &C = cold[split ? threadIdx.x : 0];
        if constexpr (split) {
            C.m = I0.m; C.p = I0.p; C.cn = I0.cn; C.fce = I0.fce; C.drk = I0.drk;
            C.gr0 = I0.gr0; C.gr1 = I0.gr1; C.gr2 = I0.gr2; C.gp0 = I0.gp0; C.gp1 = I0.gp1; C.gp2 = I0.gp2;
            for (int k = 0; k < 9; k++) C.gv[k] = I0.gv[k];
            for (int k = 0; k < 6; k++) C.t[k] = I0.t[k];
        }
        using IC = typename std::conditional<split, SplitCell<R>, Cell<R>>::type;
        const IC I = make_irec<R>(I0, C, std::integral_constant<bool, split>{});
        const double xi = F.X[i], yi = F.Y[i], zi = F.Z[i];
        const int tbi = a.timebin ? a.timebin[i] : 0;
        const int c = a.cnt[r];
        CSum<R> a0, a1, a2, ae;
        R m = 0;
        for (int q = 0; q < c; q++) {
            // LDS reads of i's record stay inside the loop (not hoisted into registers)
            if constexpr (split) __asm__ volatile("" ::: "memory");
            const int j = a.ell[ell_at(a.ld, r, q)];
            if (m < I.cs) m = I.cs;
            const Cell<R> J = rec[j];
            R dp0, dp1, dp2;
            record_dp<R>(P, I0, J, xi, yi, zi, F.X, F.Y, F.Z, j, dp0, dp1, dp2);
            const R r2 = dp0 * dp0 + dp1 * dp1 + dp2 * dp2;
            R f0, f1, f2, fp, vts;
            unsigned rung;
            int rc = force_pair<KERN, R, COMMON>(P, G, I, J, dp0, dp1, dp2, r2, f0, f1, f2, fp, vts, rung);
            if constexpr (STATS) count_rungs(F.rungs, rung);
            if constexpr (COMMON) {
                if (rc == 2) return false;
            }
            if (rc < 0) {
                atomicMax(a.err, 1234);
                if (atomicCAS(a.err + 10, -1, i) == -1) a.err[11] = j;
                continue;
            }
            if (rc == 0) continue;
            a0.add(rounded(f0)); a1.add(rounded(f1)); a2.add(rounded(f2)); ae.add(rounded(fp));
            if (vts > m) m = vts;
            if (a.active && !a.active[j] && a.wakeup) {
                if ((double)vts > kWakeup * a.msv_in[j]) {
                    atomicMax(&a.wakeup[j], tbi + 1);
                    if (BOXES) bt.b[bt.BOX[i]].need_wakeup = 1;
                    else *a.need_wakeup = 1;
                }
            }
        }
        A0 = a0.value(); A1 = a1.value(); A2 = a2.value(); dE = ae.value(); msv = m;
    }
    force_tail(P, F, i, enabled, sound_i, A0, A1, A2, dE, msv, a.ACC, a.DTU, a.MSV);
    return true;
}

// one lane per row of k_force and k_force_w3 (the kernels differ in their launch bounds only)
template <int KERN, typename R, bool BOXES, bool STATS, bool COMMON>
__device__ __forceinline__ void force_lanes(const DevParams &P, const Gam<R> &G, const ForceView &F, const Cell<R> *__restrict__ rec,
                                            const ForceRows &a, const BoxTable &bt)
{
    const int k = (int)(xcd_block() * blockDim.x + threadIdx.x);
    const int r = k < a.nrows ? (a.sched ? a.sched[k] : k) : a.nrows;
    // pairs of this launch when asked (wave sum, one atomic per wave; a launch of k_sum_counts
    // after the pass costs fewer atomics on the one address)
    if (a.pairs) {
        unsigned long long c = (r < a.nrows) ? (unsigned long long)a.cnt[r] : 0ull;
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if ((threadIdx.x & 63) == 0 && c) atomicAdd(a.pairs, c);
    }
    const bool done = (r >= a.nrows) || force_row<KERN, R, BOXES, STATS, COMMON>(P, G, F, rec, a, bt, r);
    // rows left to the whole ladder: appended a wavefront at a time (one atomic per wavefront, the
    // wavefront's rows kept together and in lane order)
    const unsigned long long left = __ballot(!done);
    if (!left) return;
    const int lane = threadIdx.x & 63, lead = __ffsll((long long)left) - 1;
    int base = 0;
    if (lane == lead) base = atomicAdd(a.err + 12, __popcll(left));
    base = __shfl(base, lead, 64);
    if (!done) a.rare[base + __popcll(left & ((1ull << lane) - 1ull))] = r;
}

// the FP64 common pass of one box at two waves per SIMD (172 registers, no scratch; its LDS
// record: 52 KB per workgroup); with the box table it takes 178 (two waves, no spill)
template <int KERN, typename R, bool BOXES, bool STATS, bool COMMON>
__global__ void __launch_bounds__(256, (COMMON && !BOXES && std::is_same<R, double>::value) ? 2 : 1) k_force(DevParams P, Gam<R> G, ForceView F, const Cell<R> *__restrict__ rec, ForceRows a,
                                               BoxTable bt)
{
    force_lanes<KERN, R, BOXES, STATS, COMMON>(P, G, F, rec, a, bt);
}

// the FP64 common pass of one box at three waves per SIMD (166 registers).  Which is faster
// depends on the field: on a 2M-row lattice pass the rows resident at three waves outgrow the
// L2 with their neighbour records (7.3 ms against 6.7 at two); on the 4M lognormal field of
// radii over many levels three waves hide more of the loads' latency (20.2 ms against 23.8)
template <int KERN>
__global__ void __launch_bounds__(256, 3) k_force_w3(DevParams P, Gam<double> G, ForceView F, const Cell<double> *__restrict__ rec,
                                                     ForceRows a, BoxTable bt)
{
    force_lanes<KERN, double, false, false, true>(P, G, F, rec, a, bt);
}

// two waves per SIMD: the whole ladder then spills a few registers (FP64 ~300 wanted) but hides
// more latency than one wave of the unspilled kernel (4M lognormal pass: FP64 93.4 -> 88.0 ms,
// FP32 68.9 -> 64.5 ms)
template <int KERN, typename R, bool BOXES>
__global__ void __launch_bounds__(256, 2) k_force_rare(DevParams P, Gam<R> G, ForceView F, const Cell<R> *__restrict__ rec, ForceRows a,
                                                    BoxTable bt)
{
    const int count = a.err[12];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x)
        force_row<KERN, R, BOXES, false, false>(P, G, F, rec, a, bt, a.rare[k]);
}

// ---------------------------------------------------------------------------------------
// The Riemann functions on given face states, one lane per case (a diagnostic with no
// counterpart in the reference; the force pass reaches the same functions through the pair
// body).  Case k reads kProbeIn values from in + k*kProbeIn: the left state (rho, p, v[3]), the
// right state, the face normal n[3], the pressure limiter, then the first-order states the
// ladder retries with (p_i, p_j, rho_i, rho_j, v_i[3], v_j[3]); it writes P_M, S_M, the momentum
// flux[3], the energy flux and 1 (0: the ladder found no positive pressure) to
// out + k*kProbeOut, and the RiemannRung bits taken to rung[k].  Modes: 0 the ladder of the pair
// body with its two retries (hydro_core_meshless_functions.h:351-427), 1 Riemann_solver
// (reimann.h:282-370), 2 get_wavespeeds_and_pressure_star alone (:485-549), 3
// Riemann_solver_KurganovTadmor_PWK alone (:445-477), 4 Riemann_solver_exact alone (:625-668).
// ---------------------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) k_riemann_probe(int mode, Gam<R> G, long n, const double *__restrict__ in,
                                                       double *__restrict__ out, unsigned *__restrict__ rung)
{
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double *c = in + k * kProbeIn;
    RState<R> L, Rs;
    L.rho = (R)c[0]; L.p = (R)c[1]; L.v0 = (R)c[2]; L.v1 = (R)c[3]; L.v2 = (R)c[4];
    Rs.rho = (R)c[5]; Rs.p = (R)c[6]; Rs.v0 = (R)c[7]; Rs.v1 = (R)c[8]; Rs.v2 = (R)c[9];
    const R nv[3] = {(R)c[10], (R)c[11], (R)c[12]};
    const R ptl = (R)c[13];
    ROut<R> o;
    o.P_M = o.S_M = o.f0 = o.f1 = o.f2 = o.fp = 0;
    unsigned rb = 0;
    bool ok = true;
    if (mode == 0) {
        ok = riemann_ladder<R>(G, L, Rs, nv, ptl, (R)c[14], (R)c[15], (R)c[16], (R)c[17], (R)c[18], (R)c[19], (R)c[20], (R)c[21],
                               (R)c[22], (R)c[23], o, rb);
    } else if (mode == 1) {
        rb = riemann(G, L, Rs, nv, ptl, o);
    } else {
        const RPrep<R> q = riemann_prep(G, L, Rs, nv);
        if (mode == 2) rb = wavespeeds(G, L, Rs, nv, q.vL, q.vR, q.csL, q.csR, q.hL, q.hR, o, ptl);
        else if (mode == 3) { kt_flux(L, Rs, nv, q.vL, q.vR, q.csL, q.csR, q.hL, q.hR, o); rb = kRungKt; }
        else rb = exact_riemann(G, L, Rs, q.vL, q.vR, q.csL, q.csR, o);
    }
    double *w = out + k * kProbeOut;
    w[0] = (double)o.P_M; w[1] = (double)o.S_M; w[2] = (double)o.f0; w[3] = (double)o.f1; w[4] = (double)o.f2; w[5] = (double)o.fp;
    w[6] = ok ? 1.0 : 0.0;
    w[7] = 0.0;
    rung[k] = rb;
}

#ifdef GZ_DEFINE_COMMON_KERNELS
// per block of 1024 rows: the sum of the list lengths and the sum of each wave's longest
__global__ void __launch_bounds__(1024) k_sched_stats(int nrows, const int *__restrict__ cnt, unsigned long long *__restrict__ partial)
{
    __shared__ unsigned long long ss[16], sm[16];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    int c = r < nrows ? cnt[r] : 0;
    int mx = c;
    unsigned long long sum = (unsigned long long)c;
    for (int o = 32; o > 0; o >>= 1) {
        mx = max(mx, __shfl_xor(mx, o, 64));
        sum += __shfl_xor(sum, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { ss[w] = sum; sm[w] = (unsigned long long)mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) { a += ss[k]; b += sm[k]; }
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = b;
    }
}

// the totals of the partial sums (one block)
__global__ void __launch_bounds__(1024) k_sched_total(int nparts, const unsigned long long *__restrict__ partial,
                                                      unsigned long long *__restrict__ total)
{
    __shared__ unsigned long long ss[16], sm[16];
    unsigned long long a = 0, b = 0;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) { a += partial[2 * k]; b += partial[2 * k + 1]; }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { ss[w] = a; sm[w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long x = 0, y = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) { x += ss[k]; y += sm[k]; }
        total[0] = x;
        total[1] = y;
    }
}

// The force pass's order of each block of span rows (span a power of two, at most
// kSchedMaxSpan): longest list first, ties in row order (lengths above 2047 tie), by a bitonic
// sort of (2047 - length, row) keys in LDS; the plain order when the waves are even.  The rows
// of a block stay close in space; their waves even out.
__global__ void __launch_bounds__(1024) k_sched_sort(int nrows, int span, const int *__restrict__ cnt,
                                                     const unsigned long long *__restrict__ total, double min_even,
                                                     int *__restrict__ sched)
{
    __shared__ unsigned int key[kSchedMaxSpan];
    const int base = blockIdx.x * span;
    const int m = min(span, nrows - base);  // rows of this block
    const double even = (double)total[0] / (64.0 * (double)(total[1] > 0 ? total[1] : 1));
    if (!(even < min_even)) {
        for (int k = threadIdx.x; k < m; k += blockDim.x) sched[base + k] = base + k;
        return;
    }
    for (int k = threadIdx.x; k < span; k += blockDim.x) {
        // rows past the end sort last
        const unsigned c = k < m ? (unsigned)min(cnt[base + k], 2047) : 0u;
        key[k] = k < m ? (((2047u - c) << 11) | (unsigned)k) : 0xffffffffu;
    }
    __syncthreads();
    for (int size = 2; size <= span; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = threadIdx.x; k < span / 2; k += blockDim.x) {
                // the pair (lo, hi) of this compare-exchange, ascending within blocks of size
                const int lo = 2 * k - (k & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const unsigned a = key[lo], b2 = key[hi];
                if ((a > b2) == up) { key[lo] = b2; key[hi] = a; }
            }
            __syncthreads();
        }
    }
    for (int k = threadIdx.x; k < m; k += blockDim.x) sched[base + k] = base + (int)(key[k] & 2047u);
}
#endif

}  // namespace gz
