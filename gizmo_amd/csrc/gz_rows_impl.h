// gz_rows_impl.h -- the hot-path passes with one wavefront per active row, for steps in which
// few particles are active (the usual case of adaptive time bins: a few hundred hot particles
// in a resting medium).  The 64 lanes of the row's wavefront evaluate 64 candidates or list
// entries at once; each lane's pair terms go to LDS, and then lane k keeps running sum k (one
// lane per accumulated quantity) over the accepted entries in ascending order, so every sum is
// the same chain of additions as in the lane-per-row kernels (and the oracle).  One wavefront
// per workgroup: the block barrier is a wave barrier.
#pragma once
#include "gz_kernels_impl.h"

namespace gz {

// The candidates of a row: the non-empty cells of its home cell's 27-stencil in ascending cell
// code (cells are consecutive ranges of the Morton order, so this is ascending start: the
// candidates come in ascending particle index whatever the grid level), laid end to end.
// Virtual index v maps to the particle of range k with pref[k] <= v < pref[k+1].
struct RowStencil {
    int start[27];
    int pref[28];
    int code[27];
    int len[27];
    int nst;
};

// of equal codes (grids under 3 cells a side) the first lane stays; ranks by code; prefix of
// the lengths.  Lanes below 27 bring their cell's code (0x7fffffff: dropped) and range, and
// have stored the code in rs.code.
__device__ __forceinline__ void stencil_finish(RowStencil &rs, int code, int st, int en)
{
    const int lane = threadIdx.x;
    __syncthreads();
    bool first = false;
    if (lane < 27 && code != 0x7fffffff) {
        first = true;
        for (int q = 0; q < lane; q++)
            if (rs.code[q] == code) { first = false; break; }
    }
    const unsigned long long firsts = __ballot(first);
    if (first) {
        int rank = 0;  // distinct kept codes below this one
        for (int q = 0; q < 27; q++)
            if (((firsts >> q) & 1ull) && rs.code[q] < code) rank++;
        rs.start[rank] = st;
        rs.len[rank] = en - st;
    }
    __syncthreads();
    if (lane == 0) {
        const int m = __popcll(firsts);
        int acc = 0;
        for (int k = 0; k < m; k++) { rs.pref[k] = acc; acc += rs.len[k]; }
        rs.pref[m] = acc;
        rs.nst = m;
    }
    __syncthreads();
}

// The stencil of the row's home cell, built by 27 lanes (one per neighbour cell): periodic
// neighbours, duplicates of small grids dropped, cells beyond reach dropped (the row lies
// farther from the cell's box than its reach plus the distance any particle has moved from its
// cell; reach = h_i, or max(h_i, the cell's largest radius) for the symmetric test), ranks by
// code, then the prefix of the lengths.
template <bool SYM>
__device__ __forceinline__ void row_stencil(const GridView &g, uint32_t home, double xi, double yi, double zi, double hrow,
                                            RowStencil &rs)
{
    const int lane = threadIdx.x;
    const int nc = 1 << g.level;
    const int cx = compact10(home >> 2), cy = compact10(home >> 1), cz = compact10(home);
    int code = 0x7fffffff, st = 0, en = 0;
    if (lane < 27) {
        const int dx = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dz = lane % 3 - 1;
        const int a = (cx + dx + nc) % nc, b = (cy + dy + nc) % nc, d = (cz + dz + nc) % nc;
        const int c = (int)((spread10(a) << 2) | (spread10(b) << 1) | spread10(d));
        st = g.cstart[g.cell_off + c];
        en = g.cend[g.cell_off + c];
        bool keep = en > st;
        if (keep && nc >= 4) {
            const double ox = cx * g.cell, oy = cy * g.cell, oz = cz * g.cell;
            const double rx = wrap(xi - ox, g.box, g.half), ry = wrap(yi - oy, g.box, g.half), rz = wrap(zi - oz, g.box, g.half);
            // gap from the row to the neighbour cell's box [d*cell, (d+1)*cell) along each axis
            const double gx = fmax(fmax(dx * g.cell - rx, rx - (dx + 1) * g.cell), 0.0);
            const double gy = fmax(fmax(dy * g.cell - ry, ry - (dy + 1) * g.cell), 0.0);
            const double gz = fmax(fmax(dz * g.cell - rz, rz - (dz + 1) * g.cell), 0.0);
            double reach = hrow;
            if (SYM) reach = fmax(reach, g.chmax ? g.chmax[g.cell_off + c] : g.hall);
            const double rr = (reach + g.margin) * (1.0 + 1e-9) + 1e-12 * g.box;
            if (gx * gx + gy * gy + gz * gz > rr * rr) keep = false;
        }
        if (keep) code = c;
        rs.code[lane] = code;
    }
    stencil_finish(rs, code, st, en);
}

// (v < the total; the range k with pref[k] <= v < pref[k+1], the last such k when empty ranges
// share a prefix value: binary search over the at most 27 prefixes)
__device__ __forceinline__ int row_candidate(const RowStencil &rs, int v)
{
    int lo = 0, hi = rs.nst - 1;  // invariant: pref[lo] <= v, and the answer is at most hi
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (rs.pref[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return rs.start[lo] + (v - rs.pref[lo]);
}

// ---------------------------------------------------------------------------------------
// search over many grid levels (WideGrid)
// ---------------------------------------------------------------------------------------

// the finest level (at most lmax) whose cells are no smaller than reach
__device__ __forceinline__ int wide_level(double box, int lmax, double reach)
{
    int L = 0;
    while (L < lmax && box / (double)(1 << (L + 1)) >= reach * (1.0 + 1e-6)) L++;
    return L;
}

// the range [a, b) of the cell with code c (key >> s) from table entry e, if the entry is
// exactly a maximal run of that code among positions [lo, hi) (map: position -> particle,
// nullptr for the order itself); otherwise the cell is empty
__device__ __forceinline__ bool wide_range(const int *__restrict__ start, const int *__restrict__ end, long long e,
                                           const uint32_t *__restrict__ key, const int *__restrict__ map, int lo, int hi, int s,
                                           uint32_t c, int &a, int &b)
{
    a = start[e];
    b = end[e];
    if (!(a >= lo && a < b && b <= hi)) return false;
    auto code = [&](int p) { return key[map ? map[p] : p] >> s; };
    if (code(a) != c || code(b - 1) != c) return false;
    if (a > lo && code(a - 1) == c) return false;
    if (b < hi && code(b) == c) return false;
    return true;
}

// The 27-cell stencil at level L around the row (home = the row's key >> s), from the tables
// of all particles (LIST false: ranges of the order) or of the particles of level L (LIST
// true: ranges of positions in lidx).  Cells beyond reach + margin are dropped when reach > 0.
template <bool LIST>
__device__ __forceinline__ void row_stencil_wide(const GridView &g, const WideGrid &W, int L, uint32_t home, double xi, double yi,
                                                 double zi, double reach, RowStencil &rs)
{
    const int lane = threadIdx.x;
    const int nc = 1 << L;
    const int s = 3 * (kFineLevel - L);
    const double cell = g.box / (double)nc;
    const int cx = compact10(home >> 2), cy = compact10(home >> 1), cz = compact10(home);
    int code = 0x7fffffff, st = 0, en = 0;
    if (lane < 27) {
        const int dx = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dz = lane % 3 - 1;
        const int a = (cx + dx + nc) % nc, b = (cy + dy + nc) % nc, d = (cz + dz + nc) % nc;
        const uint32_t c = (spread10(a) << 2) | (spread10(b) << 1) | spread10(d);
        const long long e = W.loff[L] + c;
        bool keep = LIST ? wide_range(W.lstart, W.lend, e, W.key, W.lidx, W.lofs[L], W.lofs[L + 1], s, c, st, en)
                         : wide_range(W.start, W.end, e, W.key, nullptr, 0, W.n, s, c, st, en);
        if (keep && nc >= 4 && reach > 0) {
            const double ox = cx * cell, oy = cy * cell, oz = cz * cell;
            const double rx = wrap(xi - ox, g.box, g.half), ry = wrap(yi - oy, g.box, g.half), rz = wrap(zi - oz, g.box, g.half);
            const double gx = fmax(fmax(dx * cell - rx, rx - (dx + 1) * cell), 0.0);
            const double gy = fmax(fmax(dy * cell - ry, ry - (dy + 1) * cell), 0.0);
            const double gz = fmax(fmax(dz * cell - rz, rz - (dz + 1) * cell), 0.0);
            const double rr = (reach + g.margin) * (1.0 + 1e-9) + 1e-12 * g.box;
            if (gx * gx + gy * gy + gz * gz > rr * rr) keep = false;
        }
        if (!keep) { st = 0; en = 0; }
        if (keep) code = (int)c;
        rs.code[lane] = code;
    }
    stencil_finish(rs, code, st, en);
}

// The place of this thread's entry among the flagged entries of its workgroup, in thread order
// (total: the flagged entries of the workgroup).  One barrier: smask[TPB / 64] must not be read
// by another round of the caller before the caller's next barrier.
template <int TPB>
__device__ __forceinline__ int flagged_rank(bool flag, unsigned long long *smask, int &total)
{
    const unsigned long long mask = __ballot(flag);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) smask[w] = mask;
    __syncthreads();
    int before = 0;
    total = 0;
    for (int k = 0; k < TPB / 64; k++) {
        const int c = __popcll(smask[k]);
        if (k < w) before += c;
        total += c;
    }
    return before + __popcll(mask & ((1ull << lane) - 1ull));
}

// acc + s[0] + s[1] + ... + s[n-1] (or minus each, neg), added one after another in that order;
// the loads run ahead of the chain of additions (the terms of a round sit contiguously in LDS)
template <typename R>
__device__ __forceinline__ R ordered_sum(R acc, const R *s, int n, bool neg)
{
    int k = 0;
    for (; k + 4 <= n; k += 4) {
        R a = s[k], b = s[k + 1], c = s[k + 2], d = s[k + 3];
        if (neg) { a = -a; b = -b; c = -c; d = -d; }
        acc = acc + a;
        acc = acc + b;
        acc = acc + c;
        acc = acc + d;
    }
    for (; k < n; k++) acc = acc + (neg ? -s[k] : s[k]);
    return acc;
}

// density sums of the rows still iterating (DensitySpec::pair_kernel,
// src/hydro/density_loop.h:1141-1385), the terms in k_density_cells' order
template <int KERN, typename R, int TPB>
__global__ void __launch_bounds__(TPB) k_density_rows(DevParams P, GridView g, int nrows, const int *__restrict__ act,
                                                     const uint8_t *__restrict__ need, const uint32_t *__restrict__ key,
                                                     const double *__restrict__ X, const double *__restrict__ Y,
                                                     const double *__restrict__ Z, const double *__restrict__ VX,
                                                     const double *__restrict__ VY, const double *__restrict__ VZ,
                                                     const double *__restrict__ M, const double *__restrict__ hcur,
                                                     double *__restrict__ accout, BoxTable bt, int *__restrict__ npairs,
                                                     const int *__restrict__ rows, int nlist)
{
    __shared__ RowStencil rs;
    __shared__ R sc[kDensSumsFD][TPB + 1];  // the accepted candidates' terms, in candidate order (padded: no bank conflicts)
    __shared__ unsigned long long smask[TPB / 64];
    // workgroup b takes row b, or the b-th row of the list of rows still iterating; the pair
    // count goes to npairs[b]
    const long b = xcd_block();
    if (b >= (rows ? nlist : nrows)) return;
    const long r = rows ? rows[b] : b;
    if (!need[r]) {
        if (npairs && threadIdx.x == 0) npairs[b] = 0;
        return;
    }
    const int tid = threadIdx.x;
    const int i = act[r];
    resolve_box(bt, i, P, g);
    const double xi = X[i], yi = Y[i], zi = Z[i];
    const R vxi = (R)VX[i], vyi = (R)VY[i], vzi = (R)VZ[i];
    const double h = hcur[r];
    const R hR = (R)h;
    const R h2 = hR * hR;
    R hinv, hinv3, hinv4;
    kernel_hinv(hR, hinv, hinv3, hinv4);
    if (g.wide) {
        // the row's own level: cells no smaller than its radius plus twice the displacement bound
        const WideGrid &W = *g.wide;
        const int L = wide_level(g.box, W.lmax, h + 2.0 * g.margin);
        row_stencil_wide<false>(g, W, L, key[i] >> (3 * (kFineLevel - L)), xi, yi, zi, h, rs);
    } else {
        row_stencil<false>(g, key[i] >> g.shift, xi, yi, zi, h, rs);
    }
    // FP32 form on one grid of at least 8 cells a side: differences of positions relative to the
    // home cell's origin, the cell walk's arithmetic (the two walks give the same bits)
    const bool rel = std::is_same<R, float>::value && !g.wide && (1 << g.level) >= 8;
    double ox = 0, oy = 0, oz = 0;
    if (rel) {
        const uint32_t home = key[i] >> g.shift;
        ox = compact10(home >> 2) * g.cell; oy = compact10(home >> 1) * g.cell; oz = compact10(home) * g.cell;
    }
    const R rxi = (R)wrap(xi - ox, g.box, g.half), ryi = (R)wrap(yi - oy, g.box, g.half), rzi = (R)wrap(zi - oz, g.box, g.half);
    const int total = rs.pref[rs.nst];
    const bool fd = P.fd4alpha > 0;
    const int nsum = fd ? kDensSumsFD : kDensSums;
    R acc = 0;  // thread q < nsum keeps sum q
    int accepted = 0;  // thread 17: the row's pairs (the pass's work, counted when npairs is given)
    for (int base = 0; base < total; base += TPB) {
        const int v = base + tid;
        bool ok = false;
        R t[kDensSumsFD];
        if (v < total) {
            const int j = row_candidate(rs, v);
            const double mj = M[j];
            if (mj > 0) {
                R dp0, dp1, dp2;
                if (rel) {
                    dp0 = rxi - (R)wrap(X[j] - ox, g.box, g.half);
                    dp1 = ryi - (R)wrap(Y[j] - oy, g.box, g.half);
                    dp2 = rzi - (R)wrap(Z[j] - oz, g.box, g.half);
                } else {
                    double d0 = xi - X[j], d1 = yi - Y[j], d2 = zi - Z[j];
                    d1 = wrap(d1, P.box, P.half);
                    d0 = wrap(d0, P.box, P.half);
                    d2 = wrap(d2, P.box, P.half);
                    dp0 = (R)d0; dp1 = (R)d1; dp2 = (R)d2;
                }
                const R r2 = dp0 * dp0 + dp1 * dp1 + dp2 * dp2;
                if (r2 < h2) {
                    const R rr = sqrt(r2);
                    const R u = rr * hinv;
                    R wk = 0, dwk = 0;
                    kernel_main<KERN, 0>(u, hinv3, hinv4, wk, dwk);
                    t[0] = wk;
                    t[1] = (R)mj * wk;
                    t[2] = -(R(3.0) * hinv * wk + u * dwk);
                    ok = true;
                    // the particle itself (rr = 0) adds to the first three sums only: its other
                    // terms are +0, which leave a sum that starts at +0 unchanged bit for bit
                    for (int q = 3; q < kDensSumsFD; q++) t[q] = 0;
                    if (rr > 0) {
                        t[4] = wk * (dp0 * dp0);
                        t[5] = wk * (dp1 * dp1);
                        t[6] = wk * (dp2 * dp2);
                        t[7] = wk * (dp0 * dp1);
                        t[8] = wk * (dp1 * dp2);
                        t[9] = wk * (dp0 * dp2);
                        t[10] = wk * dp0;
                        t[11] = wk * dp1;
                        t[12] = wk * dp2;
                        const R dv0 = vxi - (R)VX[j], dv1 = vyi - (R)VY[j], dv2 = vzi - (R)VZ[j];
                        t[3] = dwk * (dp0 * dv0 + dp1 * dv1 + dp2 * dv2) / rr;
                        if (fd) {  // GradH_numer, GradH_denom (src/hydro/density_loop.h:1281-1286)
                            t[13] = (dwk / rr) * dp0;
                            t[14] = (dwk / rr) * dp1;
                            t[15] = (dwk / rr) * dp2;
                            t[16] = u * dwk;
                        }
                    }
                }
            }
        }
        // the accepted candidates' terms packed in candidate order; threads 0-12 then add them
        int nok;
        const int pos = flagged_rank<TPB>(ok, smask, nok);
        if (ok) {
            // static bounds: the terms stay in registers
            for (int q = 0; q < kDensSums; q++) sc[q][pos] = t[q];
            if (fd)
                for (int q = kDensSums; q < kDensSumsFD; q++) sc[q][pos] = t[q];
        }
        __syncthreads();
        if (tid < nsum) acc = ordered_sum(acc, sc[tid], nok, tid == 3);
        else if (tid == kDensSumsFD) accepted += nok;
        __syncthreads();
    }
    if (tid < nsum) accout[(size_t)tid * nrows + r] = (double)acc;
    if (tid == kDensSumsFD && npairs) npairs[b] = accepted;
}

// the symmetric list of few rows (the same entries, in the same order, as k_build_sym_cells)
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void __launch_bounds__(kRowThreads) k_build_sym_rows(GridView g, int nrows, const int *__restrict__ act, const uint32_t *__restrict__ key,
                                                       const double *__restrict__ X, const double *__restrict__ Y,
                                                       const double *__restrict__ Z, const double *__restrict__ H, int *__restrict__ ell,
                                                       long ld, int cap, int *__restrict__ cnt, int *__restrict__ overflow,
                                                       BoxTable bt)
{
    __shared__ RowStencil rs;
    __shared__ int scount[2][kRowThreads / 64];
    const long r = xcd_block();
    if (r >= nrows) return;
    const int tid = threadIdx.x, w = tid >> 6;
    const int i = act[r];
    {
        DevParams P{};
        resolve_box(bt, i, P, g);
    }
    const double xi = X[i], yi = Y[i], zi = Z[i], hi = H[i];
    row_stencil<true>(g, key[i] >> g.shift, xi, yi, zi, hi, rs);
    const int total = rs.pref[rs.nst];
    int n = 0;
    for (int base = 0, it = 0; base < total; base += kRowThreads, it ^= 1) {
        const int v = base + tid;
        bool acc = false;
        int j = -1;
        if (v < total) {
            j = row_candidate(rs, v);
            acc = (j != i) && accept_pair<true>(xi, yi, zi, X[j], Y[j], Z[j], hi, H[j], g.box, g.half);
        }
        // place among the accepted candidates: the wavefronts before this one, then the lanes
        int cntk;
        const int pos = kept_rank(acc, cntk);
        if ((tid & 63) == 0) scount[it][w] = cntk;
        __syncthreads();  // (the two count buffers alternate: one barrier per round suffices)
        int off = 0, all = 0;
        for (int q = 0; q < kRowThreads / 64; q++) {
            const int cq = scount[it][q];
            if (q < w) off += cq;
            all += cq;
        }
        if (acc && n + off + pos < cap) ell[ell_at(ld, r, n + off + pos)] = j;
        n += all;
    }
    if (tid == 0) {
        cnt[r] = min(n, cap);  // readers stay inside the list; the overflow carries the full length
        if (n > cap) atomicMax(overflow, n);
    }
}

// ---------------------------------------------------------------------------------------
// search over many grid levels: tables, particle levels, symmetric rows
// ---------------------------------------------------------------------------------------

// the range of every cell of every level 0..lmax that holds particles (one lane per position)
__global__ void __launch_bounds__(256) k_wide_bounds(WideGrid W)
{
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= W.n) return;
    const uint32_t k = W.key[p];
    const uint32_t kp = p > 0 ? W.key[p - 1] : 0, kn = p + 1 < W.n ? W.key[p + 1] : 0;
    for (int L = 0; L <= W.lmax; L++) {
        const int s = 3 * (kFineLevel - L);
        const uint32_t c = k >> s;
        if (p == 0 || (kp >> s) != c) W.start[W.loff[L] + c] = (int)p;
        if (p + 1 == W.n || (kn >> s) != c) W.end[W.loff[L] + c] = (int)(p + 1);
    }
}

// each particle's own level (the finest, at most lmax, whose cells hold its radius plus twice
// the displacement bound), and the number of particles per level
__global__ void __launch_bounds__(256) k_particle_level(int n, const double *__restrict__ H, double box, double margin, int lmax,
                                                         int *__restrict__ plev, int *__restrict__ level_count)
{
    __shared__ int hist[kFineLevel + 1];
    if (threadIdx.x <= kFineLevel) hist[threadIdx.x] = 0;
    __syncthreads();
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        const int L = wide_level(box, lmax, H[p] + 2.0 * margin);
        plev[p] = L;
        atomicAdd(&hist[L], 1);
    }
    __syncthreads();
    if (threadIdx.x <= kFineLevel && hist[threadIdx.x]) atomicAdd(&level_count[threadIdx.x], hist[threadIdx.x]);
}

// the range of every cell of level L among the particles of level L (positions in lidx)
__global__ void __launch_bounds__(256) k_wide_list_bounds(WideGrid W, const int *__restrict__ plev)
{
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= W.n) return;
    const int j = W.lidx[p];
    const int L = plev[j];
    const int s = 3 * (kFineLevel - L);
    const uint32_t c = W.key[j] >> s;
    const int lo = W.lofs[L], hi = W.lofs[L + 1];
    if (p == lo || (W.key[W.lidx[p - 1]] >> s) != c) W.lstart[W.loff[L] + c] = (int)p;
    if (p + 1 == hi || (W.key[W.lidx[p + 1]] >> s) != c) W.lend[W.loff[L] + c] = (int)(p + 1);
}

__device__ __forceinline__ bool periodic_adjacent(int a, int b, int nc)
{
    const int d = ((a - b) % nc + nc) % nc;
    return d <= 1 || d >= nc - 1;
}

// Symmetric rows r2 < max(h_i,h_j)^2 (src/mesh/sfc_tiles_functions.h:83-102) over many levels.
// Part A: every particle in the 27 cells of the row's own level L_i (it holds every partner of
// level L_i or finer, and every partner within h_i).  Part B: for each coarser level L, the
// particles of level L in the 27 cells of level L around the row (a partner j of level L with
// r < h_j lies there) that Part A did not cover.  The row is then put in ascending order.
__global__ void __launch_bounds__(64) k_build_sym_wide(GridView g, int nrows, const int *__restrict__ act, const double *__restrict__ X,
                                                       const double *__restrict__ Y, const double *__restrict__ Z,
                                                       const double *__restrict__ H, int *__restrict__ ell, long ld, int cap,
                                                       int *__restrict__ cnt, int *__restrict__ overflow)
{
    __shared__ RowStencil rs;
    __shared__ int srt[1024];
    const long r = xcd_block();
    if (r >= nrows) return;
    const int lane = threadIdx.x;
    const int i = act[r];
    const WideGrid &W = *g.wide;
    const double xi = X[i], yi = Y[i], zi = Z[i], hi = H[i];
    const int Li = wide_level(g.box, W.lmax, hi + 2.0 * g.margin);
    const int si = 3 * (kFineLevel - Li);
    const uint32_t homei = W.key[i] >> si;
    row_stencil_wide<false>(g, W, Li, homei, xi, yi, zi, -1.0, rs);
    int total = rs.pref[rs.nst];
    int n = 0;
    for (int base = 0; base < total; base += 64) {
        const int v = base + lane;
        bool acc = false;
        int j = -1;
        if (v < total) {
            j = row_candidate(rs, v);
            acc = (j != i) && accept_pair<true>(xi, yi, zi, X[j], Y[j], Z[j], hi, H[j], g.box, g.half);
        }
        int cntk;
        const int pos = kept_rank(acc, cntk);
        if (acc && n + pos < cap) ell[ell_at(ld, r, n + pos)] = j;
        n += cntk;
    }
    const int nA = n;
    const int nci = 1 << Li;
    const int hx = compact10(homei >> 2), hy = compact10(homei >> 1), hz = compact10(homei);
    for (int L = 0; L < Li; L++) {
        if (W.lofs[L + 1] == W.lofs[L]) continue;  // no particle of this level
        __syncthreads();  // every lane is done with rs before it is built again
        const int s = 3 * (kFineLevel - L);
        row_stencil_wide<true>(g, W, L, W.key[i] >> s, xi, yi, zi, -1.0, rs);
        total = rs.pref[rs.nst];
        for (int base = 0; base < total; base += 64) {
            const int v = base + lane;
            bool acc = false;
            int j = -1;
            if (v < total) {
                j = W.lidx[row_candidate(rs, v)];
                const uint32_t cj = W.key[j] >> si;
                const bool in_a = periodic_adjacent((int)compact10(cj >> 2), hx, nci) && periodic_adjacent((int)compact10(cj >> 1), hy, nci) &&
                                  periodic_adjacent((int)compact10(cj), hz, nci);
                acc = !in_a && (j != i) && accept_pair<true>(xi, yi, zi, X[j], Y[j], Z[j], hi, H[j], g.box, g.half);
            }
            int cntk;
            const int pos = kept_rank(acc, cntk);
            if (acc && n + pos < cap) ell[ell_at(ld, r, n + pos)] = j;
            n += cntk;
        }
    }
    __syncthreads();  // the row's entries written by every lane are visible to all
    if (n > nA && n <= cap) {
        // ascending order: each entry's rank among the row's distinct entries
        if (n <= 1024) {
            for (int q = lane; q < n; q += 64) srt[q] = ell[ell_at(ld, r, q)];
            __syncthreads();
            for (int q = lane; q < n; q += 64) {
                const int e = srt[q];
                int rank = 0;
                for (int t = 0; t < n; t++) rank += srt[t] < e;
                ell[ell_at(ld, r, rank)] = e;
            }
        } else if (lane == 0) {
            for (int q = 1; q < n; q++) {
                const int e = ell[ell_at(ld, r, q)];
                int t = q - 1;
                while (t >= 0 && ell[ell_at(ld, r, t)] > e) { ell[ell_at(ld, r, t + 1)] = ell[ell_at(ld, r, t)]; t--; }
                ell[ell_at(ld, r, t + 1)] = e;
            }
        }
    }
    if (lane == 0) {
        cnt[r] = min(n, cap);  // readers stay inside the list; the overflow carries the full length
        if (n > cap) atomicMax(overflow, n);
    }
}
#endif

// ---------------------------------------------------------------------------------------
// symmetric list as a cell walk (one wavefront per occupied cell, culled LDS-staged candidates):
// rows r2 < max(h_i,h_j)^2, self excluded, ascending (src/mesh/sfc_tiles_functions.h:83-102);
// with row groups (CellRows.gsten) a group is one cell of one grid level or of one box
// ---------------------------------------------------------------------------------------
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void __launch_bounds__(64) k_build_sym_cells(GridView g, CellRows cr, const int *__restrict__ act,
                                                        const double *__restrict__ X, const double *__restrict__ Y,
                                                        const double *__restrict__ Z, const double *__restrict__ H,
                                                        int *__restrict__ ell, long ld, int cap, int *__restrict__ cnt,
                                                        int *__restrict__ overflow)
{
    __shared__ double sx[64], sy[64], sz[64], sh[64];
    __shared__ int sj[64];
    __shared__ RowStencil rs;
    const long cb = xcd_block();
    if (cb >= (cr.ncells_dev ? *cr.ncells_dev : cr.ncells)) return;
    int c, a0, a1, nst;
    const int4 *sten;
    if (cr.gsten) {
        c = cr.gcode[cb]; a0 = cr.gstart[cb]; a1 = cr.gstart[cb + 1]; nst = cr.gnst[cb]; sten = cr.gsten + (size_t)27 * cb;
        if (cr.bt.b) {
            DevParams P{};
            resolve_box(cr.bt, act[cr.grow ? cr.grow[a0] : a0], P, g);
        } else {
            g.level = cr.glev[cb];
            g.cell = g.box / (double)(1 << g.level);
        }
    } else {
        c = cr.cells[cb]; a0 = cr.ca0[c]; a1 = cr.ca1[c]; nst = g.nsten[c]; sten = g.sten + (size_t)c * 27;
    }
    // the search over many levels adds part B (its groups are cells of the rows' own levels)
    const bool wide = cr.gsten && !cr.bt.b && g.wide;
    const int t = threadIdx.x;
    for (int base = a0; base < a1; base += 64) {
        const bool on = base + t < a1;
        const int r = on ? (cr.grow ? cr.grow[base + t] : base + t) : 0;
        int i = -1;
        double xi = 0, yi = 0, zi = 0, hi = 0;
        if (on) { i = act[r]; xi = X[i]; yi = Y[i]; zi = Z[i]; hi = H[i]; }
        const LeafBox lb = leaf_box(g, c, on, xi, yi, zi);
        const double hleaf = wave_max(on ? hi : 0.0);
        int n = 0;
        for (int q = 0; q < nst; q++) {
            const int4 e = sten[q];
            for (int ts = e.x; ts < e.y; ts += 64) {
                const int j = ts + t;
                bool keep = false;
                double xj = 0, yj = 0, zj = 0, hj = 0;
                if (j < e.y) {
                    xj = X[j]; yj = Y[j]; zj = Z[j]; hj = H[j];
                    keep = leaf_reach(g, lb, xj, yj, zj, fmax(hleaf, hj));
                }
                int nt;
                const int pos = kept_rank(keep, nt);
                __syncthreads();
                if (keep) { sx[pos] = xj; sy[pos] = yj; sz[pos] = zj; sh[pos] = hj; sj[pos] = j; }
                __syncthreads();
                if (!on) continue;
                for (int k = 0; k < nt; k++) {
                    const int jj = sj[k];
                    if (jj == i) continue;
                    if (accept_pair<true>(xi, yi, zi, sx[k], sy[k], sz[k], hi, sh[k], g.box, g.half)) {
                        if (n < cap) ell[ell_at(ld, r, n)] = jj;
                        n++;
                    }
                }
            }
        }
        if (wide) {
            // Part B of the search over many levels (as k_build_sym_wide): the particles of each
            // coarser level L in the 27 cells of level L around the group's cell that lie outside
            // the 27 cells of the group's level, each accepted one inserted in ascending place
            const int nci = 1 << g.level;
            const int hx = compact10((uint32_t)c >> 2), hy = compact10((uint32_t)c >> 1), hz = compact10((uint32_t)c);
            const int si = 3 * (kFineLevel - g.level);
            const WideGrid &W = *g.wide;
            for (int L = 0; L < g.level; L++) {
                if (W.lofs[L + 1] == W.lofs[L]) continue;  // no particle of this level
                __syncthreads();  // every lane is done with rs before it is built again
                row_stencil_wide<true>(g, W, L, (uint32_t)c >> (3 * (g.level - L)), 0.0, 0.0, 0.0, -1.0, rs);
                const int total = rs.pref[rs.nst];
                for (int base2 = 0; base2 < total; base2 += 64) {
                    const int v = base2 + t;
                    bool keep = false;
                    int j = -1;
                    double xj = 0, yj = 0, zj = 0, hj = 0;
                    if (v < total) {
                        j = W.lidx[row_candidate(rs, v)];
                        const uint32_t cj = W.key[j] >> si;
                        const bool in_a = periodic_adjacent((int)compact10(cj >> 2), hx, nci) &&
                                          periodic_adjacent((int)compact10(cj >> 1), hy, nci) && periodic_adjacent((int)compact10(cj), hz, nci);
                        if (!in_a) {
                            xj = X[j]; yj = Y[j]; zj = Z[j]; hj = H[j];
                            keep = leaf_reach(g, lb, xj, yj, zj, fmax(hleaf, hj));
                        }
                    }
                    int nt;
                    const int pos = kept_rank(keep, nt);
                    __syncthreads();
                    if (keep) { sx[pos] = xj; sy[pos] = yj; sz[pos] = zj; sh[pos] = hj; sj[pos] = j; }
                    __syncthreads();
                    if (!on) continue;
                    for (int k = 0; k < nt; k++) {
                        const int jj = sj[k];
                        if (jj == i || !accept_pair<true>(xi, yi, zi, sx[k], sy[k], sz[k], hi, sh[k], g.box, g.half)) continue;
                        if (n < cap) {
                            int q = n;
                            while (q > 0 && ell[ell_at(ld, r, q - 1)] > jj) { ell[ell_at(ld, r, q)] = ell[ell_at(ld, r, q - 1)]; q--; }
                            ell[ell_at(ld, r, q)] = jj;
                        }
                        n++;
                    }
                }
            }
        }
        if (on) {
            cnt[r] = min(n, cap);  // readers stay inside the list; the overflow carries the full length
            if (n > cap) atomicMax(overflow, n);
        }
    }
}
#endif


// ---------------------------------------------------------------------------------------
// groups of rows sharing one cell of one grid, for the cell walks of the search over many
// levels (the cell of the level of each row's radius) and of batches (the cell of the row's
// box): key = the level's or the box's first table entry + the cell code
// ---------------------------------------------------------------------------------------
#ifdef GZ_DEFINE_COMMON_KERNELS
__global__ void __launch_bounds__(256) k_row_group_key(int nrows, const int *__restrict__ act, const uint8_t *__restrict__ need,
                                                        const uint32_t *__restrict__ key, const double *__restrict__ hslot,
                                                        const double *__restrict__ H, double box, double margin,
                                                        const WideGrid *__restrict__ W, BoxTable bt, uint32_t *__restrict__ gkey,
                                                        int *__restrict__ val)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    if (val) val[r] = r;
    if (need && !need[r]) { gkey[r] = 0xffffffffu; return; }  // not in this pass: after every group
    const int i = act[r];
    if (bt.b) {
        const BoxDev &B = bt.b[bt.BOX[i]];
        gkey[r] = (uint32_t)(B.cell_off + (key[i] >> B.shift));
    } else {
        const double h = hslot ? hslot[r] : H[i];
        const int L = wide_level(box, W->lmax, h + 2.0 * margin);
        gkey[r] = (uint32_t)(W->loff[L] + (key[i] >> (3 * (kFineLevel - L))));
    }
}

// group heads of the sorted keys (flag[nrows] = 0 for the scan)
__global__ void __launch_bounds__(256) k_group_heads(int nrows, const uint32_t *__restrict__ skey, int *__restrict__ flag)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > nrows) return;
    flag[p] = (p < nrows && skey[p] != 0xffffffffu && (p == 0 || skey[p - 1] != skey[p])) ? 1 : 0;
}

// group starts (positions in the sorted rows), keys and count; gstart[count] = the number of
// rows in groups
__global__ void __launch_bounds__(256) k_group_list(int nrows, const uint32_t *__restrict__ skey, const int *__restrict__ flag,
                                                     const int *__restrict__ scan, int *__restrict__ gstart, uint32_t *__restrict__ gkey,
                                                     int *__restrict__ ngroups)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nrows) return;
    if (flag[p]) { gstart[scan[p]] = p; gkey[scan[p]] = skey[p]; }
    const bool last = skey[p] != 0xffffffffu && (p + 1 == nrows || skey[p + 1] == 0xffffffffu);
    if (last) { gstart[scan[p] + flag[p]] = p + 1; *ngroups = scan[p] + flag[p]; }
    if (p == 0 && skey[0] == 0xffffffffu) { gstart[0] = 0; *ngroups = 0; }
}

// the stencil of every group (one wavefront each): the 27 cells around its cell at its level,
// duplicates dropped, ascending code (ascending start)
__global__ void __launch_bounds__(64) k_group_stencil(GridView g, int ngroups, const uint32_t *__restrict__ gkey,
                                                       const int *__restrict__ gstart, const int *__restrict__ grow,
                                                       const int *__restrict__ act, BoxTable bt, int *__restrict__ glev,
                                                       int *__restrict__ gcode, int *__restrict__ gnst, int4 *__restrict__ gsten)
{
    __shared__ RowStencil rs;
    const long gb = blockIdx.x;
    if (gb >= ngroups) return;
    const uint32_t k = gkey[gb];
    int L, c;
    if (bt.b) {
        DevParams P{};
        resolve_box(bt, act[grow ? grow[gstart[gb]] : gstart[gb]], P, g);
        L = g.level;
        c = (int)(k - g.cell_off);
        row_stencil<false>(g, (uint32_t)c, 0.0, 0.0, 0.0, 1e300, rs);
    } else {
        const WideGrid &W = *g.wide;
        L = 0;
        while (L < W.lmax && (long long)k >= W.loff[L + 1]) L++;
        c = (int)(k - W.loff[L]);
        row_stencil_wide<false>(g, W, L, (uint32_t)c, 0.0, 0.0, 0.0, -1.0, rs);
    }
    const int lane = threadIdx.x;
    if (lane < rs.nst) gsten[(size_t)27 * gb + lane] = make_int4(rs.start[lane], rs.start[lane] + rs.len[lane], 0, 0);
    if (lane == 0) { glev[gb] = L; gcode[gb] = c; gnst[gb] = rs.nst; }
}
#endif

// gradients of few rows: the terms of each list entry in LDS, accumulated in list order by lane 0
template <int KERN, typename R, int TPB>
__global__ void __launch_bounds__(TPB) k_gradients_rows(DevParams P, int nrows, const int *__restrict__ act, const double *__restrict__ X,
                                                       const double *__restrict__ Y, const double *__restrict__ Z,
                                                       const double *__restrict__ H, const GradRec<R> *__restrict__ rec, GradSrc src, const double *__restrict__ T,
                                                       const double *__restrict__ CN, long n, const int *__restrict__ ell, long ld,
                                                       const int *__restrict__ cnt, double *__restrict__ GR, double *__restrict__ GP,
                                                       double *__restrict__ GV, BoxTable bt, int *__restrict__ force_flags)
{
    // the accepted entries' values packed in list order (padded rows: no bank conflicts)
    __shared__ R sd[5][TPB + 1], sw[3][TPB + 1], srr[TPB + 1];
    __shared__ unsigned long long smask[TPB / 64];
    const long r = xcd_block();
    // the coming force pass's error and wake-up flags (its records are made in its own kernel)
    if (r == 0 && threadIdx.x == 0 && force_flags) force_flags_clear(force_flags);
    if (r >= nrows) return;
    const int lane = threadIdx.x;  // thread of the workgroup
    const int i = act[r];
    {
        GridView g{};
        resolve_box(bt, i, P, g);
    }
    R G[3][5];
    R mx[5], mn[5];
    R maxd = 0;
    R acc = 0;  // the running quantity this lane keeps
    // records packed by an earlier launch, or made here from the state (few rows among many
    // particles: no packing launch)
    const GradRec<R> I = src.H ? make_grad_rec<R>(i, src, X, Y, Z, bt) : rec[i];
    const bool enabled = (I.m > 0 && I.rho > 0);
    const R hi = I.h;
    if (enabled) {
        const double xi = X[i], yi = Y[i], zi = Z[i];
        // the row's radius in FP64 (the kernel's edge factor, pair_kernel)
        const double hid = std::is_same<R, double>::value ? (double)hi : H[i];
        const double h2i = hid * hid;
        const int c = cnt[r];
        for (int base = 0; base < c; base += TPB) {
            const int q = base + lane;
            bool ok = false;
            R w3[3] = {0, 0, 0}, d5[5], rr = 0;
            if (q < c) {
                const int j = ell[ell_at(ld, r, q)];
                const GradRec<R> J = src.H ? make_grad_rec<R>(j, src, X, Y, Z, bt) : rec[j];
                if ((J.m > 0) && (J.rho > 0) && (J.h > 0)) {
                    R dp[3];
                    record_dp<R>(P, I, J, xi, yi, zi, X, Y, Z, j, dp[0], dp[1], dp[2]);
                    const R r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
                    const double r2d = pair_r2d<R>(dp[0], dp[1], dp[2], r2);
                    if ((r2 > 0) && !((r2d >= h2i) && (r2d >= (double)J.h * (double)J.h))) {
                        rr = sqrt(r2);
                        const double rd = std::is_same<R, double>::value ? 0.0 : sqrt(r2d);
                        R wk = 0, dwk = 0;
                        ok = true;
                        // an entry beyond the row's own radius (inside the neighbour's) takes part
                        // in the extrema only: its weights stay +0, and its products (+-0) leave
                        // the gradient sums, which start at +0 and never reach -0, unchanged
                        if (std::is_same<R, double>::value ? (rr < hi) : (rd < hid)) {
                            pair_kernel<KERN, -1, R>(rr, rd, hid, wk, dwk);
                            for (int kk = 0; kk < 3; kk++) w3[kk] = -wk * dp[kk];
                        }
                        d5[0] = (R)(J.rho - I.rho);
                        d5[1] = (R)(J.p - I.p);
                        d5[2] = (R)(J.v0 - I.v0);
                        d5[3] = (R)(J.v1 - I.v1);
                        d5[4] = (R)(J.v2 - I.v2);
                    }
                }
            }
            // the accepted entries packed in list order; threads 0-14 then keep G[kk][b]
            // (kk = lane/5, b = lane%5), 15-19 the minima, 20-24 the maxima, 25 the largest
            // distance (with mn <= 0 <= mx the reference's else-if is two ifs)
            int nok;
            const int pos = flagged_rank<TPB>(ok, smask, nok);
            if (ok) {
                for (int kk = 0; kk < 3; kk++) sw[kk][pos] = w3[kk];
                for (int b = 0; b < 5; b++) sd[b][pos] = d5[b];
                srr[pos] = rr;
            }
            __syncthreads();
            if (lane < 15) {
                const R *a = sw[lane / 5], *b = sd[lane % 5];
                int k = 0;
                for (; k + 4 <= nok; k += 4) {
                    const R p0 = rounded(a[k] * b[k]), p1 = rounded(a[k + 1] * b[k + 1]);
                    const R p2 = rounded(a[k + 2] * b[k + 2]), p3 = rounded(a[k + 3] * b[k + 3]);
                    acc = acc + p0;
                    acc = acc + p1;
                    acc = acc + p2;
                    acc = acc + p3;
                }
                for (; k < nok; k++) acc = acc + rounded(a[k] * b[k]);
            } else if (lane < 26) {
                // minima (15-19, as maxima of the negated values: exact), maxima, distance
                const R *a = lane < 25 ? sd[(lane - 15) % 5] : srr;
                const bool mn = lane < 20;
                R m = mn ? -acc : acc;
                for (int k = 0; k < nok; k++) {
                    const R d = mn ? -a[k] : a[k];
                    if (d > m) m = d;
                }
                acc = mn ? -m : m;
            }
            __syncthreads();
        }
    }
    __shared__ R res[26];
    if (lane < 26) res[lane] = acc;
    __syncthreads();
    if (lane == 0) {
        for (int kk = 0; kk < 3; kk++)
            for (int b = 0; b < 5; b++) G[kk][b] = res[kk * 5 + b];
        for (int b = 0; b < 5; b++) { mn[b] = res[15 + b]; mx[b] = res[20 + b]; }
        maxd = res[25];
        gradients_tail<R>(I, enabled, G, mx, mn, maxd, T, CN, n, i, GR, GP, GV);
    }
}

// force of few rows: each lane evaluates one list entry (hydro_accumulate_neighbor), lane 0
// accumulates in list order, then the writeback and final operations
template <int KERN, typename R, bool STATS, int TPB, bool MAKE>
__global__ void __launch_bounds__(TPB) k_force_rows(DevParams P, Gam<R> G, int nrows, const int *__restrict__ act, ForceView F,
                                                   const Cell<R> *__restrict__ rec, const int *__restrict__ ell, long ld,
                                                   const int *__restrict__ cnt, const uint8_t *__restrict__ active,
                                                   const int *__restrict__ timebin, const double *__restrict__ msv_in,
                                                   int *__restrict__ wakeup, double *__restrict__ ACC, double *__restrict__ DTU,
                                                   double *__restrict__ MSV, int *__restrict__ err, int *__restrict__ need_wakeup,
                                                   unsigned long long *__restrict__ pairs, BoxTable bt, const int *__restrict__ gate)
{
    __shared__ R sv[5][TPB + 1];  // the entries' fluxes packed in list order (padded: no bank conflicts)
    __shared__ unsigned long long smask[TPB / 64];
    const long r = xcd_block();
    if (r >= nrows) return;
    if (gate && *gate) return;  // the density round behind this launch did not stand (its gate word)
    const int lane = threadIdx.x;  // thread of the workgroup
    const int i = act[r];
    {
        GridView g{};
        resolve_box(bt, i, P, g);
    }
    const double mi = F.M[i], rhoi = F.RHO[i];
    constexpr bool make = MAKE;
    double A0 = 0, A1 = 0, A2 = 0, dE = 0, msv = 0;
    const double sound_i = (rhoi > 0) ? sqrt(P.gamma * F.PR[i] / rhoi) : sqrt(P.gamma * (P.gamma - 1.) * F.U[i]);
    const bool enabled = (mi > 0 && rhoi > 0);
    const int c = cnt[r];
    if (lane == 0 && wakeup) wakeup[i] = 0;  // hydro_force_initial_operations_preloop
    if (enabled) {
        const Cell<R> I = make ? force_rec_of<R>(i, P.gamma, F, P.box, bt) : rec[i];
        const double xi = F.X[i], yi = F.Y[i], zi = F.Z[i];
        const int tbi = timebin ? timebin[i] : 0;
        // lanes 0-3 keep the momentum and energy sums, lane 4 the signal speed: max(c_s, vsig)
        // over the entries (the reference raises it to c_s before each entry, a no-op after the first)
        R acc = 0;       // lane 4
        CSum<R> sum;     // lanes 0-3
        if (lane == 4 && c > 0) acc = (acc < I.cs) ? I.cs : acc;
        for (int base = 0; base < c; base += TPB) {
            const int q = base + lane;
            int rc = 0;
            R f0, f1, f2, fp, vts;
            if (q < c) {
                const int j = ell[ell_at(ld, r, q)];
                const Cell<R> J = make ? force_rec_of<R>(j, P.gamma, F, P.box, bt) : rec[j];
                R dp0, dp1, dp2;
                record_dp<R>(P, I, J, xi, yi, zi, F.X, F.Y, F.Z, j, dp0, dp1, dp2);
                const R r2 = dp0 * dp0 + dp1 * dp1 + dp2 * dp2;
                unsigned rung;
                rc = force_pair<KERN, R>(P, G, I, J, dp0, dp1, dp2, r2, f0, f1, f2, fp, vts, rung);
                if constexpr (STATS) count_rungs(F.rungs, rung);
                if (rc < 0) {
                    atomicMax(err, 1234);
                    if (atomicCAS(err + 10, -1, i) == -1) err[11] = j;
                }
                if (rc > 0) {
                    if (active && !active[j] && wakeup) {
                        if ((double)vts > kWakeup * msv_in[j]) {
                            atomicMax(&wakeup[j], tbi + 1);
                            if (bt.b) bt.b[bt.BOX[i]].need_wakeup = 1;
                            else *need_wakeup = 1;
                        }
                    }
                }
            }
            // the entries with a flux packed in list order; lanes 0-4 then take them in turn
            int nok;
            const int pos = flagged_rank<TPB>(rc > 0, smask, nok);
            if (rc > 0) {
                sv[0][pos] = f0; sv[1][pos] = f1; sv[2][pos] = f2; sv[3][pos] = fp; sv[4][pos] = vts;
            }
            __syncthreads();
            if (lane < 4) {
                const R *a = sv[lane];
                int k = 0;
                for (; k + 4 <= nok; k += 4) {
                    const R x0 = a[k], x1 = a[k + 1], x2 = a[k + 2], x3 = a[k + 3];
                    sum.add(x0);
                    sum.add(x1);
                    sum.add(x2);
                    sum.add(x3);
                }
                for (; k < nok; k++) sum.add(a[k]);
            } else if (lane == 4) {
                for (int k = 0; k < nok; k++)
                    if (sv[4][k] > acc) acc = sv[4][k];
            }
            __syncthreads();
        }
        __shared__ double res[5];
        if (lane < 4) res[lane] = sum.value();
        if (lane == 4) res[lane] = (double)acc;
        __syncthreads();
        A0 = res[0]; A1 = res[1]; A2 = res[2]; dE = res[3]; msv = res[4];
    }
    if (lane == 0) {
        force_tail(P, F, i, enabled, sound_i, A0, A1, A2, dE, msv, ACC, DTU, MSV);
        // the pairs of the row last: no load of the row waits behind the atomic
        if (c && pairs) atomicAdd(pairs, (unsigned long long)c);
    }
}

}  // namespace gz
