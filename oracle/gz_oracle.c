/*
 * gz_oracle.c -- CPU FP64 restatement of GIZMO's meshless-finite-mass (MFM) hydro
 * neighbour-loop pipeline: neighbour search, density() with the kernel-radius
 * iteration, hydro_gradient_calc() and hydro_force().
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it, and only as the checker or the timed CPU reference: the product path
 * (gizmo_amd/) never links, loads or calls it.
 *
 * Every function below restates one reference function; the citation is on the
 * function.  Paths are relative to the reference source tree.  The restatement
 * keeps the reference's order of floating-point operations so that a GPU path
 * written to the same order can be compared bit-for-bit (both sides are compiled
 * with -ffp-contract=off and without fast-math; see src/declarations/typedefs.h:27-29
 * where the reference forbids fast-math).
 *
 * Build configuration restated (the flags that the reference's Config.sh would set
 * for the hot path): 3D, BOX_PERIODIC (cubic box of side `box`), HYDRO_MESHLESS_FINITE_MASS,
 * SELFGRAVITY_OFF, no MAGNETIC / COOLING / GALSF / EOS_GENERAL, SLOPE_LIMITER_TOLERANCE = 1,
 * non-comoving (cf_atime = 1).  The adiabatic index is a runtime parameter here
 * (GAMMA_DEFAULT is a compile-time constant in the reference).
 *
 * Parity status of each function is stated in the header comment of the function
 * and summarised in oracle/README.md.  The timestep loop around the hot path (kicks,
 * drift, timebins, wake-ups) is restated in gz_oracle_time.c.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "gz_oracle.h"

/* src/declarations/constants.h:219-225,340,445; src/declarations/constants.h:11-17 */
#define OR_MAX_REAL 1e56
#define OR_MIN_REAL 1e-56
#define OR_CN_DANGER 1.0e3
#define OR_MAXITER 150
#define OR_VOLNORM 4.188790204786
#define OR_WAKEUP 4.1
#define OR_NDIMS 3
#define OR_NACC 17 /* density accumulators: the 13 of the pass and GradH_numer[3], GradH_denom */

/* KERNEL_AWPMHD_FD_ALPHA, the first-derivative partition-of-unity coefficient of each kernel
   (src/mesh/kernel.h:185-206) */
double orc_fd_alpha(int kernel)
{
    switch (kernel) {
    case 3: return 3.0 / 16.0;
    case 4: return 23.0 / 150.0;
    case 5: return 7.0 / 54.0;
    case 6: return 1.0 / 6.0;
    case 7: return 5.0 / 39.0;
    case 9: return 5.0 / 48.0;
    default: return 3.0 / 16.0;
    }
}

#define DMAX(a, b) (((a) > (b)) ? (a) : (b))
#define DMIN(a, b) (((a) < (b)) ? (a) : (b))

/* ------------------------------------------------------------------------- */
/* kernel: src/mesh/kernel.h:245-260 (kernel_hinv), :268-440 (kernel_main);     */
/* norms 3D: cubic 8/pi (:69), Wendland C2 21/(2 pi) (:123)                     */
/* ------------------------------------------------------------------------- */
static inline void or_kernel_hinv(double h, double *hinv, double *hinv3, double *hinv4)
{
    *hinv = 1.0 / h;
    *hinv3 = *hinv * *hinv * *hinv;
    *hinv4 = *hinv3 * *hinv;
}

static inline void or_kernel_main(int kernel, double u, double hinv3, double hinv4, double *wk, double *dwk, int mode)
{
    if (u >= 1) { *wk = 0; *dwk = 0; return; }
    double norm = 0;
    if (kernel == 3) {
        norm = (8.0 / M_PI);
        if (u < 0.5) {
            if (mode >= 0) *dwk = u * (18.0 * u - 12.0);
            if (mode <= 0) *wk = (1.0 + 6.0 * (u - 1.0) * u * u);
        } else {
            double t1 = (1.0 - u);
            double t2 = t1 * t1;
            if (mode >= 0) *dwk = -6.0 * t2;
            if (mode <= 0) *wk = 2.0 * t2 * t1;
        }
    } else { /* 6: Wendland C2 (3D form) */
        norm = (21.0 / (2.0 * M_PI));
        double t1 = (1 - u);
        double t3 = t1 * t1 * t1;
        if (mode >= 0) *dwk = -20.0 * u * t3;
        if (mode <= 0) *wk = t3 * t1 * (1.0 + 4.0 * u);
    }
    if (mode >= 0) { *dwk *= norm * hinv4; }
    if (mode <= 0) { *wk *= norm * hinv3; }
}

int orc_kernel(int kernel, double r, double h, double *wk, double *dwk)
{
    if (kernel != 3 && kernel != 6) return -1;
    double hinv, hinv3, hinv4;
    or_kernel_hinv(h, &hinv, &hinv3, &hinv4);
    double u = r * hinv;
    *wk = 0; *dwk = 0;
    or_kernel_main(kernel, u, hinv3, hinv4, wk, dwk, 0);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* periodic nearest image: src/declarations/macros.h:67-106 (TMP_WRAP_*_S,      */
/* NEAREST_XYZ: Y, then X, then Z) and the absolute form NGB_PERIODIC_BOX_LONG_* */
/* used by the neighbour predicate (src/mesh/sfc_tiles_functions.h:83-102).     */
/* ------------------------------------------------------------------------- */
static inline double or_wrap(double x, double box, double half)
{
    return (x > half) ? (x - box) : ((x < -half) ? (x + box) : x);
}

static inline double or_wrap_abs(double x, double box, double half)
{
    double t = fabs(x);
    return (t > half) ? (box - t) : t;
}

/* the lazy drift of the timestep loop: the passes find their pairs on the predicted positions
   and radii and drift a particle to the current time before they read it */
static const orc_lazy *g_lazy = NULL;
void orc_set_lazy(const orc_lazy *view) { g_lazy = view; }
static inline const double *or_pos(const orc_state *s) { return g_lazy ? g_lazy->pos : s->pos; }
static inline const double *or_hpred(const orc_state *s) { return g_lazy ? g_lazy->h : s->h; }
static inline void or_read(long j) { if (g_lazy) g_lazy->ensure(j); }

/* ------------------------------------------------------------------------- */
/* uniform cell grid, for the neighbour candidate sets.  Not a reference       */
/* function: the reference walks a BVH over 64-member Morton tiles               */
/* (src/mesh/sfc_tiles_functions.h:203-258); the accepted set is what matters   */
/* and is fixed by the exact predicate below.  Candidates of a row are visited   */
/* in ascending particle index, the canonical order both the oracle and the     */
/* device path sum in.                                                           */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nc;          /* cells per dimension */
    double cell;     /* cell side */
    long *start;     /* nc^3 + 1 offsets into idx */
    int *idx;        /* particle indices, grouped by cell, ascending inside each cell */
} or_grid;

static int or_grid_build(or_grid *g, long n, const double *pos, double box, double rmax)
{
    int nc = (int)floor(box / rmax);
    if (nc < 1) nc = 1;
    if (nc > 256) nc = 256;
    g->nc = nc;
    g->cell = box / nc;
    long ncell = (long)nc * nc * nc;
    g->start = (long *)calloc((size_t)ncell + 1, sizeof(long));
    g->idx = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    int *cid = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    if (!g->start || !g->idx || !cid) { free(cid); return -1; }
    for (long i = 0; i < n; i++) {
        int c[3];
        for (int k = 0; k < 3; k++) {
            double x = pos[3 * i + k];
            x = x - box * floor(x / box);
            int ck = (int)(x / g->cell);
            if (ck < 0) ck = 0;
            if (ck >= nc) ck = nc - 1;
            c[k] = ck;
        }
        cid[i] = (c[0] * nc + c[1]) * nc + c[2];
        g->start[cid[i] + 1]++;
    }
    for (long c = 0; c < ncell; c++) g->start[c + 1] += g->start[c];
    long *fill = (long *)malloc((size_t)ncell * sizeof(long));
    memcpy(fill, g->start, (size_t)ncell * sizeof(long));
    for (long i = 0; i < n; i++) g->idx[fill[cid[i]]++] = (int)i;
    free(fill);
    free(cid);
    return 0;
}

static void or_grid_free(or_grid *g)
{
    free(g->start);
    free(g->idx);
}

static int or_cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* candidates of position p: every particle in the (deduplicated) 27 cells around p, in
   cell order (not sorted).  buf must hold n ints.  Returns the count.  Callers keep the
   accepted ones and sort those (or_sort_int): every row is visited in ascending index. */
static long or_grid_candidates(const or_grid *g, const double *p, double box, int *buf)
{
    int nc = g->nc;
    int c[3];
    for (int k = 0; k < 3; k++) {
        double x = p[k] - box * floor(p[k] / box);
        int ck = (int)(x / g->cell);
        if (ck < 0) ck = 0;
        if (ck >= nc) ck = nc - 1;
        c[k] = ck;
    }
    int cells[27];
    int ncells = 0;
    for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) {
                int a = (c[0] + dx + nc) % nc, b = (c[1] + dy + nc) % nc, e = (c[2] + dz + nc) % nc;
                int id = (a * nc + b) * nc + e;
                int dup = 0;
                for (int q = 0; q < ncells; q++) if (cells[q] == id) { dup = 1; break; }
                if (!dup) cells[ncells++] = id;
            }
    long m = 0;
    for (int q = 0; q < ncells; q++)
        for (long s = g->start[cells[q]]; s < g->start[cells[q] + 1]; s++) buf[m++] = g->idx[s];
    return m;
}

/* ascending order of a short accepted list (insertion sort; qsort beyond 64 entries) */
static void or_sort_int(int *a, long m)
{
    if (m > 64) { qsort(a, (size_t)m, sizeof(int), or_cmp_int); return; }
    for (long k = 1; k < m; k++) {
        int t = a[k];
        long q = k - 1;
        while (q >= 0 && a[q] > t) { a[q + 1] = a[q]; q--; }
        a[q + 1] = t;
    }
}

/* exact pair predicate: src/mesh/sfc_tiles_functions.h:83-102 (frame.exact branch):
   ONEWAY r2 < h_i^2, SYMMETRIC r2 < max(h_i,h_j)^2, r2 from the absolute wrapped components. */
static inline int or_accept(const double *pi, const double *pj, double h_i, double h_j, int symmetric, double box, double half)
{
    double dx_raw = pi[0] - pj[0];
    double dy_raw = pi[1] - pj[1];
    double dz_raw = pi[2] - pj[2];
    double adx = or_wrap_abs(dx_raw, box, half);
    double ady = or_wrap_abs(dy_raw, box, half);
    double adz = or_wrap_abs(dz_raw, box, half);
    double r2s;
    if (!symmetric) {
        r2s = h_i * h_i;
    } else {
        double hm = (h_i > h_j) ? h_i : h_j;
        r2s = hm * hm;
    }
    if (adx > h_i && (!symmetric || adx * adx > r2s)) return 0;
    double r2 = adx * adx + ady * ady + adz * adz;
    return r2 < r2s;
}

static double or_max_h(long n, const double *h, const unsigned char *rowmask)
{
    double m = 0;
    for (long i = 0; i < n; i++) if (!rowmask || rowmask[i]) if (h[i] > m) m = h[i];
    return m;
}

/* CSR neighbour lists (gpu_ngb_list_build, src/mesh/gpu_neighbor_list.cc:1950):
   rows = particles with rowmask set (all when NULL), in index order; each row lists the
   accepted j in ascending order; SYMMETRIC rows leave out the row itself.  Returns the
   total number of entries, or -1.  `nbr` may be NULL (count only); otherwise it holds cap ints. */
long orc_neighbors(const orc_params *par, long n, const double *pos, const double *h, int symmetric,
                   const unsigned char *rowmask, long *offsets, int *nbr, long cap)
{
    double box = par->box, half = 0.5 * par->box;
    double rmax = symmetric ? or_max_h(n, h, NULL) : or_max_h(n, h, rowmask);
    if (!(rmax > 0) || rmax > half) return -1;
    or_grid g;
    if (or_grid_build(&g, n, pos, box, rmax)) return -1;
    long *cnt = (long *)calloc((size_t)n + 1, sizeof(long));
    int err = 0;
#pragma omp parallel
    {
        int *buf = (int *)malloc((size_t)n * sizeof(int));
        if (!buf) {
#pragma omp atomic write
            err = 1;
        }
#pragma omp for schedule(dynamic, 256)
        for (long i = 0; i < n; i++) {
            if (!buf || (rowmask && !rowmask[i])) continue;
            long m = or_grid_candidates(&g, &pos[3 * i], box, buf);
            long c = 0;
            for (long q = 0; q < m; q++) {
                int j = buf[q];
                if (symmetric && j == i) continue;
                if (or_accept(&pos[3 * i], &pos[3 * j], h[i], h[j], symmetric, box, half)) c++;
            }
            cnt[i] = c;
        }
        free(buf);
    }
    if (err) { free(cnt); or_grid_free(&g); return -1; }
    long total = 0;
    for (long i = 0; i < n; i++) { if (offsets) offsets[i] = total; total += cnt[i]; }
    if (offsets) offsets[n] = total;
    if (nbr && total <= cap && offsets) {
#pragma omp parallel
        {
            int *buf = (int *)malloc((size_t)n * sizeof(int));
#pragma omp for schedule(dynamic, 256)
            for (long i = 0; i < n; i++) {
                if (!buf || (rowmask && !rowmask[i])) continue;
                long m = or_grid_candidates(&g, &pos[3 * i], box, buf);
                long o = offsets[i];
                for (long q = 0; q < m; q++) {
                    int j = buf[q];
                    if (symmetric && j == i) continue;
                    if (or_accept(&pos[3 * i], &pos[3 * j], h[i], h[j], symmetric, box, half)) nbr[o++] = j;
                }
                or_sort_int(&nbr[offsets[i]], o - offsets[i]);
            }
            free(buf);
        }
    }
    free(cnt);
    or_grid_free(&g);
    return total;
}

/* ------------------------------------------------------------------------- */
/* 3x3 inverse and condition number: matrix_invert_ndims, src/core/proto.h:75-95 */
/* with Mat3::invert / frobenius_norm_sq, src/math_types/mat3.h:20-64.          */
/* T, Tinv: row-major 3x3.                                                       */
/* ------------------------------------------------------------------------- */
static double or_frob2(const double *m)
{
    double r0 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    double r1 = m[3] * m[3] + m[4] * m[4] + m[5] * m[5];
    double r2 = m[6] * m[6] + m[7] * m[7] + m[8] * m[8];
    return r0 + r1 + r2;
}

double orc_matrix_invert(const double *T, double *Tinv)
{
    double FrobNorm = or_frob2(T);
    for (int k = 0; k < 9; k++) Tinv[k] = 0;
#define R(a, b) T[3 * (a) + (b)]
    double d = R(0, 0) * (R(1, 1) * R(2, 2) - R(1, 2) * R(2, 1))
             - R(0, 1) * (R(1, 0) * R(2, 2) - R(1, 2) * R(2, 0))
             + R(0, 2) * (R(1, 0) * R(2, 1) - R(1, 1) * R(2, 0));
    if (!(d == 0 || d != d)) {
        double inv_d = 1.0 / d;
        Tinv[0] = (R(1, 1) * R(2, 2) - R(1, 2) * R(2, 1)) * inv_d;
        Tinv[1] = (R(0, 2) * R(2, 1) - R(0, 1) * R(2, 2)) * inv_d;
        Tinv[2] = (R(0, 1) * R(1, 2) - R(0, 2) * R(1, 1)) * inv_d;
        Tinv[3] = (R(1, 2) * R(2, 0) - R(1, 0) * R(2, 2)) * inv_d;
        Tinv[4] = (R(0, 0) * R(2, 2) - R(0, 2) * R(2, 0)) * inv_d;
        Tinv[5] = (R(0, 2) * R(1, 0) - R(0, 0) * R(1, 2)) * inv_d;
        Tinv[6] = (R(1, 0) * R(2, 1) - R(1, 1) * R(2, 0)) * inv_d;
        Tinv[7] = (R(0, 1) * R(2, 0) - R(0, 0) * R(2, 1)) * inv_d;
        Tinv[8] = (R(0, 0) * R(1, 1) - R(0, 1) * R(1, 0)) * inv_d;
    }
#undef R
    double FrobNorm_inv = or_frob2(Tinv);
    double cn = sqrt(FrobNorm * FrobNorm_inv) / OR_NDIMS;
    return DMAX(cn, 1);
}

/* symmetric tensor storage [xx,yy,zz,xy,yz,xz]: src/math_types/symmetric_tensor2.h:11-85 */
static inline void or_sym_to_full(const double *s, double *m)
{
    m[0] = s[0]; m[1] = s[3]; m[2] = s[5];
    m[3] = s[3]; m[4] = s[1]; m[5] = s[4];
    m[6] = s[5]; m[7] = s[4]; m[8] = s[2];
}

static inline void or_sym_matvec(const double *s, const double *v, double *out)
{
    out[0] = s[0] * v[0] + s[3] * v[1] + s[5] * v[2];
    out[1] = s[3] * v[0] + s[1] * v[1] + s[4] * v[2];
    out[2] = s[5] * v[0] + s[4] * v[1] + s[2] * v[2];
}

/* NV_T regularised inversion shared by after_iter_body (src/hydro/density_loop.h:582-593)
   and density_finalize_post_runner (src/hydro/density_loop.cc:472-484).  Returns CN and
   the full inverse of the normalised matrix in Tinv. */
static double or_nvt_regularised_inverse(const double *nvt_raw6, double normalizer, double *Tinv)
{
    double full[9], loc[9];
    or_sym_to_full(nvt_raw6, full);
    for (int k = 0; k < 9; k++) loc[k] = full[k] / normalizer;
    const double thr = 10.0 * OR_CN_DANGER;
    const double trace_initial = loc[0] + loc[4] + loc[8];
    double add = 1.05 * (trace_initial / OR_NDIMS) / thr;
    double cn;
    while (1) {
        cn = orc_matrix_invert(loc, Tinv);
        if (cn < thr) break;
        loc[0] += add; loc[4] += add; loc[8] += add;
        add *= 1.2;
    }
    return cn;
}

/* FaceClosureError: src/hydro/density_loop.h:624-658 / src/hydro/density_loop.cc:491-512 */
static double or_face_closure(const double *nvt_raw6, const double *fw, const double *Tinv, double normalizer, double V_i)
{
    double trace_raw = nvt_raw6[0] + nvt_raw6[1] + nvt_raw6[2];
    double dx_i = sqrt(V_i * trace_raw);
    double inv[9];
    for (int k = 0; k < 9; k++) inv[k] = Tinv[k] / normalizer;
    double out[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) out[a] += 2.0 * V_i * inv[3 * a + b] * fw[b];
    double leak = 0;
    for (int a = 0; a < 3; a++) leak += fabs(out[a]) / (double)OR_NDIMS;
    return leak / (2.0 * (double)OR_NDIMS * pow(dx_i, (double)(OR_NDIMS - 1)));
}

/* ------------------------------------------------------------------------- */
/* density pair sums: DensitySpec::pair_kernel, src/hydro/density_loop.h:1141-1385 */
/* (the gas-i, no-extra-physics branches).  acc[OR_NACC] = Ngb, Rho, DrkernNgb,    */
/* DivVel, NV_T[6], face_w[3].                                                    */
/* ------------------------------------------------------------------------- */
static void or_density_sums(const orc_params *par, const orc_state *s, long i, double h,
                            int *cand, long ncand, double *acc)
{
    double box = par->box, half = 0.5 * par->box;
    for (int k = 0; k < OR_NACC; k++) acc[k] = 0;
    const double *pi = &s->pos[3 * i];
    const double *vi = &s->vel[3 * i];
    const double h2 = h * h;
    /* keep the candidates inside h (the pair test below repeats it), ascending */
    long m = 0;
    const double *P = or_pos(s);
    for (long q = 0; q < ncand; q++) {
        int j = cand[q];
        double e0 = or_wrap(pi[0] - P[3 * j + 0], box, half);
        double e1 = or_wrap(pi[1] - P[3 * j + 1], box, half);
        double e2 = or_wrap(pi[2] - P[3 * j + 2], box, half);
        if (e0 * e0 + e1 * e1 + e2 * e2 < h2) {
            or_read(j);
            cand[m++] = j;
        }
    }
    or_sort_int(cand, m);
    ncand = m;
    for (long q = 0; q < ncand; q++) {
        long j = cand[q];
        if (s->mass[j] <= 0) continue;
        double dp0 = pi[0] - s->pos[3 * j + 0];
        double dp1 = pi[1] - s->pos[3 * j + 1];
        double dp2 = pi[2] - s->pos[3 * j + 2];
        dp1 = or_wrap(dp1, box, half);
        dp0 = or_wrap(dp0, box, half);
        dp2 = or_wrap(dp2, box, half);
        double r2 = dp0 * dp0 + dp1 * dp1 + dp2 * dp2;
        if (r2 >= h2) continue;
        double r = sqrt(r2);
        double hinv, hinv3, hinv4;
        or_kernel_hinv(h, &hinv, &hinv3, &hinv4);
        double u = r * hinv;
        double wk = 0, dwk = 0;
        or_kernel_main(par->kernel, u, hinv3, hinv4, &wk, &dwk, 0);
        double mj_wk = s->mass[j] * wk;
        acc[0] += wk;
        acc[1] += mj_wk;
        acc[2] += -(OR_NDIMS * hinv * wk + u * dwk);
        if (r > 0) {
            acc[4] += wk * (dp0 * dp0);
            acc[5] += wk * (dp1 * dp1);
            acc[6] += wk * (dp2 * dp2);
            acc[7] += wk * (dp0 * dp1);
            acc[8] += wk * (dp1 * dp2);
            acc[9] += wk * (dp0 * dp2);
            acc[10] += wk * dp0;
            acc[11] += wk * dp1;
            acc[12] += wk * dp2;
            double dv0 = vi[0] - s->vel[3 * j + 0];
            double dv1 = vi[1] - s->vel[3 * j + 1];
            double dv2 = vi[2] - s->vel[3 * j + 2];
            acc[3] -= dwk * (dp0 * dv0 + dp1 * dv1 + dp2 * dv2) / r;
            if (par->fd_correction) {
                /* GradH_numer, GradH_denom (src/hydro/density_loop.h:1281-1286) */
                acc[13] += (dwk / r) * dp0;
                acc[14] += (dwk / r) * dp1;
                acc[15] += (dwk / r) * dp2;
                acc[16] += u * dwk;
            }
        }
    }
}

/* per-active iteration state: DensityIterScratch, src/hydro/density_loop.h:196-211 */
typedef struct {
    double left, right;
    int converged;
    int cn_initialized;
    double cn_current;
} or_scratch;

/* normalisation shared by after_iter_body (:532-556) and the finalize (:323-348) */
static void or_density_normalise(const double *acc, double h, double *numngb, double *drkern, double *divvel)
{
    double NumNgb_eff = acc[0], Dr = acc[2], Dv = acc[3];
    if (NumNgb_eff > 0) {
        Dr *= h / (OR_NDIMS * NumNgb_eff);
        Dv /= NumNgb_eff;
        NumNgb_eff *= OR_VOLNORM * pow(h, (double)OR_NDIMS);
    } else {
        NumNgb_eff = 0; Dr = 0; Dv = 0;
    }
    if (Dr > -0.9) Dr = 1.0 / (1.0 + Dr); else Dr = 1.0;
    Dv *= Dr;
    *numngb = NumNgb_eff; *drkern = Dr; *divvel = Dv;
}

/* DensitySpec::after_iter_body, src/hydro/density_loop.h:516-961, gas particles,
   no sinks/grains/GALSF.  Returns 0 = converged (h unchanged), 1 = adjust radius
   (*new_h), -1 = the reference's stop code 90001009. */
static int or_after_iter(const orc_params *par, double cn_stored, int iter, double h, const double *acc,
                         or_scratch *sc, double *new_h_out)
{
    double desnumngb = par->des_num_ngb;
    double desnumngbdev = par->max_num_ngb_dev;
    if (par->time == par->time_begin) {
        if (par->max_num_ngb_dev > 0.05) desnumngbdev = 0.05;
    }
    const double desnumngbdev_0 = desnumngbdev;
    double NumNgb_eff, Dr_eff, Dv_eff;
    or_density_normalise(acc, h, &NumNgb_eff, &Dr_eff, &Dv_eff);

    /* (b) NV_T inversion and face-closure error */
    const double V_i = OR_VOLNORM * pow(h, (double)OR_NDIMS) / (NumNgb_eff > 0 ? NumNgb_eff : 1.0);
    const double normalizer = pow(h, (double)(2 - OR_NDIMS));
    double Tinv[9];
    double ConditionNumber = or_nvt_regularised_inverse(&acc[4], normalizer, Tinv);
    double fce = or_face_closure(&acc[4], &acc[10], Tinv, normalizer, V_i);

    /* (c) iter==0 condition-number promotion */
    {
        const double c0 = 0.1 * (double)OR_CN_DANGER;
        if ((iter == 0) && (ConditionNumber > cn_stored) && (cn_stored > 0)) {
            double ncorr_a = 1.0, cn_a = cn_stored;
            if (cn_a > c0) ncorr_a = sqrt(1.0 + (cn_a - c0) / (double)OR_CN_DANGER);
            if (ncorr_a > 2.0) ncorr_a = 2.0;
            const double dn_ngb = fabs(NumNgb_eff - par->des_num_ngb * ncorr_a) / (desnumngbdev_0 * ncorr_a);
            double ncorr_b = 1.0, cn_b = ConditionNumber;
            if (cn_b > c0) ncorr_b = sqrt(1.0 + (cn_b - c0) / (double)OR_CN_DANGER);
            if (ncorr_b > 2.0) ncorr_b = 2.0;
            const double dn_ngb_alt = fabs(NumNgb_eff - par->des_num_ngb * ncorr_b) / (desnumngbdev_0 * ncorr_b);
            const double dn_ngb_min = DMIN(dn_ngb, dn_ngb_alt);
            if (dn_ngb_min < 10.0) {
                sc->cn_current = ConditionNumber;
                sc->cn_initialized = 1;
            }
        }
        if (!sc->cn_initialized) {
            sc->cn_current = cn_stored;
            sc->cn_initialized = 1;
        }
    }

    /* (d) neighbour-number correction */
    double ncorr_ngb = 1.0;
    {
        const double c0 = 0.1 * (double)OR_CN_DANGER;
        const double cn = sc->cn_current;
        if (cn > c0) ncorr_ngb = sqrt(1.0 + (cn - c0) / (double)OR_CN_DANGER);
        if (ncorr_ngb > 2.0) ncorr_ngb = 2.0;
        const double d00 = 0.35;
        if (fce > d00) ncorr_ngb = DMAX(ncorr_ngb, DMIN(fce / d00, 2.0));
    }
    desnumngb = par->des_num_ngb * ncorr_ngb;
    desnumngbdev = desnumngbdev_0 * ncorr_ngb;
    if (iter > 1) {
        desnumngbdev = DMIN(0.25 * desnumngb,
                            desnumngbdev * exp(0.1 * log(desnumngb / (16.0 * desnumngbdev)) * (double)iter));
    }
    double minsoft = par->min_kernel_radius;
    double maxsoft = par->max_kernel_radius;

    /* (e) convergence check and radius update */
    int redo = 0, set_min = 0, set_max = 0;
    double new_h = h;
    if ((NumNgb_eff < (desnumngb - desnumngbdev) && new_h < 0.999 * maxsoft) ||
        (NumNgb_eff > (desnumngb + desnumngbdev) && new_h > 1.001 * minsoft)) redo = 1;
    if ((new_h >= 0.999 * maxsoft) && (NumNgb_eff < (desnumngb - desnumngbdev))) {
        redo = 0;
        if (new_h == maxsoft) { set_max = 0; }
        else { redo = 1; new_h = maxsoft; set_max = 1; }
    }
    if ((new_h <= 1.001 * minsoft) && (NumNgb_eff > (desnumngb + desnumngbdev))) {
        redo = 0;
        if (new_h == minsoft) { set_min = 0; }
        else { redo = 1; new_h = minsoft; set_min = 1; }
    }
    if (redo == 0) sc->cn_current = ConditionNumber;
    if (redo) {
        if (sc->left > 0 && sc->right > 0) {
            if ((sc->right - sc->left) < 1.0e-3 * sc->left) {
                sc->converged = 1;
                sc->cn_current = ConditionNumber;
                *new_h_out = new_h;
                return 0;
            }
        }
        if ((set_max == 0) && (set_min == 0)) {
            if (NumNgb_eff < (desnumngb - desnumngbdev)) {
                sc->left = DMAX(new_h, sc->left);
            } else {
                if (sc->right != 0) { if (new_h < sc->right) sc->right = new_h; }
                else { sc->right = new_h; }
            }
            if (sc->right > 0 && sc->left > 0) {
                double maxjump = 0;
                if (iter > 1) maxjump = 0.2 * log(sc->right / sc->left);
                if (NumNgb_eff > 1) {
                    double jumpvar = Dr_eff * log(desnumngb / NumNgb_eff) / OR_NDIMS;
                    if (iter > 1) {
                        if (fabs(jumpvar) < maxjump) { if (jumpvar < 0) jumpvar = -maxjump; else jumpvar = maxjump; }
                    }
                    new_h *= exp(jumpvar);
                } else {
                    new_h *= 2.0;
                }
                if ((new_h < sc->right) && (new_h > sc->left)) {
                    if (iter > 1) {
                        const double hfac = exp(maxjump);
                        if (new_h > sc->right / hfac) new_h = sc->right / hfac;
                        if (new_h < sc->left * hfac) new_h = sc->left * hfac;
                    }
                } else {
                    if (new_h > sc->right) new_h = sc->right;
                    if (new_h < sc->left) new_h = sc->left;
                    new_h = pow(new_h * sc->left * sc->right, 1.0 / 3.0);
                }
            } else {
                if (sc->right == 0 && sc->left == 0) {
                    *new_h_out = new_h;
                    return -1;
                }
                if (sc->right == 0 && sc->left > 0) {
                    double fac_lim;
                    if (NumNgb_eff > 1) fac_lim = log(desnumngb / NumNgb_eff) / OR_NDIMS; else fac_lim = 1.4;
                    if ((NumNgb_eff < 2.0 * desnumngb) && (NumNgb_eff > 0.1 * desnumngb)) {
                        double slope = Dr_eff;
                        if (iter > 2 && slope < 1.0) slope = 0.5 * (slope + 1.0);
                        double fac = fac_lim * slope;
                        if (iter >= 4) { if (Dr_eff == 1.0) fac *= 10.0; }
                        if (fac < fac_lim + 0.231) new_h *= exp(fac); else new_h *= exp(fac_lim + 0.231);
                    } else {
                        new_h *= exp(fac_lim);
                    }
                }
                if (sc->right > 0 && sc->left == 0) {
                    double fac_lim;
                    if (NumNgb_eff > 1) fac_lim = log(desnumngb / NumNgb_eff) / OR_NDIMS; else fac_lim = 1.4;
                    if (fac_lim < -1.535) fac_lim = -1.535;
                    if ((NumNgb_eff < 2.0 * desnumngb) && (NumNgb_eff > 0.1 * desnumngb)) {
                        double slope = Dr_eff;
                        if (iter > 2 && slope < 1.0) slope = 0.5 * (slope + 1.0);
                        double fac = fac_lim * slope;
                        if (iter >= 4) { if (Dr_eff == 1.0) fac *= 10.0; }
                        if (fac > fac_lim - 0.231) new_h *= exp(fac); else new_h *= exp(fac_lim - 0.231);
                    } else {
                        new_h *= exp(fac_lim);
                    }
                }
            }
        }
        if (new_h < minsoft) new_h = minsoft;
        if (set_min == 1) new_h = minsoft;
        if (new_h > maxsoft) new_h = maxsoft;
        if (set_max == 1) new_h = maxsoft;
        *new_h_out = new_h;
        return 1;
    }
    sc->converged = 1;
    sc->cn_current = ConditionNumber;
    *new_h_out = new_h;
    return 0;
}

/* density_finalize_post_runner, src/hydro/density_loop.cc:303-733 (MFM branch:
   NV_T inversion, FaceClosureError, rho = m NumNgb/(4pi/3 h^3), EOS, NumNgb^(1/3)).
   EOS: set_eos_pressure_impl src/eos/eos_functions.h:830-873 (ideal gas). */
static void or_density_finalize(const orc_params *par, orc_state *s, long i, double h, const double *acc)
{
    s->h[i] = h;
    double NumNgb_eff, Dr_eff, Dv_eff;
    or_density_normalise(acc, h, &NumNgb_eff, &Dr_eff, &Dv_eff);
    s->numngb[i] = NumNgb_eff;
    s->drkern[i] = Dr_eff;
    s->divvel[i] = Dv_eff;
    s->rho[i] = acc[1];
    for (int k = 0; k < 3; k++) s->nvt_fw[3 * i + k] = acc[10 + k];
    {
        const double V_i = OR_VOLNORM * pow(h, (double)OR_NDIMS) / (NumNgb_eff > 0 ? NumNgb_eff : 1.0);
        const double normalizer = pow(h, (double)(2 - OR_NDIMS));
        double Tinv[9];
        double cn = or_nvt_regularised_inverse(&acc[4], normalizer, Tinv);
        /* upper triangle written back: [xx,yy,zz,xy,yz,xz] */
        s->nvt[6 * i + 0] = Tinv[0] / normalizer;
        s->nvt[6 * i + 1] = Tinv[4] / normalizer;
        s->nvt[6 * i + 2] = Tinv[8] / normalizer;
        s->nvt[6 * i + 3] = Tinv[1] / normalizer;
        s->nvt[6 * i + 4] = Tinv[5] / normalizer;
        s->nvt[6 * i + 5] = Tinv[2] / normalizer;
        s->cn[i] = cn;
        s->fce[i] = or_face_closure(&acc[4], &acc[10], Tinv, normalizer, V_i);
    }
    if (s->mass[i] > 0) {
        if ((s->h[i] > 0) && (s->numngb[i] > 0)) {
            s->rho[i] = s->mass[i] * s->numngb[i] / (OR_VOLNORM * pow(s->h[i], (double)OR_NDIMS));
        } else {
            if (s->h[i] <= 0) s->rho[i] = 0;
            else s->rho[i] = s->mass[i] / (OR_VOLNORM * pow(s->h[i], (double)OR_NDIMS));
        }
        if (par->fd_correction && acc[16] != 0) {
            /* the first-derivative partition-of-unity volume correction (src/hydro/density_loop.cc:622-633):
               the gradient of the kernel size implied by the particle distribution,
               grad H = -GradH_numer / GradH_denom, divides the density by 1 + 4 alpha |grad H|^2 */
            const double inv_denom = -1.0 / acc[16];
            const double g0 = acc[13] * inv_denom, g1 = acc[14] * inv_denom, g2 = acc[15] * inv_denom;
            const double fd = 1.0 + 4.0 * orc_fd_alpha(par->kernel) * (g0 * g0 + g1 * g1 + g2 * g2);
            s->rho[i] /= fd;
        }
        s->press[i] = (par->gamma - 1) * s->u[i] * s->rho[i];
    }
    if (s->numngb[i] > 0) s->numngb[i] = pow(s->numngb[i], 1.0 / OR_NDIMS);
    else s->numngb[i] = 0;
}

/* density(), src/hydro/density_loop.cc:764-899: per active gas particle, clamp the
   radius (:818-838), iterate pair sums + after_iter until converged
   (run_neighbor_loop_iterative, src/mesh/neighbor_loop_runner.cc:7149-7520: the final
   accumulators are those of the last evaluation, at the radius it was evaluated at),
   then finalize.  The density sums of particle i depend on no other particle's radius,
   so iterating each particle on its own reproduces the reference's lock-step rounds.
   iters_out (optional, n ints) receives the evaluation count per particle. */
int orc_density(const orc_params *par, orc_state *s, const unsigned char *active, int *iters_out)
{
    long n = s->n;
    double box = par->box;
    double maxsoft = par->max_kernel_radius;
    for (long i = 0; i < n; i++) {
        if (active && !active[i]) continue;
        if (!(s->mass[i] > 0)) continue;
        if ((s->h[i] < 0) || !isfinite(s->h[i]) || (s->h[i] > 0.99 * maxsoft)) s->h[i] = 0.99 * maxsoft;
    }
    /* one grid sized for the largest radius any particle can reach here; rebuilt when exceeded */
    double rmax = 0;
    for (long i = 0; i < n; i++) if ((!active || active[i]) && s->h[i] > rmax) rmax = s->h[i];
    rmax *= 1.3;
    if (rmax > 0.5 * box) rmax = 0.5 * box;
    int status = 0;
    unsigned char *todo = (unsigned char *)malloc((size_t)(n > 0 ? n : 1));
    double *hcur = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    or_scratch *sc = (or_scratch *)calloc((size_t)(n > 0 ? n : 1), sizeof(or_scratch));
    int *iters = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    double *accs = (double *)malloc((size_t)(n > 0 ? n : 1) * OR_NACC * sizeof(double));
    for (long i = 0; i < n; i++) {
        todo[i] = (!active || active[i]) && (s->mass[i] > 0);
        hcur[i] = s->h[i];
    }
    while (1) {
        or_grid g;
        if (or_grid_build(&g, n, or_pos(s), box, rmax)) { status = -2; break; }
        double grid_r = g.cell;
        int need_grow = 0;
#pragma omp parallel
        {
            int *buf = (int *)malloc((size_t)n * sizeof(int));
#pragma omp for schedule(dynamic, 64)
            for (long i = 0; i < n; i++) {
                if (!todo[i]) continue;
                while (1) {
                    double h = hcur[i];
                    if (h > grid_r && g.nc > 2) {
#pragma omp atomic write
                        need_grow = 1;
                        break;
                    }
                    long m = or_grid_candidates(&g, &s->pos[3 * i], box, buf);
                    or_density_sums(par, s, i, h, buf, m, &accs[OR_NACC * i]);
                    double nh = h;
                    int rc = or_after_iter(par, s->cn[i], iters[i], h, &accs[OR_NACC * i], &sc[i], &nh);
                    iters[i]++;
                    if (rc <= 0) {
                        if (rc < 0) {
#pragma omp atomic write
                            status = 90001009;
                        }
                        todo[i] = 0;
                        break;
                    }
                    hcur[i] = nh;
                    if (iters[i] >= OR_MAXITER) {
#pragma omp atomic write
                        status = 1155;
                        todo[i] = 0;
                        break;
                    }
                }
            }
            free(buf);
        }
        or_grid_free(&g);
        if (!need_grow) break;
        double nm = 0;
        for (long i = 0; i < n; i++) if (todo[i] && hcur[i] > nm) nm = hcur[i];
        rmax = DMAX(1.3 * nm, 2.0 * rmax);
        if (rmax > 0.5 * box) rmax = 0.5 * box;
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        if (!(!active || active[i]) || !(s->mass[i] > 0)) continue;
        or_density_finalize(par, s, i, hcur[i], &accs[OR_NACC * i]);
    }
    if (iters_out) for (long i = 0; i < n; i++) iters_out[i] = iters[i];
    free(todo); free(hcur); free(sc); free(iters); free(accs);
    return status;
}

/* density at a fixed, given radius: one evaluation of the pair sums plus the finalize.
   raw (optional, OR_NACC*n) receives the accumulators.  This is the fixed-h form in which
   the device path is compared to 1e-12 and bit-for-bit. */
int orc_density_fixed_h(const orc_params *par, orc_state *s, const unsigned char *active, double *raw)
{
    long n = s->n;
    double rmax = 0;
    for (long i = 0; i < n; i++) if ((!active || active[i]) && s->h[i] > rmax) rmax = s->h[i];
    if (!(rmax > 0) || rmax > 0.5 * par->box) return -1;
    or_grid g;
    if (or_grid_build(&g, n, s->pos, par->box, rmax)) return -2;
#pragma omp parallel
    {
        int *buf = (int *)malloc((size_t)n * sizeof(int));
        double acc[OR_NACC];
#pragma omp for schedule(dynamic, 64)
        for (long i = 0; i < n; i++) {
            if ((active && !active[i]) || !(s->mass[i] > 0)) continue;
            long m = or_grid_candidates(&g, &s->pos[3 * i], par->box, buf);
            or_density_sums(par, s, i, s->h[i], buf, m, acc);
            if (raw) for (int k = 0; k < OR_NACC; k++) raw[OR_NACC * i + k] = acc[k];
            or_density_finalize(par, s, i, s->h[i], acc);
        }
        free(buf);
    }
    or_grid_free(&g);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* gradients                                                                    */
/* ------------------------------------------------------------------------- */
/* local_slopelimiter, src/hydro/gradients_loop.cc:76-104 (Newtonian build) */
void orc_slopelimiter(double *g, double valmax, double valmin, double alim, double h,
                      double shoot_tol, int pos_preserve, double d_max, double val_cen)
{
    double d_abs = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    if (d_abs > 0) {
        d_abs = sqrt(d_abs);
        double cfac = 1 / (alim * h * d_abs);
        double fabs_max = fabs(valmax), fabs_min = fabs(valmin), abs_max = fabs_max, abs_min = fabs_min, f_corr;
        if (abs_max < abs_min) { abs_max = fabs_min; abs_min = fabs_max; }
        f_corr = DMIN(abs_min + shoot_tol * abs_max, abs_max);
        cfac *= f_corr;
        if (pos_preserve == 1) {
            double fmin = DMIN(val_cen, DMAX(0, DMAX(OR_MIN_REAL * val_cen, DMIN(0.5 * (val_cen + valmin), val_cen - f_corr))));
            cfac = DMIN(((val_cen - fmin) / d_max) / d_abs, cfac);
        }
        if (cfac < 1) { g[0] *= cfac; g[1] *= cfac; g[2] *= cfac; }
    }
}

/* MINMAX_CHECK, src/hydro/gradient_functions.h:29 */
#define OR_MINMAX(x, mn, mx) ((x < mn) ? (mn = x) : ((x > mx) ? (mx = x) : (1)))

/* hydro_gradient_calc(), src/hydro/gradients_loop.cc:458-1343 with the pair body
   gradient_accumulate_neighbor (src/hydro/gradient_functions.h:305-665; matrix weights,
   SHOULD_I_USE_SPH_GRADIENTS is 0 without COOLING, :31-36), construct_gradient (:129-136)
   and the slope-limiter block (:1077-1130, SELFGRAVITY_OFF: h_lim = h, stol = 0.1).
   Neighbours: the symmetric list, r2 < max(h_i,h_j)^2, ascending. */
int orc_gradients(const orc_params *par, orc_state *s, const unsigned char *active)
{
    long n = s->n;
    double box = par->box, half = 0.5 * par->box;
    const double *P = or_pos(s), *Hp = or_hpred(s);
    double rmax = or_max_h(n, Hp, NULL);
    if (!(rmax > 0) || rmax > half) return -1;
    or_grid g;
    if (or_grid_build(&g, n, P, box, rmax)) return -2;
#pragma omp parallel
    {
        int *buf = (int *)malloc((size_t)n * sizeof(int));
#pragma omp for schedule(dynamic, 64)
        for (long i = 0; i < n; i++) {
            if (active && !active[i]) continue;
            double *gr = &s->grad_rho[3 * i], *gp = &s->grad_p[3 * i], *gv = &s->grad_v[9 * i];
            for (int k = 0; k < 3; k++) { gr[k] = 0; gp[k] = 0; }
            for (int k = 0; k < 9; k++) gv[k] = 0;
            if (!(s->mass[i] > 0 && s->rho[i] > 0)) continue;
            const double *pi = &s->pos[3 * i];
            double h_i = s->h[i];
            double hinv, hinv3, hinv4;
            or_kernel_hinv(h_i, &hinv, &hinv3, &hinv4);
            double rho_i = s->rho[i], P_i = s->press[i];
            const double *vi = &s->vel[3 * i];
            double G[3][5];        /* [dir][rho, P, vx, vy, vz] */
            double mx[5], mn[5], maxd = 0;
            for (int a = 0; a < 3; a++) for (int b = 0; b < 5; b++) G[a][b] = 0;
            for (int b = 0; b < 5; b++) { mx[b] = 0; mn[b] = 0; }
            long m0 = or_grid_candidates(&g, pi, box, buf), m = 0;
            for (long q = 0; q < m0; q++) {
                int j = buf[q];
                if (j != i && or_accept(pi, &P[3 * j], h_i, Hp[j], 1, box, half)) {
                    or_read(j);
                    buf[m++] = j;
                }
            }
            or_sort_int(buf, m);
            for (long q = 0; q < m; q++) {
                long j = buf[q];
                if (!(s->mass[j] > 0) || !(s->rho[j] > 0) || !(s->h[j] > 0)) continue;
                double dp[3];
                dp[0] = pi[0] - s->pos[3 * j + 0];
                dp[1] = pi[1] - s->pos[3 * j + 1];
                dp[2] = pi[2] - s->pos[3 * j + 2];
                dp[1] = or_wrap(dp[1], box, half);
                dp[0] = or_wrap(dp[0], box, half);
                dp[2] = or_wrap(dp[2], box, half);
                double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
                double h_j = s->h[j];
                if (r2 <= 0) continue;
                double h2_i = h_i * h_i;
                if ((r2 >= h2_i) && (r2 >= h_j * h_j)) continue;
                double r = sqrt(r2);
                double wk_i = 0, dwk_i = 0;
                if (r < h_i) {
                    double u = r * hinv;
                    or_kernel_main(par->kernel, u, hinv3, hinv4, &wk_i, &dwk_i, -1);
                }
                if (r > maxd) maxd = r;
                double d[5];
                d[0] = s->rho[j] - rho_i;
                d[1] = s->press[j] - P_i;
                d[2] = s->vel[3 * j + 0] - vi[0];
                d[3] = s->vel[3 * j + 1] - vi[1];
                d[4] = s->vel[3 * j + 2] - vi[2];
                for (int b = 0; b < 5; b++) OR_MINMAX(d[b], mn[b], mx[b]);
                if (r < h_i) {
                    double wk_for_grad = wk_i;
                    for (int k = 0; k < 3; k++) {
                        double wk_xyz_i = -wk_for_grad * dp[k];
                        for (int b = 0; b < 5; b++) G[k][b] += wk_xyz_i * d[b];
                    }
                }
            }
            /* construct_gradient: g = NV_T g */
            const double *T = &s->nvt[6 * i];
            double v[3], o[3];
            for (int b = 0; b < 5; b++) {
                v[0] = G[0][b]; v[1] = G[1][b]; v[2] = G[2][b];
                or_sym_matvec(T, v, o);
                if (b == 0) { gr[0] = o[0]; gr[1] = o[1]; gr[2] = o[2]; }
                else if (b == 1) { gp[0] = o[0]; gp[1] = o[1]; gp[2] = o[2]; }
                else { gv[3 * (b - 2) + 0] = o[0]; gv[3 * (b - 2) + 1] = o[1]; gv[3 * (b - 2) + 2] = o[2]; }
            }
            /* slope limiter block */
            double stol = 0.1;
            double h_lim = h_i;
            double d_max = DMAX(h_i, maxd);
            double a_limiter = 0.25;
            if (s->cn[i] > 100) a_limiter = DMIN(0.5, 0.25 + 0.25 * (s->cn[i] - 100) / 100);
            orc_slopelimiter(gr, mx[0], mn[0], a_limiter, h_lim, 0, 1, d_max, rho_i);
            orc_slopelimiter(gp, mx[1], mn[1], a_limiter, h_lim, stol, 1, d_max, P_i);
            for (int k = 0; k < 3; k++) orc_slopelimiter(&gv[3 * k], mx[2 + k], mn[2 + k], a_limiter, h_lim, stol, 0, 0, 0);
        }
        free(buf);
    }
    or_grid_free(&g);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Riemann problem: src/hydro/reimann.h                                          */
/* ------------------------------------------------------------------------- */
typedef struct { double rho, p, v[3], u, cs; } or_rstate;
typedef struct { double P_M, S_M; double fv[3], fp; } or_rout;

typedef struct {
    double G0, G1, G2, G3, G4, G5, G6, G7, G8, G9;
} or_gamma;

static void or_gamma_init(or_gamma *g, double gamma)
{
    g->G0 = gamma;
    g->G1 = ((g->G0 - 1.0) / (2.0 * g->G0));
    g->G2 = ((g->G0 + 1.0) / (2.0 * g->G0));
    g->G3 = ((2.0 * g->G0 / (g->G0 - 1.0)));
    g->G4 = (2.0 / (g->G0 - 1.0));
    g->G5 = (2.0 / (g->G0 + 1.0));
    g->G6 = ((g->G0 - 1.0) / (g->G0 + 1.0));
    g->G7 = (0.5 * (g->G0 - 1.0));
    g->G8 = (1.0 / g->G0);
    g->G9 = (g->G0 - 1.0);
}

static inline double or_dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* The rungs of the Riemann ladder a face took, one bit each (the device's RiemannRung bits,
   in the order of orc_rung_counts): not a reference quantity, a record of which branch of
   src/hydro/reimann.h and the pair body's retries ran. */
enum {
    OR_RUNG_HLLC = 1u << 0, OR_RUNG_VACUUM = 1u << 1, OR_RUNG_ROE = 1u << 2, OR_RUNG_PVRS = 1u << 3, OR_RUNG_KT = 1u << 4,
    OR_RUNG_EXACT = 1u << 5, OR_RUNG_EXACT_VACUUM = 1u << 6, OR_RUNG_UNPHYSICAL = 1u << 7, OR_RUNG_RETRY1 = 1u << 8,
    OR_RUNG_RETRY2 = 1u << 9, OR_RUNG_FAIL = 1u << 10, OR_RUNG_FACE = 1u << 11, OR_ENTROPIC_CONSIDERED = 1u << 12,
    OR_ENTROPIC_APPLIED = 1u << 13
};
#define OR_RUNG_COUNT 14
static long long or_rung_totals[OR_RUNG_COUNT];

/* reconstruct_face_states, src/hydro/reimann.h:125-244 (non-MHD, tolerance 1, no CRK) */
void orc_reconstruct(double Q_i, const double *Gi, double Q_j, const double *Gj,
                     const double *di, const double *dj, double *Q_L, double *Q_R, int mode)
{
    if (mode == 0) { *Q_R = Q_i; *Q_L = Q_j; return; }
    if (Q_i == Q_j) { *Q_L = *Q_R = Q_i; return; }
    *Q_R = Q_i + Gi[0] * di[0] + Gi[1] * di[1] + Gi[2] * di[2];
    *Q_L = Q_j + Gj[0] * dj[0] + Gj[1] * dj[1] + Gj[2] * dj[2];
    double Qmin, Qmax, Qmed, Qmax_eff, Qmin_eff, fac, Qmed_max, Qmed_min;
    double fac_minmax = 0.5;
    double fac_meddev = 0.375;
    Qmed = 0.5 * (Q_i + Q_j);
    if (Q_i < Q_j) { Qmax = Q_j; Qmin = Q_i; } else { Qmax = Q_i; Qmin = Q_j; }
    fac = fac_minmax * (Qmax - Qmin);
    Qmax_eff = Qmax + fac;
    Qmin_eff = Qmin - fac;
    if (mode > 0) {
        if (Qmax < 0) { if (Qmax_eff > 0) Qmax_eff = Qmax * Qmax / (Qmax - (Qmax_eff - Qmax)); }
        if (Qmin > 0) { if (Qmin_eff < 0) Qmin_eff = Qmin * Qmin / (Qmin + (Qmin - Qmin_eff)); }
    }
    fac = fac_meddev * (Qmax - Qmin);
    Qmed_max = Qmed + fac;
    Qmed_min = Qmed - fac;
    if (Qmed_max > Qmax_eff) Qmed_max = Qmax_eff;
    if (Qmed_min < Qmin_eff) Qmed_min = Qmin_eff;
    if (Q_i < Q_j) {
        if (*Q_R < Qmin_eff) *Q_R = Qmin_eff;
        if (*Q_R > Qmed_max) *Q_R = Qmed_max;
        if (*Q_L > Qmax_eff) *Q_L = Qmax_eff;
        if (*Q_L < Qmed_min) *Q_L = Qmed_min;
    } else {
        if (*Q_R > Qmax_eff) *Q_R = Qmax_eff;
        if (*Q_R < Qmed_min) *Q_R = Qmed_min;
        if (*Q_L < Qmin_eff) *Q_L = Qmin_eff;
        if (*Q_L > Qmed_max) *Q_L = Qmed_max;
    }
}

/* get_wavespeeds_and_pressure_star, src/hydro/reimann.h:485-549; returns the estimate taken */
static unsigned or_wavespeeds(const or_gamma *G, const or_rstate *L, const or_rstate *R, const double *n,
                              double vL, double vR, double csL, double csR, double hL, double hR,
                              or_rout *out, double ptl)
{
    double S_L, S_R;
    unsigned rung;
    if ((vR - vL) > DMAX(csL, csR)) {
        out->P_M = OR_MIN_REAL; out->S_M = S_L = S_R = 0;
        rung = OR_RUNG_VACUUM;
    } else {
        double PT_L = L->p, PT_R = R->p;
        S_L = DMIN(vL, vR) - DMAX(csL, csR);
        S_R = DMAX(vL, vR) + DMAX(csL, csR);
        double rho_wt_L = L->rho * (S_L - vL);
        double rho_wt_R = R->rho * (S_R - vR);
        out->S_M = ((PT_R - PT_L) + rho_wt_L * vL - rho_wt_R * vR) / (rho_wt_L - rho_wt_R);
        out->P_M = (PT_L * rho_wt_R - PT_R * rho_wt_L + rho_wt_L * rho_wt_R * (vR - vL)) / (rho_wt_R - rho_wt_L);
        if (out->P_M <= OR_MIN_REAL) { out->P_M = OR_MIN_REAL; out->S_M = S_L = S_R = 0; }
        rung = OR_RUNG_HLLC;
        if ((out->P_M <= 0) || (isnan(out->P_M)) || (out->P_M > ptl)) {
            double sqrt_rho_L = sqrt(L->rho);
            double sqrt_rho_R = sqrt(R->rho);
            double sqrt_rho_inv = 1 / (sqrt_rho_L + sqrt_rho_R);
            double v_roe[3];
            for (int k = 0; k < 3; k++) v_roe[k] = (sqrt_rho_L * L->v[k] + sqrt_rho_R * R->v[k]) * sqrt_rho_inv;
            double v_line_roe = or_dot3(v_roe, n);
            double h_roe = (sqrt_rho_L * hL + sqrt_rho_R * hR) * sqrt_rho_inv;
            double cs_roe = sqrt(DMAX(1.e-30, G->G9 * (h_roe - 0.5 * (v_roe[0] * v_roe[0] + v_roe[1] * v_roe[1] + v_roe[2] * v_roe[2]))));
            S_R = DMAX(vR + csR, v_line_roe + cs_roe);
            S_L = DMIN(vL - csL, v_line_roe - cs_roe);
            rho_wt_R = R->rho * (S_R - vR);
            rho_wt_L = -L->rho * (S_L - vL);
            out->S_M = ((rho_wt_R * vR + rho_wt_L * vL) + (PT_L - PT_R)) / (rho_wt_R + rho_wt_L);
            out->P_M = L->rho * (vL - S_L) * (vL - out->S_M) + PT_L;
            if (out->P_M <= OR_MIN_REAL) { out->P_M = OR_MIN_REAL; out->S_M = S_L = S_R = 0; }
            rung = OR_RUNG_ROE;
            if ((out->P_M <= 0) || (isnan(out->P_M)) || (out->P_M > ptl)) {
                out->P_M = 0.5 * ((PT_L + PT_R) + (vL - vR) * 0.25 * (L->rho + R->rho) * (csL + csR));
                out->S_M = 0.5 * (vR + vL) + 2.0 * (PT_L - PT_R) / ((L->rho + R->rho) * (csL + csR));
                double S_plus = DMAX(DMAX(fabs(vL - csL), fabs(vR - csR)), DMAX(fabs(vL + csL), fabs(vR + csR)));
                S_L = -S_plus; S_R = S_plus;
                if (out->S_M < S_L) out->S_M = S_L;
                if (out->S_M > S_R) out->S_M = S_R;
                if (out->P_M <= OR_MIN_REAL) { out->P_M = OR_MIN_REAL; out->S_M = S_L = S_R = 0; }
                rung = OR_RUNG_PVRS;
            }
        }
    }
    (void)S_L; (void)S_R;
    return rung;
}

/* Riemann_solver_KurganovTadmor_PWK, src/hydro/reimann.h:445-477 (MFM branch) */
static void or_kt(const or_rstate *L, const or_rstate *R, const double *n, double vL, double vR,
                  double csL, double csR, double hL, double hR, or_rout *out)
{
    double S_L, S_R, S_M, alpha;
    double delta_threshold = 0.001 * 0.5 * (L->rho + R->rho) * 0.5 * (csL + csR);
    double dv[3];
    for (int k = 0; k < 3; k++) dv[k] = R->rho * R->v[k] - L->rho * L->v[k];
    double dv2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
    double dvL = fabs(R->rho * vR - L->rho * vL);
    alpha = dvL / sqrt(delta_threshold * delta_threshold + dv2);
    S_L = alpha * csL; if (vL > 0) { S_L += vL; } else { S_L -= vL; }
    S_R = alpha * csR; if (vR > 0) { S_R += vR; } else { S_R -= vR; }
    S_M = DMAX(S_L, S_R);
    out->S_M = DMAX(csL + fabs(vL), csR + fabs(vR));
    double f_rho_left = L->rho * (vL + S_M);
    double f_rho_right = R->rho * (vR - S_M);
    double denom = 1 / (L->rho * vL - R->rho * vR + S_M * (L->rho + R->rho));
    out->P_M = (f_rho_left * R->p - f_rho_right * L->p) * denom;
    for (int k = 0; k < 3; k++) out->fv[k] = (R->v[k] - L->v[k]) * (f_rho_left * f_rho_right * denom) + (out->P_M) * n[k];
    out->fp = (S_M * (f_rho_left * R->p + f_rho_right * L->p) + (hR - hL) * f_rho_left * f_rho_right) * denom;
}

/* guess_for_pressure, src/hydro/reimann.h:1028-1068 */
static double or_guess_p(const or_gamma *G, const or_rstate *L, const or_rstate *R, double vL, double vR, double csL, double csR)
{
    double pmin, pmax;
    double pv = 0.5 * (L->p + R->p) - 0.125 * (vR - vL) * (L->p + R->p) * (csL + csR);
    pmin = DMIN(L->p, R->p);
    pmax = DMAX(L->p, R->p);
    if (pmin <= 0) return 0.5 * (pmin + pmax);
    double qrat = pmax / pmin;
    if (qrat <= 2.0 && (pmin <= pv && pv <= pmax)) return pv;
    if (pv < pmin) {
        double pnu = (csL + csR) - G->G7 * (vR - vL);
        double pde = csL / pow(L->p, G->G1) + csR / pow(R->p, G->G1);
        return pow(pnu / pde, G->G3);
    } else {
        double gel = sqrt((G->G5 / L->rho) / (G->G6 * L->p + pv));
        double ger = sqrt((G->G5 / R->rho) / (G->G6 * R->p + pv));
        double x = (gel * L->p + ger * R->p - (vR - vL)) / (gel + ger);
        if (x < pmin || x > pmax) x = pmin;
        return x;
    }
}

/* iterative_Riemann_solver, src/hydro/reimann.h:958-1026 (TOL_ITER 1e-6, NMAX_ITER 1000) */
static int or_iterative(const or_gamma *G, const or_rstate *L, const or_rstate *R, double vL, double vR,
                        double csL, double csR, or_rout *out)
{
    double Pg, Pg_prev, W_L, W_R, Z_L, Z_R, tol, pratio;
    int niter = 0;
    double a0, a1, a2, dvel, check_vel;
    dvel = vR - vL;
    check_vel = G->G4 * (csR + csL) - dvel;
    if (check_vel < 0) return 0;
    tol = 100.0;
    Pg = or_guess_p(G, L, R, vL, vR, csL, csR);
    while ((tol > 1.e-6) && (niter < 1000)) {
        Pg_prev = Pg;
        if (Pg > L->p) {
            a0 = G->G5 / L->rho;
            a1 = G->G6 * L->p;
            a2 = sqrt(a0 / (Pg + a1));
            W_L = (Pg - L->p) * a2;
            Z_L = a2 * (1.0 - 0.5 * (Pg - L->p) / (a1 + Pg));
        } else {
            pratio = Pg / L->p;
            W_L = G->G4 * csL * (pow(pratio, G->G1) - 1);
            Z_L = 1 / (L->rho * csL) * pow(Pg / L->p, -G->G2);
        }
        if (Pg > R->p) {
            a0 = G->G5 / R->rho;
            a1 = G->G6 * R->p;
            a2 = sqrt(a0 / (Pg + a1));
            W_R = (Pg - R->p) * a2;
            Z_R = a2 * (1.0 - 0.5 * (Pg - R->p) / (a1 + Pg));
        } else {
            pratio = Pg / R->p;
            W_R = G->G4 * csR * (pow(pratio, G->G1) - 1);
            Z_R = 1 / (R->rho * csR) * pow(pratio, -G->G2);
        }
        if (niter < 1000 / 2) Pg -= (W_L + W_R + dvel) / (Z_L + Z_R);
        else Pg -= 0.5 * (W_L + W_R + dvel) / (Z_L + Z_R);
        if (Pg < 0.1 * Pg_prev) Pg = 0.1 * Pg_prev;
        tol = 2.0 * fabs((Pg - Pg_prev) / (Pg + Pg_prev));
        niter++;
    }
    if (niter < 1000) {
        out->P_M = Pg;
        out->S_M = 0.5 * (vL + vR) + 0.5 * (W_R - W_L);
        return 1;
    }
    return 0;
}

/* Riemann_solver_exact, src/hydro/reimann.h:625-668, MFM: only P_M and S_M change
   (sample_reimann_standard returns at once, the vacuum samplers return P_M = 0, :672-958). */
static unsigned or_exact(const or_gamma *G, const or_rstate *L, const or_rstate *R, double vL, double vR,
                         double csL, double csR, or_rout *out)
{
    if ((L->p == 0 && R->p == 0) || (L->rho == 0 && R->rho == 0)) {
        out->P_M = out->S_M = 0;
        return OR_RUNG_EXACT | OR_RUNG_EXACT_VACUUM;
    }
    if ((L->rho > 0) && (R->rho > 0)) {
        if (!or_iterative(G, L, R, vL, vR, csL, csR, out)) {
            /* sample_reimann_vaccum_internal at S = 0 */
            double S_L = vL + G->G4 * csL;
            double S_R = vR - G->G4 * csR;
            if (0.0 <= S_L) { out->P_M = 0; out->S_M = vL + G->G4 * csL; }
            else if (0.0 >= S_R) { out->P_M = 0; out->S_M = vR - G->G4 * csR; }
            else { out->P_M = 0; out->S_M = 0.0; }
            return OR_RUNG_EXACT | OR_RUNG_EXACT_VACUUM;
        }
        return OR_RUNG_EXACT;
    }
    if (L->rho > 0) { out->P_M = 0; out->S_M = vL + G->G4 * csL; }
    if (R->rho > 0) { out->P_M = 0; out->S_M = vR - G->G4 * csR; }
    return OR_RUNG_EXACT | OR_RUNG_EXACT_VACUUM;
}

/* Riemann_solver, src/hydro/reimann.h:282-370 (non-MHD, no EOS_GENERAL, no COOLING/GALSF),
   with HLLC_Riemann_solver (:382-396, MFM fluxes).  Returns the rungs taken; on unphysical
   input (OR_RUNG_UNPHYSICAL) P_M is set to 0 and the fluxes are left as they were. */
static void or_riemann_prep(const or_gamma *G, or_rstate *L, or_rstate *R, const double *n, double *hL, double *hR, double *vL,
                            double *vR)
{
    L->cs = sqrt(G->G0 * L->p / L->rho);
    R->cs = sqrt(G->G0 * R->p / R->rho);
    L->u = L->p / (G->G9 * L->rho);
    R->u = R->p / (G->G9 * R->rho);
    *hL = L->p / L->rho + L->u + 0.5 * (L->v[0] * L->v[0] + L->v[1] * L->v[1] + L->v[2] * L->v[2]);
    *hR = R->p / R->rho + R->u + 0.5 * (R->v[0] * R->v[0] + R->v[1] * R->v[1] + R->v[2] * R->v[2]);
    *vL = or_dot3(L->v, n);
    *vR = or_dot3(R->v, n);
}

static unsigned or_riemann(const or_gamma *G, or_rstate L, or_rstate R, const double *n, double ptl, or_rout *out)
{
    if ((L.p < 0 && R.p < 0) || (L.rho < 0) || (R.rho < 0)) { out->P_M = 0; return OR_RUNG_UNPHYSICAL; }
    double hL, hR, vL, vR;
    or_riemann_prep(G, &L, &R, n, &hL, &hR, &vL, &vR);
    double csL = L.cs, csR = R.cs;
    unsigned rung = or_wavespeeds(G, &L, &R, n, vL, vR, csL, csR, hL, hR, out, ptl);
    for (int k = 0; k < 3; k++) out->fv[k] = out->P_M * n[k];
    out->fp = out->P_M * out->S_M;
    if ((out->P_M < 0) || (isnan(out->P_M))) {
        or_kt(&L, &R, n, vL, vR, csL, csR, hL, hR, out);
        rung |= OR_RUNG_KT;
    }
    if ((out->P_M < 0) || (isnan(out->P_M)) || (out->P_M > ptl)) rung |= or_exact(G, &L, &R, vL, vR, csL, csR, out);
    return rung;
}

/* The Riemann solve of a face with the pair body's two retries
   (src/hydro/hydro_core_meshless_functions.h:351-427): the reconstructed states, then the
   first-order states (R = particle i, L = particle j, velocities in the face frame) at 1.4 times
   the pressure limiter, then zero velocities at twice it.  Returns 1, or 0 when no positive
   pressure was found (the reference's endrun(1234)); *rung receives the rungs taken. */
static int or_ladder(const or_gamma *G, or_rstate Lst, or_rstate Rst, const double *n_unit, double ptl, double P_i, double P_j,
                     double rho_i, double rho_j, const double *vi_f, const double *vj_f, or_rout *ro, unsigned *rung)
{
    memset(ro, 0, sizeof(*ro));
    *rung |= or_riemann(G, Lst, Rst, n_unit, ptl, ro);
    if ((ro->P_M < 0) || (isnan(ro->P_M)) || (ro->P_M > 1.4 * ptl)) {
        Rst.p = P_i; Lst.p = P_j;
        Rst.rho = rho_i; Lst.rho = rho_j;
        for (int k = 0; k < 3; k++) { Rst.v[k] = vi_f[k]; Lst.v[k] = vj_f[k]; }
        *rung |= OR_RUNG_RETRY1 | or_riemann(G, Lst, Rst, n_unit, 1.4 * ptl, ro);
        if ((ro->P_M < 0) || (isnan(ro->P_M))) {
            Rst.p = P_i; Lst.p = P_j;
            Rst.rho = rho_i; Lst.rho = rho_j;
            for (int k = 0; k < 3; k++) { Rst.v[k] = 0; Lst.v[k] = 0; }
            *rung |= OR_RUNG_RETRY2 | or_riemann(G, Lst, Rst, n_unit, 2.0 * ptl, ro);
            if ((ro->P_M < 0) || (isnan(ro->P_M))) { *rung |= OR_RUNG_FAIL; return 0; }
        }
    }
    return 1;
}

int orc_riemann(double gamma, const double *left /*rho,p,vx,vy,vz*/, const double *right, const double *n,
                double ptl, double *P_M, double *S_M, double *fv, double *fp)
{
    or_gamma G;
    or_gamma_init(&G, gamma);
    or_rstate L, R;
    memset(&L, 0, sizeof(L)); memset(&R, 0, sizeof(R));
    L.rho = left[0]; L.p = left[1]; L.v[0] = left[2]; L.v[1] = left[3]; L.v[2] = left[4];
    R.rho = right[0]; R.p = right[1]; R.v[0] = right[2]; R.v[1] = right[3]; R.v[2] = right[4];
    or_rout out;
    memset(&out, 0, sizeof(out));
    int rc = (or_riemann(&G, L, R, n, ptl, &out) & OR_RUNG_UNPHYSICAL) ? -1 : 0;
    *P_M = out.P_M; *S_M = out.S_M;
    fv[0] = out.fv[0]; fv[1] = out.fv[1]; fv[2] = out.fv[2];
    *fp = out.fp;
    return rc;
}

/* exact solver alone (for tests against Toro's tables) */
int orc_riemann_exact(double gamma, const double *left, const double *right, double *P_M, double *S_M)
{
    or_gamma G;
    or_gamma_init(&G, gamma);
    or_rstate L, R;
    memset(&L, 0, sizeof(L)); memset(&R, 0, sizeof(R));
    L.rho = left[0]; L.p = left[1]; L.v[0] = left[2];
    R.rho = right[0]; R.p = right[1]; R.v[0] = right[2];
    L.cs = sqrt(G.G0 * L.p / L.rho);
    R.cs = sqrt(G.G0 * R.p / R.rho);
    or_rout out;
    memset(&out, 0, sizeof(out));
    or_exact(&G, &L, &R, L.v[0], R.v[0], L.cs, R.cs, &out);
    *P_M = out.P_M; *S_M = out.S_M;
    return 0;
}

/* The Riemann functions on given face states (the device's k_riemann_probe, same layout):
   case k reads 24 values at in + 24k (left rho, p, v[3]; right rho, p, v[3]; n[3]; limiter;
   p_i, p_j, rho_i, rho_j, v_i[3], v_j[3]) and writes P_M, S_M, fv[3], fp, ok, 0 at out + 8k and
   the rung bits to rung[k].  mode 0 the ladder, 1 or_riemann, 2 or_wavespeeds, 3 or_kt,
   4 or_exact. */
int orc_riemann_probe(int mode, double gamma, long n, const double *in, double *out, unsigned *rung)
{
    or_gamma G;
    or_gamma_init(&G, gamma);
    if (mode < 0 || mode > 4) return -1;
    for (long k = 0; k < n; k++) {
        const double *c = in + 24 * k;
        or_rstate L, R;
        memset(&L, 0, sizeof(L)); memset(&R, 0, sizeof(R));
        L.rho = c[0]; L.p = c[1]; L.v[0] = c[2]; L.v[1] = c[3]; L.v[2] = c[4];
        R.rho = c[5]; R.p = c[6]; R.v[0] = c[7]; R.v[1] = c[8]; R.v[2] = c[9];
        const double nv[3] = {c[10], c[11], c[12]};
        const double ptl = c[13];
        or_rout o;
        memset(&o, 0, sizeof(o));
        unsigned rb = 0;
        int ok = 1;
        if (mode == 0) {
            ok = or_ladder(&G, L, R, nv, ptl, c[14], c[15], c[16], c[17], &c[18], &c[21], &o, &rb);
        } else if (mode == 1) {
            rb = or_riemann(&G, L, R, nv, ptl, &o);
        } else {
            double hL, hR, vL, vR;
            or_riemann_prep(&G, &L, &R, nv, &hL, &hR, &vL, &vR);
            if (mode == 2) rb = or_wavespeeds(&G, &L, &R, nv, vL, vR, L.cs, R.cs, hL, hR, &o, ptl);
            else if (mode == 3) { or_kt(&L, &R, nv, vL, vR, L.cs, R.cs, hL, hR, &o); rb = OR_RUNG_KT; }
            else rb = or_exact(&G, &L, &R, vL, vR, L.cs, R.cs, &o);
        }
        double *w = out + 8 * k;
        w[0] = o.P_M; w[1] = o.S_M; w[2] = o.fv[0]; w[3] = o.fv[1]; w[4] = o.fv[2]; w[5] = o.fp; w[6] = ok; w[7] = 0;
        rung[k] = rb;
    }
    return 0;
}

/* the rung counts of the force passes since the last reset (OR_RUNG_* order), then reset them if asked */
void orc_rung_counts(long long *out, int reset)
{
    for (int b = 0; b < OR_RUNG_COUNT; b++) {
        if (out) out[b] = or_rung_totals[b];
        if (reset) or_rung_totals[b] = 0;
    }
}

/* ------------------------------------------------------------------------- */
/* force                                                                         */
/* ------------------------------------------------------------------------- */
static inline double or_soundspeed(const orc_params *par, const orc_state *s, long i)
{
    /* effective_soundspeed, src/declarations/cell_data.h:703-712 */
    if (s->rho[i] > 0) return sqrt(par->gamma * s->press[i] / s->rho[i]);
    return sqrt(par->gamma * (par->gamma - 1.) * s->u[i]);
}

/* the pair flux (hydro_accumulate_neighbor, src/hydro/hydro_functions.h:57-546, with
   hydro_core_meshless_compute_pair, src/hydro/hydro_core_meshless_functions.h:49-616 and
   compute_finitevol_faces, src/hydro/compute_finitevol_faces_functions.h:48-168).
   Returns 0 when the pair is skipped, 1 when fluxes were produced (possibly zero), -1234
   when every Riemann fallback failed (the reference's endrun(1234)). */
static int or_force_pair(const orc_params *par, const or_gamma *G, const orc_state *s, long i, long j,
                         double sound_i, double *Fv, double *Fp, double *vsig_ts, unsigned *rung)
{
    *rung = 0;
    double box = par->box, half = 0.5 * par->box;
    Fv[0] = Fv[1] = Fv[2] = 0; *Fp = 0; *vsig_ts = 0;
    if (s->mass[j] <= 0) return 0;
    if (s->rho[j] <= 0) return 0;
    double dp[3];
    dp[0] = s->pos[3 * i + 0] - s->pos[3 * j + 0];
    dp[1] = s->pos[3 * i + 1] - s->pos[3 * j + 1];
    dp[2] = s->pos[3 * i + 2] - s->pos[3 * j + 2];
    dp[1] = or_wrap(dp[1], box, half);
    dp[0] = or_wrap(dp[0], box, half);
    dp[2] = or_wrap(dp[2], box, half);
    double r2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
    double h_i = s->h[i], h_j = s->h[j];
    if ((r2 >= h_i * h_i) && (r2 >= h_j * h_j)) return 0;
    if (r2 <= 0) return 0;
    double r = sqrt(r2);
    double rinv = 1.0 / r;
    const double *vi = &s->vel[3 * i], *vj = &s->vel[3 * j];
    double dv[3] = {vi[0] - vj[0], vi[1] - vj[1], vi[2] - vj[2]};
    double sound_j = or_soundspeed(par, s, j);
    double vsig = sound_i + sound_j;
    double hinv_i, hinv3_i, hinv4_i, hinv_j, hinv3_j, hinv4_j, u;
    double wk_i = 0, dwk_i = 0, wk_j = 0, dwk_j = 0;
    or_kernel_hinv(h_i, &hinv_i, &hinv3_i, &hinv4_i);
    if (r < h_i) { u = r * hinv_i; or_kernel_main(par->kernel, u, hinv3_i, hinv4_i, &wk_i, &dwk_i, 0); }
    or_kernel_hinv(h_j, &hinv_j, &hinv3_j, &hinv4_j);
    if (r < h_j) { u = r * hinv_j; or_kernel_main(par->kernel, u, hinv3_j, hinv4_j, &wk_j, &dwk_j, 0); }
    double cnumcrit2 = ((double)OR_CN_DANGER) * ((double)OR_CN_DANGER) - s->cn[i] * s->cn[i];
    double vdotr2 = dp[0] * dv[0] + dp[1] * dv[1] + dp[2] * dv[2];
    if (vdotr2 < 0) vsig -= 1.0 * vdotr2 * rinv;
    /* the timestep signal (src/hydro/hydro_functions.h:478-486) is taken whatever the core returns */
    *vsig_ts = vsig + fabs(sound_i - sound_j);
    double V_i = s->mass[i] / s->rho[i];
    double V_j = s->mass[j] / s->rho[j];

    /* ---- hydro_core_meshless_compute_pair ---- */
    double leak_vs_tol = 0.5 * (s->fce[i] + s->fce[j]);
    double Pressure_i = s->press[i], Pressure_j = s->press[j];
    /* compute_finitevol_faces */
    double A[3], Anorm;
    {
        double wt_i = V_i, wt_j = V_j;
        if ((fabs(V_i - V_j) / DMIN(V_i, V_j)) / OR_NDIMS > 1.25)
            wt_i = wt_j = V_i * V_j * (wk_i + wk_j) / (V_i * wk_i + V_j * wk_j);
        double mi[3], mj[3];
        or_sym_matvec(&s->nvt[6 * i], dp, mi);
        or_sym_matvec(&s->nvt[6 * j], dp, mj);
        double ci = wk_i * wt_i, cj = wk_j * wt_j;
        for (int k = 0; k < 3; k++) A[k] = (ci * mi[k] + cj * mj[k]) * (1.0 * 1.0);
        Anorm = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
        double fdp = A[0] * dp[0] + A[1] * dp[1] + A[2] * dp[2];
        if ((s->cn[j] * s->cn[j] > 1.0e12 + cnumcrit2) || (fdp < 0)) {
            double face_kernel = -(wt_i * V_i * dwk_i + wt_j * V_j * dwk_j) / r;
            face_kernel *= 1.0 * 1.0;
            for (int k = 0; k < 3; k++) A[k] = dp[k] * face_kernel;
            Anorm = face_kernel * face_kernel * r2;
        }
        Anorm = sqrt(Anorm);
    }
    if (Anorm == 0) return 1;
    double n_unit[3];
    {
        double inv = 1.0 / Anorm;
        for (int k = 0; k < 3; k++) n_unit[k] = A[k] * inv;
    }
    double s_i = 0.5 * r, s_j = -0.5 * r;
    s_i = 0 - s_i;
    s_j = 0 - s_j;
    double dist_i[3], dist_j[3];
    for (int k = 0; k < 3; k++) dist_i[k] = dp[k] * rinv;
    for (int k = 0; k < 3; k++) { dist_j[k] = dist_i[k] * s_j; dist_i[k] *= s_i; }
    double v_frame[3];
    for (int k = 0; k < 3; k++) v_frame[k] = rinv * ((-s_i) * vj[k] + s_j * vi[k]);
    double face_vel_i = or_dot3(vi, n_unit), face_vel_j = or_dot3(vj, n_unit);
    face_vel_i /= 1.0; face_vel_j /= 1.0;
    double face_area_dot_vel = rinv * (-s_i * face_vel_j + s_j * face_vel_i);
    double v2_approach = 0;
    double vdotr2_phys = vdotr2;
    vdotr2_phys *= 1.0 / (r * 1.0);
    if (vdotr2_phys < 0) v2_approach = vdotr2_phys * vdotr2_phys;
    double vdotf2_phys = face_vel_i - face_vel_j;
    if (vdotf2_phys < 0) v2_approach = DMAX(v2_approach, vdotf2_phys * vdotf2_phys);
    int recon_mode = 1;
    if (leak_vs_tol > 1) recon_mode = 0;
    double rho_i = s->rho[i], rho_j = s->rho[j];
    double P_i = Pressure_i, P_j = Pressure_j;
    or_rstate Lst, Rst;
    memset(&Lst, 0, sizeof(Lst)); memset(&Rst, 0, sizeof(Rst));
    orc_reconstruct(rho_i, &s->grad_rho[3 * i], rho_j, &s->grad_rho[3 * j], dist_i, dist_j, &Lst.rho, &Rst.rho, recon_mode);
    orc_reconstruct(P_i, &s->grad_p[3 * i], P_j, &s->grad_p[3 * j], dist_i, dist_j, &Lst.p, &Rst.p, recon_mode);
    for (int k = 0; k < 3; k++)
        orc_reconstruct(vi[k] - v_frame[k], &s->grad_v[9 * i + 3 * k], vj[k] - v_frame[k], &s->grad_v[9 * j + 3 * k],
                        dist_i, dist_j, &Lst.v[k], &Rst.v[k], recon_mode);
    double press_i_tot = Pressure_i + rho_i * v2_approach;
    double press_j_tot = Pressure_j + rho_j * v2_approach;
    double ptl = 1.1 * 1.0 * DMAX(press_i_tot, press_j_tot);
    if (recon_mode == 0)
        ptl = DMAX(ptl, DMAX(DMAX(Pressure_i, Pressure_j), 2. * DMAX(rho_i, rho_j) * v2_approach));
    or_rout ro;
    double vi_f[3], vj_f[3];
    for (int k = 0; k < 3; k++) { vi_f[k] = vi[k] - v_frame[k]; vj_f[k] = vj[k] - v_frame[k]; }
    *rung = OR_RUNG_FACE;
    if (!or_ladder(G, Lst, Rst, n_unit, ptl, Pressure_i, Pressure_j, rho_i, rho_j, vi_f, vj_f, &ro, rung)) return -1234;
    if (!((ro.P_M > 0) && (!isnan(ro.P_M)))) return 1; /* no valid solution: zero fluxes */
    ro.fp += or_dot3(v_frame, ro.fv);
    *Fp = Anorm * ro.fp;
    for (int k = 0; k < 3; k++) Fv[k] = Anorm * ro.fv[k];
    /* face entropic switch (src/hydro/hydro_core_meshless_functions.h:515-573) */
    {
        double SM_over_ceff = fabs(ro.S_M) / DMIN(sound_i, sound_j);
        if ((SM_over_ceff < 0.5) || (leak_vs_tol > 1)) {
            *rung |= OR_ENTROPIC_CONSIDERED;
            int use = 1;
            double facenorm_pm = ro.P_M * Anorm;
            double PdV_fac = ro.P_M * vdotr2_phys / 1.0;
            double PdV_i = dwk_i * V_i * V_i * s->drkern[i] * PdV_fac;
            double PdV_j = dwk_j * V_j * V_j * s->drkern[j] * PdV_fac;
            double du_old = facenorm_pm * (ro.S_M + face_area_dot_vel);
            double du_new = 0.5 * (PdV_i - PdV_j + facenorm_pm * (face_vel_i + face_vel_j));
            double cnum2 = s->cn[j] * s->cn[j];
            double ratio = 1.001;
            double to_i = Pressure_i / rho_i, to_j = Pressure_j / rho_j;
            if (use && SM_over_ceff > 1.e-3 && cnum2 < cnumcrit2) {
                if (to_i > ratio * to_j) {
                    double dtoj = -du_old + facenorm_pm * face_vel_j;
                    if (dtoj > 0) { use = 0; }
                    else { if (dtoj < 0) { if (dtoj > -du_new + facenorm_pm * face_vel_j) use = 0; } }
                } else if (ratio * to_i < to_j) {
                    double dtoi = +du_old - facenorm_pm * face_vel_i;
                    if (dtoi > 0) { use = 0; }
                    else { if (dtoi < 0) { if (dtoi > +du_new - facenorm_pm * face_vel_i) use = 0; } }
                } else {
                    double dtoj = -du_old + facenorm_pm * face_vel_j;
                    double dtoi = +du_old - facenorm_pm * face_vel_i;
                    if (dtoj < 0) { if (dtoj > -du_new + facenorm_pm * face_vel_j) use = 0; }
                    if (dtoi < 0) { if (dtoi > +du_new - facenorm_pm * face_vel_i) use = 0; }
                }
            }
            if (cnum2 >= cnumcrit2) use = 1;
            if (use) {
                *Fp += du_new - du_old;
                *rung |= OR_ENTROPIC_APPLIED;
            }
        }
    }
    return 1;
}

/* hydro_force(), src/hydro/hydro_force_loop.cc:423-480: preloop zeroing
   (hydro_force_initial_operations_preloop, src/hydro/hydro_toplevel.cc:457-565), the pair
   loop over the symmetric list (two-sided: each active particle accumulates its own side
   of every pair, src/hydro/hydro_functions.h:14-16), the writeback
   (src/hydro/hydro_force_loop.cc:227-321) and hydro_final_operations_and_cleanup
   (src/hydro/hydro_toplevel.cc:133-250).  Wakeup (src/hydro/hydro_functions.h:536-544):
   an inactive j is flagged with (timebin_i + 1) when the pair's signal speed exceeds
   WAKEUP * its MaxSignalVel.  Returns 0, or 1234 when a Riemann problem had no solution. */
int orc_force(const orc_params *par, orc_state *s, const unsigned char *active)
{
    return orc_force_ex(par, s, active, NULL);
}

/* need_wakeup (optional) is set to 1 when any inactive neighbour was flagged, as the
   reference's NeedToWakeupParticles_local (src/hydro/hydro_force_loop.cc:160-163). */
int orc_force_ex(const orc_params *par, orc_state *s, const unsigned char *active, int *need_wakeup)
{
    long n = s->n;
    double box = par->box, half = 0.5 * par->box;
    or_gamma G;
    or_gamma_init(&G, par->gamma);
    const double *P = or_pos(s), *Hp = or_hpred(s);
    double rmax = or_max_h(n, Hp, NULL);
    if (!(rmax > 0) || rmax > half) return -1;
    or_grid g;
    if (or_grid_build(&g, n, P, box, rmax)) return -2;
    int status = 0;
    /* the signal speed an inactive neighbour is tested against is its value before this call */
    double *msv_in = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    memcpy(msv_in, s->maxsig, (size_t)n * sizeof(double));
#pragma omp parallel
    {
        int *buf = (int *)malloc((size_t)n * sizeof(int));
        long long rungs[OR_RUNG_COUNT] = {0};
#pragma omp for schedule(dynamic, 64)
        for (long i = 0; i < n; i++) {
            if (active && !active[i]) continue;
            /* hydro_force_initial_operations_preloop, src/hydro/hydro_toplevel.cc:457-565 */
            if (s->wakeup) s->wakeup[i] = 0;
            s->maxsig[i] = OR_MIN_REAL;
            s->dtu[i] = 0;
            s->accel[3 * i + 0] = s->accel[3 * i + 1] = s->accel[3 * i + 2] = 0;
            double sound_i = or_soundspeed(par, s, i);
            if (!(s->mass[i] > 0 && s->rho[i] > 0)) {
                /* dead row: no pair contributes; the writeback still floors MaxSignalVel at c_s */
                if (s->maxsig[i] < sound_i) s->maxsig[i] = sound_i;
                continue;
            }
            double Acc[3] = {0, 0, 0}, dE = 0, msv = 0;
            long m0 = or_grid_candidates(&g, &s->pos[3 * i], box, buf), m = 0;
            for (long q = 0; q < m0; q++) {
                int j = buf[q];
                if (j != i && or_accept(&s->pos[3 * i], &P[3 * j], s->h[i], Hp[j], 1, box, half)) {
                    or_read(j);
                    buf[m++] = j;
                }
            }
            or_sort_int(buf, m);
            for (long q = 0; q < m; q++) {
                long j = buf[q];
                if (msv < sound_i) msv = sound_i;
                double Fv[3], Fp, vts;
                unsigned rung;
                int rc = or_force_pair(par, &G, s, i, j, sound_i, Fv, &Fp, &vts, &rung);
                for (int b = 0; b < OR_RUNG_COUNT; b++) rungs[b] += (rung >> b) & 1u;
                if (rc == -1234) {
#pragma omp atomic write
                    status = 1234;
                    continue;
                }
                if (rc == 0) continue;
                Acc[0] += Fv[0]; Acc[1] += Fv[1]; Acc[2] += Fv[2];
                dE += Fp;
                if (vts > msv) msv = vts;
                if (active && !active[j] && s->wakeup) {
                    if (vts > OR_WAKEUP * msv_in[j]) {
                        int w = (s->timebin ? s->timebin[i] : 0) + 1;
#pragma omp critical(or_wakeup)
                        { if (s->wakeup[j] < w) s->wakeup[j] = w; if (need_wakeup) *need_wakeup = 1; }
                    }
                }
            }
            s->accel[3 * i + 0] += Acc[0];
            s->accel[3 * i + 1] += Acc[1];
            s->accel[3 * i + 2] += Acc[2];
            s->dtu[i] += dE;
            {
                double max_sig = fmax(msv, sound_i);
                if (s->maxsig[i] < max_sig) s->maxsig[i] = max_sig;
            }
            /* hydro_final_operations_and_cleanup */
            {
                double vel_phys[3];
                double inv1 = 1.0 / 1.0;
                for (int k = 0; k < 3; k++) vel_phys[k] = s->vel[3 * i + k] * inv1;
                s->dtu[i] -= or_dot3(vel_phys, &s->accel[3 * i]);
                double invm = 1.0 / s->mass[i];
                for (int k = 0; k < 3; k++) s->accel[3 * i + k] *= invm;
                s->dtu[i] /= s->mass[i];
                if (s->h[i] >= 0.99 * par->max_kernel_radius) s->dtu[i] = 0;
            }
        }
        for (int b = 0; b < OR_RUNG_COUNT; b++) {
#pragma omp atomic
            or_rung_totals[b] += rungs[b];
        }
        free(buf);
    }
    free(msv_in);
    or_grid_free(&g);
    return status;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
