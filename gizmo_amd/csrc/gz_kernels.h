// gz_kernels.h -- kernel declarations and the small views the kernels take by value.
#pragma once
#include "gz_device.h"

namespace gz {

constexpr int kFineLevel = 10;                 // 1024^3 Morton grid
constexpr int kFine = 1 << kFineLevel;
constexpr int kXcds = 8;                       // MI355X: 8 XCDs, each with its own L2

// Workgroups are handed to the XCDs round-robin (workgroup b runs on XCD b % 8).  A grid of
// 8*q workgroups is remapped so that XCD x processes the contiguous logical range
// [x*q, (x+1)*q): neighbouring particles (adjacent in Morton order) then share one L2.
// Launch with xcd_grid(nblocks) workgroups; logical blocks past the real count exit.
__device__ __forceinline__ long xcd_block()
{
    const long b = blockIdx.x, q = gridDim.x / kXcds;
    return (b % kXcds) * q + b / kXcds;
}
inline unsigned xcd_grid(long nblocks) { return (unsigned)(((nblocks + kXcds - 1) / kXcds) * kXcds); }

// Search over kernel radii that span many grid levels.  Every cell of every level 0..lmax is a
// contiguous range of the Morton order; `start`/`end` hold each cell's range (level L from
// entry loff[L]).  An entry is used only if it is exactly a maximal run of the cell's code, so
// entries of cells that are empty now need no clearing.  Each particle also has its own level
// (the finest whose cells hold its kernel radius plus twice the displacement bound); the
// particles of one level, in ascending index (list `lidx`, slice [lofs[L], lofs[L+1])), have
// tables of their own (`lstart`/`lend`, positions in lidx), so a row can reach the few
// particles of coarser levels whose radius covers it without walking the fine ones.
struct WideGrid {
    const uint32_t *key;       // Morton keys of the order
    int n, lmax;
    int *start, *end;          // (written by the table builds only)
    const int *lidx;
    int *lstart, *lend;
    long long loff[kFineLevel + 1];
    int lofs[kFineLevel + 2];
};

struct GridView {
    int level, shift;         // nc = 2^level cells per side; cell code = key >> shift
    double cell, box, half;
    const int4 *sten;         // 27 entries per cell: (start, end, offset code, cell code)
    const int *nsten;
    const double *chmax;      // largest kernel radius in each cell (when computed; else hall bounds every radius)
    double hall;
    double margin;            // largest distance a particle has moved from its cell since the ordering
    const int *cstart, *cend; // [start, end) of each cell
    long long cell_off;       // first cell of the box in the cell arrays (batches of boxes)
    const WideGrid *wide;     // device copy; non-null: per-row levels over the tables above
};

// One periodic box of a batch: geometry, grid and timeline.  A batch holds many independent
// boxes in one particle array (box b in its own index range, each in its own Morton order);
// the row and timeline kernels look the box of each particle up here.
struct BoxDev {
    double box, half, min_h, max_h;      // side, All.MinKernelRadius, All.MaxKernelRadius
    double cell, margin, rowlim;         // grid cell side, displacement bound, radius limit of density rows
    int level, shift;
    long long cell_off;                  // first cell of the box in the cell arrays
    double time, time_begin, tbi;        // All.Time, All.TimeBegin, All.Timebase_interval
    long long ti_current, ti_next;       // All.Ti_Current and the next sync point
    unsigned long long mask, mask_next;  // TimeBinActive now and at the next sync point
    int stepping;                        // 1: the box takes part in the current batch step
    int need_wakeup;                     // NeedToWakeupParticles of the box (set by its force pass)
    double courant_fac, err_tol, max_dt, min_dt, min_egy;  // the box's All.CourantFac ... All.MinEgySpec
};

struct BoxTable {
    BoxDev *b;                // nullptr: one box, held by the kernel arguments
    const int *BOX;           // box of each particle
    unsigned long long *slot; // per box, kBoxSlots running maxima (bits of positive doubles)
    int nbox;
};
// the per-box maxima: displacement since the ordering, radius after the drift, radius of the
// active rows, radius of density rows that outgrew their cell, final radius of density rows
// (and the particles woken, and the controlled-stop code of the box's timeline)
enum { kSlotDisp = 0, kSlotHmax = 1, kSlotHrows = 2, kSlotHgrow = 3, kSlotHfinal = 4, kSlotWoken = 5, kSlotStop = 6, kBoxSlots = 8 };

// the parameters of particle i's box over the single-box ones
__device__ __forceinline__ void resolve_box(const BoxTable &bt, int i, DevParams &P, GridView &g)
{
    if (!bt.b) return;
    const BoxDev &B = bt.b[bt.BOX[i]];
    P.box = B.box; P.half = B.half; P.min_kernel_radius = B.min_h; P.max_kernel_radius = B.max_h;
    P.time = B.time; P.time_begin = B.time_begin;
    g.level = B.level; g.shift = B.shift; g.cell = B.cell; g.box = B.box; g.half = B.half;
    g.margin = B.margin; g.cell_off = B.cell_off;
}

struct CellRows {             // occupied cells of the active rows
    const int *cells;         // cell codes, ascending
    int ncells;               // launch bound: the count, or an upper bound when ncells_dev is set
    const int *ncells_dev;    // the count on the device (nullptr: ncells is exact)
    const int *ca0, *ca1;     // per cell code: [first, last+1) slot of its active rows
    int nrows;                // number of active rows (column stride of per-row sums)
    // Groups (gsten non-null; search over many levels, batches of boxes): wavefront cb walks
    // group cb, the rows grow[gstart[cb] .. gstart[cb+1]) (slots) sharing one cell of one grid:
    // level glev[cb], code gcode[cb], stencil gsten[27*cb ..) of gnst[cb] ranges ascending
    const int *grow, *gstart, *glev, *gcode, *gnst;
    const int4 *gsten;
    BoxTable bt;              // batches: the group's box geometry
};

struct DensityIter {          // per active slot
    double *hcur, *hbuild, *left, *right, *cn_cur;
    int *cn_init, *iter;
};

struct DensityOut {           // per particle, device order; vectors SoA with column stride n
    long n;
    double *h, *nn, *drk, *divv, *rho, *p, *t, *fw, *cn, *fce;
};

struct ForceView {            // per particle, device order; vectors SoA with column stride n
    long n;
    const double *X, *Y, *Z, *VX, *VY, *VZ, *M, *H, *U, *RHO, *PR, *CN, *FCE, *DRK, *GR, *GP, *GV, *T;
    unsigned long long *rungs;  // kRungCount face counts (the force kernels built to count them)
};

// Adds each RiemannRung bit of the calling lanes' faces to its count: one atomic per rung
// that any of them took (the lanes of the wavefront that call it together)
__device__ __forceinline__ void count_rungs(unsigned long long *__restrict__ cnt, unsigned bits)
{
    const unsigned long long live = __ballot(1);
    const int leader = __ffsll((long long)live) - 1;
    for (int b = 0; b < kRungCount; b++) {
        const unsigned long long m = __ballot((bits >> b) & 1u);
        if ((int)(threadIdx.x & 63) == leader && m) atomicAdd(&cnt[b], (unsigned long long)__popcll(m));
    }
}

// the timestep loop: parameters of one step and the particle fields it reads and writes
constexpr int kTimebins = 60;  // TIMEBINS, src/declarations/typedefs.h:8

struct TimeParams {
    double courant_fac, err_tol, max_dt, min_dt, min_egy;  // All.CourantFac ... All.MinEgySpec
    double core_size;                                      // KERNEL_CORE_SIZE
    double tbi;                                            // All.Timebase_interval
    long long ti_current;                                  // All.Ti_Current
    unsigned long long active_mask;                        // TimeBinActive, bit per bin
    double min_h, max_h, gamma, box;
    // the drift of a step (single box): every particle when more than full_frac of them are in
    // the active bins (All.DomainBuild_ActiveFraction, src/core/run.cc:360,380-390), when more
    // than lazy_rows are, or when lazy_ok is 0; otherwise only the active bins, the others when
    // a pass reads them
    double full_frac;
    long long n_total, lazy_rows;
    int lazy_ok;
};

// the timeline of particle i's box over the single-box one
__device__ __forceinline__ void resolve_time(const BoxTable &bt, int i, TimeParams &T)
{
    if (!bt.b) return;
    const BoxDev &B = bt.b[bt.BOX[i]];
    T.tbi = B.tbi; T.ti_current = B.ti_current; T.active_mask = B.mask;
    T.min_h = B.min_h; T.max_h = B.max_h; T.box = B.box;
    T.courant_fac = B.courant_fac; T.err_tol = B.err_tol; T.max_dt = B.max_dt; T.min_dt = B.min_dt; T.min_egy = B.min_egy;
}

struct ParticleView {         // device order; vectors SoA with column stride n
    long n;
    double *X, *Y, *Z, *VX, *VY, *VZ, *VKX, *VKY, *VKZ, *M, *H, *U, *UK, *NN, *DIVV, *RHO, *PR, *ACC, *DTU, *MSV, *SOFT;
    int *TB, *WAKE;
    long long *TIB, *DTS;
    // P.Ti_current of each particle (the time its predicted fields stand at; nullptr: every
    // particle at All.Ti_Current, the batches' eager drift), and the bits of a bound on every
    // kicked velocity component |P.Vel| (raised by the kicks; nullptr: not kept)
    long long *TIC = nullptr;
    unsigned long long *vkmax = nullptr;
};

__global__ void k_find_timesteps(TimeParams T, ParticleView S, int nrows, const int *act, int *err, unsigned long long *count,
                                 BoxTable bt, int kick_first);
__global__ void k_bin_hist(long n, const int *TB, unsigned long long *count, BoxTable bt);
__global__ void k_wakeup_limits(TimeParams T, const unsigned long long *count, int *lim, BoxTable bt);
__global__ void k_wakeups(TimeParams T, ParticleView S, const int *lim, int wakeup_bin_offset, unsigned long long *count,
                          int *nwoken, BoxTable bt);
// prep (may be null): the counters and maxima of the coming sync point, cleared by block 0 as
// k_step_prep would (a run of several sync points then needs no launch for it)
__global__ void k_halfstep_kick(TimeParams T, ParticleView S, int nrows, const int *act, int second, BoxTable bt,
                                int *prep_counters = nullptr, unsigned long long *prep_u64c = nullptr);
__global__ void k_drift(TimeParams T, ParticleView S, long long time0, const unsigned long long *tl, const double *X0,
                        const double *Y0, const double *Z0, unsigned long long *disp, unsigned long long *hmax, BoxTable bt);
__global__ void k_step_prep(int *counters, unsigned long long *u64c);
__global__ void k_next_sync(TimeParams T, const unsigned long long *count, unsigned long long *tl, BoxTable bt);
__global__ void k_active_flags(long n, const unsigned long long *tl, const int *TB, const double *M, const double *H, double hclamp,
                               uint8_t *ACT, double *SOFT, int *flag, unsigned long long *hmax_act, BoxTable bt);
__global__ void k_compact_count(long n, const int *flag, const int *scan, int *act, int *count_out);
constexpr int kActChunk = 2048;  // particles per block of the active-list kernels (and of the drift they are fused with)
// count (may be null): the bin occupation, from which every block takes the next sync point
// itself (k_next_sync's work; block 0 writes it to tl) instead of reading tl
__global__ void k_drift_active(TimeParams T, ParticleView S, long long time0, unsigned long long *tl,
                               unsigned long long *disp, unsigned long long *hmax,
                               double hclamp, uint8_t *ACT, unsigned long long *bits, int *bcount, unsigned long long *hmax_act,
                               const unsigned long long *count = nullptr);
// the start of density()'s radius iteration for each row of a new active list (as k_density_init
// does it; H null: not done here)
struct DensityStart {
    double *H;
    double hclamp;
    double *hcur, *left, *right, *cn_cur;
    int *cn_init, *iter;
    uint8_t *need;
};
__global__ void k_active_compact(long n, int nblocks, const unsigned long long *bits, const int *bcount, int *act, int *count_out,
                                 DensityStart ds);
// the reference's lazy drift (src/core/predict.cc:175-210, src/mesh/gpu_neighbor_list.cc:2443-2510):
// every particle of the 27 grid cells around each active row's home cell that stands behind
// time1 is drifted to it before a pass reads the cells (the cells listed once each, then one
// workgroup of 256 per cell); every particle behind time1 (before a re-ordering or a
// snapshot); the bound on the kicked velocities
__global__ void k_stencil_cells(GridView g, int nrows, const int *act, const uint32_t *key, int *cmark, int *cells, int *ncells);
__global__ void k_drift_cells(TimeParams T, ParticleView S, GridView g, const int *cells, const int *ncells, int *cmark,
                              long long time1, unsigned long long *hmax);
__global__ void k_drift_behind(TimeParams T, ParticleView S, long long time1, unsigned long long *hmax);
__global__ void k_vk_max(ParticleView S);
__global__ void k_advance_boxes(BoxTable bt);
__global__ void k_active_from_bins(long n, unsigned long long mask, const int *TB, const double *M, const double *H,
                                   uint8_t *ACT, double *SOFT);
__global__ void k_sum_counts(int nrows, const int *cnt, unsigned long long *out, const int *gate = nullptr);
// groups of rows sharing one cell of one grid (the search over many levels, batches of boxes)
__global__ void k_row_group_key(int nrows, const int *act, const uint8_t *need, const uint32_t *key, const double *hslot,
                                const double *H, double box, double margin, const WideGrid *W, BoxTable bt, uint32_t *gkey, int *val);
__global__ void k_group_heads(int nrows, const uint32_t *skey, int *flag);
__global__ void k_group_list(int nrows, const uint32_t *skey, const int *flag, const int *scan, int *gstart, uint32_t *gkey, int *ngroups);
__global__ void k_group_stencil(GridView g, int ngroups, const uint32_t *gkey, const int *gstart, const int *grow, const int *act,
                                BoxTable bt, int *glev, int *gcode, int *gnst, int4 *gsten);

// search over many grid levels
__global__ void k_wide_bounds(WideGrid W);
__global__ void k_particle_level(int n, const double *H, double box, double margin, int lmax, int *plev, int *level_count);
__global__ void k_wide_list_bounds(WideGrid W, const int *plev);
__global__ void k_build_sym_wide(GridView g, int nrows, const int *act, const double *X, const double *Y, const double *Z,
                                 const double *H, int *ell, long ld, int cap, int *cnt, int *overflow);

// batches of independent boxes
__global__ void k_batch_keys(long n, const double *X, const double *Y, const double *Z, const int *BOX, const BoxDev *tab,
                             const int *rank, const uint8_t *due, uint32_t *key, unsigned long long *key64, int *val);
__global__ void k_batch_cell_bounds(long n, const uint32_t *key, const int *BOX, const BoxDev *tab, int retired, int *cstart, int *cend);
__global__ void k_batch_flag_boxes(long n, const int *BOX, const uint8_t *sel, const double *M, int *flag, uint8_t *ACT);
__global__ void k_batch_flag_active(long n, const int *BOX, const BoxDev *tab, const uint8_t *ACT, const double *M, int *flag);
__global__ void k_batch_box_hmax(long n, const double *H, const uint8_t *sel, BoxTable bt);
__global__ void k_batch_start(long n, const int *BOX, const uint8_t *sel, ParticleView S);
__global__ void k_batch_soft(long n, const int *BOX, const uint8_t *sel, const double *H, double *SOFT);
__global__ void k_batch_retire(long n, int *BOX, const uint8_t *sel, int retired, double *M, uint8_t *ACT);
__global__ void k_fill_int(long n, int *a, int v);
__global__ void k_gather_ll(long n, const int *perm, const long long *in, long long *out);
__global__ void k_scatter_ll(long n, const int *perm, const long long *in, long long *out);

__global__ void k_morton(long n, const double *pos, double box, uint32_t *key, int *val);
__global__ void k_morton_soa(long n, const double *X, const double *Y, const double *Z, double box, uint32_t *key, int *val);
__global__ void k_permute_cols(long n, int width, const int *order, const unsigned long long *in, unsigned long long *out);
__global__ void k_permute_int(long n, const int *order, const int *in, int *out);
__global__ void k_permute_u8(long n, const int *order, const uint8_t *in, uint8_t *out);
constexpr int kPermCols = 16;
struct PermCols {
    const unsigned long long *in[kPermCols];
    unsigned long long *out[kPermCols];
    int ncol;
};
__global__ void k_permute_many(long n, const int *order, PermCols P);
__global__ void k_gather(long n, int width, const int *perm, const double *in, double *out, int soa);
__global__ void k_gather_int(long n, const int *perm, const int *in, int *out);
__global__ void k_gather_u8(long n, const int *perm, const uint8_t *in, uint8_t *out);
__global__ void k_scatter(long n, int width, const int *perm, const double *in, double *out, int soa);
__global__ void k_scatter_int(long n, const int *perm, const int *in, int *out);
__global__ void k_flag_active(long n, const uint8_t *active, const double *mass, int *flag);
__global__ void k_compact(long n, const int *flag, const int *scan, int *act);
__global__ void k_cell_clear(long ncell, int *cstart, int *cend, double *chmax);
__global__ void k_cell_bounds(long n, const uint32_t *key, int shift, int *cstart, int *cend);
__global__ void k_cell_hmax(long ncell, const int *cstart, const int *cend, const double *h, double *chmax);
__global__ void k_cell_stencil(int level, const int *cstart, const int *cend, int4 *sten, int *nsten);

template <bool SYMMETRIC>
__global__ void k_build_ell(int nrows, const int *rows, const int *act, const double *hrow, double scale, const double *X,
                            const double *Y, const double *Z, const double *H, const uint32_t *key, GridView g, int *ell,
                            long ld, int cap, int *cnt, int *overflow);

__global__ void k_cell_rows_mark(int nact, const int *act, const uint32_t *key, int shift, int *ca0, int *ca1, int *flag);
__global__ void k_cell_rows_list(int nact, const int *act, const uint32_t *key, int shift, const int *flag, const int *scan,
                                 int *cells);
__global__ void k_build_sym_cells(GridView g, CellRows cr, const int *act, const double *X, const double *Y, const double *Z,
                                  const double *H, int *ell, long ld, int cap, int *cnt, int *overflow);

template <int KERN, typename R>
__global__ void k_density_cells(DevParams P, GridView g, CellRows cr, const int *act, const uint8_t *need, const double *X,
                                const double *Y, const double *Z, const double *VX, const double *VY, const double *VZ,
                                const double *M, const double *hcur, double *accout);
__global__ void k_density_init(int na, const int *act, double *H, double hclamp, double *hcur, double *left, double *right,
                               double *cn_cur, int *cn_init, int *iter, uint8_t *need, int *counters, unsigned long long *hmax,
                               unsigned long long *hfinal, BoxTable bt);
// gate (may be null): the last block to finish writes the speculative pass's gate word from the
// round's counters (done: a zero counter the blocks take turns on, left zero)
__global__ void k_density_after(DevParams P, int nrows, const int *act, uint8_t *need, const double *accin, const double *M,
                                const double *U, double cell, DensityIter it, DensityOut out, int *nnext, int *rebuild, int *err,
                                unsigned long long *hfinal, BoxTable bt, double hlim = 0, int *gate = nullptr,
                                unsigned int *done = nullptr, const int *rows = nullptr, int nlist = 0, int *next_rows = nullptr);

template <typename R>
__global__ void k_pack_grad(long n, const double *H, const double *M, const double *RHO, const double *PR, const double *VX,
                            const double *VY, const double *VZ, GradRec<R> *rec, const double *X, const double *Y, const double *Z,
                            double box, BoxTable bt);
template <typename R>
__global__ void k_pack_force(double gamma, ForceView F, Cell<R> *rec, int *zero2, double box, BoxTable bt);
template <typename R>
__global__ void k_pack_grad_list(int nrows, const int *act, const int *ell, long ld, const int *cnt, const double *H, const double *M,
                                 const double *RHO, const double *PR, const double *VX, const double *VY, const double *VZ,
                                 GradRec<R> *rec, const double *X, const double *Y, const double *Z, double box, BoxTable bt);
template <typename R>
__global__ void k_pack_force_list(int nrows, const int *act, const int *ell, long ld, const int *cnt, double gamma, ForceView F,
                                  Cell<R> *rec, int *zero2, double box, BoxTable bt);

// BOXES: each row's box geometry from bt (batches); else the single box of the arguments
template <int KERN, typename R, bool BOXES>
__global__ void k_gradients(DevParams P, int nrows, const int *act, const double *X, const double *Y, const double *Z,
                            const double *H, const GradRec<R> *rec, const double *T, const double *CN, long n, const int *ell, long ld,
                            const int *cnt, double *GR, double *GP, double *GV, BoxTable bt);

// one workgroup per active row, for passes over few rows: kRowThreads lanes (4 wavefronts)
// evaluate that many candidates or list entries at once, then lanes of the first wavefront add
// the accepted terms in order
constexpr int kRowThreads = 256;
// TPB threads per row: kRowThreads when few rows run (each row's rounds are latency-bound),
// 64 when many do (lanes past a row's last candidate would idle)
template <int KERN, typename R, int TPB>
__global__ void k_density_rows(DevParams P, GridView g, int nrows, const int *act, const uint8_t *need, const uint32_t *key,
                               const double *X, const double *Y, const double *Z, const double *VX, const double *VY,
                               const double *VZ, const double *M, const double *hcur, double *accout, BoxTable bt,
                               int *npairs = nullptr, const int *rows = nullptr, int nlist = 0);
__global__ void k_build_sym_rows(GridView g, int nrows, const int *act, const uint32_t *key, const double *X, const double *Y,
                                 const double *Z, const double *H, int *ell, long ld, int cap, int *cnt, int *overflow, BoxTable bt);
// the fields a gradient record is made of (H null: the pass reads packed records)
struct GradSrc {
    const double *H, *M, *RHO, *PR, *VX, *VY, *VZ;
    double box;
};
// The gradient and force passes over few rows walk lists of ~1-2x the neighbour number:
// kListThreads lanes per row take a list of up to that many entries in one round, and twice as
// many rows fit on the chip at once as with kRowThreads (the row kernels' registers allow 2-3
// wavefronts per SIMD)
constexpr int kListThreads = 128;
template <int KERN, typename R, int TPB = kListThreads>
__global__ void k_gradients_rows(DevParams P, int nrows, const int *act, const double *X, const double *Y, const double *Z,
                                 const double *H, const GradRec<R> *rec, GradSrc src, const double *T, const double *CN, long n, const int *ell,
                                 long ld, const int *cnt, double *GR, double *GP, double *GV, BoxTable bt,
                                 int *force_flags = nullptr);
// STATS: the force kernels also count the RiemannRung bits of their faces into F.rungs.  MAKE:
// the kernel makes the records of its row and list entries from the state instead of reading
// rec (few rows among many particles: no packing launch; more registers, fewer waves)
template <int KERN, typename R, bool STATS = false, int TPB = kListThreads, bool MAKE = false>
__global__ void k_force_rows(DevParams P, Gam<R> G, int nrows, const int *act, ForceView F, const Cell<R> *rec, const int *ell,
                             long ld, const int *cnt, const uint8_t *active, const int *timebin, const double *msv_in, int *wakeup,
                             double *ACC, double *DTU, double *MSV, int *err, int *need_wakeup, unsigned long long *pairs,
                             BoxTable bt, const int *gate = nullptr);

// the rows, lists and outputs of a force pass's lane-per-row kernels
struct ForceRows {
    int nrows;
    const int *act;              // the rows' particles
    const int *ell;              // symmetric list, column-major with leading dimension ld
    long ld;
    const int *cnt;              // list lengths
    const uint8_t *active;       // active flags (nullptr: every particle active)
    const int *timebin;          // time bins (the wake-up bin of a flagged neighbour)
    const double *msv_in;        // MaxSignalVel before the pass (the wake-up test)
    int *wakeup;                 // wake-up flags (nullptr: none)
    double *ACC, *DTU, *MSV;     // outputs
    int *err;                    // err[0] stop code; err[10], err[11] the first pair without a solution; err[12] rare rows
    int *need_wakeup;            // set when a neighbour is flagged (one box)
    unsigned long long *pairs;   // the pass's pair count (nullptr: not counted)
    const int *sched;            // row order (nullptr: 0..nrows-1)
    int *rare;                   // rows the common pass leaves to the whole ladder
};

// Lane per row.  STATS: count the faces' RiemannRung bits into F.rungs.  COMMON: every face takes
// only the Riemann ladder's common case (riemann_common); a row with a face that needs more is
// not written but appended to a.rare (count in a.err[12]), for k_force_rare.  The common pass
// then carries no register cost of the rare rungs.
template <int KERN, typename R, bool BOXES, bool STATS = false, bool COMMON = false>
__global__ void k_force(DevParams P, Gam<R> G, ForceView F, const Cell<R> *rec, ForceRows a, BoxTable bt);
// k_force<KERN, double, false, false, true> at three waves per SIMD instead of two (option "force_waves")
template <int KERN>
__global__ void k_force_w3(DevParams P, Gam<double> G, ForceView F, const Cell<double> *rec, ForceRows a, BoxTable bt);

// the rows that k_force<..., COMMON> listed, each with the whole ladder (a fixed grid strides over
// the count the device holds; the rows' results are the single-pass kernel's, bit for bit)
constexpr unsigned kRareBlocks = 1024;  // workgroups of 256 at most that k_force_rare launches
// FP32 force pass over many rows: above this share of rows left to k_force_rare by the last
// pass, the next pass runs one kernel with the whole ladder (the two kernels would evaluate
// most rows twice)
constexpr double kRareShareSingle = 0.25;
// few-row force passes over at most this many rows make their records in the row kernel (its
// MAKE form holds one wave per SIMD: this many rows of two wavefronts fill the chip once)
constexpr int kMakeForceRows = 512;
template <int KERN, typename R, bool BOXES>
__global__ void k_force_rare(DevParams P, Gam<R> G, ForceView F, const Cell<R> *rec, ForceRows a, BoxTable bt);

// the Riemann functions on given face states (diagnostic; layout of in/out at the definition)
constexpr int kProbeIn = 24, kProbeOut = 8;
template <typename R>
__global__ void k_riemann_probe(int mode, Gam<R> G, long n, const double *in, double *out, unsigned *rung);

// the order in which the force pass's lanes take the rows: within each block of span rows,
// longest list first (a wavefront runs as long as its longest row), when the waves of the plain
// order are uneven enough (list sum over 64 x the sum of the waves' longest lists below
// min_even; the scattered reads of the sorted order cost more than even waves save otherwise)
__global__ void k_sched_stats(int nrows, const int *cnt, unsigned long long *partial);
__global__ void k_sched_total(int nparts, const unsigned long long *partial, unsigned long long *total);
constexpr int kSchedMaxSpan = 2048;
__global__ void k_sched_sort(int nrows, int span, const int *cnt, const unsigned long long *total, double min_even, int *sched);

}  // namespace gz
