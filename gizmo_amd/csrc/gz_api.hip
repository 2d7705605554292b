// gz_api.hip -- host orchestration and the C ABI (include/gizmo_hydro.h).
//
// One context = one periodic box resident on one GPU.  The host side follows the
// reference's stage order: density() builds a one-way list at 1.3x the kernel radius and
// iterates the radius on the device, rebuilding only the rows whose radius outgrows the
// buffer (src/hydro/density_loop.h:316, src/mesh/neighbor_loop_runner.cc:6060-6110);
// hydro_gradient_calc() builds the symmetric list once and hydro_force() reuses it, as the
// reference's hydro corridor does (src/hydro/hydro_corridor.cc:195).
#include <hip/hip_runtime.h>
#include "gz_prim.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/gizmo_hydro.h"
#include "gz_kernels.h"
#include "gz_groups.h"

using namespace gz;

__global__ void k_iota(int n, int *out);

namespace {

constexpr int kBlock = 256;
inline unsigned nblk(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }
// steps with more active particles than this drift every particle (the cells around that many
// rows cover most of a box: one sweep costs less than walking them)
constexpr long long kLazyMaxRows = 65536;
// counters[kPassGate]: the gate word of a speculative first density round (written by the round's closing launch)
constexpr int kPassGate = 15;

}  // namespace

struct gz_context {
    gz_params par{};
    DevParams dp{};
    int device = 0;
    hipStream_t stream = nullptr;
    long cap = 0, n = 0;
    char err[512] = {0};  // the last error message (a fixed buffer: reporting an error allocates nothing)

    // particle state, device order (SoA; vectors with column stride n)
    double *X = nullptr, *Y = nullptr, *Z = nullptr, *VX = nullptr, *VY = nullptr, *VZ = nullptr;
    double *M = nullptr, *H = nullptr, *U = nullptr, *NN = nullptr, *DRK = nullptr, *DIVV = nullptr;
    double *RHO = nullptr, *PR = nullptr, *T = nullptr, *FW = nullptr, *CN = nullptr, *FCE = nullptr;
    double *GR = nullptr, *GP = nullptr, *GV = nullptr, *ACC = nullptr, *DTU = nullptr, *MSV = nullptr;
    int *TB = nullptr, *WAKE = nullptr;
    uint8_t *ACT = nullptr;   // active flag per particle (device order)
    // timestep-loop fields: kicked velocity and energy, softening cache, step start and length
    double *VKX = nullptr, *VKY = nullptr, *VKZ = nullptr, *UK = nullptr, *SOFT = nullptr;
    long long *TIB = nullptr, *DTS = nullptr;
    long long *TIC = nullptr;  // P.Ti_current: the time each particle's predicted fields stand at
    // positions when the particles were last ordered, and the largest periodic distance any
    // particle has moved since (the search grid widens its cells by twice that)
    // dmax: a bound on how far any particle moved since the last ordering (the sum of each
    // step's longest drift along an axis)
    double dmax = 0, last_rmax = 0;
    bool overflow_clear = false;  // counters[3] (the symmetric list's overflow) known to be zero on the stream
    // the step in progress drifts lazily: the passes drift the cells around the active rows that
    // they read (the particles there that stand behind All.Ti_Current) after building their grid
    bool lazy = false;
    bool last_full = true;  // the last step drifted every particle
    bool density_ready = false;   // the active list's launch set density()'s iteration state for its rows
    // run-loop steps with few rows launch the passes behind the first density round without
    // waiting for it (speculate: allowed; spec: in flight, settled by the force pass's read)
    bool speculate = false, spec = false;
    bool spec_enabled = true;  // option "speculate"
    bool fd = false;           // option "fd_correction": the first-derivative volume correction of the density
    int force_split = -1;      // option "force_split": -1 by the last pass's rare share, 0 one kernel, 1 two kernels
    int force_waves = 0;       // option "force_waves": the FP64 common force kernel's waves per SIMD (0: by wide_search)
    double rare_share = 0;     // share of the rows the last two-kernel force pass left to k_force_rare
    bool split_ran = false;
    bool force_flags_clear = false;  // the last gradient launch cleared the force pass's flags
    bool in_run_steps = false;       // inside gz_run_steps' loop
    bool prepped = false;            // the last step's closing kick cleared this step's counters
    double resort_fraction = 0.125;  // order again once dmax exceeds this fraction of the search radius
    int few_rows = 8192;             // at most this many active rows: one wavefront per row
    // the cell ranges and stencils depend only on the order (key_version) and the level: they
    // are kept across calls while both stay the same
    long key_version = 0, grid_version = -1;
    // an upper bound on every kernel radius (< 0: unknown, reduce over all particles)
    double hmax_bound = -1;
    double hmax_rows = -1;  // the largest clamped radius of the active rows, when already known
    // passes over few rows pack only the records of the rows and their list entries
    bool pack_list = false;
    bool all_active = true;
    int *perm = nullptr;      // device order -> caller index
    uint32_t *key = nullptr;  // fine Morton key, device order

    // search over many grid levels (kernel radii spanning orders of magnitude): per-level cell
    // tables of the order and of each level's own particles
    int wide = 0;
    WideGrid wide_h{};
    WideGrid *wide_dev = nullptr;
    int *wstart = nullptr, *wend = nullptr, *wlstart = nullptr, *wlend = nullptr, *lidx = nullptr, *plev = nullptr, *lcount = nullptr;
    long long wcells_alloc = 0;
    long wide_key_version = -1;
    int wide_lmax = -1;
    // the symmetric list of many rows over many levels: rows grouped by the cell of their own level
    RowGroups grp;
    // the force pass's row order over the lanes (longest lists first within blocks of rows)
    int *sched = nullptr;
    int *rare = nullptr;  // rows the force pass's common kernel leaves to the whole Riemann ladder
    // the rows still iterating in the density rounds of the row kernels: rl_a holds the coming
    // round's rows (hcounters[0] of them) when rl_valid, rl_b receives the next round's
    int *rl_a = nullptr, *rl_b = nullptr;
    bool rl_valid = false;
    unsigned long long *sched_part = nullptr;
    int sched_span = 2048;  // rows per sorted block (0: plain order)
    bool sched_valid = false;

    // staging and scratch
    double *stage = nullptr;  // n*9 doubles
    std::vector<double *> spare[10];  // by width: capacity x width buffers a re-ordering gathers into
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    uint32_t *key2 = nullptr;
    int *ival = nullptr, *ival2 = nullptr, *iflag = nullptr, *iscan = nullptr;

    // active rows
    int *act = nullptr;       // slot -> device index
    int nact = 0;
    unsigned long long *act_bits = nullptr;  // the timestep loop's active list: row flags per block
    int *act_bcount = nullptr;               // and rows per block

    // grid
    int level = -1;
    double cell = 0;
    int *cstart = nullptr, *cend = nullptr, *nsten = nullptr;
    int4 *sten = nullptr;
    double *chmax = nullptr;
    long grid_cells_alloc = 0;
    int *cmark = nullptr, *dcells = nullptr;  // the lazy drift's cell marks (zero between calls) and its cell list

    // lists
    int *ell_d = nullptr, *cnt_d = nullptr;  // density (one-way, buffered)
    int cap_d = 0;
    long ld_d = 0;
    int *ell_s = nullptr, *cnt_s = nullptr;  // symmetric
    size_t ell_s_alloc = 0;                  // entries allocated at ell_s (at least cap_s * ld_s)
    // the lazy drift of the rows' stencils: (time, active list, grid level, order) of the last
    // one, and whether the listing counter is known to be clear
    long long stencil_drifted[4] = {-1, -1, -1, -1};
    bool ncells_clear = false;
    int cap_s = 0;
    long ld_s = 0;
    bool sym_valid = false;
    // inside run_step's hydro pass the symmetric list's overflow is read back with the force
    // pass's counters (one wait fewer); on overflow both passes run again on a wider list
    bool defer_sym_check = false, sym_check_pending = false;
    bool density_done = false, gradients_done = false;

    // occupied cells of the active rows at the current grid level
    int *ca0 = nullptr, *ca1 = nullptr, *cells = nullptr, *cflag = nullptr, *cscan = nullptr;
    int ncells = 0, rows_level = -1;
    bool ncells_exact = true;
    long act_version = 0, rows_act_version = -1, rows_key_version = -1;
    uint8_t *need = nullptr;

    // per-pass neighbour records
    void *rec_grad = nullptr, *rec_force = nullptr;
    double *dacc = nullptr;    // density pair sums of a round: 13 columns of stride nact

    // density iteration state (per slot)
    double *it_h = nullptr, *it_hb = nullptr, *it_l = nullptr, *it_r = nullptr, *it_cn = nullptr;
    int *it_ci = nullptr, *it_it = nullptr, *todo = nullptr, *todo2 = nullptr;
    int *counters = nullptr;   // device: [0] next count, [1] rebuild flag, [2] error, [3] overflow,
                               // [4] wake-up flagged, [5] woken count, [6] timestep stop code,
                               // [8..9] wake-up bin limits, [10] active count
    int *hcounters = nullptr;  // pinned host mirror
    unsigned long long *rungs = nullptr;  // device: kRungCount face counts of the force passes (when counted)
    bool rung_stats = false;              // the force passes count their Riemann rungs
    unsigned long long *u64c = nullptr;   // device: [0..63] bin histogram, [64] displacement bits, [65] force pairs
    unsigned long long *hu64c = nullptr;  // pinned host mirror

    // the integer timeline (host) of gz_run_*
    gz_integ_params ip{};
    bool run_started = false;
    long long ti_current = 0, num_steps = 0, n_woken = 0;
    double time = 0, tbi = 0;
    unsigned long long mask = 0;
    long long bin_count[kTimebins] = {};
    int need_wakeup = 0, stop_code = 0;

    // stats and timing
    // named counters (gz_stat): fixed fields, so that updating them allocates nothing
    struct Stats {
        double density_iterations = 0, ngb_rebuilds = 0, n_active = 0, ell_widen = 0, active_updates = 0, resorts = 0,
               full_drifts = 0, force_rare_rows = 0;
    } st;
    bool timing = false;
    // contexts of gz_hydro_pass_noalloc: every device buffer is carved from `pool` (the static
    // scratch array of the device) and nothing is freed; host counters live in host_u64
    unsigned char *pool = nullptr;
    size_t pool_size = 0, pool_used = 0;
    unsigned long long host_u64[80 + 8] = {};
    struct TEv { std::string name; hipEvent_t a, b; };
    std::vector<TEv> pending;
    std::map<std::string, std::pair<long, double>> timings;
    std::vector<hipEvent_t> event_pool;
};

namespace {

int fail(gz_context *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) snprintf(c->err, sizeof c->err, "%s", buf);
    return code;
}

#define HIPCHK(call)                                                                                      \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess) return fail(ctx, GZ_E_HIP, "%s failed: %s (%s:%d)", #call,                  \
                                          hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

#define CHK(call)                          \
    do {                                   \
        int rc_ = (call);                  \
        if (rc_) return rc_;               \
    } while (0)

// count elements from the context's pool (no allocation, nothing freed: a buffer that grows
// takes new space), 256-byte aligned
int pool_take(gz_context *ctx, void **p, size_t bytes)
{
    const size_t at = (ctx->pool_used + 255) & ~(size_t)255;
    if (at > ctx->pool_size || bytes > ctx->pool_size - at) {
        *p = nullptr;
        return fail(ctx, GZ_E_MEMORY, "the static scratch of %zu bytes is exhausted (%zu in use, %zu more asked for)", ctx->pool_size,
                    at, bytes);
    }
    *p = ctx->pool + at;
    ctx->pool_used = at + bytes;
    return 0;
}

template <typename T>
int dalloc(gz_context *ctx, T **p, size_t count)
{
    if (count == 0) count = 1;
    if (ctx->pool) return pool_take(ctx, (void **)p, count * sizeof(T));
    if (*p) { hipFree(*p); *p = nullptr; }
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) { *p = nullptr; return fail(ctx, GZ_E_MEMORY, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e)); }
    return 0;
}

template <typename T>
void dfree(T *&p)
{
    if (p) hipFree(p);
    p = nullptr;
}

hipEvent_t get_event(gz_context *c)
{
    if (!c->event_pool.empty()) { hipEvent_t e = c->event_pool.back(); c->event_pool.pop_back(); return e; }
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
}

struct Timed {
    gz_context *c;
    const char *name;
    hipEvent_t a{}, b{};
    Timed(gz_context *c_, const char *n_) : c(c_), name(n_)
    {
        if (c->timing) { a = get_event(c); b = get_event(c); hipEventRecord(a, c->stream); }
    }
    ~Timed()
    {
        if (c->timing) { hipEventRecord(b, c->stream); c->pending.push_back({name, a, b}); }
    }
};

void collect_timings(gz_context *c)
{
    if (c->pending.empty()) return;
    hipStreamSynchronize(c->stream);
    for (auto &t : c->pending) {
        float ms = 0;
        hipEventElapsedTime(&ms, t.a, t.b);
        auto &e = c->timings[t.name];
        e.first += 1;
        e.second += ms;
        c->event_pool.push_back(t.a);
        c->event_pool.push_back(t.b);
    }
    c->pending.clear();
}

int set_params(gz_context *ctx, const gz_params *p)
{
    if (!p) return fail(ctx, GZ_E_ARG, "params is NULL");
    if (!(p->box_size > 0)) return fail(ctx, GZ_E_ARG, "box_size must be > 0");
    if (p->kernel != 3 && p->kernel != 6) return fail(ctx, GZ_E_ARG, "kernel must be 3 (cubic spline) or 6 (Wendland C2)");
    if (!(p->gamma > 1)) return fail(ctx, GZ_E_ARG, "gamma must be > 1");
    if (!(p->des_num_ngb > 0)) return fail(ctx, GZ_E_ARG, "des_num_ngb must be > 0");
    if (!(p->max_kernel_radius > 0) || p->max_kernel_radius > 0.5 * p->box_size / 1.3)
        return fail(ctx, GZ_E_ARG, "max_kernel_radius must be in (0, box/2.6]");
    if (p->precision != GZ_FP64 && p->precision != GZ_FP32_MIXED) return fail(ctx, GZ_E_ARG, "precision must be GZ_FP64 or GZ_FP32_MIXED");
    ctx->par = *p;
    DevParams &d = ctx->dp;
    d.box = p->box_size;
    d.half = 0.5 * p->box_size;
    d.gamma = p->gamma;
    d.G0 = p->gamma;
    d.G1 = ((d.G0 - 1.0) / (2.0 * d.G0));
    d.G2 = ((d.G0 + 1.0) / (2.0 * d.G0));
    d.G3 = ((2.0 * d.G0 / (d.G0 - 1.0)));
    d.G4 = (2.0 / (d.G0 - 1.0));
    d.G5 = (2.0 / (d.G0 + 1.0));
    d.G6 = ((d.G0 - 1.0) / (d.G0 + 1.0));
    d.G7 = (0.5 * (d.G0 - 1.0));
    d.G8 = (1.0 / d.G0);
    d.G9 = (d.G0 - 1.0);
    d.des_num_ngb = p->des_num_ngb;
    d.max_num_ngb_dev = p->max_num_ngb_dev;
    d.min_kernel_radius = p->min_kernel_radius;
    d.max_kernel_radius = p->max_kernel_radius;
    d.time = p->time;
    d.time_begin = p->time_begin;
    d.fd4alpha = ctx->fd ? 4.0 * fd_alpha(p->kernel) : 0.0;
    return 0;
}

template <typename R>
Gam<R> gam(const DevParams &d)
{
    Gam<R> g;
    g.G0 = (R)d.G0; g.G1 = (R)d.G1; g.G2 = (R)d.G2; g.G3 = (R)d.G3; g.G4 = (R)d.G4;
    g.G5 = (R)d.G5; g.G6 = (R)d.G6; g.G7 = (R)d.G7; g.G8 = (R)d.G8; g.G9 = (R)d.G9;
    return g;
}

int ensure_tmp(gz_context *ctx, size_t bytes)
{
    if (bytes <= ctx->tmp_bytes) return 0;
    size_t b = std::max(bytes, (size_t)1 << 20);
    if (ctx->pool) {
        CHK(pool_take(ctx, &ctx->tmp, b));
        ctx->tmp_bytes = b;
        return 0;
    }
    if (ctx->tmp) hipFree(ctx->tmp);
    ctx->tmp = nullptr;
    HIPCHK(hipMalloc(&ctx->tmp, b));
    ctx->tmp_bytes = b;
    return 0;
}

int alloc_state(gz_context *ctx, long cap)
{
    const size_t n = (size_t)cap;
    CHK(dalloc(ctx, &ctx->X, n)); CHK(dalloc(ctx, &ctx->Y, n)); CHK(dalloc(ctx, &ctx->Z, n));
    CHK(dalloc(ctx, &ctx->VX, n)); CHK(dalloc(ctx, &ctx->VY, n)); CHK(dalloc(ctx, &ctx->VZ, n));
    CHK(dalloc(ctx, &ctx->M, n)); CHK(dalloc(ctx, &ctx->H, n)); CHK(dalloc(ctx, &ctx->U, n));
    CHK(dalloc(ctx, &ctx->NN, n)); CHK(dalloc(ctx, &ctx->DRK, n)); CHK(dalloc(ctx, &ctx->DIVV, n));
    CHK(dalloc(ctx, &ctx->RHO, n)); CHK(dalloc(ctx, &ctx->PR, n)); CHK(dalloc(ctx, &ctx->T, 6 * n));
    CHK(dalloc(ctx, &ctx->FW, 3 * n)); CHK(dalloc(ctx, &ctx->CN, n)); CHK(dalloc(ctx, &ctx->FCE, n));
    CHK(dalloc(ctx, &ctx->GR, 3 * n)); CHK(dalloc(ctx, &ctx->GP, 3 * n)); CHK(dalloc(ctx, &ctx->GV, 9 * n));
    CHK(dalloc(ctx, &ctx->ACC, 3 * n)); CHK(dalloc(ctx, &ctx->DTU, n)); CHK(dalloc(ctx, &ctx->MSV, n));
    CHK(dalloc(ctx, &ctx->TB, n)); CHK(dalloc(ctx, &ctx->WAKE, n)); CHK(dalloc(ctx, &ctx->ACT, n));
    CHK(dalloc(ctx, &ctx->perm, n)); CHK(dalloc(ctx, &ctx->key, n)); CHK(dalloc(ctx, &ctx->key2, n));
    CHK(dalloc(ctx, &ctx->ival, n)); CHK(dalloc(ctx, &ctx->ival2, n)); CHK(dalloc(ctx, &ctx->iflag, n + 1));
    CHK(dalloc(ctx, &ctx->iscan, n + 1)); CHK(dalloc(ctx, &ctx->act, n));
    CHK(dalloc(ctx, &ctx->stage, 9 * n));
    CHK(dalloc(ctx, &ctx->it_h, n)); CHK(dalloc(ctx, &ctx->it_hb, n)); CHK(dalloc(ctx, &ctx->it_l, n));
    CHK(dalloc(ctx, &ctx->it_r, n)); CHK(dalloc(ctx, &ctx->it_cn, n)); CHK(dalloc(ctx, &ctx->it_ci, n));
    CHK(dalloc(ctx, &ctx->it_it, n)); CHK(dalloc(ctx, &ctx->todo, n)); CHK(dalloc(ctx, &ctx->todo2, n));
    CHK(dalloc(ctx, &ctx->cnt_d, n)); CHK(dalloc(ctx, &ctx->cnt_s, n));
    CHK(dalloc(ctx, &ctx->cells, n)); CHK(dalloc(ctx, &ctx->cflag, n + 1)); CHK(dalloc(ctx, &ctx->cscan, n + 1));
    CHK(dalloc(ctx, &ctx->need, n));
    CHK(dalloc(ctx, &ctx->dacc, kDensSumsFD * n));
    static_assert(sizeof(GradRec<float>) <= sizeof(GradRec<double>) && sizeof(Cell<float>) <= sizeof(Cell<double>),
                  "the record buffers hold either form");
    CHK(dalloc(ctx, (char **)&ctx->rec_grad, n * sizeof(GradRec<double>)));
    CHK(dalloc(ctx, (char **)&ctx->rec_force, n * sizeof(Cell<double>)));
    CHK(dalloc(ctx, &ctx->VKX, n)); CHK(dalloc(ctx, &ctx->VKY, n)); CHK(dalloc(ctx, &ctx->VKZ, n));
    CHK(dalloc(ctx, &ctx->UK, n)); CHK(dalloc(ctx, &ctx->SOFT, n)); CHK(dalloc(ctx, &ctx->TIB, n)); CHK(dalloc(ctx, &ctx->DTS, n));
    CHK(dalloc(ctx, &ctx->TIC, n));
    // the 64-bit counters and maxima, then the 16 int counters, in one allocation: one copy
    // reads a stage's results back
    CHK(dalloc(ctx, &ctx->u64c, 80 + 8));
    HIPCHK(hipMemsetAsync(ctx->u64c, 0, (80 + 8) * sizeof(unsigned long long), ctx->stream));  // the running pair count starts at zero
    ctx->counters = (int *)(ctx->u64c + 80);
    CHK(dalloc(ctx, &ctx->rungs, 16));
    HIPCHK(hipMemsetAsync(ctx->rungs, 0, 16 * sizeof(unsigned long long), ctx->stream));
    if (ctx->pool) {
        ctx->hu64c = ctx->host_u64;
    } else if (!ctx->hu64c) {
        hipError_t e = hipHostMalloc((void **)&ctx->hu64c, (80 + 8) * sizeof(unsigned long long), hipHostMallocDefault);
        if (e != hipSuccess) return fail(ctx, GZ_E_MEMORY, "hipHostMalloc failed");
    }
    ctx->hcounters = (int *)(ctx->hu64c + 80);
    ctx->cap = cap;
    return 0;
}

void free_state(gz_context *c)
{
    dfree(c->X); dfree(c->Y); dfree(c->Z); dfree(c->VX); dfree(c->VY); dfree(c->VZ); dfree(c->M); dfree(c->H);
    dfree(c->U); dfree(c->NN); dfree(c->DRK); dfree(c->DIVV); dfree(c->RHO); dfree(c->PR); dfree(c->T); dfree(c->FW);
    dfree(c->CN); dfree(c->FCE); dfree(c->GR); dfree(c->GP); dfree(c->GV); dfree(c->ACC); dfree(c->DTU); dfree(c->MSV);
    dfree(c->TB); dfree(c->WAKE); dfree(c->ACT); dfree(c->perm); dfree(c->key); dfree(c->key2);
    dfree(c->wstart); dfree(c->wend); dfree(c->wlstart); dfree(c->wlend); dfree(c->lidx); dfree(c->plev); dfree(c->lcount);
    dfree(c->wide_dev);
    {
        dfree(c->act_bits); dfree(c->act_bcount);
        dfree(c->sched); dfree(c->sched_part); dfree(c->rare); dfree(c->rl_a); dfree(c->rl_b);
        c->sched_valid = false;
        c->rl_valid = false;
        RowGroups &G = c->grp;
        dfree(G.key); dfree(G.skey); dfree(G.gkey); dfree(G.val); dfree(G.row); dfree(G.flag); dfree(G.scan); dfree(G.start);
        dfree(G.count); dfree(G.lev); dfree(G.code); dfree(G.nst); dfree(G.sten);
        G = RowGroups{};
    }
    c->wcells_alloc = 0;
    c->wide_key_version = -1;
    dfree(c->ival); dfree(c->ival2); dfree(c->iflag); dfree(c->iscan); dfree(c->act); dfree(c->stage);
    for (auto &v : c->spare) {
        for (double *p : v) hipFree(p);
        v.clear();
    }
    dfree(c->it_h); dfree(c->it_hb); dfree(c->it_l); dfree(c->it_r); dfree(c->it_cn); dfree(c->it_ci); dfree(c->it_it);
    dfree(c->todo); dfree(c->todo2); dfree(c->cnt_d); dfree(c->cnt_s); dfree(c->ell_d); dfree(c->ell_s);
    c->counters = nullptr;  // inside u64c
    dfree(c->cstart); dfree(c->cend); dfree(c->nsten); dfree(c->sten); dfree(c->chmax);
    dfree(c->ca0); dfree(c->ca1); dfree(c->cmark); dfree(c->dcells); dfree(c->cells); dfree(c->cflag); dfree(c->cscan); dfree(c->need);
    dfree(c->rec_grad); dfree(c->rec_force); dfree(c->dacc);
    dfree(c->VKX); dfree(c->VKY); dfree(c->VKZ); dfree(c->UK); dfree(c->SOFT); dfree(c->TIB); dfree(c->DTS); dfree(c->TIC);
    dfree(c->u64c);
    dfree(c->rungs);
    if (c->hu64c) hipHostFree(c->hu64c);
    c->hu64c = nullptr;
    if (c->tmp) hipFree(c->tmp);
    c->tmp = nullptr;
    c->tmp_bytes = 0;
    c->hcounters = nullptr;  // inside hu64c
}

// upload one host field (width doubles per particle, caller order) into device order
int upload_field(gz_context *ctx, const double *host, int width, double *dst, bool soa)
{
    const long n = ctx->n;
    HIPCHK(hipMemcpyAsync(ctx->stage, host, (size_t)n * width * sizeof(double), hipMemcpyDefault, ctx->stream));
    k_gather<<<nblk(n), kBlock, 0, ctx->stream>>>(n, width, ctx->perm, ctx->stage, dst, soa ? 1 : 0);
    HIPCHK(hipGetLastError());
    return 0;
}

int download_field(gz_context *ctx, double *host, int width, const double *src, bool soa)
{
    const long n = ctx->n;
    k_scatter<<<nblk(n), kBlock, 0, ctx->stream>>>(n, width, ctx->perm, src, ctx->stage, soa ? 1 : 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, ctx->stage, (size_t)n * width * sizeof(double), hipMemcpyDefault, ctx->stream));
    return 0;
}

// a host vector field (n x 3, caller order) into three device columns, and back
int upload_vec3(gz_context *ctx, const double *host, double *dx, double *dy, double *dz)
{
    const long n = ctx->n;
    HIPCHK(hipMemcpyAsync(ctx->stage, host, (size_t)n * 3 * sizeof(double), hipMemcpyDefault, ctx->stream));
    k_gather<<<nblk(n), kBlock, 0, ctx->stream>>>(n, 3, ctx->perm, ctx->stage, ctx->stage + 3 * n, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dx, ctx->stage + 3 * n, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dy, ctx->stage + 4 * n, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dz, ctx->stage + 5 * n, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return 0;
}

int download_vec3(gz_context *ctx, double *host, const double *dx, const double *dy, const double *dz)
{
    const long n = ctx->n;
    HIPCHK(hipMemcpyAsync(ctx->stage + 3 * n, dx, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->stage + 4 * n, dy, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->stage + 5 * n, dz, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    k_scatter<<<nblk(n), kBlock, 0, ctx->stream>>>(n, 3, ctx->perm, ctx->stage + 3 * n, ctx->stage, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, ctx->stage, (size_t)n * 3 * sizeof(double), hipMemcpyDefault, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int copy_dev(gz_context *ctx, void *dst, const void *src, size_t bytes)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return 0;
}

// grid of level L over the current device order; the stencil of each cell
int build_grid(gz_context *ctx, double rmax, bool need_chmax = false)
{
    const double box = ctx->par.box_size;
    // cell side >= the largest search radius plus twice the distance any particle has moved
    // since the particles were ordered (cells are those of the ordering), with a margin for
    // positions that round onto a cell face (the 27-cell stencil must then still hold every
    // partner)
    ctx->last_rmax = rmax;
    const double reach = rmax + 2.0 * ctx->dmax;
    int L = 0;
    while (L < kFineLevel && box / (double)(1 << (L + 1)) >= reach * (1.0 + 1e-6)) L++;
    const long ncell = 1L << (3 * L);
    if (L == ctx->level && ctx->grid_version == ctx->key_version && !need_chmax) return 0;
    if (ncell > ctx->grid_cells_alloc) {
        CHK(dalloc(ctx, &ctx->cstart, ncell)); CHK(dalloc(ctx, &ctx->cend, ncell)); CHK(dalloc(ctx, &ctx->nsten, ncell));
        CHK(dalloc(ctx, &ctx->sten, ncell * 27)); CHK(dalloc(ctx, &ctx->chmax, ncell));
        CHK(dalloc(ctx, &ctx->ca0, ncell)); CHK(dalloc(ctx, &ctx->ca1, ncell));
        CHK(dalloc(ctx, &ctx->cmark, ncell)); CHK(dalloc(ctx, &ctx->dcells, ncell));
        HIPCHK(hipMemsetAsync(ctx->cmark, 0, ncell * sizeof(int), ctx->stream));  // the lazy drift's marks stay clear between calls
        ctx->grid_cells_alloc = ncell;
        ctx->grid_version = -1;
    }
    Timed t(ctx, "ngb_build");
    if (L != ctx->level || ctx->grid_version != ctx->key_version) {
        ctx->level = L;
        ctx->rows_level = -1;
        ctx->cell = box / (double)(1 << L);
        k_cell_clear<<<nblk(ncell), kBlock, 0, ctx->stream>>>(ncell, ctx->cstart, ctx->cend, ctx->chmax);
        k_cell_bounds<<<nblk(ctx->n), kBlock, 0, ctx->stream>>>(ctx->n, ctx->key, 3 * (kFineLevel - L), ctx->cstart, ctx->cend);
        k_cell_stencil<<<nblk(ncell), kBlock, 0, ctx->stream>>>(L, ctx->cstart, ctx->cend, ctx->sten, ctx->nsten);
        HIPCHK(hipGetLastError());
        ctx->grid_version = ctx->key_version;
    }
    if (need_chmax) {
        k_cell_hmax<<<nblk(ncell), kBlock, 0, ctx->stream>>>(ncell, ctx->cstart, ctx->cend, ctx->H, ctx->chmax);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

GridView grid_view(gz_context *ctx)
{
    GridView g{};
    g.wide = ctx->wide ? ctx->wide_dev : nullptr;
    g.level = ctx->level;
    g.shift = 3 * (kFineLevel - ctx->level);
    g.cell = ctx->cell;
    g.box = ctx->par.box_size;
    g.half = 0.5 * ctx->par.box_size;
    g.sten = ctx->sten;
    g.nsten = ctx->nsten;
    g.chmax = ctx->chmax;
    g.margin = ctx->dmax;
    g.cstart = ctx->cstart;
    g.cend = ctx->cend;
    g.cell_off = 0;
    return g;
}

TimeParams time_params(gz_context *ctx);
ParticleView particle_view(gz_context *c);

// In a step that drifts lazily: the particles of the cells a pass over the active rows reads on
// the current grid, drifted to All.Ti_Current where they stand behind it (idempotent: once per
// particle and step).  The largest radius drifted goes to u64c[75].
int drift_stencil(gz_context *ctx)
{
    if (!ctx->lazy || ctx->nact == 0) return 0;
    // the same rows on the same grid at the same time: their cells were drifted already (the
    // symmetric list's walk after the density rounds on an unchanged grid)
    const long long mark[4] = {ctx->ti_current, (long long)ctx->act_version, (long long)ctx->level, (long long)ctx->key_version};
    if (std::equal(mark, mark + 4, ctx->stencil_drifted)) return 0;
    std::copy(mark, mark + 4, ctx->stencil_drifted);
    Timed t(ctx, "drift");
    // the cells listed (counted by k_stencil_cells): cleared by the step's k_step_prep for the
    // step's first listing, here for any other
    int *ncells = (int *)(ctx->u64c + 76);
    if (!ctx->ncells_clear) HIPCHK(hipMemsetAsync(ncells, 0, sizeof(int), ctx->stream));
    ctx->ncells_clear = false;
    const long pairs = 27L * ctx->nact;
    k_stencil_cells<<<nblk(pairs), kBlock, 0, ctx->stream>>>(grid_view(ctx), ctx->nact, ctx->act, ctx->key, ctx->cmark, ctx->dcells,
                                                             ncells);
    const long most = std::min<long>(pairs, 1L << (3 * ctx->level));
    k_drift_cells<<<(unsigned)std::min<long>(most, 4096), kBlock, 0, ctx->stream>>>(time_params(ctx), particle_view(ctx), grid_view(ctx),
                                                                                    ctx->dcells, ncells, ctx->cmark, ctx->ti_current,
                                                                                    ctx->u64c + 75);
    HIPCHK(hipGetLastError());
    return 0;
}

// occupied cells of the active rows (one wavefront each in the cell-walk kernels)
int build_cell_rows(gz_context *ctx)
{
    const int na = ctx->nact;
    if (ctx->rows_level == ctx->level && ctx->rows_act_version == ctx->act_version && ctx->rows_key_version == ctx->key_version)
        return 0;
    const int shift = 3 * (kFineLevel - ctx->level);
    k_cell_rows_mark<<<nblk(na), kBlock, 0, ctx->stream>>>(na, ctx->act, ctx->key, shift, ctx->ca0, ctx->ca1, ctx->cflag);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemsetAsync(ctx->cflag + na, 0, sizeof(int), ctx->stream));
    size_t bytes = 0;
    HIPCHK(gz_exclusive_sum(nullptr, bytes, ctx->cflag, ctx->cscan, na + 1, ctx->stream));
    CHK(ensure_tmp(ctx, bytes));
    HIPCHK(gz_exclusive_sum(ctx->tmp, bytes, ctx->cflag, ctx->cscan, na + 1, ctx->stream));
    k_cell_rows_list<<<nblk(na), kBlock, 0, ctx->stream>>>(na, ctx->act, ctx->key, shift, ctx->cflag, ctx->cscan, ctx->cells);
    HIPCHK(hipGetLastError());
    ctx->rows_level = ctx->level;
    ctx->rows_act_version = ctx->act_version;
    ctx->rows_key_version = ctx->key_version;
    if (na <= ctx->few_rows) {
        // few rows: launch one wavefront per row at most; the walks read the cell count on the device
        ctx->ncells = na;
        ctx->ncells_exact = false;
        return 0;
    }
    HIPCHK(hipMemcpyAsync(ctx->hcounters + 11, ctx->cscan + na, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->ncells = ctx->hcounters[11];
    ctx->ncells_exact = true;
    return 0;
}

CellRows cell_rows(gz_context *ctx)
{
    return CellRows{ctx->cells, ctx->ncells, ctx->ncells_exact ? nullptr : ctx->cscan + ctx->nact, ctx->ca0, ctx->ca1, ctx->nact};
}

int read_counters(gz_context *ctx)
{
    // u64c[64..71] and the int counters behind them (u64c[80..87])
    HIPCHK(hipMemcpyAsync(ctx->hu64c + 64, ctx->u64c + 64, 24 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int max_on_device(gz_context *ctx, const double *v, long n, double *out)
{
    size_t bytes = 0;
    double *dout = (double *)ctx->stage;  // stage is free between uploads
    HIPCHK(gz_reduce_max(nullptr, bytes, v, dout, (int)n, ctx->stream));
    CHK(ensure_tmp(ctx, bytes));
    HIPCHK(gz_reduce_max(ctx->tmp, bytes, v, dout, (int)n, ctx->stream));
    HIPCHK(hipMemcpyAsync(out, dout, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int min_on_device(gz_context *ctx, const double *v, long n, double *out)
{
    size_t bytes = 0;
    double *dout = (double *)ctx->stage;
    HIPCHK(gz_reduce_min(nullptr, bytes, v, dout, (int)n, ctx->stream));
    CHK(ensure_tmp(ctx, bytes));
    HIPCHK(gz_reduce_min(ctx->tmp, bytes, v, dout, (int)n, ctx->stream));
    HIPCHK(hipMemcpyAsync(out, dout, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// the finest level (at most lmax) whose cells are no smaller than reach (host twin of the
// device wide_level)
int host_wide_level(double box, int lmax, double reach)
{
    int L = 0;
    while (L < lmax && box / (double)(1 << (L + 1)) >= reach * (1.0 + 1e-6)) L++;
    return L;
}

// Tables of the search over many levels: cell ranges of the order at every level down to the
// one of the smallest radius (kept while the order and that level stay), and with lists, every
// particle's own level and the per-level tables of each level's particles (from the radii now).
int build_wide(gz_context *ctx, bool lists)
{
    const long n = ctx->n;
    const double box = ctx->par.box_size;
    double hmin = 0;
    CHK(min_on_device(ctx, ctx->H, n, &hmin));
    if (!(hmin > 0)) hmin = 1e-300;
    // at most level 9 (512^3 cells, 1.2 G table entries over all levels; finer radii walk a
    // few more candidates there)
    const int lmax = host_wide_level(box, kFineLevel - 1, hmin + 2.0 * ctx->dmax);
    long long cells = 0;
    WideGrid &W = ctx->wide_h;
    for (int L = 0; L <= lmax; L++) { W.loff[L] = cells; cells += 1LL << (3 * L); }
    if (cells > ctx->wcells_alloc) {
        // entries are validated on read: fresh tables need no clearing
        CHK(dalloc(ctx, &ctx->wstart, cells)); CHK(dalloc(ctx, &ctx->wend, cells));
        CHK(dalloc(ctx, &ctx->wlstart, cells)); CHK(dalloc(ctx, &ctx->wlend, cells));
        ctx->wcells_alloc = cells;
        ctx->wide_key_version = -1;
    }
    if (!ctx->lidx) {
        CHK(dalloc(ctx, &ctx->lidx, ctx->cap)); CHK(dalloc(ctx, &ctx->plev, ctx->cap)); CHK(dalloc(ctx, &ctx->lcount, kFineLevel + 2));
        CHK(dalloc(ctx, &ctx->wide_dev, 1));
    }
    W.key = ctx->key;
    W.n = (int)n;
    W.lmax = lmax;
    W.start = ctx->wstart; W.end = ctx->wend; W.lidx = ctx->lidx; W.lstart = ctx->wlstart; W.lend = ctx->wlend;
    Timed t(ctx, "ngb_build");
    if (ctx->wide_key_version != ctx->key_version || lmax > ctx->wide_lmax) {
        k_wide_bounds<<<nblk(n), kBlock, 0, ctx->stream>>>(W);
        HIPCHK(hipGetLastError());
        ctx->wide_key_version = ctx->key_version;
        ctx->wide_lmax = lmax;
    }
    if (lists) {
        HIPCHK(hipMemsetAsync(ctx->lcount, 0, (kFineLevel + 2) * sizeof(int), ctx->stream));
        k_particle_level<<<nblk(n), kBlock, 0, ctx->stream>>>((int)n, ctx->H, box, ctx->dmax, lmax, ctx->plev, ctx->lcount);
        k_iota<<<nblk(n), kBlock, 0, ctx->stream>>>((int)n, ctx->ival);
        HIPCHK(hipGetLastError());
        size_t bytes = 0;
        HIPCHK(gz_sort_pairs(nullptr, bytes, ctx->plev, (int *)ctx->key2, ctx->ival, ctx->lidx, (int)n, 0, 4, ctx->stream));
        CHK(ensure_tmp(ctx, bytes));
        HIPCHK(gz_sort_pairs(ctx->tmp, bytes, ctx->plev, (int *)ctx->key2, ctx->ival, ctx->lidx, (int)n, 0, 4, ctx->stream));
        int counts[kFineLevel + 2];
        HIPCHK(hipMemcpyAsync(counts, ctx->lcount, sizeof counts, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        int acc = 0;
        for (int L = 0; L <= kFineLevel + 1; L++) {
            W.lofs[L] = acc;
            if (L <= kFineLevel) acc += counts[L];
        }
        k_wide_list_bounds<<<nblk(n), kBlock, 0, ctx->stream>>>(W, ctx->plev);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(ctx->wide_dev, &W, sizeof W, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // the displacement that calls for a new order is measured against the smallest radius
    ctx->last_rmax = hmin;
    return 0;
}

// The active rows grouped by the cell of their own level (from their radii) for the cell walk of
// the symmetric list over many levels
int build_groups(gz_context *ctx)
{
    RowGroups &G = ctx->grp;
    const int na = ctx->nact;
    if (G.rows_alloc < (long)na + 1) {
        const size_t m = (size_t)ctx->cap + 1;
        CHK(dalloc(ctx, &G.key, m)); CHK(dalloc(ctx, &G.skey, m)); CHK(dalloc(ctx, &G.gkey, m)); CHK(dalloc(ctx, &G.val, m));
        CHK(dalloc(ctx, &G.row, m)); CHK(dalloc(ctx, &G.flag, m)); CHK(dalloc(ctx, &G.scan, m)); CHK(dalloc(ctx, &G.start, m));
        CHK(dalloc(ctx, &G.count, 1));
        G.rows_alloc = (long)m;
    }
    size_t bytes = 0;
    HIPCHK(group_tmp_bytes(na, bytes));
    CHK(ensure_tmp(ctx, bytes));
    Timed t(ctx, "ngb_build");
    HIPCHK(group_rows(G, ctx->stream, ctx->tmp, ctx->tmp_bytes, na, ctx->act, nullptr, ctx->key, nullptr, ctx->H, ctx->par.box_size,
                      ctx->dmax, ctx->wide_dev, BoxTable{}, true));
    HIPCHK(hipMemcpyAsync(ctx->hcounters + 12, G.count, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    G.ngroups = ctx->hcounters[12];
    if (G.groups_alloc < G.ngroups) {
        const long m = std::min((long)na, (long)G.ngroups + G.ngroups / 4 + 64);
        CHK(dalloc(ctx, &G.lev, m)); CHK(dalloc(ctx, &G.code, m)); CHK(dalloc(ctx, &G.nst, m)); CHK(dalloc(ctx, &G.sten, (size_t)27 * m));
        G.groups_alloc = m;
    }
    HIPCHK(group_stencils(G, ctx->stream, grid_view(ctx), ctx->act, BoxTable{}));
    return 0;
}

// build ELL rows (all slots when rows == nullptr) into a list that grows its width on overflow
template <bool SYM>
int build_list(gz_context *ctx, int nrows, const int *rows, const double *hrow, double scale, int **ell, long *ld, int *cap,
               int *cnt, int cap_hint)
{
    const long need_ld = ((long)ctx->nact + 63) / 64 * 64;
    if (*ell == nullptr || *ld != need_ld || *cap < cap_hint) {
        *cap = std::max(*cap, cap_hint);
        *ld = need_ld;
        CHK(dalloc(ctx, ell, (size_t)ell_width(*cap) * (size_t)(*ld)));
    }
    for (int attempt = 0; attempt < 8; attempt++) {
        HIPCHK(hipMemsetAsync(ctx->counters + 3, 0, sizeof(int), ctx->stream));
        ctx->overflow_clear = false;
        {
            Timed t(ctx, "ngb_build");
            k_build_ell<SYM><<<xcd_grid(nblk(nrows)), kBlock, 0, ctx->stream>>>(nrows, rows, ctx->act, hrow, scale, ctx->X, ctx->Y, ctx->Z, ctx->H,
                                                                     ctx->key, grid_view(ctx), *ell, *ld, *cap, cnt, ctx->counters + 3);
            HIPCHK(hipGetLastError());
        }
        CHK(read_counters(ctx));
        const int over = ctx->hcounters[3];
        if (over == 0) return 0;
        // widen and rebuild every row of this list (a narrower list cannot keep its rows)
        *cap = std::max(over + 16, *cap * 3 / 2);
        CHK(dalloc(ctx, ell, (size_t)ell_width(*cap) * (size_t)(*ld)));
        nrows = ctx->nact;
        rows = nullptr;
        ctx->st.ell_widen += 1;
    }
    return fail(ctx, GZ_E_STATE, "neighbour list did not fit after widening");
}

int build_active(gz_context *ctx)
{
    const long n = ctx->n;
    ctx->density_ready = false;  // a new list: density() starts its rows itself
    k_flag_active<<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->all_active ? nullptr : ctx->ACT, ctx->M, ctx->iflag);
    HIPCHK(hipGetLastError());
    size_t bytes = 0;
    HIPCHK(gz_exclusive_sum(nullptr, bytes, ctx->iflag, ctx->iscan, (int)(n + 1), ctx->stream));
    CHK(ensure_tmp(ctx, bytes));
    HIPCHK(hipMemsetAsync(ctx->iflag + n, 0, sizeof(int), ctx->stream));
    HIPCHK(gz_exclusive_sum(ctx->tmp, bytes, ctx->iflag, ctx->iscan, (int)(n + 1), ctx->stream));
    k_compact<<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->iflag, ctx->iscan, ctx->act);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ctx->hcounters + 10, ctx->iscan + n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int na = ctx->hcounters[10];
    ctx->nact = na;
    ctx->act_version++;
    ctx->st.n_active = na;
    return 0;
}

// ---------------------------------------------------------------------------------------
// stages
// ---------------------------------------------------------------------------------------
}  // namespace

// iota on device
__global__ void k_iota(int n, int *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i;
}

namespace {

template <int KERN, typename R>
int density_rounds(gz_context *ctx, int iters, bool speculate);

template <int KERN, typename R>
int density_stage(gz_context *ctx)
{
    const int na = ctx->nact;
    ctx->st.density_iterations = 0;
    ctx->st.ngb_rebuilds = 0;
    if (na == 0) { ctx->density_done = true; ctx->sym_valid = false; return 0; }
    // one launch: the radius clamp, the iteration state, the counters, and the largest radius
    double hmax = ctx->hmax_rows;  // known when the step's timeline pass found it
    ctx->hmax_rows = -1;
    const bool ready = ctx->density_ready && hmax > 0;  // the step's active-list launch started the iteration
    ctx->density_ready = false;
    if (!ready) {
        if (!(hmax > 0)) HIPCHK(hipMemsetAsync(ctx->u64c + 67, 0, sizeof(unsigned long long), ctx->stream));
        k_density_init<<<nblk(na), kBlock, 0, ctx->stream>>>(na, ctx->act, ctx->H, 0.99 * ctx->par.max_kernel_radius, ctx->it_h,
                                                             ctx->it_l, ctx->it_r, ctx->it_cn, ctx->it_ci, ctx->it_it, ctx->need,
                                                             ctx->counters, hmax > 0 ? nullptr : ctx->u64c + 67, ctx->u64c + 68, BoxTable{});
        HIPCHK(hipGetLastError());
    }
    ctx->overflow_clear = true;  // counters[0..7] cleared, the list overflow among them
    if (!(hmax > 0)) {
        HIPCHK(hipMemcpyAsync(ctx->hu64c + 67, ctx->u64c + 67, sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        memcpy(&hmax, &ctx->hu64c[67], sizeof hmax);
    }
    const bool wide = ctx->wide != 0;
    if (wide) CHK(build_wide(ctx, false));
    else CHK(build_grid(ctx, std::min(hmax, 0.5 * ctx->par.box_size)));
    CHK(drift_stencil(ctx));
    const bool few = wide || na <= ctx->few_rows;
    if (!few) CHK(build_cell_rows(ctx));
    // a step of the run loop with few rows may launch the passes behind the first round before
    // its counters are read (they are gated on the round's gate word); the radius bound the symmetric
    // list's grid will take must then be known
    const bool speculate = ctx->speculate && few && !wide && ready && ctx->hmax_bound > 0;
    return density_rounds<KERN, R>(ctx, 0, speculate);
}

// The density rounds from round `iters` on (0: the first; iters > 0: resuming after the
// counters of round iters were read back by a later read of the step).  speculate: the first
// round's closing launch writes a gate word and is not waited for; the passes follow, and the
// force pass's read of the counters settles the step (force_stage).
constexpr int kListedRounds = 16384;  // density rounds over more rows than this take the list of rows still iterating

template <int KERN, typename R>
int density_rounds(gz_context *ctx, int iters, bool speculate)
{
    const int na = ctx->nact;
    const bool wide = ctx->wide != 0;
    const bool few = wide || na <= ctx->few_rows;
    DensityIter it{ctx->it_h, ctx->it_hb, ctx->it_l, ctx->it_r, ctx->it_cn, ctx->it_ci, ctx->it_it};
    DensityOut out{ctx->n, ctx->H, ctx->NN, ctx->DRK, ctx->DIVV, ctx->RHO, ctx->PR, ctx->T, ctx->FW, ctx->CN, ctx->FCE};
    double hmax = 0;
    int remaining = iters > 0 ? ctx->hcounters[0] : na;  // rows of the coming round
    bool read = iters > 0;  // the counters of round `iters` are on the host already
    // the row kernels' rounds after the first take only the rows still iterating (a list the
    // closing kernel writes), not a workgroup per row that mostly exits at once; over a few
    // thousand rows those launches cost little, and the list's offsets would add a wait on an
    // atomic's result to every closing kernel
    const bool listed = few && na > kListedRounds;
    if (listed && !ctx->rl_a) {
        CHK(dalloc(ctx, &ctx->rl_a, ctx->cap));
        CHK(dalloc(ctx, &ctx->rl_b, ctx->cap));
    }
    if (iters == 0) ctx->rl_valid = false;
    while (true) {
        if (!read) {
            // the round's counters (k_density_init, or the step's start, cleared them for the first round)
            if (iters > 0) HIPCHK(hipMemsetAsync(ctx->counters, 0, 3 * sizeof(int), ctx->stream));
            const int *rows = (listed && ctx->rl_valid) ? ctx->rl_a : nullptr;
            const int nround = rows ? remaining : na;
            int *next_rows = listed ? ctx->rl_b : nullptr;
            // over many levels the rounds take one workgroup per row whatever their number (the
            // cells of the rows' own levels hold too few rows to fill a wavefront): four
            // wavefronts per row while few rows iterate, one while many do; with timing on, the
            // rows kernels count each row's pairs (the pass's work)
            int *npairs = (ctx->timing && few) ? ctx->cnt_d : nullptr;
            {
                Timed t(ctx, "density");
                if (few && remaining <= ctx->few_rows)
                    k_density_rows<KERN, R, kRowThreads><<<xcd_grid(nround), kRowThreads, 0, ctx->stream>>>(
                        ctx->dp, grid_view(ctx), na, ctx->act, ctx->need, ctx->key, ctx->X, ctx->Y, ctx->Z, ctx->VX, ctx->VY, ctx->VZ,
                        ctx->M, ctx->it_h, ctx->dacc, BoxTable{}, npairs, rows, nround);
                else if (few)
                    k_density_rows<KERN, R, 64><<<xcd_grid(nround), 64, 0, ctx->stream>>>(
                        ctx->dp, grid_view(ctx), na, ctx->act, ctx->need, ctx->key, ctx->X, ctx->Y, ctx->Z, ctx->VX, ctx->VY, ctx->VZ,
                        ctx->M, ctx->it_h, ctx->dacc, BoxTable{}, npairs, rows, nround);
                else
                    k_density_cells<KERN, R><<<xcd_grid(ctx->ncells), 64, 0, ctx->stream>>>(
                        ctx->dp, grid_view(ctx), cell_rows(ctx), ctx->act, ctx->need, ctx->X, ctx->Y, ctx->Z, ctx->VX, ctx->VY, ctx->VZ,
                        ctx->M, ctx->it_h, ctx->dacc);
            }
            if (npairs) k_sum_counts<<<std::min(nblk(nround), 256u), kBlock, 0, ctx->stream>>>(nround, npairs, ctx->u64c + 77);
            // the speculative first round: its closing launch also writes the pass's gate word
            const bool gate_now = speculate && iters == 0;
            {
                Timed t(ctx, "density_after");
                k_density_after<<<nblk(nround), kBlock, 0, ctx->stream>>>(
                    ctx->dp, na, ctx->act, ctx->need, ctx->dacc, ctx->M, ctx->U, wide ? 1e300 : ctx->cell - 2.0 * ctx->dmax, it, out,
                    ctx->counters + 0, ctx->counters + 1, ctx->counters + 2, ctx->u64c + 68, BoxTable{}, ctx->hmax_bound,
                    gate_now ? ctx->counters + kPassGate : nullptr, (unsigned int *)(ctx->u64c + 79), rows, nround, next_rows);
            }
            HIPCHK(hipGetLastError());
            if (listed) {
                std::swap(ctx->rl_a, ctx->rl_b);
                ctx->rl_valid = true;
            }
            if (gate_now) {
                ctx->spec = true;
                ctx->st.density_iterations = 1;
                ctx->density_done = true;
                ctx->gradients_done = false;
                ctx->sym_valid = false;
                return 0;
            }
            CHK(read_counters(ctx));
            iters++;
            remaining = ctx->hcounters[0];
        }
        // resuming: the round's stop code was settled by its gate word (counters[2] then belongs
        // to the force pass)
        const int rebuild = ctx->hcounters[1], err = read ? 0 : ctx->hcounters[2];
        read = false;
        if (err) {
            ctx->st.density_iterations = iters;
            if (err == GZ_STOP_MAXITER) return fail(ctx, GZ_STOP_MAXITER, "density(): kernel radius iteration did not converge in %d rounds", kMaxIter);
            if (err == GZ_STOP_BRACKET) return fail(ctx, GZ_STOP_BRACKET, "density(): empty radius bracket (Right==0 && Left==0)");
            return fail(ctx, err, "density(): device stop code %d", err);
        }
        if (remaining == 0) break;
        if (rebuild && !wide) {
            // a radius outgrew the cell side: a coarser grid for the rows still iterating
            CHK(max_on_device(ctx, ctx->it_h, na, &hmax));
            CHK(build_grid(ctx, std::min(hmax, 0.5 * ctx->par.box_size)));
            CHK(drift_stencil(ctx));
            if (!few) CHK(build_cell_rows(ctx));
            ctx->st.ngb_rebuilds += 1;
        }
    }
    ctx->st.density_iterations = iters;
    ctx->density_done = true;
    ctx->gradients_done = false;
    ctx->sym_valid = false;
    if (ctx->hmax_bound > 0) {
        // the largest final radius of the rows, read back with the last round's counters
        double hf;
        memcpy(&hf, &ctx->hu64c[68], sizeof hf);
        ctx->hmax_bound = std::max(ctx->hmax_bound, hf);
    }
    return 0;
}

constexpr int kRowPairAtomics = 1024;  // at most this many rows count their pairs with one atomic each

// the order in which the lanes of the force pass take the rows: longest symmetric list first
// within each block of sched_span rows (the gradient pass, cheaper per pair, keeps the plain
// order: its scattered writes would cost more than the even waves save)
int build_sched(gz_context *ctx)
{
    const int na = ctx->nact;
    const int nparts = (int)((na + 1023) / 1024);
    if (!ctx->sched) {
        CHK(dalloc(ctx, &ctx->sched, ctx->cap));
        CHK(dalloc(ctx, &ctx->sched_part, 2 * ((ctx->cap + 1023) / 1024) + 2));
    }
    // FP64 pairs are compute-bound: always sorted; FP32 pairs only when the waves are uneven
    const double min_even = ctx->par.precision == GZ_FP64 ? 2.0 : 0.75;
    unsigned long long *total = ctx->sched_part + 2 * nparts;
    const int span = ctx->sched_span;
    Timed t(ctx, "ngb_build");
    k_sched_stats<<<nparts, 1024, 0, ctx->stream>>>(na, ctx->cnt_s, ctx->sched_part);
    k_sched_total<<<1, 1024, 0, ctx->stream>>>(nparts, ctx->sched_part, total);
    k_sched_sort<<<(unsigned)((na + span - 1) / span), 1024, 0, ctx->stream>>>(na, span, ctx->cnt_s, total, min_even, ctx->sched);
    HIPCHK(hipGetLastError());
    ctx->sched_valid = true;
    return 0;
}

// room for the symmetric list's cap_s x ld_s entries: a larger earlier allocation is kept (the
// row count, and with it ld_s, changes every step of the run loop; freeing and allocating again
// would wait for the device each time)
int sym_alloc(gz_context *ctx)
{
    const size_t need = (size_t)ell_width(ctx->cap_s) * (size_t)ctx->ld_s;
    if (ctx->ell_s && ctx->ell_s_alloc >= need) return 0;
    CHK(dalloc(ctx, &ctx->ell_s, need));
    ctx->ell_s_alloc = need;
    return 0;
}

int ensure_sym(gz_context *ctx)
{
    if (ctx->sym_valid) return 0;
    double hmax = ctx->hmax_bound;
    if (!(hmax > 0)) {
        CHK(max_on_device(ctx, ctx->H, ctx->n, &hmax));
        ctx->hmax_bound = hmax;
    }
    // in a lazy step a particle the walk reads may still be drifted to now, its radius growing by
    // at most exp(0.3/3) (kernel_radius_drift_log_change caps one drift at 0.3 in log density)
    if (ctx->lazy) hmax = std::min(hmax * std::exp(0.1) * (1.0 + 1e-12), ctx->par.max_kernel_radius);
    const bool wide = ctx->wide != 0;
    const bool few = ctx->nact <= ctx->few_rows;
    // the row walks drop stencil cells beyond reach: every radius is at most hmax
    if (wide) CHK(build_wide(ctx, true));
    else CHK(build_grid(ctx, std::min(hmax, 0.5 * ctx->par.box_size)));
    CHK(drift_stencil(ctx));
    GridView gs = grid_view(ctx);
    gs.chmax = nullptr;
    gs.hall = hmax;
    if (wide && !few) CHK(build_groups(ctx));
    if (!wide && !few) CHK(build_cell_rows(ctx));
    const long need_ld = ((long)ctx->nact + 63) / 64 * 64;
    const int cap_hint = std::max(64, (int)std::ceil(ctx->par.des_num_ngb * 2.0) + 32);
    if (ctx->ell_s == nullptr || ctx->ld_s != need_ld || ctx->cap_s < cap_hint) {
        ctx->cap_s = std::max(ctx->cap_s, cap_hint);
        ctx->ld_s = need_ld;
        CHK(sym_alloc(ctx));
    }
    for (int attempt = 0; attempt < 8; attempt++) {
        if (attempt > 0 || !ctx->overflow_clear) HIPCHK(hipMemsetAsync(ctx->counters + 3, 0, sizeof(int), ctx->stream));
        ctx->overflow_clear = false;
        {
            Timed t(ctx, "ngb_build");
            if (wide && !few)
                k_build_sym_cells<<<xcd_grid(ctx->grp.ngroups), 64, 0, ctx->stream>>>(grid_view(ctx), ctx->grp.view(BoxTable{}), ctx->act,
                                                                                       ctx->X, ctx->Y, ctx->Z, ctx->H, ctx->ell_s, ctx->ld_s,
                                                                                       ctx->cap_s, ctx->cnt_s, ctx->counters + 3);
            else if (wide)
                k_build_sym_wide<<<xcd_grid(ctx->nact), 64, 0, ctx->stream>>>(grid_view(ctx), ctx->nact, ctx->act, ctx->X, ctx->Y, ctx->Z,
                                                                               ctx->H, ctx->ell_s, ctx->ld_s, ctx->cap_s, ctx->cnt_s,
                                                                               ctx->counters + 3);
            else if (few)
                k_build_sym_rows<<<xcd_grid(ctx->nact), kRowThreads, 0, ctx->stream>>>(gs, ctx->nact, ctx->act, ctx->key, ctx->X, ctx->Y,
                                                                               ctx->Z, ctx->H, ctx->ell_s, ctx->ld_s, ctx->cap_s, ctx->cnt_s,
                                                                               ctx->counters + 3, BoxTable{});
            else
                k_build_sym_cells<<<xcd_grid(ctx->ncells), 64, 0, ctx->stream>>>(grid_view(ctx), cell_rows(ctx), ctx->act, ctx->X, ctx->Y,
                                                                                  ctx->Z, ctx->H, ctx->ell_s, ctx->ld_s, ctx->cap_s,
                                                                                  ctx->cnt_s, ctx->counters + 3);
            HIPCHK(hipGetLastError());
        }
        int over = 0;
        if (!ctx->defer_sym_check) {
            CHK(read_counters(ctx));
            over = ctx->hcounters[3];
        }
        if (over == 0) {
            ctx->sym_valid = true;
            ctx->sym_check_pending = ctx->defer_sym_check;
            ctx->sched_valid = false;
            if (!few && ctx->sched_span > 0) CHK(build_sched(ctx));
            // the records the gradient and force passes read: every record when many rows are
            // active, else only those of the rows and their neighbours, packed from the list (a
            // record once per list entry that names it: cheaper while the lists hold fewer
            // entries than there are particles)
            ctx->pack_list = (long)ctx->nact * 64 < ctx->n;
            return 0;
        }
        ctx->cap_s = std::max(over + 16, ctx->cap_s * 3 / 2);
        CHK(sym_alloc(ctx));
        ctx->st.ell_widen += 1;
    }
    return fail(ctx, GZ_E_STATE, "symmetric neighbour list did not fit after widening");
}

template <int KERN, typename R>
int gradients_stage(gz_context *ctx)
{
    if (!ctx->density_done) return fail(ctx, GZ_E_STATE, "hydro_gradient_calc() before density()");
    if (ctx->nact == 0) { ctx->gradients_done = true; return 0; }
    CHK(ensure_sym(ctx));
    GradRec<R> *rec = (GradRec<R> *)ctx->rec_grad;
    // few rows among many particles: the row kernel makes the records it reads (no packing launch)
    const bool make = ctx->nact <= ctx->few_rows && ctx->pack_list;
    if (!make) {
        Timed t(ctx, "pack_grad");
        if (ctx->pack_list)
            k_pack_grad_list<R><<<xcd_grid(ctx->nact), 64, 0, ctx->stream>>>(ctx->nact, ctx->act, ctx->ell_s, ctx->ld_s, ctx->cnt_s, ctx->H,
                                                                             ctx->M, ctx->RHO, ctx->PR, ctx->VX, ctx->VY, ctx->VZ, rec,
                                                                             ctx->X, ctx->Y, ctx->Z, ctx->par.box_size, BoxTable{});
        else
            k_pack_grad<R><<<nblk(ctx->n), kBlock, 0, ctx->stream>>>(ctx->n, ctx->H, ctx->M, ctx->RHO, ctx->PR, ctx->VX, ctx->VY, ctx->VZ,
                                                                     rec, ctx->X, ctx->Y, ctx->Z, ctx->par.box_size, BoxTable{});
    }
    Timed t(ctx, "gradients");
    // with its records made in the row kernels too, the force pass's flags are cleared here
    ctx->force_flags_clear = make;
    if (ctx->nact <= ctx->few_rows)
        k_gradients_rows<KERN, R><<<xcd_grid(ctx->nact), kListThreads, 0, ctx->stream>>>(
            ctx->dp, ctx->nact, ctx->act, ctx->X, ctx->Y, ctx->Z, ctx->H, rec,
            make ? GradSrc{ctx->H, ctx->M, ctx->RHO, ctx->PR, ctx->VX, ctx->VY, ctx->VZ, ctx->par.box_size} : GradSrc{}, ctx->T, ctx->CN,
            ctx->n, ctx->ell_s, ctx->ld_s, ctx->cnt_s, ctx->GR, ctx->GP, ctx->GV, BoxTable{}, make ? ctx->counters + 2 : nullptr);
    else
        k_gradients<KERN, R, false><<<xcd_grid(nblk(ctx->nact)), kBlock, 0, ctx->stream>>>(ctx->dp, ctx->nact, ctx->act, ctx->X, ctx->Y, ctx->Z, ctx->H,
                                                                                    rec, ctx->T, ctx->CN, ctx->n, ctx->ell_s, ctx->ld_s,
                                                                                    ctx->cnt_s, ctx->GR, ctx->GP, ctx->GV, BoxTable{});
    HIPCHK(hipGetLastError());
    ctx->gradients_done = true;
    return 0;
}

template <int KERN, typename R>
int force_stage(gz_context *ctx)
{
    if (!ctx->gradients_done) return fail(ctx, GZ_E_STATE, "hydro_force() before hydro_gradient_calc()");
    if (ctx->nact == 0) return 0;
    CHK(ensure_sym(ctx));
    ForceView F{ctx->n, ctx->X, ctx->Y, ctx->Z, ctx->VX, ctx->VY, ctx->VZ, ctx->M, ctx->H, ctx->U, ctx->RHO, ctx->PR, ctx->CN,
                ctx->FCE, ctx->DRK, ctx->GR, ctx->GP, ctx->GV, ctx->T, ctx->rungs};
    {
        Cell<R> *rec = (Cell<R> *)ctx->rec_force;
        // few rows among many particles: the row kernel makes the records of its row and list
        // entries (no packing launch) when the gradient launch cleared the pass's flags
        const bool make = ctx->nact <= kMakeForceRows && ctx->nact <= ctx->few_rows && ctx->pack_list && ctx->force_flags_clear &&
                          !ctx->rung_stats;
        ctx->force_flags_clear = false;
        if (!make) {
            Timed t(ctx, "pack_force");
            if (ctx->pack_list)
                k_pack_force_list<R><<<xcd_grid(ctx->nact), 64, 0, ctx->stream>>>(ctx->nact, ctx->act, ctx->ell_s, ctx->ld_s, ctx->cnt_s,
                                                                                  ctx->par.gamma, F, rec, ctx->counters + 2,
                                                                                  ctx->par.box_size, BoxTable{});
            else
                k_pack_force<R><<<nblk(ctx->n), kBlock, 0, ctx->stream>>>(ctx->par.gamma, F, rec, ctx->counters + 2, ctx->par.box_size,
                                                                          BoxTable{});
        }
        Timed t(ctx, "force");
        // a speculative first density round: the pass's kernels with lasting effects wait on its gate
        const int *gate = ctx->spec ? ctx->counters + kPassGate : nullptr;
        if (ctx->nact <= ctx->few_rows) {
            // the kernel that counts the faces' Riemann rungs when asked (the same code otherwise)
            auto rows_kernel = ctx->rung_stats ? k_force_rows<KERN, R, true>
                               : make          ? k_force_rows<KERN, R, false, kListThreads, true>
                                               : k_force_rows<KERN, R, false>;
            rows_kernel<<<xcd_grid(ctx->nact), kListThreads, 0, ctx->stream>>>(
                ctx->dp, gam<R>(ctx->dp), ctx->nact, ctx->act, F, rec, ctx->ell_s, ctx->ld_s, ctx->cnt_s,
                ctx->all_active ? nullptr : ctx->ACT, ctx->TB, ctx->MSV, ctx->WAKE, ctx->ACC, ctx->DTU, ctx->MSV, ctx->counters + 2,
                ctx->counters + 4, ctx->nact <= kRowPairAtomics ? ctx->u64c + 65 : nullptr, BoxTable{}, gate);
        } else {
            if (!ctx->rare) CHK(dalloc(ctx, &ctx->rare, ctx->cap));
            const ForceRows a{ctx->nact, ctx->act, ctx->ell_s, ctx->ld_s, ctx->cnt_s, ctx->all_active ? nullptr : ctx->ACT, ctx->TB,
                              ctx->MSV, ctx->WAKE, ctx->ACC, ctx->DTU, ctx->MSV, ctx->counters + 2, ctx->counters + 4, nullptr,
                              ctx->sched_valid ? ctx->sched : nullptr, ctx->rare};
            const dim3 grid = xcd_grid(nblk(ctx->nact));
            // one kernel with the whole ladder when most rows took a later rung in the last pass
            // (the two kernels then evaluate most rows twice): by the option "force_split", or
            // on its default by the last pass's share of such rows in the FP32 form (where the
            // whole ladder costs few registers more; the FP64 whole-ladder kernel spills)
            const bool single = ctx->force_split == 0 ||
                                (ctx->force_split < 0 && std::is_same<R, float>::value && ctx->rare_share > kRareShareSingle);
            ctx->split_ran = !ctx->rung_stats && !single;
            if (ctx->rung_stats) {
                k_force<KERN, R, false, true, false><<<grid, kBlock, 0, ctx->stream>>>(ctx->dp, gam<R>(ctx->dp), F, rec, a, BoxTable{});
            } else if (single) {
                k_force<KERN, R, false, false, false><<<grid, kBlock, 0, ctx->stream>>>(ctx->dp, gam<R>(ctx->dp), F, rec, a, BoxTable{});
            } else {
                // the rows whose faces all take the Riemann ladder's common case, then the others
                // with the whole ladder
                // FP64: two or three waves per SIMD by the option "force_waves" (default: three
                // with the search over many levels, else two)
                const int waves = ctx->force_waves ? ctx->force_waves : (ctx->wide ? 3 : 2);
                if constexpr (std::is_same<R, double>::value) {
                    if (waves == 3)
                        k_force_w3<KERN><<<grid, kBlock, 0, ctx->stream>>>(ctx->dp, gam<R>(ctx->dp), F, rec, a, BoxTable{});
                    else
                        k_force<KERN, R, false, false, true><<<grid, kBlock, 0, ctx->stream>>>(ctx->dp, gam<R>(ctx->dp), F, rec, a,
                                                                                                BoxTable{});
                } else {
                    k_force<KERN, R, false, false, true><<<grid, kBlock, 0, ctx->stream>>>(ctx->dp, gam<R>(ctx->dp), F, rec, a, BoxTable{});
                }
                k_force_rare<KERN, R, false><<<std::min(nblk(ctx->nact), kRareBlocks), kBlock, 0, ctx->stream>>>(
                    ctx->dp, gam<R>(ctx->dp), F, rec, a, BoxTable{});
            }
        }
        // the pairs of the pass: each row's workgroup adds its own while a few hundred rows are
        // active (atomics on one address execute one after another)
        if (ctx->nact > ctx->few_rows || ctx->nact > kRowPairAtomics)
            k_sum_counts<<<std::min(nblk(ctx->nact), 256u), kBlock, 0, ctx->stream>>>(ctx->nact, ctx->cnt_s, ctx->u64c + 65, gate);
        HIPCHK(hipGetLastError());
    }
    CHK(read_counters(ctx));
    if (ctx->split_ran) {
        // rows the common kernel left to k_force_rare (err[12] = counters[14])
        ctx->rare_share = (double)ctx->hcounters[14] / (double)std::max(ctx->nact, 1);
        ctx->split_ran = false;
    }
    if (ctx->spec) {
        ctx->spec = false;
        const int g = ctx->hcounters[kPassGate];
        if (g > 0) {
            ctx->st.density_iterations = 1;
            if (g == GZ_STOP_MAXITER) return fail(ctx, GZ_STOP_MAXITER, "density(): kernel radius iteration did not converge in %d rounds", kMaxIter);
            if (g == GZ_STOP_BRACKET) return fail(ctx, GZ_STOP_BRACKET, "density(): empty radius bracket (Right==0 && Left==0)");
            return fail(ctx, g, "density(): device stop code %d", g);
        }
        if (g < 0) {
            // rows still iterate after the first round, or a radius outgrew the bound the list's
            // grid took: the rounds go on from the counters just read, then the list, the
            // gradients and this pass again (the gated kernels did nothing)
            {
                const int rc_d = density_rounds<KERN, R>(ctx, 1, false);
                if (rc_d) return rc_d;
            }
            int rc = gradients_stage<KERN, R>(ctx);
            if (!rc) rc = force_stage<KERN, R>(ctx);
            return rc;
        }
    }
    // rows the common kernel left to the whole Riemann ladder (counters[14]: zero unless it ran)
    if (ctx->nact > ctx->few_rows && !ctx->rung_stats) ctx->st.force_rare_rows += ctx->hcounters[14];
    if (ctx->lazy && ctx->hmax_bound > 0) {
        // the radii the passes' lazy drifts produced (u64c[75], read back with the counters)
        double hl;
        memcpy(&hl, &ctx->hu64c[75], sizeof hl);
        ctx->hmax_bound = std::max(ctx->hmax_bound, hl);
    }
    if (ctx->sym_check_pending) {
        ctx->sym_check_pending = false;
        const int over = ctx->hcounters[3];
        if (over) {
            // the list the two passes used was cut short: a wider one, and both passes again
            ctx->cap_s = std::max(over + 16, ctx->cap_s * 3 / 2);
            CHK(sym_alloc(ctx));
            ctx->st.ell_widen += 1;
            ctx->sym_valid = false;
            const bool d = ctx->defer_sym_check;
            ctx->defer_sym_check = false;
            int rc = gradients_stage<KERN, R>(ctx);
            if (!rc) rc = force_stage<KERN, R>(ctx);
            ctx->defer_sym_check = d;
            return rc;
        }
    }
    if (ctx->hcounters[4]) ctx->need_wakeup = 1;
    if (ctx->hcounters[2]) {
        // name the first pair without a solution (caller's indices)
        int pair[2] = {-1, -1}, pi = -1, pj = -1;
        if (hipMemcpy(pair, ctx->counters + 12, sizeof pair, hipMemcpyDeviceToHost) == hipSuccess && pair[0] >= 0 && pair[1] >= 0) {
            (void)hipMemcpy(&pi, ctx->perm + pair[0], sizeof(int), hipMemcpyDeviceToHost);
            (void)hipMemcpy(&pj, ctx->perm + pair[1], sizeof(int), hipMemcpyDeviceToHost);
        }
        return fail(ctx, GZ_STOP_RIEMANN, "hydro_force(): Riemann solver failed to find a positive pressure (first pair: particles %d and %d)",
                    pi, pj);
    }
    return 0;
}

template <template <int, typename> class F>
int dispatch(gz_context *ctx)
{
    const bool k3 = ctx->par.kernel == 3;
    if (ctx->par.precision == GZ_FP64) return k3 ? F<3, double>::run(ctx) : F<6, double>::run(ctx);
    return k3 ? F<3, float>::run(ctx) : F<6, float>::run(ctx);
}

template <int K, typename R> struct DensityF { static int run(gz_context *c) { return density_stage<K, R>(c); } };
template <int K, typename R> struct GradF { static int run(gz_context *c) { return gradients_stage<K, R>(c); } };
template <int K, typename R> struct ForceF { static int run(gz_context *c) { return force_stage<K, R>(c); } };

// ---------------------------------------------------------------------------------------
// re-ordering: particles that moved are put back in Morton order (keys of the current
// positions), every per-particle field follows, and the displacement bound restarts at zero
// ---------------------------------------------------------------------------------------
// the spare buffers a re-ordering of the columns below takes (16 of width 1, 2 of width 3, one of
// widths 6 and 9), allocated ahead so that the timestep loop's first re-ordering does not wait
// for allocations (a re-ordering still allocates any that are missing)
int reserve_spares(gz_context *ctx)
{
    const int need[10] = {0, 16, 0, 2, 0, 0, 1, 0, 0, 1};
    for (int w = 1; w < 10; w++)
        while ((int)ctx->spare[w].size() < need[w]) {
            double *p = nullptr;
            CHK(dalloc(ctx, &p, (size_t)ctx->cap * w));
            ctx->spare[w].push_back(p);
        }
    return 0;
}

int resort(gz_context *ctx)
{
    const long n = ctx->n;
    ctx->density_ready = false;
    Timed t(ctx, "sort");
    k_morton_soa<<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->X, ctx->Y, ctx->Z, ctx->par.box_size, ctx->key2, ctx->ival);
    HIPCHK(hipGetLastError());
    size_t bytes = 0;
    HIPCHK(gz_sort_pairs(nullptr, bytes, ctx->key2, ctx->key, ctx->ival, ctx->ival2, (int)n, 0, 3 * kFineLevel, ctx->stream));
    CHK(ensure_tmp(ctx, bytes));
    HIPCHK(gz_sort_pairs(ctx->tmp, bytes, ctx->key2, ctx->key, ctx->ival, ctx->ival2, (int)n, 0, 3 * kFineLevel, ctx->stream));
    const int *order = ctx->ival2;
    // every per-particle column gathered into a spare buffer of its width, up to kPermCols
    // columns per launch; the spare then takes the field's place and the field's old buffer
    // becomes a spare (stream order: a later launch writes it only after this one read it)
    struct F { double **src; double **dst; int w; } fields[] = {
        {&ctx->X, nullptr, 1}, {&ctx->Y, nullptr, 1}, {&ctx->Z, nullptr, 1}, {&ctx->VX, nullptr, 1}, {&ctx->VY, nullptr, 1},
        {&ctx->VZ, nullptr, 1}, {&ctx->VKX, nullptr, 1}, {&ctx->VKY, nullptr, 1}, {&ctx->VKZ, nullptr, 1}, {&ctx->M, nullptr, 1},
        {&ctx->H, nullptr, 1}, {&ctx->U, nullptr, 1}, {&ctx->UK, nullptr, 1}, {&ctx->NN, nullptr, 1}, {&ctx->DRK, nullptr, 1},
        {&ctx->DIVV, nullptr, 1}, {&ctx->RHO, nullptr, 1}, {&ctx->PR, nullptr, 1}, {&ctx->CN, nullptr, 1}, {&ctx->FCE, nullptr, 1},
        {&ctx->DTU, nullptr, 1}, {&ctx->MSV, nullptr, 1}, {&ctx->SOFT, nullptr, 1}, {reinterpret_cast<double **>(&ctx->TIB), nullptr, 1},
        {reinterpret_cast<double **>(&ctx->TIC), nullptr, 1},
        {reinterpret_cast<double **>(&ctx->DTS), nullptr, 1}, {&ctx->ACC, nullptr, 3}, {&ctx->FW, nullptr, 3}, {&ctx->GR, nullptr, 3},
        {&ctx->GP, nullptr, 3}, {&ctx->T, nullptr, 6}, {&ctx->GV, nullptr, 9}};
    PermCols P{};
    struct Swap { double **field; double *fresh; int w; };
    std::vector<Swap> swaps;
    auto flush = [&]() -> int {
        if (P.ncol == 0) return 0;
        k_permute_many<<<nblk(n), kBlock, 0, ctx->stream>>>(n, order, P);
        HIPCHK(hipGetLastError());
        for (const Swap &w : swaps) {
            ctx->spare[w.w].push_back(*w.field);
            *w.field = w.fresh;
        }
        swaps.clear();
        P = PermCols{};
        return 0;
    };
    for (const F &f : fields) {
        if (P.ncol + f.w > kPermCols) CHK(flush());
        double *out = f.dst ? *f.dst : nullptr;
        if (!out) {
            auto &pool = ctx->spare[f.w];
            if (pool.empty()) {
                double *p = nullptr;
                CHK(dalloc(ctx, &p, (size_t)ctx->cap * f.w));
                pool.push_back(p);
            }
            out = pool.back();
            pool.pop_back();
            swaps.push_back({f.src, out, f.w});
        }
        for (int c = 0; c < f.w; c++) {
            P.in[P.ncol] = (const unsigned long long *)(*f.src + (size_t)c * n);
            P.out[P.ncol] = (unsigned long long *)(out + (size_t)c * n);
            P.ncol++;
        }
    }
    CHK(flush());
    // the int columns: gathered into ival, which then takes the column's place
    int **ints[3] = {&ctx->TB, &ctx->WAKE, &ctx->perm};
    for (int **a : ints) {
        k_permute_int<<<nblk(n), kBlock, 0, ctx->stream>>>(n, order, *a, ctx->ival);
        HIPCHK(hipGetLastError());
        std::swap(*a, ctx->ival);
    }
    k_permute_u8<<<nblk(n), kBlock, 0, ctx->stream>>>(n, order, ctx->ACT, (uint8_t *)ctx->stage);
    HIPCHK(hipGetLastError());
    CHK(copy_dev(ctx, ctx->ACT, ctx->stage, n));
    ctx->dmax = 0;
    ctx->key_version++;
    ctx->sym_valid = false;
    ctx->rows_level = -1;
    ctx->st.resorts += 1;
    return 0;
}

// ---------------------------------------------------------------------------------------
// the timestep loop (host side of the integer timeline)
// ---------------------------------------------------------------------------------------
constexpr long long kTimebaseHost = 1LL << kTimebins;

TimeParams time_params(gz_context *ctx)
{
    TimeParams T;
    T.courant_fac = ctx->ip.courant_fac;
    T.err_tol = ctx->ip.err_tol_int_accuracy;
    T.max_dt = ctx->ip.max_size_timestep;
    T.min_dt = ctx->ip.min_size_timestep;
    T.min_egy = ctx->ip.min_egy_spec;
    T.core_size = ctx->par.kernel == 3 ? (1.0 / 2.0) : (2.0 / 5.0);  // KERNEL_CORE_SIZE, src/mesh/kernel.h:59,113
    T.tbi = ctx->tbi;
    T.ti_current = ctx->ti_current;
    T.active_mask = ctx->mask;
    T.min_h = ctx->par.min_kernel_radius;
    T.max_h = ctx->par.max_kernel_radius;
    T.gamma = ctx->par.gamma;
    T.box = ctx->par.box_size;
    T.full_frac = ctx->ip.domain_build_active_fraction;
    T.n_total = ctx->n;
    T.lazy_rows = kLazyMaxRows;
    T.lazy_ok = ctx->wide ? 0 : 1;
    return T;
}

ParticleView particle_view(gz_context *c)
{
    ParticleView S;
    S.n = c->n;
    S.X = c->X; S.Y = c->Y; S.Z = c->Z; S.VX = c->VX; S.VY = c->VY; S.VZ = c->VZ; S.VKX = c->VKX; S.VKY = c->VKY; S.VKZ = c->VKZ;
    S.M = c->M; S.H = c->H; S.U = c->U; S.UK = c->UK; S.NN = c->NN; S.DIVV = c->DIVV; S.RHO = c->RHO; S.PR = c->PR;
    S.ACC = c->ACC; S.DTU = c->DTU; S.MSV = c->MSV; S.SOFT = c->SOFT;
    S.TB = c->TB; S.WAKE = c->WAKE; S.TIB = c->TIB; S.DTS = c->DTS;
    S.TIC = c->TIC;
    S.vkmax = c->u64c + 74;
    return S;
}

int read_bins(gz_context *ctx)
{
    const long n = ctx->n;
    HIPCHK(hipMemsetAsync(ctx->u64c, 0, 64 * sizeof(unsigned long long), ctx->stream));
    k_bin_hist<<<std::min(nblk(n), 1024u), kBlock, 0, ctx->stream>>>(n, ctx->TB, ctx->u64c, BoxTable{});
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ctx->hu64c, ctx->u64c, 64 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int b = 0; b < kTimebins; b++) ctx->bin_count[b] = (long long)ctx->hu64c[b];
    return 0;
}

// the earliest next kick over the occupied bins (src/core/run.cc:870-890, src/core/timestep.cc:1577-1591)
long long next_kick(const gz_context *ctx)
{
    long long ti_next_kick = kTimebaseHost;
    for (int b = 0; b < kTimebins; b++) {
        if (!ctx->bin_count[b]) continue;
        long long t;
        if (b > 0) {
            const long long dt_bin = 1LL << b;
            t = (ctx->ti_current / dt_bin) * dt_bin + dt_bin;
        } else {
            t = ctx->ti_current;
        }
        if (t < ti_next_kick) ti_next_kick = t;
    }
    return ti_next_kick;
}

// the active list of the current bins (make_list_of_active_particles, src/core/run.cc:997-1009)
int active_from_bins(gz_context *ctx)
{
    const long n = ctx->n;
    k_active_from_bins<<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->mask, ctx->TB, ctx->M, ctx->H, ctx->ACT, ctx->SOFT);
    HIPCHK(hipGetLastError());
    ctx->all_active = false;
    ctx->sym_valid = false;
    return build_active(ctx);
}

int hydro_pass(gz_context *ctx)
{
    ctx->dp.time = ctx->time;
    ctx->par.time = ctx->time;
    CHK(dispatch<DensityF>(ctx));
    ctx->defer_sym_check = true;
    int rc = dispatch<GradF>(ctx);
    if (!rc) rc = dispatch<ForceF>(ctx);
    ctx->defer_sym_check = false;
    ctx->sym_check_pending = false;
    if (rc) return rc;
    ctx->st.active_updates += ctx->nact;
    return 0;
}

int run_step(gz_context *ctx)
{
    const long n = ctx->n;
    TimeParams T = time_params(ctx);
    ParticleView S = particle_view(ctx);
    const long long ti_prev = ctx->ti_current;
    // The timeline part of the step runs on the device and is read back once: find_timesteps
    // (src/core/timestep.cc:39-355) over the active list of the last sync point, the bin
    // occupation, process_wake_ups (:1565-1774), do_first_halfstep_kick (src/core/kicks.cc:34-97),
    // the next sync point and the drift of every particle to it (src/core/run.cc:857-994), and
    // make_list_of_active_particles (:997-1009) with the softening cache
    {
        Timed t(ctx, "timeline");
        // the step's counters: cleared by the last step's closing kick when it ran in the same
        // run of steps, else here
        if (!ctx->prepped) k_step_prep<<<1, 64, 0, ctx->stream>>>(ctx->counters, ctx->u64c);
        ctx->prepped = false;
        ctx->ncells_clear = true;
        // the rows' new steps with their first half-kicks (one launch)
        if (ctx->nact > 0)
            k_find_timesteps<<<nblk(ctx->nact), kBlock, 0, ctx->stream>>>(T, S, ctx->nact, ctx->act, ctx->counters + 6, ctx->u64c,
                                                                          BoxTable{}, 1);
        if (ctx->need_wakeup) {
            int wakeup_bin_offset = 0;
            while ((1LL << wakeup_bin_offset) < (long long)kWakeup) wakeup_bin_offset++;
            k_wakeup_limits<<<1, 64, 0, ctx->stream>>>(T, ctx->u64c, ctx->counters + 8, BoxTable{});
            k_wakeups<<<nblk(n), kBlock, 0, ctx->stream>>>(T, S, ctx->counters + 8, wakeup_bin_offset, ctx->u64c, ctx->counters + 5, BoxTable{});
        }
    }
    // the drift of every particle to the next sync point fused with the row flags and counts
    // per block of particles of make_list_of_active_particles; then the list in ascending order
    // (no scan over every particle)
    const int nb_act = (int)((n + kActChunk - 1) / kActChunk);
    if (!ctx->act_bits) {
        const long m = (ctx->cap + kActChunk - 1) / kActChunk;
        CHK(dalloc(ctx, &ctx->act_bits, (size_t)m * (kActChunk / 64))); CHK(dalloc(ctx, &ctx->act_bcount, m));
    }
    {
        Timed t(ctx, "drift");
        // every block takes the next sync point from the occupation itself (no k_next_sync launch)
        k_drift_active<<<nb_act, 256, 0, ctx->stream>>>(T, S, ti_prev, ctx->u64c + 70, ctx->u64c + 64,
                                                         ctx->u64c + 66, 0.99 * ctx->par.max_kernel_radius, ctx->ACT, ctx->act_bits,
                                                         ctx->act_bcount, ctx->u64c + 67, ctx->u64c);
        HIPCHK(hipGetLastError());
    }
    {
        Timed t(ctx, "timeline");
        // with the list, the start of the coming density() iteration for its rows (the counters
        // it needs zero were cleared by k_step_prep)
        const DensityStart ds{ctx->H, 0.99 * ctx->par.max_kernel_radius, ctx->it_h, ctx->it_l, ctx->it_r, ctx->it_cn, ctx->it_ci,
                              ctx->it_it, ctx->need};
        k_active_compact<<<nb_act, 256, 0, ctx->stream>>>(n, nb_act, ctx->act_bits, ctx->act_bcount, ctx->act, ctx->counters + 10, ds);
        ctx->density_ready = true;
        HIPCHK(hipGetLastError());
    }
    ctx->need_wakeup = 0;
    HIPCHK(hipMemcpyAsync(ctx->hu64c, ctx->u64c, (80 + 8) * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int b = 0; b < kTimebins; b++) ctx->bin_count[b] = (long long)ctx->hu64c[b];
    ctx->n_woken = ctx->hcounters[5];
    if (ctx->hcounters[6] && !ctx->stop_code) ctx->stop_code = ctx->hcounters[6];
    ctx->num_steps++;
    ctx->ti_current = (long long)ctx->hu64c[70];
    ctx->mask = ctx->hu64c[71];
    ctx->time = ctx->ip.time_begin + ctx->ti_current * ctx->tbi;
    ctx->nact = ctx->hcounters[10];
    ctx->act_version++;
    ctx->all_active = false;
    ctx->sym_valid = false;
    ctx->st.n_active = ctx->nact;
    // a step that drifted only the active bins: the passes drift the cells they read
    ctx->last_full = ctx->hu64c[72] != 0;
    ctx->lazy = !ctx->last_full;
    if (ctx->last_full) ctx->st.full_drifts += 1;
    {
        double d, hm, ha;
        memcpy(&d, &ctx->hu64c[64], sizeof d);
        memcpy(&hm, &ctx->hu64c[66], sizeof hm);
        memcpy(&ha, &ctx->hu64c[67], sizeof ha);
        ctx->dmax += d;  // the bound grows by the step's bound on any particle's motion
        // every radius after the drift (a particle not drifted keeps its radius under the old bound)
        ctx->hmax_bound = ctx->last_full ? hm : std::max(ctx->hmax_bound, hm);
        ctx->hmax_rows = ha;     // the largest radius density() starts its rows from
    }
    // particles that moved far from where they were ordered are ordered again, after every
    // particle is drifted to now (the reference's full drift before a decomposition,
    // src/core/run.cc:380-390)
    if (ctx->last_rmax > 0 && ctx->dmax > ctx->resort_fraction * ctx->last_rmax) {
        if (ctx->lazy) {
            Timed t(ctx, "drift");
            k_drift_behind<<<nblk(n), kBlock, 0, ctx->stream>>>(time_params(ctx), particle_view(ctx), ctx->ti_current, nullptr);
            HIPCHK(hipGetLastError());
            ctx->lazy = false;
            ctx->hmax_bound = -1;  // taken again over every radius
        }
        CHK(resort(ctx));
        CHK(build_active(ctx));
    }
    // compute_hydro_densities_and_forces, then do_second_halfstep_kick (the columns' buffers
    // change places in a re-ordering: the view is taken again)
    T = time_params(ctx);
    S = particle_view(ctx);
    ctx->speculate = ctx->spec_enabled;
    const int rc_pass = hydro_pass(ctx);
    ctx->speculate = false;
    ctx->spec = false;
    ctx->density_ready = false;
    if (rc_pass) ctx->lazy = false;
    CHK(rc_pass);
    ctx->lazy = false;
    if (ctx->nact > 0) {
        Timed t(ctx, "timeline");
        // within a run of several steps, the closing kick also clears the next step's counters
        // (read back above for this step already)
        const bool prep = ctx->in_run_steps;
        k_halfstep_kick<<<nblk(ctx->nact), kBlock, 0, ctx->stream>>>(T, S, ctx->nact, ctx->act, 1, BoxTable{},
                                                                     prep ? ctx->counters : nullptr, prep ? ctx->u64c : nullptr);
        HIPCHK(hipGetLastError());
        ctx->prepped = prep;
    }
    if (ctx->stop_code)
        return fail(ctx, ctx->stop_code, "timestep loop: controlled stop %d (%s)", ctx->stop_code,
                    ctx->stop_code == 818 ? "a step of size zero on the integer timeline" : "step beyond the end of the timeline");
    return 0;
}

// the Riemann functions on n given cases (k_riemann_probe), staged through device buffers the
// caller provides (din: n*kProbeIn doubles, dout: n*kProbeOut doubles, drung: n)
int probe_run(gz_context *ctx, int mode, long n, const double *in, double *out, unsigned *rung, double *din, double *dout,
              unsigned *drung)
{
    HIPCHK(hipMemcpyAsync(din, in, (size_t)n * kProbeIn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (ctx->par.precision == GZ_FP64)
        k_riemann_probe<double><<<nblk(n), kBlock, 0, ctx->stream>>>(mode, gam<double>(ctx->dp), n, din, dout, drung);
    else
        k_riemann_probe<float><<<nblk(n), kBlock, 0, ctx->stream>>>(mode, gam<float>(ctx->dp), n, din, dout, drung);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dout, (size_t)n * kProbeOut * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(rung, drung, (size_t)n * sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int set_device(gz_context *ctx)
{
    HIPCHK(hipSetDevice(ctx->device));
    return 0;
}

// ---------------------------------------------------------------------------------------
// gz_hydro_pass_noalloc: the working memory of a pass carved from one static device array
// per device (device code sees `scratch`; host code uses only the address that
// hipGetSymbolAddress returned for the current device, never the array's name)
// ---------------------------------------------------------------------------------------
constexpr int kNoAllocDevices = 64;
unsigned char *g_scratch[kNoAllocDevices] = {};  // the array's address on each device, once set up
gz_context g_noalloc_ctx[kNoAllocDevices];       // the host side of each device's pass

}  // namespace

static __device__ __align__(256) unsigned char scratch[GZ_NOALLOC_SCRATCH_BYTES];

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------
extern "C" {

gz_context *gz_create(const gz_params *params, long capacity, int device)
{
    if (!params || capacity <= 0 || capacity > 0x7fffffffL) return nullptr;
    gz_context *ctx = new gz_context();
    ctx->device = device;
    if (set_device(ctx) || set_params(ctx, params)) { delete ctx; return nullptr; }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return nullptr; }
    if (alloc_state(ctx, capacity)) { free_state(ctx); hipStreamDestroy(ctx->stream); delete ctx; return nullptr; }
    return ctx;
}

void gz_destroy(gz_context *ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto &t : ctx->pending) { hipEventDestroy(t.a); hipEventDestroy(t.b); }
    for (auto e : ctx->event_pool) hipEventDestroy(e);
    free_state(ctx);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *gz_last_error(const gz_context *ctx) { return ctx ? ctx->err : "no context"; }

int gz_set_params(gz_context *ctx, const gz_params *params)
{
    if (!ctx) return GZ_E_ARG;
    gz_params old = ctx->par;
    int rc = set_params(ctx, params);
    if (rc) return rc;
    if (old.box_size != params->box_size || old.kernel != params->kernel) {
        ctx->density_done = ctx->gradients_done = ctx->sym_valid = false;
    }
    return 0;
}

int gz_load(gz_context *ctx, long n, const gz_fields *f)
{
    if (!ctx || !f) return GZ_E_ARG;
    if (n <= 0 || n > ctx->cap) return fail(ctx, GZ_E_ARG, "n=%ld outside (0, capacity=%ld]", n, ctx->cap);
    if (!f->pos || !f->vel || !f->mass || !f->u || !f->h) return fail(ctx, GZ_E_ARG, "pos, vel, mass, u and h are required");
    CHK(set_device(ctx));
    ctx->n = n;
    ctx->density_ready = false;
    // order along the Morton curve: keys from the caller's positions, stable radix sort
    HIPCHK(hipMemcpyAsync(ctx->stage, f->pos, (size_t)n * 3 * sizeof(double), hipMemcpyDefault, ctx->stream));
    {
        Timed t(ctx, "sort");
        k_morton<<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->stage, ctx->par.box_size, ctx->key2, ctx->ival);
        HIPCHK(hipGetLastError());
        size_t bytes = 0;
        HIPCHK(gz_sort_pairs(nullptr, bytes, ctx->key2, ctx->key, ctx->ival, ctx->perm, (int)n, 0, 3 * kFineLevel, ctx->stream));
        CHK(ensure_tmp(ctx, bytes));
        HIPCHK(gz_sort_pairs(ctx->tmp, bytes, ctx->key2, ctx->key, ctx->ival, ctx->perm, (int)n, 0, 3 * kFineLevel, ctx->stream));
    }
    // positions and velocities split into SoA columns
    k_gather<<<nblk(n), kBlock, 0, ctx->stream>>>(n, 3, ctx->perm, ctx->stage, ctx->stage + 3 * n, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ctx->X, ctx->stage + 3 * n, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->Y, ctx->stage + 4 * n, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->Z, ctx->stage + 5 * n, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->stage, f->vel, (size_t)n * 3 * sizeof(double), hipMemcpyDefault, ctx->stream));
    k_gather<<<nblk(n), kBlock, 0, ctx->stream>>>(n, 3, ctx->perm, ctx->stage, ctx->stage + 3 * n, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ctx->VX, ctx->stage + 3 * n, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->VY, ctx->stage + 4 * n, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->VZ, ctx->stage + 5 * n, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    CHK(upload_field(ctx, f->mass, 1, ctx->M, false));
    CHK(upload_field(ctx, f->u, 1, ctx->U, false));
    CHK(upload_field(ctx, f->h, 1, ctx->H, false));
    struct Opt { const double *src; int w; double *dst; } opts[] = {
        {f->numngb, 1, ctx->NN}, {f->drkern, 1, ctx->DRK}, {f->divvel, 1, ctx->DIVV}, {f->rho, 1, ctx->RHO},
        {f->press, 1, ctx->PR}, {f->nvt, 6, ctx->T}, {f->nvt_fw, 3, ctx->FW}, {f->cn, 1, ctx->CN}, {f->fce, 1, ctx->FCE},
        {f->grad_rho, 3, ctx->GR}, {f->grad_p, 3, ctx->GP}, {f->grad_v, 9, ctx->GV}, {f->accel, 3, ctx->ACC},
        {f->dtu, 1, ctx->DTU}, {f->maxsig, 1, ctx->MSV}};
    for (auto &o : opts) {
        if (o.src) { CHK(upload_field(ctx, o.src, o.w, o.dst, o.w > 1)); }
        else { HIPCHK(hipMemsetAsync(o.dst, 0, (size_t)n * o.w * sizeof(double), ctx->stream)); }
    }
    int *ti[2] = {f->timebin, f->wakeup};
    int *td[2] = {ctx->TB, ctx->WAKE};
    for (int q = 0; q < 2; q++) {
        if (ti[q]) {
            HIPCHK(hipMemcpyAsync(ctx->ival2, ti[q], n * sizeof(int), hipMemcpyDefault, ctx->stream));
            k_gather_int<<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->perm, ctx->ival2, td[q]);
            HIPCHK(hipGetLastError());
        } else {
            HIPCHK(hipMemsetAsync(td[q], 0, n * sizeof(int), ctx->stream));
        }
    }
    // timestep-loop fields; absent ones start from the predicted fields (as init() sets them)
    if (f->vel_kick) { CHK(upload_vec3(ctx, f->vel_kick, ctx->VKX, ctx->VKY, ctx->VKZ)); }
    else { CHK(copy_dev(ctx, ctx->VKX, ctx->VX, n * 8)); CHK(copy_dev(ctx, ctx->VKY, ctx->VY, n * 8)); CHK(copy_dev(ctx, ctx->VKZ, ctx->VZ, n * 8)); }
    if (f->u_kick) { CHK(upload_field(ctx, f->u_kick, 1, ctx->UK, false)); } else { CHK(copy_dev(ctx, ctx->UK, ctx->U, n * 8)); }
    if (f->soft) { CHK(upload_field(ctx, f->soft, 1, ctx->SOFT, false)); } else { CHK(copy_dev(ctx, ctx->SOFT, ctx->H, n * 8)); }
    long long *tl[3] = {f->ti_begstep, f->dt_step, f->ti_current};
    long long *tdst[3] = {ctx->TIB, ctx->DTS, ctx->TIC};
    for (int q = 0; q < 3; q++) {
        if (tl[q]) {
            HIPCHK(hipMemcpyAsync(ctx->stage, tl[q], n * sizeof(long long), hipMemcpyDefault, ctx->stream));
            k_gather_ll<<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->perm, (const long long *)ctx->stage, tdst[q]);
            HIPCHK(hipGetLastError());
        } else {
            HIPCHK(hipMemsetAsync(tdst[q], 0, n * sizeof(long long), ctx->stream));
        }
    }
    ctx->dmax = 0;
    ctx->key_version++;
    ctx->hmax_bound = -1;
    ctx->run_started = false;
    ctx->all_active = true;
    // a caller that hands over the state density() leaves (or also the gradients) may start at
    // the later stage, as the reference's inactive particles keep that state between steps
    ctx->density_done = f->rho && f->press && f->nvt && f->cn && f->fce && f->drkern;
    ctx->gradients_done = ctx->density_done && f->grad_rho && f->grad_p && f->grad_v;
    ctx->sym_valid = false;
    CHK(build_active(ctx));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int gz_store(gz_context *ctx, gz_fields *f)
{
    if (!ctx || !f) return GZ_E_ARG;
    if (ctx->n <= 0) return fail(ctx, GZ_E_STATE, "nothing loaded");
    CHK(set_device(ctx));
    const long n = ctx->n;
    if (f->pos) {
        HIPCHK(hipMemcpyAsync(ctx->stage + 3 * n, ctx->X, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->stage + 4 * n, ctx->Y, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->stage + 5 * n, ctx->Z, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        k_scatter<<<nblk(n), kBlock, 0, ctx->stream>>>(n, 3, ctx->perm, ctx->stage + 3 * n, ctx->stage, 1);
        HIPCHK(hipMemcpyAsync(f->pos, ctx->stage, (size_t)n * 3 * sizeof(double), hipMemcpyDefault, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    if (f->vel) {
        HIPCHK(hipMemcpyAsync(ctx->stage + 3 * n, ctx->VX, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->stage + 4 * n, ctx->VY, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->stage + 5 * n, ctx->VZ, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        k_scatter<<<nblk(n), kBlock, 0, ctx->stream>>>(n, 3, ctx->perm, ctx->stage + 3 * n, ctx->stage, 1);
        HIPCHK(hipMemcpyAsync(f->vel, ctx->stage, (size_t)n * 3 * sizeof(double), hipMemcpyDefault, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    struct Opt { double *dst; int w; const double *src; } opts[] = {
        {f->mass, 1, ctx->M}, {f->h, 1, ctx->H}, {f->u, 1, ctx->U}, {f->numngb, 1, ctx->NN}, {f->drkern, 1, ctx->DRK},
        {f->divvel, 1, ctx->DIVV}, {f->rho, 1, ctx->RHO}, {f->press, 1, ctx->PR}, {f->nvt, 6, ctx->T}, {f->nvt_fw, 3, ctx->FW},
        {f->cn, 1, ctx->CN}, {f->fce, 1, ctx->FCE}, {f->grad_rho, 3, ctx->GR}, {f->grad_p, 3, ctx->GP}, {f->grad_v, 9, ctx->GV},
        {f->accel, 3, ctx->ACC}, {f->dtu, 1, ctx->DTU}, {f->maxsig, 1, ctx->MSV}};
    for (auto &o : opts) {
        if (!o.dst) continue;
        CHK(download_field(ctx, o.dst, o.w, o.src, o.w > 1));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    int *ti[2] = {f->timebin, f->wakeup};
    const int *td[2] = {ctx->TB, ctx->WAKE};
    for (int q = 0; q < 2; q++) {
        if (!ti[q]) continue;
        k_scatter_int<<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->perm, td[q], ctx->ival2);
        HIPCHK(hipMemcpyAsync(ti[q], ctx->ival2, n * sizeof(int), hipMemcpyDefault, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    if (f->vel_kick) CHK(download_vec3(ctx, f->vel_kick, ctx->VKX, ctx->VKY, ctx->VKZ));
    struct Opt2 { double *dst; const double *src; } o2[] = {{f->u_kick, ctx->UK}, {f->soft, ctx->SOFT}};
    for (auto &o : o2) {
        if (!o.dst) continue;
        CHK(download_field(ctx, o.dst, 1, o.src, false));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    long long *tl[3] = {f->ti_begstep, f->dt_step, f->ti_current};
    const long long *tsrc[3] = {ctx->TIB, ctx->DTS, ctx->TIC};
    for (int q = 0; q < 3; q++) {
        if (!tl[q]) continue;
        k_scatter_ll<<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->perm, tsrc[q], (long long *)ctx->stage);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(tl[q], ctx->stage, n * sizeof(long long), hipMemcpyDefault, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return 0;
}

int gz_set_active(gz_context *ctx, const unsigned char *active)
{
    if (!ctx) return GZ_E_ARG;
    if (ctx->n <= 0) return fail(ctx, GZ_E_STATE, "nothing loaded");
    CHK(set_device(ctx));
    const long n = ctx->n;
    if (!active) {
        ctx->all_active = true;
    } else {
        HIPCHK(hipMemcpyAsync(ctx->stage, active, n, hipMemcpyHostToDevice, ctx->stream));
        k_gather_u8<<<nblk(n), kBlock, 0, ctx->stream>>>(n, ctx->perm, (const uint8_t *)ctx->stage, ctx->ACT);
        HIPCHK(hipGetLastError());
        ctx->all_active = false;
    }
    ctx->sym_valid = false;
    CHK(build_active(ctx));
    return 0;
}

int gz_density(gz_context *ctx)
{
    if (!ctx) return GZ_E_ARG;
    if (ctx->n <= 0) return fail(ctx, GZ_E_STATE, "nothing loaded");
    CHK(set_device(ctx));
    int rc = dispatch<DensityF>(ctx);
    collect_timings(ctx);
    return rc;
}

int gz_hydro_gradient_calc(gz_context *ctx)
{
    if (!ctx) return GZ_E_ARG;
    CHK(set_device(ctx));
    int rc = dispatch<GradF>(ctx);
    collect_timings(ctx);
    return rc;
}

int gz_hydro_force(gz_context *ctx)
{
    if (!ctx) return GZ_E_ARG;
    CHK(set_device(ctx));
    int rc = dispatch<ForceF>(ctx);
    collect_timings(ctx);
    return rc;
}

int gz_synchronize(gz_context *ctx)
{
    if (!ctx) return GZ_E_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int gz_permutation(gz_context *ctx, long *perm)
{
    if (!ctx || !perm) return GZ_E_ARG;
    std::vector<int> p(ctx->n);
    HIPCHK(hipMemcpy(p.data(), ctx->perm, ctx->n * sizeof(int), hipMemcpyDeviceToHost));
    for (long k = 0; k < ctx->n; k++) perm[k] = p[k];
    return 0;
}

int gz_stat(gz_context *ctx, const char *name, double *value)
{
    if (!ctx || !name || !value) return GZ_E_ARG;
    std::string k(name);
    if (k == "pairs_density" || k == "pairs_symmetric") {
        const bool sym = (k == "pairs_symmetric");
        const int *cnt = sym ? ctx->cnt_s : ctx->cnt_d;
        if ((sym && !ctx->sym_valid) || (!sym && !ctx->density_done) || ctx->nact == 0) { *value = 0; return 0; }
        std::vector<int> c(ctx->nact);
        HIPCHK(hipMemcpy(c.data(), cnt, ctx->nact * sizeof(int), hipMemcpyDeviceToHost));
        double s = 0;
        for (int v : c) s += v;
        *value = s;
        return 0;
    }
    if (k == "density_pairs_total") {
        // pairs of the few-row density rounds since gz_timing_reset, counted while timing is on
        unsigned long long v = 0;
        HIPCHK(hipMemcpy(&v, ctx->u64c + 77, sizeof v, hipMemcpyDeviceToHost));
        *value = (double)v;
        return 0;
    }
    if (k == "force_pairs_total") {
        unsigned long long v = 0;
        HIPCHK(hipMemcpy(&v, ctx->u64c + 65, sizeof v, hipMemcpyDeviceToHost));
        *value = (double)v;
        return 0;
    }
    {
        // the Riemann rung counts of the force passes since "riemann_stats" was last set
        static const char *rung_names[kRungCount] = {"riemann_hllc", "riemann_vacuum", "riemann_roe", "riemann_pvrs", "riemann_kt",
                                                     "riemann_exact", "riemann_exact_vacuum", "riemann_unphysical",
                                                     "riemann_retry_first_order", "riemann_retry_zero_velocity", "riemann_fail",
                                                     "riemann_faces", "entropic_considered", "entropic_applied"};
        for (int b = 0; b < kRungCount; b++)
            if (k == rung_names[b]) {
                unsigned long long v = 0;
                HIPCHK(hipMemcpy(&v, ctx->rungs + b, sizeof v, hipMemcpyDeviceToHost));
                *value = (double)v;
                return 0;
            }
    }
    using S = gz_context::Stats;
    static const struct { const char *name; double S::*v; } known[] = {
        {"density_iterations", &S::density_iterations}, {"ngb_rebuilds", &S::ngb_rebuilds}, {"n_active", &S::n_active},
        {"ell_widen", &S::ell_widen}, {"active_updates", &S::active_updates}, {"resorts", &S::resorts},
        {"full_drifts", &S::full_drifts}, {"force_rare_rows", &S::force_rare_rows}};
    for (const auto &kn : known)
        if (k == kn.name) { *value = ctx->st.*kn.v; return 0; }
    return fail(ctx, GZ_E_ARG, "unknown stat %s", name);
}

int gz_timing_enable(gz_context *ctx, int on)
{
    if (!ctx) return GZ_E_ARG;
    ctx->timing = on != 0;
    return 0;
}

int gz_timing_reset(gz_context *ctx)
{
    if (!ctx) return GZ_E_ARG;
    collect_timings(ctx);
    ctx->timings.clear();
    ctx->st.active_updates = 0;
    ctx->st.resorts = 0;
    ctx->st.force_rare_rows = 0;
    ctx->st.full_drifts = 0;
    HIPCHK(hipMemsetAsync(ctx->u64c + 65, 0, sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->u64c + 77, 0, sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int gz_timing(gz_context *ctx, const char *name, long *launches, double *total_ms)
{
    if (!ctx || !name) return GZ_E_ARG;
    collect_timings(ctx);
    auto it = ctx->timings.find(name);
    if (launches) *launches = it == ctx->timings.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == ctx->timings.end() ? 0 : it->second.second;
    return 0;
}

int gz_run_begin(gz_context *ctx, const gz_integ_params *ip)
{
    if (!ctx || !ip) return GZ_E_ARG;
    if (ctx->n <= 0) return fail(ctx, GZ_E_STATE, "nothing loaded");
    if (!(ip->time_max > ip->time_begin)) return fail(ctx, GZ_E_ARG, "time_max must exceed time_begin");
    if (!(ip->courant_fac > 0) || !(ip->err_tol_int_accuracy > 0) || !(ip->max_size_timestep > 0) || ip->min_size_timestep < 0 ||
        ip->min_egy_spec < 0 || !(ip->domain_build_active_fraction >= 0))
        return fail(ctx, GZ_E_ARG,
                    "courant_fac, err_tol_int_accuracy, max_size_timestep must be > 0; min_size_timestep, min_egy_spec, "
                    "domain_build_active_fraction >= 0");
    CHK(set_device(ctx));
    CHK(reserve_spares(ctx));
    const long n = ctx->n;
    ctx->ip = *ip;
    ctx->tbi = (ip->time_max - ip->time_begin) / (double)kTimebaseHost;  // src/core/init.cc:113
    ctx->ti_current = 0;
    ctx->time = ip->time_begin;
    ctx->num_steps = 0;
    ctx->n_woken = 0;
    ctx->need_wakeup = 0;
    ctx->stop_code = 0;
    ctx->mask = (1ull << kTimebins) - 1;  // every bin active at the start (src/core/init.cc:583)
    ctx->dp.time_begin = ip->time_begin;
    ctx->par.time_begin = ip->time_begin;
    // predicted = current, every particle in bin 0 at the start of the timeline
    CHK(copy_dev(ctx, ctx->VX, ctx->VKX, n * 8)); CHK(copy_dev(ctx, ctx->VY, ctx->VKY, n * 8)); CHK(copy_dev(ctx, ctx->VZ, ctx->VKZ, n * 8));
    CHK(copy_dev(ctx, ctx->U, ctx->UK, n * 8));
    HIPCHK(hipMemsetAsync(ctx->TB, 0, n * sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->WAKE, 0, n * sizeof(int), ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->TIB, 0, n * sizeof(long long), ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->TIC, 0, n * sizeof(long long), ctx->stream));
    ctx->lazy = false;
    ctx->last_full = true;
    HIPCHK(hipMemsetAsync(ctx->DTS, 0, n * sizeof(long long), ctx->stream));
    // the bound on the kicked velocities, over every particle (the kicks raise it from here on)
    HIPCHK(hipMemsetAsync(ctx->u64c + 74, 0, sizeof(unsigned long long), ctx->stream));
    k_vk_max<<<nblk(n), kBlock, 0, ctx->stream>>>(particle_view(ctx));
    HIPCHK(hipGetLastError());
    CHK(active_from_bins(ctx));
    ctx->dp.time = ctx->time;
    ctx->par.time = ctx->time;
    int rc = dispatch<DensityF>(ctx);                    // init(): density()
    if (!rc) rc = copy_dev(ctx, ctx->SOFT, ctx->H, n * 8);  // compute_all_force_softening(1)
    if (!rc) rc = hydro_pass(ctx);                       // run(): compute_hydro_densities_and_forces()
    if (!rc) rc = read_bins(ctx);
    collect_timings(ctx);
    if (rc) return rc;
    ctx->run_started = true;
    return 0;
}

int gz_run_step(gz_context *ctx)
{
    if (!ctx) return GZ_E_ARG;
    if (!ctx->run_started) return fail(ctx, GZ_E_STATE, "gz_run_step before gz_run_begin");
    CHK(set_device(ctx));
    int rc = run_step(ctx);
    collect_timings(ctx);
    return rc;
}

int gz_run_steps(gz_context *ctx, int nsteps, long long *n_active, long long *n_woken)
{
    if (!ctx || nsteps < 0) return GZ_E_ARG;
    if (!ctx->run_started) return fail(ctx, GZ_E_STATE, "gz_run_steps before gz_run_begin");
    CHK(set_device(ctx));
    int rc = 0;
    for (int k = 0; k < nsteps && !rc; k++) {
        ctx->in_run_steps = k + 1 < nsteps;  // another step of this call follows
        rc = run_step(ctx);
        if (!rc) {
            if (n_active) n_active[k] = ctx->nact;
            if (n_woken) n_woken[k] = ctx->n_woken;
        }
    }
    ctx->in_run_steps = false;
    ctx->prepped = false;
    collect_timings(ctx);
    return rc;
}

int gz_set_option(gz_context *ctx, const char *name, double value)
{
    if (!ctx || !name) return GZ_E_ARG;
    std::string k(name);
    if (k == "resort_fraction") {
        if (!(value > 0)) return fail(ctx, GZ_E_ARG, "resort_fraction must be > 0");
        ctx->resort_fraction = value;
        return 0;
    }
    if (k == "few_rows") {
        if (!(value >= 0) || value > 2e9) return fail(ctx, GZ_E_ARG, "few_rows must be in [0, 2e9]");
        ctx->few_rows = (int)value;
        return 0;
    }
    if (k == "row_schedule") {
        const long v = (long)value;
        if ((double)v != value || v < 0 || v > kSchedMaxSpan || (v & (v - 1)) || (v > 0 && v < 64))
            return fail(ctx, GZ_E_ARG, "row_schedule must be 0 or a power of two in [64, %d]", kSchedMaxSpan);
        ctx->sched_span = (int)v;
        ctx->sym_valid = false;
        return 0;
    }
    if (k == "riemann_stats") {
        if (value != 0 && value != 1) return fail(ctx, GZ_E_ARG, "riemann_stats must be 0 or 1");
        ctx->rung_stats = value != 0;
        HIPCHK(hipMemsetAsync(ctx->rungs, 0, 16 * sizeof(unsigned long long), ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return 0;
    }
    if (k == "speculate") {
        if (value != 0 && value != 1) return fail(ctx, GZ_E_ARG, "speculate must be 0 or 1");
        ctx->spec_enabled = value != 0;
        return 0;
    }
    if (k == "force_waves") {
        if (value != 0 && value != 2 && value != 3) return fail(ctx, GZ_E_ARG, "force_waves must be 0, 2 or 3");
        ctx->force_waves = (int)value;
        return 0;
    }
    if (k == "force_split") {
        if (value != -1 && value != 0 && value != 1) return fail(ctx, GZ_E_ARG, "force_split must be -1, 0 or 1");
        ctx->force_split = (int)value;
        return 0;
    }
    if (k == "fd_correction") {
        if (value != 0 && value != 1) return fail(ctx, GZ_E_ARG, "fd_correction must be 0 or 1");
        ctx->fd = value != 0;
        ctx->dp.fd4alpha = ctx->fd ? 4.0 * fd_alpha(ctx->par.kernel) : 0.0;
        ctx->density_done = false;
        return 0;
    }
    if (k == "wide_search") {
        if (value != 0 && value != 1) return fail(ctx, GZ_E_ARG, "wide_search must be 0 or 1");
        ctx->wide = (int)value;
        ctx->sym_valid = false;
        return 0;
    }
    return fail(ctx, GZ_E_ARG, "unknown option %s", name);
}

int gz_timeline_get(gz_context *ctx, gz_timeline *out)
{
    if (!ctx || !out) return GZ_E_ARG;
    memset(out, 0, sizeof *out);
    out->ti_current = ctx->ti_current;
    out->time = ctx->time;
    out->timebase_interval = ctx->tbi;
    out->num_steps = ctx->num_steps;
    out->n_active = ctx->nact;
    out->n_woken = ctx->n_woken;
    for (int b = 0; b < kTimebins; b++) out->bin_count[b] = ctx->bin_count[b];
    out->bin_active = ctx->mask;
    out->stop_code = ctx->stop_code;
    out->full_drift = ctx->last_full ? 1 : 0;
    return 0;
}

int gz_drift_all(gz_context *ctx)
{
    if (!ctx) return GZ_E_ARG;
    if (!ctx->run_started) return 0;  // outside a run every particle stands at the loaded time
    CHK(set_device(ctx));
    k_drift_behind<<<nblk(ctx->n), kBlock, 0, ctx->stream>>>(time_params(ctx), particle_view(ctx), ctx->ti_current, nullptr);
    HIPCHK(hipGetLastError());
    ctx->hmax_bound = -1;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int gz_riemann_probe(gz_context *ctx, int mode, long n, const double *in, double *out, unsigned *rung)
{
    if (!ctx) return GZ_E_ARG;
    if (mode < 0 || mode > 4) return fail(ctx, GZ_E_ARG, "mode must be 0..4");
    if (n < 0 || n > (1L << 24)) return fail(ctx, GZ_E_ARG, "n=%ld outside [0, 2^24]", n);
    if (n > 0 && (!in || !out || !rung)) return fail(ctx, GZ_E_ARG, "in, out and rung are required");
    if (n == 0) return 0;
    CHK(set_device(ctx));
    double *din = nullptr, *dout = nullptr;
    unsigned *drung = nullptr;
    int rc = dalloc(ctx, &din, (size_t)n * kProbeIn);
    if (!rc) rc = dalloc(ctx, &dout, (size_t)n * kProbeOut);
    if (!rc) rc = dalloc(ctx, &drung, (size_t)n);
    if (!rc) rc = probe_run(ctx, mode, n, in, out, rung, din, dout, drung);
    dfree(din);
    dfree(dout);
    dfree(drung);
    return rc;
}

long gz_neighbors(gz_context *ctx, int symmetric, long *offsets, int *nbr, long cap)
{
    if (!ctx || !offsets) return -GZ_E_ARG;
    if (set_device(ctx)) return -GZ_E_HIP;
    const long n = ctx->n;
    // a fresh exact list over every particle (no buffer): one-way at h_i or symmetric
    std::vector<int> act_h(ctx->nact);
    if (hipMemcpy(act_h.data(), ctx->act, ctx->nact * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return -GZ_E_HIP;
    int *ell = nullptr, *cnt = nullptr;
    long ld = 0;
    int capl = 0;
    double hmax = 0;
    if (max_on_device(ctx, ctx->H, n, &hmax)) return -GZ_E_HIP;
    if (build_grid(ctx, std::min(hmax, 0.5 * ctx->par.box_size), true)) return -GZ_E_HIP;
    if (dalloc(ctx, &cnt, ctx->nact)) return -GZ_E_MEMORY;
    int rc = 0;
    if (symmetric && ctx->wide) {
        // the search over many levels (the list the gradient and force passes use)
        rc = build_wide(ctx, true);
        if (!rc && ctx->nact > ctx->few_rows) rc = build_groups(ctx);
        ld = ((long)ctx->nact + 63) / 64 * 64;
        capl = 64;
        for (int attempt = 0; !rc && attempt < 12; attempt++) {
            rc = dalloc(ctx, &ell, (size_t)ell_width(capl) * ld);
            if (rc) break;
            ctx->overflow_clear = false;
            if (hipMemsetAsync(ctx->counters + 3, 0, sizeof(int), ctx->stream) != hipSuccess) { rc = GZ_E_HIP; break; }
            if (ctx->nact > ctx->few_rows)
                k_build_sym_cells<<<xcd_grid(ctx->grp.ngroups), 64, 0, ctx->stream>>>(grid_view(ctx), ctx->grp.view(BoxTable{}), ctx->act,
                                                                                       ctx->X, ctx->Y, ctx->Z, ctx->H, ell, ld, capl, cnt,
                                                                                       ctx->counters + 3);
            else
                k_build_sym_wide<<<xcd_grid(ctx->nact), 64, 0, ctx->stream>>>(grid_view(ctx), ctx->nact, ctx->act, ctx->X, ctx->Y, ctx->Z,
                                                                               ctx->H, ell, ld, capl, cnt, ctx->counters + 3);
            if (hipGetLastError() != hipSuccess) { rc = GZ_E_HIP; break; }
            rc = read_counters(ctx);
            if (rc || ctx->hcounters[3] == 0) break;
            capl = std::max(ctx->hcounters[3] + 16, capl * 3 / 2);
        }
        if (!rc && ctx->hcounters[3]) rc = GZ_E_STATE;
    } else {
        rc = symmetric ? build_list<true>(ctx, ctx->nact, nullptr, nullptr, 1.0, &ell, &ld, &capl, cnt, 64)
                       : build_list<false>(ctx, ctx->nact, nullptr, nullptr, 1.0, &ell, &ld, &capl, cnt, 64);
    }
    if (rc) { dfree(ell); dfree(cnt); return -rc; }
    ctx->sym_valid = false;  // the grid changed
    std::vector<int> c(ctx->nact), e((size_t)ell_width(capl) * ld);
    std::vector<int> p(n);
    hipMemcpy(c.data(), cnt, ctx->nact * sizeof(int), hipMemcpyDeviceToHost);
    hipMemcpy(e.data(), ell, e.size() * sizeof(int), hipMemcpyDeviceToHost);
    hipMemcpy(p.data(), ctx->perm, n * sizeof(int), hipMemcpyDeviceToHost);
    dfree(ell);
    dfree(cnt);
    // rows in caller order
    std::vector<long> rowlen(n, 0);
    std::vector<int> slot_of(n, -1);
    for (int r = 0; r < ctx->nact; r++) { rowlen[p[act_h[r]]] = c[r]; slot_of[p[act_h[r]]] = r; }
    long tot = 0;
    for (long u = 0; u < n; u++) { offsets[u] = tot; tot += rowlen[u]; }
    offsets[n] = tot;
    if (nbr && tot <= cap) {
        for (long u = 0; u < n; u++) {
            int r = slot_of[u];
            if (r < 0) continue;
            for (int q = 0; q < c[r]; q++) nbr[offsets[u] + q] = p[e[ell_at(ld, r, q)]];
        }
    }
    return tot;
}

int gz_hydro_pass_noalloc(const gz_params *params, long n, const gz_fields *in, gz_fields *out, void *stream, char *err,
                          size_t err_len)
{
    if (err && err_len) err[0] = 0;
    auto report = [&](int rc, const char *msg) {
        if (err && err_len) snprintf(err, err_len, "%s", msg);
        return rc;
    };
    if (!params || !in || !out) return report(GZ_E_ARG, "gz_hydro_pass_noalloc: params, in and out are required");
    if (n <= 0 || n > GZ_NOALLOC_MAX_N) {
        char m[160];
        snprintf(m, sizeof m, "gz_hydro_pass_noalloc: n=%ld outside (0, %ld] (the static scratch holds at most that many)", n,
                 (long)GZ_NOALLOC_MAX_N);
        return report(GZ_E_ARG, m);
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocDevices)
        return report(GZ_E_HIP, "gz_hydro_pass_noalloc: no current device (or its index is 64 or more)");
    if (!g_scratch[dev]) {
        // set-up, once per device, with that device current: the array's address on it
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char m[200];
            snprintf(m, sizeof m, "gz_hydro_pass_noalloc: hipGetSymbolAddress of the scratch array failed: %s", hipGetErrorString(e));
            return report(GZ_E_HIP, m);
        }
        g_scratch[dev] = (unsigned char *)p;
    }
    // a fresh context over the array (its members allocate nothing until used, and a pass uses
    // none of the ones that would: timing stays off)
    gz_context *ctx = &g_noalloc_ctx[dev];
    ctx->~gz_context();
    new (ctx) gz_context();
    ctx->device = dev;
    ctx->stream = (hipStream_t)stream;
    ctx->pool = g_scratch[dev];
    ctx->pool_size = GZ_NOALLOC_SCRATCH_BYTES;
    int rc = set_params(ctx, params);
    if (!rc) rc = alloc_state(ctx, n);
    if (!rc) rc = gz_load(ctx, n, in);
    if (!rc) rc = gz_density(ctx);
    if (!rc) rc = gz_hydro_gradient_calc(ctx);
    if (!rc) rc = gz_hydro_force(ctx);
    if (!rc) rc = gz_store(ctx, out);
    if (!rc) {
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, GZ_E_HIP, "hipStreamSynchronize failed: %s", hipGetErrorString(e));
    }
    if (rc) report(rc, ctx->err);
    return rc;
}

}  // extern "C"
