// gz_batch.hip -- host side of a batch of independent periodic boxes (include/gizmo_batch.h).
//
// Device state: the per-particle fields of the single-box context for `capacity` particles,
// the box of every particle, and a table of boxes (BoxDev: geometry, grid, timeline).  Live
// boxes lie at the front in rank order, each in its own Morton order; the tail belongs to the
// retired box (no mass, never stepping, never drifted).  Vector fields are SoA with column
// stride `capacity`, so every whole-array kernel runs over `capacity` particles.
//
// A step of the batch is the single-box step of every started box with shared launches: the
// timeline kernels look each particle's box up in the table, the hot path runs the
// one-wavefront-per-row kernels over the active rows of all boxes, and each box keeps its own
// grid (level from its own radii), its own ordering (re-keyed when its own particles moved
// far), its own wake-up flag and stop code.  Every sum of a box runs over that box's particles
// in their ascending order, which a shift of the box's index range does not change: a box
// computes the same numbers alone or in any mix.
//
// The device table is the authority for the timeline part of each entry (times, masks, the
// wake-up flag, which the kernels write); the host is the authority for geometry, grid and the
// stepping flag, and writes only those (the grid columns of every entry, or one whole entry of
// a box that is not stepping).
#include <hip/hip_runtime.h>
#include "gz_prim.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/gizmo_batch.h"
#include "gz_kernels.h"
#include "gz_groups.h"

using namespace gz;

namespace {
constexpr int kB = 256;
inline unsigned nb(long n) { return (unsigned)((n + kB - 1) / kB); }
constexpr long long kTimebaseH = 1LL << kTimebins;
constexpr double kResortFraction = 0.125;

struct BoxHost {
    bool alive = false, begun = false;
    long n = 0, off = 0;
    gz_box_params bp{};
    int level = -1;
    double cell = 0;
    double dmax = 0, hbound = 0, last_rmax = 0;
    long long num_steps = 0, n_active = 0, n_woken = 0;
    long long bins[kTimebins] = {};
    int stop_code = 0;
};
}  // namespace

struct gz_batch {
    gz_params par{};
    DevParams dp{};
    int device = 0, max_boxes = 0, retired = 0;
    hipStream_t stream = nullptr;
    long cap = 0, n_live = 0;
    std::string err;

    double *X = nullptr, *Y = nullptr, *Z = nullptr, *VX = nullptr, *VY = nullptr, *VZ = nullptr;
    double *VKX = nullptr, *VKY = nullptr, *VKZ = nullptr, *M = nullptr, *H = nullptr, *U = nullptr, *UK = nullptr;
    double *NN = nullptr, *DRK = nullptr, *DIVV = nullptr, *RHO = nullptr, *PR = nullptr, *T = nullptr, *FW = nullptr;
    double *CN = nullptr, *FCE = nullptr, *GR = nullptr, *GP = nullptr, *GV = nullptr, *ACC = nullptr, *DTU = nullptr;
    double *MSV = nullptr, *SOFT = nullptr, *X0 = nullptr, *Y0 = nullptr, *Z0 = nullptr;
    long long *TIB = nullptr, *DTS = nullptr;
    int *TB = nullptr, *WAKE = nullptr, *BOX = nullptr, *perm = nullptr;
    uint8_t *ACT = nullptr, *need = nullptr, *sel = nullptr, *due = nullptr;
    uint32_t *key = nullptr;
    unsigned long long *key64 = nullptr, *key64b = nullptr;
    int *ival = nullptr, *ival2 = nullptr, *iflag = nullptr, *iscan = nullptr, *act = nullptr, *rank = nullptr;
    double *stage = nullptr, *dacc = nullptr;
    double *it_h = nullptr, *it_hb = nullptr, *it_l = nullptr, *it_r = nullptr, *it_cn = nullptr;
    int *it_ci = nullptr, *it_it = nullptr;
    void *rec_grad = nullptr, *rec_force = nullptr;
    int *ell = nullptr, *cnt = nullptr;
    int *rare = nullptr;  // rows the force pass's common kernel leaves to the whole Riemann ladder
    int ell_cap = 0;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    int *cstart = nullptr, *cend = nullptr;
    double *chmax = nullptr;
    long long cells_alloc = 0, cells_total = 0;
    bool grid_valid = false;

    BoxDev *tab = nullptr;                                   // max_boxes + 1 entries, the last the retired box
    std::vector<BoxDev> htab;
    unsigned long long *hist = nullptr, *slot = nullptr;     // [box][64] bin occupation, [box][kBoxSlots]
    unsigned long long *hhist = nullptr, *hslot = nullptr;   // pinned mirrors
    unsigned long long *u64 = nullptr, *hu64 = nullptr;      // [0] force pairs, [1] unused radius maximum
    int *counters = nullptr, *hcounters = nullptr;
    int *lim = nullptr;

    std::map<void *, size_t> alloc_bytes;  // live device allocations
    size_t dev_bytes = 0;

    std::vector<BoxHost> boxes;
    std::vector<int> order;  // live boxes in storage (rank) order
    int nact = 0;
    // passes over more active rows than this walk the rows grouped by box cell (one wavefront
    // per group), fewer take one wavefront per row
    int few_rows = 8192;
    RowGroups grp;
    bool grp_valid = false;
    bool act_rows_valid = false;
    std::map<std::string, double> stats;
};

namespace {

int fail(gz_batch *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define BHIP(call)                                                                                                 \
    do {                                                                                                           \
        hipError_t e_ = (call);                                                                                    \
        if (e_ != hipSuccess)                                                                                      \
            return fail(b, GZ_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__);  \
    } while (0)
#define BCHK(call)           \
    do {                     \
        int rc_ = (call);    \
        if (rc_) return rc_; \
    } while (0)

// device allocation of count elements (any previous one freed), counted in the batch's footprint
template <typename T>
int balloc(gz_batch *b, T **p, size_t count)
{
    if (*p) {
        auto it = b->alloc_bytes.find((void *)*p);
        if (it != b->alloc_bytes.end()) { b->dev_bytes -= it->second; b->alloc_bytes.erase(it); }
        (void)hipFree(*p);
    }
    *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc((void **)p, count * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return fail(b, GZ_E_MEMORY, "hipMalloc of %zu bytes failed", count * sizeof(T));
    }
    b->alloc_bytes[(void *)*p] = count * sizeof(T);
    b->dev_bytes += count * sizeof(T);
    return 0;
}

template <typename T>
void bfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

int ensure_tmp(gz_batch *b, size_t bytes)
{
    if (bytes <= b->tmp_bytes) return 0;
    const size_t nbytes = std::max(bytes, (size_t)1 << 20);
    char *t = (char *)b->tmp;
    BCHK(balloc(b, &t, nbytes));
    b->tmp = t;
    b->tmp_bytes = nbytes;
    return 0;
}

size_t nbox_all(const gz_batch *b) { return (size_t)b->max_boxes + 1; }
BoxTable table(gz_batch *b) { return BoxTable{b->tab, b->BOX, b->slot, b->max_boxes + 1}; }
long ld_rows(long na) { return (na + 63) / 64 * 64; }

ParticleView pview(gz_batch *c)
{
    ParticleView S;
    S.n = c->cap;
    S.X = c->X; S.Y = c->Y; S.Z = c->Z; S.VX = c->VX; S.VY = c->VY; S.VZ = c->VZ; S.VKX = c->VKX; S.VKY = c->VKY; S.VKZ = c->VKZ;
    S.M = c->M; S.H = c->H; S.U = c->U; S.UK = c->UK; S.NN = c->NN; S.DIVV = c->DIVV; S.RHO = c->RHO; S.PR = c->PR;
    S.ACC = c->ACC; S.DTU = c->DTU; S.MSV = c->MSV; S.SOFT = c->SOFT;
    S.TB = c->TB; S.WAKE = c->WAKE; S.TIB = c->TIB; S.DTS = c->DTS;
    return S;
}

// the parameters every box shares; each particle's box supplies the rest through the table
TimeParams tparams(const gz_batch *b)
{
    TimeParams T{};
    T.core_size = b->par.kernel == 3 ? (1.0 / 2.0) : (2.0 / 5.0);  // KERNEL_CORE_SIZE, src/mesh/kernel.h:59,113
    T.gamma = b->par.gamma;
    return T;
}

GridView gview(gz_batch *b)
{
    GridView g{};
    g.chmax = b->chmax;
    g.cstart = b->cstart;
    g.cend = b->cend;
    return g;
}

double bits_to_double(unsigned long long u)
{
    double d;
    memcpy(&d, &u, sizeof d);
    return d;
}

int alloc_all(gz_batch *b)
{
    const size_t n = (size_t)b->cap;
    double **d1[] = {&b->X, &b->Y, &b->Z, &b->VX, &b->VY, &b->VZ, &b->VKX, &b->VKY, &b->VKZ, &b->M, &b->H, &b->U, &b->UK,
                     &b->NN, &b->DRK, &b->DIVV, &b->RHO, &b->PR, &b->CN, &b->FCE, &b->DTU, &b->MSV, &b->SOFT, &b->X0, &b->Y0,
                     &b->Z0, &b->it_h, &b->it_hb, &b->it_l, &b->it_r, &b->it_cn};
    for (auto p : d1) BCHK(balloc(b, p, n));
    BCHK(balloc(b, &b->T, 6 * n)); BCHK(balloc(b, &b->FW, 3 * n)); BCHK(balloc(b, &b->GR, 3 * n)); BCHK(balloc(b, &b->GP, 3 * n));
    BCHK(balloc(b, &b->GV, 9 * n)); BCHK(balloc(b, &b->ACC, 3 * n)); BCHK(balloc(b, &b->stage, 9 * n)); BCHK(balloc(b, &b->dacc, kDensSumsFD * n));
    BCHK(balloc(b, &b->TIB, n)); BCHK(balloc(b, &b->DTS, n));
    int **i1[] = {&b->TB, &b->WAKE, &b->BOX, &b->perm, &b->ival, &b->ival2, &b->act, &b->it_ci, &b->it_it, &b->cnt, &b->rare};
    for (auto p : i1) BCHK(balloc(b, p, n));
    BCHK(balloc(b, &b->iflag, n + 1)); BCHK(balloc(b, &b->iscan, n + 1));
    BCHK(balloc(b, &b->ACT, n)); BCHK(balloc(b, &b->need, n));
    BCHK(balloc(b, &b->key, n)); BCHK(balloc(b, &b->key64, n)); BCHK(balloc(b, &b->key64b, n));
    BCHK(balloc(b, (char **)&b->rec_grad, n * sizeof(GradRec<double>)));
    BCHK(balloc(b, (char **)&b->rec_force, n * sizeof(Cell<double>)));
    const size_t nbx = nbox_all(b);
    BCHK(balloc(b, &b->tab, nbx)); BCHK(balloc(b, &b->hist, 64 * nbx)); BCHK(balloc(b, &b->slot, kBoxSlots * nbx));
    BCHK(balloc(b, &b->sel, nbx)); BCHK(balloc(b, &b->due, nbx)); BCHK(balloc(b, &b->rank, nbx)); BCHK(balloc(b, &b->lim, 2 * nbx));
    BCHK(balloc(b, &b->counters, 16)); BCHK(balloc(b, &b->u64, 8));
    BHIP(hipMemset(b->counters, 0, 16 * sizeof(int)));
    BHIP(hipMemset(b->u64, 0, 8 * sizeof(unsigned long long)));  // the running pair count starts at zero
    BHIP(hipHostMalloc((void **)&b->hhist, 64 * nbx * sizeof(unsigned long long), hipHostMallocDefault));
    BHIP(hipHostMalloc((void **)&b->hslot, kBoxSlots * nbx * sizeof(unsigned long long), hipHostMallocDefault));
    BHIP(hipHostMalloc((void **)&b->hu64, 8 * sizeof(unsigned long long), hipHostMallocDefault));
    BHIP(hipHostMalloc((void **)&b->hcounters, 16 * sizeof(int), hipHostMallocDefault));
    return 0;
}

void free_all(gz_batch *c)
{
    double **d1[] = {&c->X, &c->Y, &c->Z, &c->VX, &c->VY, &c->VZ, &c->VKX, &c->VKY, &c->VKZ, &c->M, &c->H, &c->U, &c->UK,
                     &c->NN, &c->DRK, &c->DIVV, &c->RHO, &c->PR, &c->CN, &c->FCE, &c->DTU, &c->MSV, &c->SOFT, &c->X0, &c->Y0,
                     &c->Z0, &c->it_h, &c->it_hb, &c->it_l, &c->it_r, &c->it_cn, &c->T, &c->FW, &c->GR, &c->GP, &c->GV,
                     &c->ACC, &c->stage, &c->dacc, &c->chmax};
    for (auto p : d1) bfree(*p);
    int **i1[] = {&c->TB, &c->WAKE, &c->BOX, &c->perm, &c->ival, &c->ival2, &c->act, &c->it_ci, &c->it_it, &c->cnt, &c->iflag,
                  &c->iscan, &c->rank, &c->lim, &c->counters, &c->ell, &c->cstart, &c->cend, &c->rare};
    for (auto p : i1) bfree(*p);
    bfree(c->TIB); bfree(c->DTS); bfree(c->ACT); bfree(c->need); bfree(c->sel); bfree(c->due);
    bfree(c->key); bfree(c->key64); bfree(c->key64b); bfree(c->rec_grad); bfree(c->rec_force); bfree(c->tab);
    bfree(c->hist); bfree(c->slot); bfree(c->u64);
    bfree(c->tmp);
    {
        RowGroups &G = c->grp;
        bfree(G.key); bfree(G.skey); bfree(G.gkey); bfree(G.val); bfree(G.row); bfree(G.flag); bfree(G.scan); bfree(G.start);
        bfree(G.count); bfree(G.lev); bfree(G.code); bfree(G.nst); bfree(G.sten);
        G = RowGroups{};
    }
    unsigned long long **hp[] = {&c->hhist, &c->hslot, &c->hu64};
    for (auto p : hp) {
        if (*p) (void)hipHostFree(*p);
        *p = nullptr;
    }
    if (c->hcounters) (void)hipHostFree(c->hcounters);
    c->hcounters = nullptr;
}

int sync(gz_batch *b)
{
    BHIP(hipStreamSynchronize(b->stream));
    return 0;
}

// Host-to-device uploads of the small host tables wait for their copy: the host may change
// the table again right after, and the source of an asynchronous copy must stay unchanged
// until the copy has run.  They are issued when no other work is queued, so the wait is short.

// one whole table entry (a box that is not stepping: no kernel writes it meanwhile)
int upload_entry(gz_batch *b, int id)
{
    BHIP(hipMemcpyAsync(b->tab + id, &b->htab[id], sizeof(BoxDev), hipMemcpyHostToDevice, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    return 0;
}

// the grid columns (cell .. cell_off) of every entry
int upload_grid(gz_batch *b)
{
    const size_t lo = offsetof(BoxDev, cell), hi = offsetof(BoxDev, time);
    BHIP(hipMemcpy2DAsync((char *)b->tab + lo, sizeof(BoxDev), (const char *)b->htab.data() + lo, sizeof(BoxDev), hi - lo,
                          nbox_all(b), hipMemcpyHostToDevice, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    return 0;
}

int upload_bytes(gz_batch *b, const std::vector<uint8_t> &s, uint8_t *dst)
{
    BHIP(hipMemcpyAsync(dst, s.data(), s.size(), hipMemcpyHostToDevice, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    return 0;
}

// slots, counters and the pair count back to the host (one wait)
int read_back(gz_batch *b, bool with_hist_and_table)
{
    const size_t nbx = nbox_all(b);
    if (with_hist_and_table) {
        BHIP(hipMemcpyAsync(b->hhist, b->hist, 64 * nbx * 8, hipMemcpyDeviceToHost, b->stream));
        BHIP(hipMemcpyAsync(b->htab.data(), b->tab, nbx * sizeof(BoxDev), hipMemcpyDeviceToHost, b->stream));
    }
    BHIP(hipMemcpyAsync(b->hslot, b->slot, kBoxSlots * nbx * 8, hipMemcpyDeviceToHost, b->stream));
    BHIP(hipMemcpyAsync(b->hcounters, b->counters, 16 * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    return sync(b);
}

double slot_value(const gz_batch *b, int id, int s) { return bits_to_double(b->hslot[(size_t)id * kBoxSlots + s]); }

// scanned row flags to the active list; the count into counters[10]
int compact_rows(gz_batch *b)
{
    const long n = b->cap;
    size_t bytes = 0;
    BHIP(gz_exclusive_sum(nullptr, bytes, b->iflag, b->iscan, (int)(n + 1), b->stream));
    BCHK(ensure_tmp(b, bytes));
    BHIP(gz_exclusive_sum(b->tmp, bytes, b->iflag, b->iscan, (int)(n + 1), b->stream));
    k_compact_count<<<nb(n + 1), kB, 0, b->stream>>>(n, b->iflag, b->iscan, b->act, b->counters + 10);
    BHIP(hipGetLastError());
    return 0;
}

// the active list of every stepping box from the active bytes
int rows_from_active(gz_batch *b)
{
    const long n = b->cap;
    k_batch_flag_active<<<nb(n + 1), kB, 0, b->stream>>>(n, b->BOX, b->tab, b->ACT, b->M, b->iflag);
    BHIP(hipGetLastError());
    BCHK(compact_rows(b));
    BCHK(read_back(b, false));
    b->nact = b->hcounters[10];
    b->act_rows_valid = true;
    return 0;
}

// ---------------------------------------------------------------------------------------
// order and grid
// ---------------------------------------------------------------------------------------

// every per-particle field in the new order (columns of stride cap)
int permute_all(gz_batch *b, const int *order)
{
    const long n = b->cap;
    struct F { void *p; int w; } f8[] = {
        {b->X, 1}, {b->Y, 1}, {b->Z, 1}, {b->VX, 1}, {b->VY, 1}, {b->VZ, 1}, {b->VKX, 1}, {b->VKY, 1}, {b->VKZ, 1}, {b->M, 1},
        {b->H, 1}, {b->U, 1}, {b->UK, 1}, {b->NN, 1}, {b->DRK, 1}, {b->DIVV, 1}, {b->RHO, 1}, {b->PR, 1}, {b->T, 6},
        {b->FW, 3}, {b->CN, 1}, {b->FCE, 1}, {b->GR, 3}, {b->GP, 3}, {b->GV, 9}, {b->ACC, 3}, {b->DTU, 1}, {b->MSV, 1},
        {b->SOFT, 1}, {b->TIB, 1}, {b->DTS, 1}, {b->X0, 1}, {b->Y0, 1}, {b->Z0, 1}};
    for (auto &f : f8) {
        k_permute_cols<<<nb(n), kB, 0, b->stream>>>(n, f.w, order, (const unsigned long long *)f.p, (unsigned long long *)b->stage);
        BHIP(hipGetLastError());
        BHIP(hipMemcpyAsync(f.p, b->stage, (size_t)n * f.w * 8, hipMemcpyDeviceToDevice, b->stream));
    }
    int *i4[] = {b->TB, b->WAKE, b->BOX, b->perm, (int *)b->key};
    for (int *a : i4) {
        k_permute_int<<<nb(n), kB, 0, b->stream>>>(n, order, a, b->ival);
        BHIP(hipGetLastError());
        BHIP(hipMemcpyAsync(a, b->ival, n * sizeof(int), hipMemcpyDeviceToDevice, b->stream));
    }
    k_permute_u8<<<nb(n), kB, 0, b->stream>>>(n, order, b->ACT, (uint8_t *)b->stage);
    BHIP(hipGetLastError());
    BHIP(hipMemcpyAsync(b->ACT, b->stage, n, hipMemcpyDeviceToDevice, b->stream));
    return 0;
}

// One stable sort on (rank of the box, Morton key): the boxes marked due get the keys of their
// current positions (re-ordered, displacement bound back to zero), the others keep their
// order; live boxes packed to the front in rank order, retired particles to the tail.
int reorder(gz_batch *b, const std::vector<uint8_t> &due_h)
{
    const long n = b->cap;
    std::vector<int> rk(nbox_all(b), b->max_boxes + 1);
    for (size_t r = 0; r < b->order.size(); r++) rk[b->order[r]] = (int)r;
    BHIP(hipMemcpyAsync(b->rank, rk.data(), rk.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    BCHK(upload_bytes(b, due_h, b->due));  // (waits for both copies)
    k_batch_keys<<<nb(n), kB, 0, b->stream>>>(n, b->X, b->Y, b->Z, b->BOX, b->tab, b->rank, b->due, b->key, b->key64, b->ival);
    BHIP(hipGetLastError());
    int bits = 30;
    while ((1LL << (bits - 30)) <= (long long)b->max_boxes + 1) bits++;
    size_t bytes = 0;
    BHIP(gz_sort_pairs(nullptr, bytes, b->key64, b->key64b, b->ival, b->ival2, (int)n, 0, bits, b->stream));
    BCHK(ensure_tmp(b, bytes));
    BHIP(gz_sort_pairs(b->tmp, bytes, b->key64, b->key64b, b->ival, b->ival2, (int)n, 0, bits, b->stream));
    BCHK(permute_all(b, b->ival2));
    long off = 0;
    for (int id : b->order) {
        BoxHost &h = b->boxes[id];
        h.off = off;
        off += h.n;
        if (due_h[id]) {
            h.dmax = 0;
            BHIP(hipMemcpyAsync(b->X0 + h.off, b->X + h.off, h.n * 8, hipMemcpyDeviceToDevice, b->stream));
            BHIP(hipMemcpyAsync(b->Y0 + h.off, b->Y + h.off, h.n * 8, hipMemcpyDeviceToDevice, b->stream));
            BHIP(hipMemcpyAsync(b->Z0 + h.off, b->Z + h.off, h.n * 8, hipMemcpyDeviceToDevice, b->stream));
            b->stats["resorts"] += 1;
        }
    }
    b->n_live = off;
    b->grid_valid = false;
    b->act_rows_valid = false;
    return 0;
}

// Level of every live box: cells no smaller than its largest radius plus twice its
// displacement bound (the 27 cells around a row then hold every partner), and no more cells
// than a few per particle (coarser cells only cost candidates).  Cell arrays of the boxes laid
// end to end, the retired box one cell.
int build_grid(gz_batch *b)
{
    bool changed = !b->grid_valid;
    long long off = 0;
    for (int id : b->order) {
        BoxHost &h = b->boxes[id];
        BoxDev &d = b->htab[id];
        const double box = h.bp.box_size;
        const double reach = std::min(h.hbound, 0.5 * box) + 2.0 * h.dmax;
        int L = 0;
        while (L < kFineLevel && box / (double)(1 << (L + 1)) >= reach * (1.0 + 1e-6) && (1LL << (3 * (L + 1))) <= 4 * (long long)h.n + 64)
            L++;
        if (L != h.level) changed = true;
        h.level = L;
        h.cell = box / (double)(1 << L);
        d.level = L;
        d.shift = 3 * (kFineLevel - L);
        d.cell = h.cell;
        d.cell_off = off;
        d.margin = h.dmax;
        d.rowlim = h.cell - 2.0 * h.dmax;
        off += 1LL << (3 * L);
    }
    BoxDev &r = b->htab[b->retired];
    r.level = 0; r.shift = 3 * kFineLevel; r.cell = 1; r.cell_off = off; r.margin = 0; r.rowlim = 1;
    off += 1;
    if (off != b->cells_total) changed = true;
    BCHK(upload_grid(b));
    if (changed) b->grp_valid = false;
    if (!changed) return 0;
    if (off > b->cells_alloc) {
        BCHK(balloc(b, &b->cstart, off)); BCHK(balloc(b, &b->cend, off)); BCHK(balloc(b, &b->chmax, off));
        b->cells_alloc = off;
    }
    b->cells_total = off;
    k_cell_clear<<<nb(off), kB, 0, b->stream>>>(off, b->cstart, b->cend, b->chmax);
    k_batch_cell_bounds<<<nb(b->cap), kB, 0, b->stream>>>(b->cap, b->key, b->BOX, b->tab, b->retired, b->cstart, b->cend);
    BHIP(hipGetLastError());
    b->grid_valid = true;
    return 0;
}

// The active rows grouped by the cell of their box's grid (rows come in storage order, so the
// groups need no sort); kept until the rows, the order or a grid change
int build_groups(gz_batch *b)
{
    if (b->grp_valid) return 0;
    RowGroups &G = b->grp;
    const int na = b->nact;
    if (G.rows_alloc < (long)na + 1) {
        const size_t m = (size_t)b->cap + 1;
        BCHK(balloc(b, &G.key, m)); BCHK(balloc(b, &G.gkey, m)); BCHK(balloc(b, &G.flag, m)); BCHK(balloc(b, &G.scan, m));
        BCHK(balloc(b, &G.start, m)); BCHK(balloc(b, &G.count, 1));
        G.rows_alloc = (long)m;
    }
    size_t bytes = 0;
    BHIP(group_tmp_bytes(na, bytes));
    BCHK(ensure_tmp(b, bytes));
    const BoxTable bt = table(b);
    BHIP(group_rows(G, b->stream, b->tmp, b->tmp_bytes, na, b->act, nullptr, b->key, nullptr, b->H, 0.0, 0.0, nullptr, bt, false));
    BHIP(hipMemcpyAsync(b->hcounters + 12, G.count, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    BCHK(sync(b));
    G.ngroups = b->hcounters[12];
    if (G.groups_alloc < G.ngroups) {
        const long m = std::min((long)na, (long)G.ngroups + G.ngroups / 4 + 64);
        BCHK(balloc(b, &G.lev, m)); BCHK(balloc(b, &G.code, m)); BCHK(balloc(b, &G.nst, m)); BCHK(balloc(b, &G.sten, (size_t)27 * m));
        G.groups_alloc = m;
    }
    BHIP(group_stencils(G, b->stream, gview(b), b->act, bt));
    b->grp_valid = true;
    return 0;
}

// ---------------------------------------------------------------------------------------
// the hot path over the active rows of every box in the pass
// ---------------------------------------------------------------------------------------

// density() with the radius iteration (the single-box stage with per-box grids)
template <int KERN, typename R>
int density(gz_batch *b)
{
    const int na = b->nact;
    if (na == 0) return 0;
    const BoxTable bt = table(b);
    b->grp_valid = false;  // this pass's rows
    // cells of this step's order, radius bound and displacement bound
    BCHK(build_grid(b));
    BHIP(hipMemsetAsync(b->slot, 0, kBoxSlots * nbox_all(b) * 8, b->stream));
    k_density_init<<<nb(na), kB, 0, b->stream>>>(na, b->act, b->H, 0, b->it_h, b->it_l, b->it_r, b->it_cn, b->it_ci, b->it_it, b->need,
                                                 b->counters, nullptr, b->u64 + 1, bt);
    BHIP(hipGetLastError());
    DensityIter it{b->it_h, b->it_hb, b->it_l, b->it_r, b->it_cn, b->it_ci, b->it_it};
    DensityOut out{b->cap, b->H, b->NN, b->DRK, b->DIVV, b->RHO, b->PR, b->T, b->FW, b->CN, b->FCE};
    int iters = 0;
    while (true) {
        BHIP(hipMemsetAsync(b->counters, 0, 3 * sizeof(int), b->stream));
        if (na > b->few_rows) {
            BCHK(build_groups(b));
            k_density_cells<KERN, R><<<xcd_grid(b->grp.ngroups), 64, 0, b->stream>>>(b->dp, gview(b), b->grp.view(bt), b->act, b->need,
                                                                                      b->X, b->Y, b->Z, b->VX, b->VY, b->VZ, b->M,
                                                                                      b->it_h, b->dacc);
        } else {
            k_density_rows<KERN, R, kRowThreads><<<xcd_grid(na), kRowThreads, 0, b->stream>>>(b->dp, gview(b), na, b->act, b->need, b->key, b->X, b->Y, b->Z,
                                                                         b->VX, b->VY, b->VZ, b->M, b->it_h, b->dacc, bt);
        }
        k_density_after<<<nb(na), kB, 0, b->stream>>>(b->dp, na, b->act, b->need, b->dacc, b->M, b->U, 0, it, out, b->counters + 0,
                                                      b->counters + 1, b->counters + 2, b->u64 + 1, bt);
        BHIP(hipGetLastError());
        BCHK(read_back(b, false));
        iters++;
        const int remaining = b->hcounters[0], rebuild = b->hcounters[1], e = b->hcounters[2];
        if (e == GZ_STOP_MAXITER) return fail(b, e, "density(): kernel radius iteration did not converge in %d rounds", kMaxIter);
        if (e == GZ_STOP_BRACKET) return fail(b, e, "density(): empty radius bracket (Right==0 && Left==0)");
        if (e) return fail(b, e, "density(): device stop code %d", e);
        if (remaining == 0) break;
        if (rebuild) {
            // radii that outgrew their box's cells: coarser cells for those boxes
            for (int id : b->order) b->boxes[id].hbound = std::max(b->boxes[id].hbound, slot_value(b, id, kSlotHgrow));
            BCHK(build_grid(b));
            b->stats["ngb_rebuilds"] += 1;
        }
    }
    for (int id : b->order) b->boxes[id].hbound = std::max(b->boxes[id].hbound, slot_value(b, id, kSlotHfinal));
    b->stats["density_iterations"] += iters;
    return 0;
}

// the symmetric neighbour list of the rows and the records the gradient and force passes read
int sym_list(gz_batch *b, const std::vector<uint8_t> &in_pass)
{
    const int na = b->nact;
    const BoxTable bt = table(b);
    for (int id : b->order)
        if (in_pass[id]) b->boxes[id].last_rmax = std::min(b->boxes[id].hbound, 0.5 * b->boxes[id].bp.box_size);
    BCHK(build_grid(b));
    k_cell_hmax<<<nb(b->cells_total), kB, 0, b->stream>>>(b->cells_total, b->cstart, b->cend, b->H, b->chmax);
    BHIP(hipGetLastError());
    const long ld = ld_rows(na);
    const long ld_max = ld_rows(b->cap);
    const int hint = std::max(64, (int)std::ceil(b->par.des_num_ngb * 2.0) + 32);
    if (!b->ell || b->ell_cap < hint) {
        b->ell_cap = std::max(b->ell_cap, hint);
        BCHK(balloc(b, &b->ell, (size_t)ell_width(b->ell_cap) * ld_max));
    }
    for (int attempt = 0;; attempt++) {
        BHIP(hipMemsetAsync(b->counters + 3, 0, sizeof(int), b->stream));
        if (na > b->few_rows) {
            BCHK(build_groups(b));
            k_build_sym_cells<<<xcd_grid(b->grp.ngroups), 64, 0, b->stream>>>(gview(b), b->grp.view(bt), b->act, b->X, b->Y, b->Z, b->H,
                                                                               b->ell, ld, b->ell_cap, b->cnt, b->counters + 3);
        } else {
            k_build_sym_rows<<<xcd_grid(na), kRowThreads, 0, b->stream>>>(gview(b), na, b->act, b->key, b->X, b->Y, b->Z, b->H, b->ell, ld,
                                                                  b->ell_cap, b->cnt, b->counters + 3, bt);
        }
        BHIP(hipGetLastError());
        BCHK(read_back(b, false));
        const int over = b->hcounters[3];
        if (over == 0) break;
        if (attempt == 7) return fail(b, GZ_E_STATE, "symmetric neighbour list did not fit after widening");
        b->ell_cap = std::max(over + 16, b->ell_cap * 3 / 2);
        BCHK(balloc(b, &b->ell, (size_t)ell_width(b->ell_cap) * ld_max));
        b->stats["ell_widen"] += 1;
    }
    return 0;
}

// hydro_gradient_calc() and hydro_force() over the list
template <int KERN, typename R>
int gradients_force(gz_batch *b)
{
    const int na = b->nact;
    const long ld = ld_rows(na);
    const BoxTable bt = table(b);
    GradRec<R> *rg = (GradRec<R> *)b->rec_grad;
    // few rows among many particles: only the records of the rows and their list entries (a
    // record once per list entry that names it: cheaper while the lists hold fewer entries than
    // there are particles)
    const bool list = (long)na * 64 < b->cap;
    if (list)
        k_pack_grad_list<R><<<xcd_grid(na), 64, 0, b->stream>>>(na, b->act, b->ell, ld, b->cnt, b->H, b->M, b->RHO, b->PR, b->VX, b->VY,
                                                                b->VZ, rg, b->X, b->Y, b->Z, 1.0, bt);
    else
        k_pack_grad<R><<<nb(b->cap), kB, 0, b->stream>>>(b->cap, b->H, b->M, b->RHO, b->PR, b->VX, b->VY, b->VZ, rg, b->X, b->Y, b->Z, 1.0,
                                                          bt);
    const bool lanes = na > b->few_rows;  // many rows: one lane per row
    if (lanes)
        k_gradients<KERN, R, true><<<xcd_grid(nb(na)), kB, 0, b->stream>>>(b->dp, na, b->act, b->X, b->Y, b->Z, b->H, rg, b->T, b->CN, b->cap, b->ell, ld,
                                                                      b->cnt, b->GR, b->GP, b->GV, bt);
    else
        k_gradients_rows<KERN, R><<<xcd_grid(na), kListThreads, 0, b->stream>>>(b->dp, na, b->act, b->X, b->Y, b->Z, b->H, rg, GradSrc{}, b->T, b->CN, b->cap,
                                                                       b->ell, ld, b->cnt, b->GR, b->GP, b->GV, bt);
    BHIP(hipGetLastError());
    ForceView F{b->cap, b->X, b->Y, b->Z, b->VX, b->VY, b->VZ, b->M, b->H, b->U, b->RHO, b->PR, b->CN, b->FCE, b->DRK, b->GR, b->GP,
                b->GV, b->T};
    Cell<R> *rf = (Cell<R> *)b->rec_force;
    Gam<R> G;
    G.G0 = (R)b->dp.G0; G.G1 = (R)b->dp.G1; G.G2 = (R)b->dp.G2; G.G3 = (R)b->dp.G3; G.G4 = (R)b->dp.G4;
    G.G5 = (R)b->dp.G5; G.G6 = (R)b->dp.G6; G.G7 = (R)b->dp.G7; G.G8 = (R)b->dp.G8; G.G9 = (R)b->dp.G9;
    if (list)
        k_pack_force_list<R><<<xcd_grid(na), 64, 0, b->stream>>>(na, b->act, b->ell, ld, b->cnt, b->par.gamma, F, rf, b->counters + 2, 1.0,
                                                                 bt);
    else
        k_pack_force<R><<<nb(b->cap), kB, 0, b->stream>>>(b->par.gamma, F, rf, b->counters + 2, 1.0, bt);
    if (lanes) {
        // the rows whose faces all take the Riemann ladder's common case, then the others with the
        // whole ladder
        const ForceRows a{na, b->act, b->ell, ld, b->cnt, b->ACT, b->TB, b->MSV, b->WAKE, b->ACC, b->DTU, b->MSV, b->counters + 2,
                          b->counters + 4, nullptr, nullptr, b->rare};
        k_force<KERN, R, true, false, true><<<xcd_grid(nb(na)), kB, 0, b->stream>>>(b->dp, G, F, rf, a, bt);
        k_force_rare<KERN, R, true><<<std::min(nb(na), kRareBlocks), kB, 0, b->stream>>>(b->dp, G, F, rf, a, bt);
    } else {
        k_force_rows<KERN, R><<<xcd_grid(na), kListThreads, 0, b->stream>>>(b->dp, G, na, b->act, F, rf, b->ell, ld, b->cnt, b->ACT, b->TB, b->MSV,
                                                                   b->WAKE, b->ACC, b->DTU, b->MSV, b->counters + 2, b->counters + 4,
                                                                   nullptr, bt);
    }
    k_sum_counts<<<std::min(nb(na), 256u), kB, 0, b->stream>>>(na, b->cnt, b->u64);
    BHIP(hipGetLastError());
    BCHK(read_back(b, false));
    if (b->hcounters[2]) return fail(b, GZ_STOP_RIEMANN, "hydro_force(): Riemann solver failed to find a positive pressure");
    return 0;
}

template <int KERN, typename R>
int hot_path(gz_batch *b, const std::vector<uint8_t> &in_pass)
{
    if (b->nact == 0) return 0;
    BCHK((density<KERN, R>(b)));
    BCHK(sym_list(b, in_pass));
    BCHK((gradients_force<KERN, R>(b)));
    b->stats["active_updates"] += b->nact;
    return 0;
}

int hot_path_dispatch(gz_batch *b, const std::vector<uint8_t> &in_pass)
{
    const bool k3 = b->par.kernel == 3;
    if (b->par.precision == GZ_FP64) return k3 ? hot_path<3, double>(b, in_pass) : hot_path<6, double>(b, in_pass);
    return k3 ? hot_path<3, float>(b, in_pass) : hot_path<6, float>(b, in_pass);
}

int density_dispatch(gz_batch *b)
{
    const bool k3 = b->par.kernel == 3;
    if (b->par.precision == GZ_FP64) return k3 ? density<3, double>(b) : density<6, double>(b);
    return k3 ? density<3, float>(b) : density<6, float>(b);
}

int set_device(gz_batch *b)
{
    BHIP(hipSetDevice(b->device));
    return 0;
}

bool valid_box(const gz_batch *b, int id) { return id >= 0 && id < b->max_boxes && b->boxes[id].alive; }

BoxDev empty_entry()
{
    BoxDev t{};
    t.box = 1; t.half = 0.5; t.max_h = 1; t.cell = 1; t.rowlim = 1; t.shift = 3 * kFineLevel;
    return t;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------
extern "C" {

gz_batch *gz_batch_create(const gz_params *p, long capacity, int max_boxes, int device)
{
    if (!p || capacity <= 0 || capacity > 0x7fffffffL || max_boxes <= 0 || max_boxes > 65535) return nullptr;
    if (p->kernel != 3 && p->kernel != 6) return nullptr;
    if (!(p->gamma > 1) || !(p->des_num_ngb > 0)) return nullptr;
    if (p->precision != GZ_FP64 && p->precision != GZ_FP32_MIXED) return nullptr;
    gz_batch *b = new gz_batch();
    b->par = *p;
    b->device = device;
    b->cap = capacity;
    b->max_boxes = max_boxes;
    b->retired = max_boxes;
    DevParams &d = b->dp;
    d.box = 1;
    d.half = 0.5;
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
    bool ok = hipSetDevice(device) == hipSuccess && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
    if (ok) ok = alloc_all(b) == 0;
    if (ok) {
        b->boxes.resize(nbox_all(b));
        b->htab.assign(nbox_all(b), empty_entry());
        // every particle starts in the retired box, inert
        const long n = b->cap;
        k_fill_int<<<nb(n), kB, 0, b->stream>>>(n, b->BOX, b->retired);
        ok = hipGetLastError() == hipSuccess;
        double *zf[] = {b->M, b->H, b->X, b->Y, b->Z, b->VKX, b->VKY, b->VKZ, b->DIVV, b->X0, b->Y0, b->Z0};
        for (double *f : zf) ok = ok && hipMemsetAsync(f, 0, n * 8, b->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(b->TB, 0, n * 4, b->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(b->WAKE, 0, n * 4, b->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(b->key, 0, n * 4, b->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(b->ACT, 0, n, b->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(b->u64, 0, 8 * 8, b->stream) == hipSuccess;
        ok = ok && hipMemcpyAsync(b->tab, b->htab.data(), nbox_all(b) * sizeof(BoxDev), hipMemcpyHostToDevice, b->stream) == hipSuccess;
        ok = ok && hipStreamSynchronize(b->stream) == hipSuccess;
    }
    if (!ok) {
        free_all(b);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return nullptr;
    }
    return b;
}

void gz_batch_destroy(gz_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    free_all(b);
    (void)hipStreamDestroy(b->stream);
    delete b;
}

const char *gz_batch_last_error(const gz_batch *b) { return b ? b->err.c_str() : "no batch"; }

int gz_batch_add(gz_batch *b, long n, const gz_fields *f, const gz_box_params *bp, int *box_id)
{
    if (!b) return GZ_E_ARG;
    if (!f || !bp || !box_id) return fail(b, GZ_E_ARG, "fields, box parameters and box_id are required");
    if (!f->pos || !f->vel || !f->mass || !f->u || !f->h) return fail(b, GZ_E_ARG, "pos, vel, mass, u and h are required");
    if (n <= 0) return fail(b, GZ_E_ARG, "n must be > 0");
    if (!(bp->box_size > 0)) return fail(b, GZ_E_ARG, "box_size must be > 0");
    if (!(bp->max_kernel_radius > 0) || bp->max_kernel_radius > 0.5 * bp->box_size / 1.3)
        return fail(b, GZ_E_ARG, "max_kernel_radius must be in (0, box/2.6]");
    const gz_integ_params &ip = bp->integ;
    if (!(ip.time_max > ip.time_begin)) return fail(b, GZ_E_ARG, "time_max must exceed time_begin");
    if (!(ip.courant_fac > 0) || !(ip.err_tol_int_accuracy > 0) || !(ip.max_size_timestep > 0) || ip.min_size_timestep < 0 ||
        ip.min_egy_spec < 0)
        return fail(b, GZ_E_ARG, "courant_fac, err_tol_int_accuracy, max_size_timestep must be > 0; min_size_timestep, min_egy_spec >= 0");
    if (b->n_live + n > b->cap) return fail(b, GZ_E_MEMORY, "batch full: %ld + %ld particles over capacity %ld", b->n_live, n, b->cap);
    int id = -1;
    for (int k = 0; k < b->max_boxes; k++)
        if (!b->boxes[k].alive) { id = k; break; }
    if (id < 0) return fail(b, GZ_E_MEMORY, "batch full: %d boxes", b->max_boxes);
    BCHK(set_device(b));
    const long off = b->n_live, cap = b->cap;
    // the box's own Morton order (stable sort of the keys of the given positions)
    BHIP(hipMemcpyAsync(b->stage, f->pos, (size_t)n * 3 * 8, hipMemcpyHostToDevice, b->stream));
    k_morton<<<nb(n), kB, 0, b->stream>>>(n, b->stage, bp->box_size, (uint32_t *)b->ival2, b->ival);
    BHIP(hipGetLastError());
    uint32_t *ksrt = b->key + off;
    int *psrt = b->perm + off;
    size_t bytes = 0;
    BHIP(gz_sort_pairs(nullptr, bytes, (const uint32_t *)b->ival2, ksrt, b->ival, psrt, (int)n, 0, 3 * kFineLevel,
                                            b->stream));
    BCHK(ensure_tmp(b, bytes));
    BHIP(gz_sort_pairs(b->tmp, bytes, (const uint32_t *)b->ival2, ksrt, b->ival, psrt, (int)n, 0, 3 * kFineLevel,
                                            b->stream));
    // a host field (width doubles per particle, given order) into the box's range of its columns
    auto up = [&](const double *src, int w, double *dst) -> int {
        BHIP(hipMemcpyAsync(b->stage, src, (size_t)n * w * 8, hipMemcpyHostToDevice, b->stream));
        k_gather<<<nb(n), kB, 0, b->stream>>>(n, w, psrt, b->stage, b->dacc, 1);
        BHIP(hipGetLastError());
        for (int c = 0; c < w; c++)
            BHIP(hipMemcpyAsync(dst + (size_t)c * cap + off, b->dacc + (size_t)c * n, n * 8, hipMemcpyDeviceToDevice, b->stream));
        return 0;
    };
    auto up3 = [&](const double *src, double *dx, double *dy, double *dz) -> int {
        BHIP(hipMemcpyAsync(b->stage, src, (size_t)n * 3 * 8, hipMemcpyHostToDevice, b->stream));
        k_gather<<<nb(n), kB, 0, b->stream>>>(n, 3, psrt, b->stage, b->dacc, 1);
        BHIP(hipGetLastError());
        BHIP(hipMemcpyAsync(dx + off, b->dacc, n * 8, hipMemcpyDeviceToDevice, b->stream));
        BHIP(hipMemcpyAsync(dy + off, b->dacc + n, n * 8, hipMemcpyDeviceToDevice, b->stream));
        BHIP(hipMemcpyAsync(dz + off, b->dacc + 2 * n, n * 8, hipMemcpyDeviceToDevice, b->stream));
        return 0;
    };
    BCHK(up3(f->pos, b->X, b->Y, b->Z));
    BCHK(up3(f->vel, b->VX, b->VY, b->VZ));
    BCHK(up3(f->vel_kick ? f->vel_kick : f->vel, b->VKX, b->VKY, b->VKZ));
    BCHK(up(f->mass, 1, b->M));
    BCHK(up(f->u, 1, b->U));
    BCHK(up(f->h, 1, b->H));
    BCHK(up(f->u_kick ? f->u_kick : f->u, 1, b->UK));
    BCHK(up(f->soft ? f->soft : f->h, 1, b->SOFT));
    struct O { const double *src; int w; double *dst; } opt[] = {
        {f->numngb, 1, b->NN}, {f->drkern, 1, b->DRK}, {f->divvel, 1, b->DIVV}, {f->rho, 1, b->RHO}, {f->press, 1, b->PR},
        {f->nvt, 6, b->T}, {f->nvt_fw, 3, b->FW}, {f->cn, 1, b->CN}, {f->fce, 1, b->FCE}, {f->grad_rho, 3, b->GR},
        {f->grad_p, 3, b->GP}, {f->grad_v, 9, b->GV}, {f->accel, 3, b->ACC}, {f->dtu, 1, b->DTU}, {f->maxsig, 1, b->MSV}};
    for (auto &o : opt) {
        if (o.src) {
            BCHK(up(o.src, o.w, o.dst));
        } else {
            for (int c = 0; c < o.w; c++) BHIP(hipMemsetAsync(o.dst + (size_t)c * cap + off, 0, n * 8, b->stream));
        }
    }
    BHIP(hipMemcpyAsync(b->X0 + off, b->X + off, n * 8, hipMemcpyDeviceToDevice, b->stream));
    BHIP(hipMemcpyAsync(b->Y0 + off, b->Y + off, n * 8, hipMemcpyDeviceToDevice, b->stream));
    BHIP(hipMemcpyAsync(b->Z0 + off, b->Z + off, n * 8, hipMemcpyDeviceToDevice, b->stream));
    k_fill_int<<<nb(n), kB, 0, b->stream>>>(n, b->BOX + off, id);
    BHIP(hipGetLastError());
    BHIP(hipMemsetAsync(b->TB + off, 0, n * 4, b->stream));
    BHIP(hipMemsetAsync(b->WAKE + off, 0, n * 4, b->stream));
    BHIP(hipMemsetAsync(b->TIB + off, 0, n * 8, b->stream));
    BHIP(hipMemsetAsync(b->DTS + off, 0, n * 8, b->stream));
    BHIP(hipMemsetAsync(b->ACT + off, 0, n, b->stream));
    // the table entry: geometry, the start of the timeline (src/core/init.cc:113, :583), the
    // box's integration parameters; not stepping until it is started
    BoxDev &d = b->htab[id];
    d = empty_entry();
    d.box = bp->box_size;
    d.half = 0.5 * bp->box_size;
    d.min_h = bp->min_kernel_radius;
    d.max_h = bp->max_kernel_radius;
    d.cell = bp->box_size;
    d.rowlim = bp->box_size;
    d.time_begin = ip.time_begin;
    d.time = ip.time_begin;
    d.tbi = (ip.time_max - ip.time_begin) / (double)kTimebaseH;
    d.ti_current = 0;
    d.mask = (1ull << kTimebins) - 1;
    d.stepping = 0;
    d.courant_fac = ip.courant_fac;
    d.err_tol = ip.err_tol_int_accuracy;
    d.max_dt = ip.max_size_timestep;
    d.min_dt = ip.min_size_timestep;
    d.min_egy = ip.min_egy_spec;
    BCHK(upload_entry(b, id));
    BoxHost &h = b->boxes[id];
    h = BoxHost{};
    h.alive = true;
    h.n = n;
    h.off = off;
    h.bp = *bp;
    // the box's largest radius
    std::vector<uint8_t> sel(nbox_all(b), 0);
    sel[id] = 1;
    BCHK(upload_bytes(b, sel, b->sel));
    BHIP(hipMemsetAsync(b->slot, 0, kBoxSlots * nbox_all(b) * 8, b->stream));
    k_batch_box_hmax<<<nb(cap), kB, 0, b->stream>>>(cap, b->H, b->sel, table(b));
    BHIP(hipGetLastError());
    BCHK(read_back(b, false));
    h.hbound = slot_value(b, id, kSlotHmax);
    b->order.push_back(id);
    b->n_live += n;
    b->grid_valid = false;
    *box_id = id;
    return 0;
}

int gz_batch_begin(gz_batch *b)
{
    if (!b) return GZ_E_ARG;
    std::vector<uint8_t> sel(nbox_all(b), 0);
    int count = 0;
    for (int id : b->order)
        if (!b->boxes[id].begun) { sel[id] = 1; count++; }
    if (!count) return 0;
    BCHK(set_device(b));
    const long n = b->cap;
    BCHK(upload_bytes(b, sel, b->sel));
    BCHK(build_grid(b));
    // init(): predicted = current, bin 0, every particle of the new boxes active
    k_batch_start<<<nb(n), kB, 0, b->stream>>>(n, b->BOX, b->sel, pview(b));
    k_batch_flag_boxes<<<nb(n + 1), kB, 0, b->stream>>>(n, b->BOX, b->sel, b->M, b->iflag, b->ACT);
    BHIP(hipGetLastError());
    BCHK(compact_rows(b));
    BCHK(read_back(b, false));
    b->nact = b->hcounters[10];
    b->act_rows_valid = false;
    // init()'s density() (src/core/init.cc:1051) and softening cache (:1061), then run()'s first
    // compute_hydro_densities_and_forces() (src/core/run.cc:128-150)
    BCHK(density_dispatch(b));
    // every radius of the new boxes after it (the bound the symmetric list is built for)
    BHIP(hipMemsetAsync(b->slot, 0, kBoxSlots * nbox_all(b) * 8, b->stream));
    k_batch_box_hmax<<<nb(n), kB, 0, b->stream>>>(n, b->H, b->sel, table(b));
    k_batch_soft<<<nb(n), kB, 0, b->stream>>>(n, b->BOX, b->sel, b->H, b->SOFT);
    BHIP(hipGetLastError());
    BCHK(read_back(b, false));
    for (int id : b->order)
        if (sel[id]) b->boxes[id].hbound = slot_value(b, id, kSlotHmax);
    BCHK(hot_path_dispatch(b, sel));
    for (int id : b->order) {
        if (!sel[id]) continue;
        BoxHost &h = b->boxes[id];
        h.begun = true;
        h.n_active = h.n;
        for (int k = 0; k < kTimebins; k++) h.bins[k] = 0;
        h.bins[0] = h.n;
        b->htab[id].stepping = 1;
        BCHK(upload_entry(b, id));
    }
    BCHK(sync(b));
    b->act_rows_valid = false;
    b->stats["begins"] += count;
    return 0;
}

int gz_batch_step(gz_batch *b)
{
    if (!b) return GZ_E_ARG;
    std::vector<uint8_t> in_pass(nbox_all(b), 0);
    bool any = false;
    for (int id : b->order) {
        const BoxHost &h = b->boxes[id];
        if (h.begun && !h.stop_code) { in_pass[id] = 1; any = true; }
    }
    if (!any) return 0;
    BCHK(set_device(b));
    const long n = b->cap;
    const size_t nbx = nbox_all(b);
    const BoxTable bt = table(b);
    const ParticleView S = pview(b);
    const TimeParams T = tparams(b);
    if (!b->act_rows_valid) BCHK(rows_from_active(b));
    // The timeline part of the step on the device, read back once: find_timesteps
    // (src/core/timestep.cc:39-355) over the rows of the last sync point, the bin occupation,
    // process_wake_ups (:1565-1774), do_first_halfstep_kick (src/core/kicks.cc:34-97), the next
    // sync point of every box and the drift of its particles to it (src/core/run.cc:857-994),
    // and make_list_of_active_particles (:997-1009) with the softening cache
    BHIP(hipMemsetAsync(b->hist, 0, 64 * nbx * 8, b->stream));
    BHIP(hipMemsetAsync(b->slot, 0, kBoxSlots * nbx * 8, b->stream));
    BHIP(hipMemsetAsync(b->counters, 0, 16 * sizeof(int), b->stream));
    if (b->nact > 0) k_find_timesteps<<<nb(b->nact), kB, 0, b->stream>>>(T, S, b->nact, b->act, b->counters + 6, nullptr, bt, 0);
    k_bin_hist<<<std::min(nb(n), 1024u), kB, 0, b->stream>>>(n, b->TB, b->hist, bt);
    int wakeup_bin_offset = 0;
    while ((1LL << wakeup_bin_offset) < (long long)kWakeup) wakeup_bin_offset++;
    const unsigned nbb = (unsigned)((nbx + 63) / 64);
    k_wakeup_limits<<<(unsigned)nbx, 64, 0, b->stream>>>(T, b->hist, b->lim, bt);  // one wavefront per box
    k_wakeups<<<nb(n), kB, 0, b->stream>>>(T, S, b->lim, wakeup_bin_offset, b->hist, b->counters + 5, bt);
    if (b->nact > 0) k_halfstep_kick<<<nb(b->nact), kB, 0, b->stream>>>(T, S, b->nact, b->act, 0, bt);
    k_next_sync<<<(unsigned)nbx, 64, 0, b->stream>>>(T, b->hist, nullptr, bt);
    k_drift<<<std::min(nb(n), 2048u), kB, 0, b->stream>>>(T, S, 0, nullptr, b->X0, b->Y0, b->Z0, nullptr, nullptr, bt);
    k_active_flags<<<nb(n + 1), kB, 0, b->stream>>>(n, nullptr, b->TB, b->M, b->H, 0, b->ACT, b->SOFT, b->iflag, nullptr, bt);
    BHIP(hipGetLastError());
    BCHK(compact_rows(b));
    k_advance_boxes<<<nbb, 64, 0, b->stream>>>(bt);
    BHIP(hipGetLastError());
    BCHK(read_back(b, true));
    b->nact = b->hcounters[10];
    b->act_rows_valid = true;
    // per box: occupation, displacement and radius bounds, stop code; boxes whose particles
    // moved far from where they were ordered are ordered again
    std::vector<uint8_t> due(nbx, 0);
    bool any_due = false;
    for (int id : b->order) {
        if (!in_pass[id]) continue;
        BoxHost &h = b->boxes[id];
        const BoxDev &d = b->htab[id];
        long long na = 0;
        for (int k = 0; k < kTimebins; k++) {
            h.bins[k] = (long long)b->hhist[(size_t)64 * id + k];
            if ((d.mask >> k) & 1ull) na += h.bins[k];
        }
        h.n_active = na;
        h.num_steps++;
        h.n_woken = (long long)b->hslot[(size_t)id * kBoxSlots + kSlotWoken];
        h.dmax = slot_value(b, id, kSlotDisp);
        h.hbound = slot_value(b, id, kSlotHmax);
        const int stop = (int)b->hslot[(size_t)id * kBoxSlots + kSlotStop];
        if (stop) h.stop_code = stop;
        if (h.last_rmax > 0 && h.dmax > kResortFraction * h.last_rmax) { due[id] = 1; any_due = true; }
    }
    if (any_due) {
        BCHK(reorder(b, due));
        BCHK(rows_from_active(b));
    }
    // compute_hydro_densities_and_forces, then do_second_halfstep_kick (src/core/run.cc:128-150)
    BCHK(hot_path_dispatch(b, in_pass));
    if (b->nact > 0) {
        k_halfstep_kick<<<nb(b->nact), kB, 0, b->stream>>>(T, S, b->nact, b->act, 1, bt);
        BHIP(hipGetLastError());
    }
    BCHK(sync(b));
    // a box whose timeline ended stops stepping (its caller reads the stop code and removes it)
    for (int id : b->order) {
        if (in_pass[id] && b->boxes[id].stop_code) {
            b->htab[id].stepping = 0;
            BCHK(upload_entry(b, id));
            b->act_rows_valid = false;
        }
    }
    b->stats["steps"] += 1;
    return sync(b);
}

int gz_batch_remove(gz_batch *b, int box_id)
{
    if (!b) return GZ_E_ARG;
    if (!valid_box(b, box_id)) return fail(b, GZ_E_ARG, "no box %d", box_id);
    BCHK(set_device(b));
    std::vector<uint8_t> sel(nbox_all(b), 0);
    sel[box_id] = 1;
    BCHK(upload_bytes(b, sel, b->sel));
    k_batch_retire<<<nb(b->cap), kB, 0, b->stream>>>(b->cap, b->BOX, b->sel, b->retired, b->M, b->ACT);
    BHIP(hipGetLastError());
    b->htab[box_id] = empty_entry();
    BCHK(upload_entry(b, box_id));
    b->order.erase(std::remove(b->order.begin(), b->order.end(), box_id), b->order.end());
    b->boxes[box_id] = BoxHost{};
    // the live boxes packed to the front again, each in its own order
    std::vector<uint8_t> due(nbox_all(b), 0);
    BCHK(reorder(b, due));
    return sync(b);
}

int gz_batch_store(gz_batch *b, int box_id, gz_fields *f)
{
    if (!b) return GZ_E_ARG;
    if (!f) return fail(b, GZ_E_ARG, "fields are required");
    if (!valid_box(b, box_id)) return fail(b, GZ_E_ARG, "no box %d", box_id);
    BCHK(set_device(b));
    const BoxHost &h = b->boxes[box_id];
    const long n = h.n, off = h.off, cap = b->cap;
    const int *pm = b->perm + off;
    auto down = [&](double *dst, int w, const double *src) -> int {
        for (int c = 0; c < w; c++)
            BHIP(hipMemcpyAsync(b->dacc + (size_t)c * n, src + (size_t)c * cap + off, n * 8, hipMemcpyDeviceToDevice, b->stream));
        k_scatter<<<nb(n), kB, 0, b->stream>>>(n, w, pm, b->dacc, b->stage, 1);
        BHIP(hipGetLastError());
        BHIP(hipMemcpyAsync(dst, b->stage, (size_t)n * w * 8, hipMemcpyDeviceToHost, b->stream));
        return sync(b);
    };
    auto down3 = [&](double *dst, const double *dx, const double *dy, const double *dz) -> int {
        BHIP(hipMemcpyAsync(b->dacc, dx + off, n * 8, hipMemcpyDeviceToDevice, b->stream));
        BHIP(hipMemcpyAsync(b->dacc + n, dy + off, n * 8, hipMemcpyDeviceToDevice, b->stream));
        BHIP(hipMemcpyAsync(b->dacc + 2 * n, dz + off, n * 8, hipMemcpyDeviceToDevice, b->stream));
        k_scatter<<<nb(n), kB, 0, b->stream>>>(n, 3, pm, b->dacc, b->stage, 1);
        BHIP(hipGetLastError());
        BHIP(hipMemcpyAsync(dst, b->stage, (size_t)n * 3 * 8, hipMemcpyDeviceToHost, b->stream));
        return sync(b);
    };
    if (f->pos) BCHK(down3(f->pos, b->X, b->Y, b->Z));
    if (f->vel) BCHK(down3(f->vel, b->VX, b->VY, b->VZ));
    if (f->vel_kick) BCHK(down3(f->vel_kick, b->VKX, b->VKY, b->VKZ));
    struct O { double *dst; int w; const double *src; } opt[] = {
        {f->mass, 1, b->M}, {f->h, 1, b->H}, {f->u, 1, b->U}, {f->u_kick, 1, b->UK}, {f->soft, 1, b->SOFT}, {f->numngb, 1, b->NN},
        {f->drkern, 1, b->DRK}, {f->divvel, 1, b->DIVV}, {f->rho, 1, b->RHO}, {f->press, 1, b->PR}, {f->nvt, 6, b->T},
        {f->nvt_fw, 3, b->FW}, {f->cn, 1, b->CN}, {f->fce, 1, b->FCE}, {f->grad_rho, 3, b->GR}, {f->grad_p, 3, b->GP},
        {f->grad_v, 9, b->GV}, {f->accel, 3, b->ACC}, {f->dtu, 1, b->DTU}, {f->maxsig, 1, b->MSV}};
    for (auto &o : opt)
        if (o.dst) BCHK(down(o.dst, o.w, o.src));
    int *ti[2] = {f->timebin, f->wakeup};
    const int *td[2] = {b->TB + off, b->WAKE + off};
    for (int q = 0; q < 2; q++) {
        if (!ti[q]) continue;
        k_scatter_int<<<nb(n), kB, 0, b->stream>>>(n, pm, td[q], b->ival2);
        BHIP(hipGetLastError());
        BHIP(hipMemcpyAsync(ti[q], b->ival2, n * 4, hipMemcpyDeviceToHost, b->stream));
        BCHK(sync(b));
    }
    long long *tl[2] = {f->ti_begstep, f->dt_step};
    const long long *ts[2] = {b->TIB + off, b->DTS + off};
    for (int q = 0; q < 2; q++) {
        if (!tl[q]) continue;
        k_scatter_ll<<<nb(n), kB, 0, b->stream>>>(n, pm, ts[q], (long long *)b->stage);
        BHIP(hipGetLastError());
        BHIP(hipMemcpyAsync(tl[q], b->stage, n * 8, hipMemcpyDeviceToHost, b->stream));
        BCHK(sync(b));
    }
    return 0;
}

int gz_batch_timeline(gz_batch *b, int box_id, gz_timeline *out)
{
    if (!b) return GZ_E_ARG;
    if (!out) return fail(b, GZ_E_ARG, "out is required");
    if (!valid_box(b, box_id)) return fail(b, GZ_E_ARG, "no box %d", box_id);
    const BoxHost &h = b->boxes[box_id];
    const BoxDev &d = b->htab[box_id];
    memset(out, 0, sizeof *out);
    out->ti_current = d.ti_current;
    out->time = d.time;
    out->timebase_interval = d.tbi;
    out->num_steps = h.num_steps;
    out->n_active = h.n_active;
    out->n_woken = h.n_woken;
    for (int k = 0; k < kTimebins; k++) out->bin_count[k] = h.bins[k];
    out->bin_active = d.mask;
    out->stop_code = h.stop_code;
    out->full_drift = 1;  // a batch drifts every particle of its stepping boxes
    return 0;
}

long gz_batch_particles(const gz_batch *b) { return b ? b->n_live : 0; }
int gz_batch_boxes(const gz_batch *b) { return b ? (int)b->order.size() : 0; }

int gz_batch_synchronize(gz_batch *b)
{
    if (!b) return GZ_E_ARG;
    return sync(b);
}

int gz_batch_stat(gz_batch *b, const char *name, double *value)
{
    if (!b || !name || !value) return GZ_E_ARG;
    if (!strcmp(name, "force_pairs_total")) {
        BCHK(set_device(b));
        BHIP(hipMemcpyAsync(b->hu64, b->u64, 8, hipMemcpyDeviceToHost, b->stream));
        BCHK(sync(b));
        *value = (double)b->hu64[0];
        return 0;
    }
    if (!strcmp(name, "device_bytes")) {
        *value = (double)b->dev_bytes;
        return 0;
    }
    auto it = b->stats.find(name);
    *value = it == b->stats.end() ? 0 : it->second;
    return 0;
}

int gz_batch_set_option(gz_batch *b, const char *name, double value)
{
    if (!b || !name) return GZ_E_ARG;
    if (!strcmp(name, "few_rows")) {
        if (!(value >= 0) || value > 2e9) return fail(b, GZ_E_ARG, "few_rows must be in [0, 2e9]");
        b->few_rows = (int)value;
        return 0;
    }
    if (!strcmp(name, "fd_correction")) {
        // the first-derivative volume correction of the density, for the boxes' coming passes
        if (value != 0 && value != 1) return fail(b, GZ_E_ARG, "fd_correction must be 0 or 1");
        b->dp.fd4alpha = value != 0 ? 4.0 * fd_alpha(b->par.kernel) : 0.0;
        return 0;
    }
    return fail(b, GZ_E_ARG, "unknown option %s", name);
}

int gz_batch_timing_reset(gz_batch *b)
{
    if (!b) return GZ_E_ARG;
    b->stats.clear();
    BCHK(set_device(b));
    BHIP(hipMemsetAsync(b->u64, 0, 8, b->stream));
    return sync(b);
}

}  // extern "C"
