// gz_batch_impl.h -- kernels of a batch of independent periodic boxes (order, cells, ranges).
// Particles of box b lie in one index range, in the Morton order of their own box; the batch
// is ordered by (rank of the box, Morton key), so one stable sort re-orders the boxes that
// moved, packs the live boxes to the front and leaves every other box's order as it was.
// Unused particles belong to the "retired" box (inert: no mass, never stepping).
#pragma once
#include "gz_kernels.h"

namespace gz {

// the batch sort key: rank of the box above the box's own 30-bit Morton key (recomputed from
// the current position for the boxes due, kept for the others)
__global__ void __launch_bounds__(256) k_batch_keys(long n, const double *__restrict__ X, const double *__restrict__ Y,
                                                     const double *__restrict__ Z, const int *__restrict__ BOX, const BoxDev *__restrict__ tab,
                                                     const int *__restrict__ rank, const uint8_t *__restrict__ due, uint32_t *__restrict__ key,
                                                     unsigned long long *__restrict__ key64, int *__restrict__ val)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = BOX[i];
    uint32_t k = key[i];
    if (due[b]) {
        const double box = tab[b].box;
        const uint32_t cx = fine_coord(X[i], box), cy = fine_coord(Y[i], box), cz = fine_coord(Z[i], box);
        k = (spread10(cx) << 2) | (spread10(cy) << 1) | spread10(cz);
        key[i] = k;
    }
    key64[i] = ((unsigned long long)rank[b] << 30) | k;
    val[i] = (int)i;
}

// [start, end) of every cell of every live box (cell index = the box's first cell + code); the
// retired particles stay out of the cells (no pass visits them)
__global__ void __launch_bounds__(256) k_batch_cell_bounds(long n, const uint32_t *__restrict__ key, const int *__restrict__ BOX,
                                                            const BoxDev *__restrict__ tab, int retired, int *__restrict__ cstart,
                                                            int *__restrict__ cend)
{
    const long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int b = BOX[s];
    if (b == retired) return;
    const int shift = tab[b].shift;
    const uint32_t c = key[s] >> shift;
    const long long cell = tab[b].cell_off + c;
    if (s == 0 || BOX[s - 1] != b || (key[s - 1] >> shift) != c) cstart[cell] = (int)s;
    if (s == n - 1 || BOX[s + 1] != b || (key[s + 1] >> shift) != c) cend[cell] = (int)(s + 1);
}

// rows of a pass over whole boxes (their start: every particle of the boxes marked in sel is
// active, init()'s all-bins-active state); other particles keep their active byte
__global__ void __launch_bounds__(256) k_batch_flag_boxes(long n, const int *__restrict__ BOX, const uint8_t *__restrict__ sel,
                                                           const double *__restrict__ M, int *__restrict__ flag, uint8_t *__restrict__ ACT)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { flag[n] = 0; return; }
    const bool on = sel[BOX[i]] != 0;
    if (on) ACT[i] = 1;
    flag[i] = (on && M[i] > 0) ? 1 : 0;
}

// rows of the last pass of every stepping box, from the active bytes (the active list that
// find_timesteps walks at the next sync point)
__global__ void __launch_bounds__(256) k_batch_flag_active(long n, const int *__restrict__ BOX, const BoxDev *__restrict__ tab,
                                                            const uint8_t *__restrict__ ACT, const double *__restrict__ M,
                                                            int *__restrict__ flag)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { flag[n] = 0; return; }
    flag[i] = (ACT[i] && M[i] > 0 && tab[BOX[i]].stepping) ? 1 : 0;
}

// the largest kernel radius of every box marked in sel (slot kSlotHmax)
__global__ void __launch_bounds__(256) k_batch_box_hmax(long n, const double *__restrict__ H, const uint8_t *__restrict__ sel, BoxTable bt)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const bool valid = i < n && sel[bt.BOX[i]];
    const int box = valid ? bt.BOX[i] : 0;
    double h = valid ? H[i] : 0.0;
    const unsigned long long vm = __ballot(valid);
    if (!vm) return;
    const int b0 = __shfl(box, __ffsll((long long)vm) - 1, 64);
    if (__all(!valid || box == b0)) {
        for (int o = 32; o > 0; o >>= 1) h = fmax(h, __shfl_xor(h, o, 64));
        if ((threadIdx.x & 63) == 0 && h > 0) atomic_max_bits(&bt.slot[(size_t)b0 * kBoxSlots + kSlotHmax], h);
    } else if (valid && h > 0) {
        atomic_max_bits(&bt.slot[(size_t)box * kBoxSlots + kSlotHmax], h);
    }
}

// the start of a box's timeline for the particles of the boxes marked in sel: bin 0, no step
// yet, no wake-up flag, predicted velocity and energy equal to the current ones
__global__ void __launch_bounds__(256) k_batch_start(long n, const int *__restrict__ BOX, const uint8_t *__restrict__ sel, ParticleView S)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n || !sel[BOX[i]]) return;
    S.TB[i] = 0; S.WAKE[i] = 0; S.TIB[i] = 0; S.DTS[i] = 0;
    S.VX[i] = S.VKX[i]; S.VY[i] = S.VKY[i]; S.VZ[i] = S.VKZ[i];
    S.U[i] = S.UK[i];
}

// softening cache of the particles of the boxes marked in sel (compute_all_force_softening)
__global__ void __launch_bounds__(256) k_batch_soft(long n, const int *__restrict__ BOX, const uint8_t *__restrict__ sel,
                                                     const double *__restrict__ H, double *__restrict__ SOFT)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n || !sel[BOX[i]]) return;
    SOFT[i] = H[i];
}

// retire the particles of the boxes marked in sel: the retired box, no mass
__global__ void __launch_bounds__(256) k_batch_retire(long n, int *__restrict__ BOX, const uint8_t *__restrict__ sel, int retired,
                                                       double *__restrict__ M, uint8_t *__restrict__ ACT)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n || !sel[BOX[i]]) return;
    BOX[i] = retired;
    M[i] = 0;
    ACT[i] = 0;
}

// fill a range with one value (no memset nodes)
__global__ void k_fill_int(long n, int *__restrict__ a, int v)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

}  // namespace gz
