// gz_kernels.hip -- FP64 instantiations and the form-independent kernels (strict FP: built with
// -ffp-contract=off and correctly rounded division/sqrt, the oracle's rounding order).
#define GZ_DEFINE_COMMON_KERNELS
#include "gz_kernels_impl.h"
#include "gz_rows_impl.h"
#include "gz_time_impl.h"
#include "gz_batch_impl.h"

namespace gz {

// ---------------------------------------------------------------------------------------
// instantiations
// ---------------------------------------------------------------------------------------
template __global__ void k_build_ell<false>(int, const int *, const int *, const double *, double, const double *, const double *,
                                            const double *, const double *, const uint32_t *, GridView, int *, long, int, int *, int *);
template __global__ void k_build_ell<true>(int, const int *, const int *, const double *, double, const double *, const double *,
                                           const double *, const double *, const uint32_t *, GridView, int *, long, int, int *, int *);

#define GZ_INST_LANES(KERN, R, BOXES)                                                                                           \
    template __global__ void k_gradients<KERN, R, BOXES>(DevParams, int, const int *, const double *, const double *,           \
                                                         const double *, const double *, const GradRec<R> *, const double *, const double *,    \
                                                         long, const int *, long, const int *, double *, double *, double *,    \
                                                         BoxTable);                                                              \
    template __global__ void k_force<KERN, R, BOXES, false, true>(DevParams, Gam<R>, ForceView, const Cell<R> *, ForceRows,   \
                                                               BoxTable);                                                  \
    template __global__ void k_force_rare<KERN, R, BOXES>(DevParams, Gam<R>, ForceView, const Cell<R> *, ForceRows, BoxTable);

#define GZ_INST_DROWS(KERN, R, TPB)                                                                                             \
    template __global__ void k_density_rows<KERN, R, TPB>(DevParams, GridView, int, const int *, const uint8_t *,               \
                                                          const uint32_t *, const double *, const double *, const double *,     \
                                                          const double *, const double *, const double *, const double *,       \
                                                          const double *, double *, BoxTable, int *, const int *, int);

#define GZ_INST(KERN, R)                                                                                                         \
    template __global__ void k_density_cells<KERN, R>(DevParams, GridView, CellRows, const int *, const uint8_t *,              \
                                                      const double *, const double *, const double *, const double *,           \
                                                      const double *, const double *, const double *, const double *, double *);\
    GZ_INST_LANES(KERN, R, false)                                                                                                \
    GZ_INST_LANES(KERN, R, true)                                                                                                 \
    GZ_INST_DROWS(KERN, R, 64)                                                                                                   \
    GZ_INST_DROWS(KERN, R, kRowThreads)                                                                                          \
    template __global__ void k_gradients_rows<KERN, R>(DevParams, int, const int *, const double *, const double *,             \
                                                       const double *, const double *, const GradRec<R> *, GradSrc, const double *,             \
                                                       const double *, long, const int *, long, const int *, double *, double *,\
                                                       double *, BoxTable, int *);                                              \
    template __global__ void k_force_rows<KERN, R>(DevParams, Gam<R>, int, const int *, ForceView, const Cell<R> *,            \
                                                   const int *, long, const int *, const uint8_t *, const int *,                \
                                                   const double *, int *, double *, double *, double *, int *, int *,           \
                                                   unsigned long long *, BoxTable, const int *);                                \
    template __global__ void k_force_rows<KERN, R, false, kListThreads, true>(DevParams, Gam<R>, int, const int *, ForceView,  \
                                                   const Cell<R> *, const int *, long, const int *, const uint8_t *,            \
                                                   const int *, const double *, int *, double *, double *, double *, int *,     \
                                                   int *, unsigned long long *, BoxTable, const int *);
GZ_INST(3, double)
GZ_INST(6, double)
template __global__ void k_force_w3<3>(DevParams, Gam<double>, ForceView, const Cell<double> *, ForceRows, BoxTable);
template __global__ void k_force_w3<6>(DevParams, Gam<double>, ForceView, const Cell<double> *, ForceRows, BoxTable);
// the force kernels that count the Riemann rungs of their faces (one box)
#define GZ_INST_STATS(KERN, R)                                                                                                   \
    template __global__ void k_force<KERN, R, false, true, false>(DevParams, Gam<R>, ForceView, const Cell<R> *, ForceRows,   \
                                                               BoxTable);                                                  \
    template __global__ void k_force<KERN, R, false, false, false>(DevParams, Gam<R>, ForceView, const Cell<R> *, ForceRows,  \
                                                                BoxTable);                        \
    template __global__ void k_force_rows<KERN, R, true>(DevParams, Gam<R>, int, const int *, ForceView, const Cell<R> *,      \
                                                         const int *, long, const int *, const uint8_t *, const int *,          \
                                                         const double *, int *, double *, double *, double *, int *, int *,     \
                                                         unsigned long long *, BoxTable, const int *);
GZ_INST_STATS(3, double)
GZ_INST_STATS(6, double)
template __global__ void k_riemann_probe<double>(int, Gam<double>, long, const double *, double *, unsigned *);
template __global__ void k_pack_grad<double>(long, const double *, const double *, const double *, const double *, const double *,
                                             const double *, const double *, GradRec<double> *, const double *, const double *,
                                             const double *, double, BoxTable);
template __global__ void k_pack_force<double>(double, ForceView, Cell<double> *, int *, double, BoxTable);
template __global__ void k_pack_grad_list<double>(int, const int *, const int *, long, const int *, const double *, const double *,
                                               const double *, const double *, const double *, const double *, const double *,
                                               GradRec<double> *, const double *, const double *, const double *, double, BoxTable);
template __global__ void k_pack_force_list<double>(int, const int *, const int *, long, const int *, double, ForceView, Cell<double> *,
                                                int *, double, BoxTable);
#undef GZ_INST

}  // namespace gz
