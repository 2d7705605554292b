// gz_device.h -- device-side physics of the MFM hydro pipeline (gfx950).
//
// Pair bodies and per-particle closures of GIZMO's density(), hydro_gradient_calc() and
// hydro_force(), written for one lane per active particle.  Every function is templated on
// the pair arithmetic type R: R = double is the plain FP64 form and keeps the reference's
// order of operations term by term (the library is compiled with -ffp-contract=off, so
// nothing is fused behind its back); R = float is the FP32-mixed form.  Reference citations
// are paths relative to the reference source tree.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <type_traits>

namespace gz {

// src/declarations/constants.h:11-17,219-225,340,445
constexpr double kMaxReal = 1e56;
constexpr double kMinReal = 1e-56;
constexpr double kCnDanger = 1.0e3;
constexpr int kMaxIter = 150;
constexpr double kVolNorm = 4.188790204786;
constexpr double kWakeup = 4.1;

template <typename R> struct Tiny;                       // MIN_REAL_NUMBER in the pair type
template <> struct Tiny<double> { static constexpr double v = 1e-56; };
template <> struct Tiny<float> { static constexpr float v = 1e-37f; };  // 1e-56 underflows in FP32

struct DevParams {
    double box, half;
    double gamma;
    double G0, G1, G2, G3, G4, G5, G6, G7, G8, G9;  // src/hydro/reimann.h:1-10 (GAMMA_G*)
    double des_num_ngb, max_num_ngb_dev, min_kernel_radius, max_kernel_radius;
    double time, time_begin;
    // HYDRO_PARTITION_UNITY_IMPROVE_FD (src/hydro/density_loop.cc:622-633): 4 x the kernel's
    // KERNEL_AWPMHD_FD_ALPHA (src/mesh/kernel.h:185-206) when on, 0 when off
    double fd4alpha = 0;
};

// Neighbour lists (ELL): entry q of row r lives in chunks of kEllChunk consecutive entries of one
// row; chunk k of every row (ld rows apart), then chunk k+1.  kEllChunk = 1 is column-major: the
// lanes of a wave, one row each, read entry q from whole lines shared by 16 rows.  Chunks of 16
// (one line per row chunk) write each line once as a row appends, but lane-per-row readers then
// touch 64 lines per load and fetch them again as the gathered records evict them (measured on
// the 2M all-active pass: list writes 3.8 -> 1.1 GB, gradient and force fetches up to 10x, the
// pass 0.7 ms slower; the list build was no faster).
constexpr int kEllChunk = 1;
__host__ __device__ __forceinline__ size_t ell_at(long ld, long r, int q)
{
    return ((size_t)(q / kEllChunk) * (size_t)ld + (size_t)r) * kEllChunk + (size_t)(q % kEllChunk);
}
// entries to allocate per row for a list of width cap
__host__ __device__ __forceinline__ int ell_width(int cap) { return (cap + kEllChunk - 1) / kEllChunk * kEllChunk; }

// density accumulators per row: Ngb, Rho, DrkernNgb, DivVel, NV_T (6), face weights (3), and
// with the first-derivative volume correction GradH_numer (3) and GradH_denom
constexpr int kDensSums = 13;
constexpr int kDensSumsFD = 17;

// KERNEL_AWPMHD_FD_ALPHA of a kernel (src/mesh/kernel.h:185-206)
__host__ __device__ inline double fd_alpha(int kernel)
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

// ---------------------------------------------------------------------------------------
// kernel: kernel_hinv / kernel_main, src/mesh/kernel.h:245-260, :268-440
// ---------------------------------------------------------------------------------------
template <int KERN> struct KernelNorm;
template <> struct KernelNorm<3> { static constexpr double v = (8.0 / M_PI); };          // :69
template <> struct KernelNorm<6> { static constexpr double v = (21.0 / (2.0 * M_PI)); }; // :123

template <typename R>
__device__ __forceinline__ void kernel_hinv(R h, R &hinv, R &hinv3, R &hinv4)
{
    hinv = R(1.0) / h;
    hinv3 = hinv * hinv * hinv;
    hinv4 = hinv3 * hinv;
}

// MODE 0: wk and dwk; MODE -1: wk only (as kernel_main's mode argument).  kernel_main_q takes
// 1 - u as well (q; the edge factor of both kernels), kernel_main forms it from u.
template <int KERN, int MODE, typename R>
__device__ __forceinline__ void kernel_main_q(R u, R q, R hinv3, R hinv4, R &wk, R &dwk)
{
    if (u >= R(1)) { wk = 0; dwk = 0; return; }
    if constexpr (KERN == 3) {
        if (u < R(0.5)) {
            if (MODE >= 0) dwk = u * (R(18.0) * u - R(12.0));
            if (MODE <= 0) wk = (R(1.0) + R(6.0) * (u - R(1.0)) * u * u);
        } else {
            R t1 = q;
            R t2 = t1 * t1;
            if (MODE >= 0) dwk = R(-6.0) * t2;
            if (MODE <= 0) wk = R(2.0) * t2 * t1;
        }
    } else {
        R t1 = q;
        R t3 = t1 * t1 * t1;
        if (MODE >= 0) dwk = R(-20.0) * u * t3;
        if (MODE <= 0) wk = t3 * t1 * (R(1.0) + R(4.0) * u);
    }
    const R norm = R(KernelNorm<KERN>::v);
    if (MODE >= 0) dwk *= norm * hinv4;
    if (MODE <= 0) wk *= norm * hinv3;
}
template <int KERN, int MODE, typename R>
__device__ __forceinline__ void kernel_main(R u, R hinv3, R hinv4, R &wk, R &dwk)
{
    kernel_main_q<KERN, MODE, R>(u, R(1) - u, hinv3, hinv4, wk, dwk);
}

// The kernel of a pair at distance r from a particle of radius h (wk, dwk left as they are
// beyond h).  FP64 form: the reference's kernel_hinv and kernel_main at u = r / h.  FP32 form:
// the edge factor 1 - u from the FP64 distance rd and radius h, the rest in FP32.  A partner
// near the edge of h sits where the kernel is steep (W ~ (1 - u)^4), so the FP32 rounding of r
// and h would enter its weight many times over: a hot partner at u = 0.993 of a cold particle
// carries most of its pressure gradient, and an FP32 u moves that gradient by 5e-5.
template <int KERN, int MODE, typename R>
__device__ __forceinline__ void pair_kernel(R r, double rd, double h, R &wk, R &dwk)
{
    R hinv, hinv3, hinv4;
    kernel_hinv((R)h, hinv, hinv3, hinv4);
    if constexpr (std::is_same<R, double>::value) {
        (void)rd;
        if (r < h) kernel_main<KERN, MODE>(r * hinv, hinv3, hinv4, wk, dwk);
    } else {
        (void)r;
        if (rd < h) {
            const R q = (R)((h - rd) * (double)hinv);
            kernel_main_q<KERN, MODE>(R(1) - q, q, hinv3, hinv4, wk, dwk);
        }
    }
}

// |dp|^2 of a pair in FP64 from its components (the FP32 form's differences are exact to a few
// 1e-10 of the distance: fine cells and FP32 offsets); the FP64 form's is its own r2
template <typename R>
__device__ __forceinline__ double pair_r2d(R dp0, R dp1, R dp2, R r2)
{
    if constexpr (std::is_same<R, double>::value) return r2;
    return (double)dp0 * (double)dp0 + (double)dp1 * (double)dp1 + (double)dp2 * (double)dp2;
}

// periodic nearest image, TMP_WRAP_*_S (src/declarations/macros.h:67-95)
__device__ __forceinline__ double wrap(double x, double box, double half)
{
    return (x > half) ? (x - box) : ((x < -half) ? (x + box) : x);
}
// absolute form NGB_PERIODIC_BOX_LONG_* used by the neighbour predicate
__device__ __forceinline__ double wrap_abs(double x, double box, double half)
{
    double t = fabs(x);
    return (t > half) ? (box - t) : t;
}

// exact pair predicate, src/mesh/sfc_tiles_functions.h:83-102
template <bool SYMMETRIC>
__device__ __forceinline__ bool accept_pair(double xi, double yi, double zi, double xj, double yj, double zj,
                                            double h_i, double h_j, double box, double half)
{
    double adx = wrap_abs(xi - xj, box, half);
    double ady = wrap_abs(yi - yj, box, half);
    double adz = wrap_abs(zi - zj, box, half);
    double r2s;
    if (!SYMMETRIC) {
        r2s = h_i * h_i;
    } else {
        double hm = (h_i > h_j) ? h_i : h_j;
        r2s = hm * hm;
    }
    if (adx > h_i && (!SYMMETRIC || adx * adx > r2s)) return false;
    double r2 = adx * adx + ady * ady + adz * adz;
    return r2 < r2s;
}

template <typename T> __device__ __forceinline__ T dmax(T a, T b) { return (a > b) ? a : b; }
template <typename T> __device__ __forceinline__ T dmin(T a, T b) { return (a < b) ? a : b; }

// atomicMax of the bits of a positive double (they order like the double) that skips the
// read-modify-write when the value held is not smaller.  Atomics on one address are executed
// one after another at the memory side (~10 ns each), so a maximum over many waves costs only
// the loads; the maxima only grow while a kernel runs, so a stale read costs only an atomic.
__device__ __forceinline__ void atomic_max_bits(unsigned long long *p, double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < b) atomicMax(p, b);
}

// ---------------------------------------------------------------------------------------
// 3x3 inverse with condition number: matrix_invert_ndims, src/core/proto.h:75-95 with
// Mat3::invert, src/math_types/mat3.h:50-64.  Row-major.
// ---------------------------------------------------------------------------------------
__device__ inline double frob2(const double *m)
{
    double r0 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    double r1 = m[3] * m[3] + m[4] * m[4] + m[5] * m[5];
    double r2 = m[6] * m[6] + m[7] * m[7] + m[8] * m[8];
    return r0 + r1 + r2;
}

__device__ inline double matrix_invert(const double *T, double *Ti)
{
    double fn = frob2(T);
    for (int k = 0; k < 9; k++) Ti[k] = 0;
    double d = T[0] * (T[4] * T[8] - T[5] * T[7]) - T[1] * (T[3] * T[8] - T[5] * T[6]) + T[2] * (T[3] * T[7] - T[4] * T[6]);
    if (!(d == 0 || d != d)) {
        double id = 1.0 / d;
        Ti[0] = (T[4] * T[8] - T[5] * T[7]) * id;
        Ti[1] = (T[2] * T[7] - T[1] * T[8]) * id;
        Ti[2] = (T[1] * T[5] - T[2] * T[4]) * id;
        Ti[3] = (T[5] * T[6] - T[3] * T[8]) * id;
        Ti[4] = (T[0] * T[8] - T[2] * T[6]) * id;
        Ti[5] = (T[2] * T[3] - T[0] * T[5]) * id;
        Ti[6] = (T[3] * T[7] - T[4] * T[6]) * id;
        Ti[7] = (T[1] * T[6] - T[0] * T[7]) * id;
        Ti[8] = (T[0] * T[4] - T[1] * T[3]) * id;
    }
    double fi = frob2(Ti);
    double cn = sqrt(fn * fi) / 3;
    return dmax(cn, 1.0);
}

// NV_T regularised inverse: src/hydro/density_loop.h:582-593; src/hydro/density_loop.cc:472-484
__device__ inline double nvt_inverse(const double *s6, double normalizer, double *Ti)
{
    double loc[9] = {s6[0], s6[3], s6[5], s6[3], s6[1], s6[4], s6[5], s6[4], s6[2]};
    for (int k = 0; k < 9; k++) loc[k] = loc[k] / normalizer;
    const double thr = 10.0 * kCnDanger;
    const double tr = loc[0] + loc[4] + loc[8];
    double add = 1.05 * (tr / 3) / thr;
    double cn;
    // the reference loops until the condition number drops below the threshold; the growing
    // diagonal term gets there in a few dozen rounds for any positive semi-definite sum, and the
    // cap only keeps a non-finite input from holding a wave forever
    for (int guard = 0; guard < 400; guard++) {
        cn = matrix_invert(loc, Ti);
        if (cn < thr) break;
        loc[0] += add; loc[4] += add; loc[8] += add;
        add *= 1.2;
    }
    return cn;
}

// FaceClosureError: src/hydro/density_loop.h:624-658
__device__ inline double face_closure(const double *s6, const double *fw, const double *Ti, double normalizer, double V_i)
{
    double tr = s6[0] + s6[1] + s6[2];
    double dx_i = sqrt(V_i * tr);
    double inv[9];
    for (int k = 0; k < 9; k++) inv[k] = Ti[k] / normalizer;
    double o[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) o[a] += 2.0 * V_i * inv[3 * a + b] * fw[b];
    double leak = 0;
    for (int a = 0; a < 3; a++) leak += fabs(o[a]) / 3.0;
    return leak / (2.0 * 3.0 * pow(dx_i, 2.0));
}

// ---------------------------------------------------------------------------------------
// density: normalisation (src/hydro/density_loop.h:532-556) and after_iter_body
// (:516-961, gas particles, no sinks/grains/GALSF)
// ---------------------------------------------------------------------------------------
__device__ inline void density_normalise(const double *acc, double h, double &nn, double &dr, double &dv)
{
    nn = acc[0]; dr = acc[2]; dv = acc[3];
    if (nn > 0) {
        dr *= h / (3 * nn);
        dv /= nn;
        nn *= kVolNorm * pow(h, 3.0);
    } else {
        nn = 0; dr = 0; dv = 0;
    }
    if (dr > -0.9) dr = 1.0 / (1.0 + dr); else dr = 1.0;
    dv *= dr;
}

struct IterState {
    double left, right, cn_cur;
    int cn_init;
};

// a row's sums normalised at radius h (src/hydro/density_loop.h:516-620): what the radius
// iteration and the finalisation both read, computed once per round
struct DensityNorm {
    double nn, dr, dv, vol, normalizer, cn, fce;
    double Ti[9];
};

__device__ inline void density_norm(const double *acc, double h, DensityNorm &d)
{
    density_normalise(acc, h, d.nn, d.dr, d.dv);
    d.vol = kVolNorm * pow(h, 3.0) / (d.nn > 0 ? d.nn : 1.0);
    d.normalizer = pow(h, -1.0);
    d.cn = nvt_inverse(&acc[4], d.normalizer, d.Ti);
    d.fce = face_closure(&acc[4], &acc[10], d.Ti, d.normalizer, d.vol);
}

// returns 0 converged, 1 new radius in new_h, -1 empty bracket (stop code 90001009)
__device__ inline int density_after_iter(const DevParams &P, double cn_stored, int iter, double h, const DensityNorm &dn,
                                         IterState &sc, double &new_h_out)
{
    double desnumngb = P.des_num_ngb;
    double desnumngbdev = P.max_num_ngb_dev;
    if (P.time == P.time_begin) {
        if (P.max_num_ngb_dev > 0.05) desnumngbdev = 0.05;
    }
    const double dev0 = desnumngbdev;
    const double NumNgb = dn.nn, Dr = dn.dr, CN = dn.cn, fce = dn.fce;
    {
        const double c0 = 0.1 * kCnDanger;
        if ((iter == 0) && (CN > cn_stored) && (cn_stored > 0)) {
            double na = 1.0, ca = cn_stored;
            if (ca > c0) na = sqrt(1.0 + (ca - c0) / kCnDanger);
            if (na > 2.0) na = 2.0;
            const double d1 = fabs(NumNgb - P.des_num_ngb * na) / (dev0 * na);
            double nb = 1.0, cb = CN;
            if (cb > c0) nb = sqrt(1.0 + (cb - c0) / kCnDanger);
            if (nb > 2.0) nb = 2.0;
            const double d2 = fabs(NumNgb - P.des_num_ngb * nb) / (dev0 * nb);
            if (dmin(d1, d2) < 10.0) { sc.cn_cur = CN; sc.cn_init = 1; }
        }
        if (!sc.cn_init) { sc.cn_cur = cn_stored; sc.cn_init = 1; }
    }
    double ncorr = 1.0;
    {
        const double c0 = 0.1 * kCnDanger;
        if (sc.cn_cur > c0) ncorr = sqrt(1.0 + (sc.cn_cur - c0) / kCnDanger);
        if (ncorr > 2.0) ncorr = 2.0;
        if (fce > 0.35) ncorr = dmax(ncorr, dmin(fce / 0.35, 2.0));
    }
    desnumngb = P.des_num_ngb * ncorr;
    desnumngbdev = dev0 * ncorr;
    if (iter > 1)
        desnumngbdev = dmin(0.25 * desnumngb, desnumngbdev * exp(0.1 * log(desnumngb / (16.0 * desnumngbdev)) * (double)iter));
    const double minsoft = P.min_kernel_radius, maxsoft = P.max_kernel_radius;
    int redo = 0, set_min = 0, set_max = 0;
    double nh = h;
    if ((NumNgb < (desnumngb - desnumngbdev) && nh < 0.999 * maxsoft) ||
        (NumNgb > (desnumngb + desnumngbdev) && nh > 1.001 * minsoft)) redo = 1;
    if ((nh >= 0.999 * maxsoft) && (NumNgb < (desnumngb - desnumngbdev))) {
        redo = 0;
        if (nh == maxsoft) set_max = 0; else { redo = 1; nh = maxsoft; set_max = 1; }
    }
    if ((nh <= 1.001 * minsoft) && (NumNgb > (desnumngb + desnumngbdev))) {
        redo = 0;
        if (nh == minsoft) set_min = 0; else { redo = 1; nh = minsoft; set_min = 1; }
    }
    if (redo == 0) sc.cn_cur = CN;
    if (redo) {
        if (sc.left > 0 && sc.right > 0) {
            if ((sc.right - sc.left) < 1.0e-3 * sc.left) { sc.cn_cur = CN; new_h_out = nh; return 0; }
        }
        if ((set_max == 0) && (set_min == 0)) {
            if (NumNgb < (desnumngb - desnumngbdev)) {
                sc.left = dmax(nh, sc.left);
            } else {
                if (sc.right != 0) { if (nh < sc.right) sc.right = nh; }
                else sc.right = nh;
            }
            if (sc.right > 0 && sc.left > 0) {
                double maxjump = 0;
                if (iter > 1) maxjump = 0.2 * log(sc.right / sc.left);
                if (NumNgb > 1) {
                    double jumpvar = Dr * log(desnumngb / NumNgb) / 3;
                    if (iter > 1) {
                        if (fabs(jumpvar) < maxjump) { if (jumpvar < 0) jumpvar = -maxjump; else jumpvar = maxjump; }
                    }
                    nh *= exp(jumpvar);
                } else {
                    nh *= 2.0;
                }
                if ((nh < sc.right) && (nh > sc.left)) {
                    if (iter > 1) {
                        const double hfac = exp(maxjump);
                        if (nh > sc.right / hfac) nh = sc.right / hfac;
                        if (nh < sc.left * hfac) nh = sc.left * hfac;
                    }
                } else {
                    if (nh > sc.right) nh = sc.right;
                    if (nh < sc.left) nh = sc.left;
                    nh = pow(nh * sc.left * sc.right, 1.0 / 3.0);
                }
            } else {
                if (sc.right == 0 && sc.left == 0) { new_h_out = nh; return -1; }
                if (sc.right == 0 && sc.left > 0) {
                    double fl = (NumNgb > 1) ? log(desnumngb / NumNgb) / 3 : 1.4;
                    if ((NumNgb < 2.0 * desnumngb) && (NumNgb > 0.1 * desnumngb)) {
                        double slope = Dr;
                        if (iter > 2 && slope < 1.0) slope = 0.5 * (slope + 1.0);
                        double fac = fl * slope;
                        if (iter >= 4) { if (Dr == 1.0) fac *= 10.0; }
                        if (fac < fl + 0.231) nh *= exp(fac); else nh *= exp(fl + 0.231);
                    } else {
                        nh *= exp(fl);
                    }
                }
                if (sc.right > 0 && sc.left == 0) {
                    double fl = (NumNgb > 1) ? log(desnumngb / NumNgb) / 3 : 1.4;
                    if (fl < -1.535) fl = -1.535;
                    if ((NumNgb < 2.0 * desnumngb) && (NumNgb > 0.1 * desnumngb)) {
                        double slope = Dr;
                        if (iter > 2 && slope < 1.0) slope = 0.5 * (slope + 1.0);
                        double fac = fl * slope;
                        if (iter >= 4) { if (Dr == 1.0) fac *= 10.0; }
                        if (fac > fl - 0.231) nh *= exp(fac); else nh *= exp(fl - 0.231);
                    } else {
                        nh *= exp(fl);
                    }
                }
            }
        }
        if (nh < minsoft) nh = minsoft;
        if (set_min == 1) nh = minsoft;
        if (nh > maxsoft) nh = maxsoft;
        if (set_max == 1) nh = maxsoft;
        new_h_out = nh;
        return 1;
    }
    sc.cn_cur = CN;
    new_h_out = nh;
    return 0;
}

// ---------------------------------------------------------------------------------------
// gradients: local_slopelimiter, src/hydro/gradients_loop.cc:76-104
// ---------------------------------------------------------------------------------------
template <typename R>
__device__ inline void slopelimiter(R *g, R valmax, R valmin, R alim, R h, R shoot_tol, int pos_preserve, R d_max, R val_cen)
{
    R d_abs = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    if (d_abs > 0) {
        d_abs = sqrt(d_abs);
        R cfac = R(1) / (alim * h * d_abs);
        R fmx = fabs(valmax), fmn = fabs(valmin), amax = fmx, amin = fmn;
        if (amax < amin) { amax = fmn; amin = fmx; }
        R fc = dmin(amin + shoot_tol * amax, amax);
        cfac *= fc;
        if (pos_preserve == 1) {
            R fmin = dmin(val_cen, dmax(R(0), dmax(R(Tiny<R>::v) * val_cen, dmin(R(0.5) * (val_cen + valmin), val_cen - fc))));
            cfac = dmin(((val_cen - fmin) / d_max) / d_abs, cfac);
        }
        if (cfac < R(1)) { g[0] *= cfac; g[1] *= cfac; g[2] *= cfac; }
    }
}

// ---------------------------------------------------------------------------------------
// Riemann problem (src/hydro/reimann.h), non-MHD, no EOS_GENERAL/COOLING/GALSF
// ---------------------------------------------------------------------------------------
template <typename R> struct Gam { R G0, G1, G2, G3, G4, G5, G6, G7, G8, G9; };

template <typename R> struct RState { R rho, p, v0, v1, v2; };
template <typename R> struct ROut { R P_M, S_M, f0, f1, f2, fp; };

// reconstruct_face_states, src/hydro/reimann.h:125-244 (fac_minmax 0.5, fac_meddev 0.375)
template <typename R>
__device__ __forceinline__ void reconstruct(R Qi, R gi0, R gi1, R gi2, R Qj, R gj0, R gj1, R gj2,
                                            const R *di, const R *dj, R &QL, R &QR, int mode)
{
    if (mode == 0) { QR = Qi; QL = Qj; return; }
    if (Qi == Qj) { QL = QR = Qi; return; }
    QR = Qi + gi0 * di[0] + gi1 * di[1] + gi2 * di[2];
    QL = Qj + gj0 * dj[0] + gj1 * dj[1] + gj2 * dj[2];
    R Qmin, Qmax, Qmed, Qmax_eff, Qmin_eff, fac, Qmed_max, Qmed_min;
    Qmed = R(0.5) * (Qi + Qj);
    if (Qi < Qj) { Qmax = Qj; Qmin = Qi; } else { Qmax = Qi; Qmin = Qj; }
    fac = R(0.5) * (Qmax - Qmin);
    Qmax_eff = Qmax + fac;
    Qmin_eff = Qmin - fac;
    if (mode > 0) {
        if (Qmax < 0) { if (Qmax_eff > 0) Qmax_eff = Qmax * Qmax / (Qmax - (Qmax_eff - Qmax)); }
        if (Qmin > 0) { if (Qmin_eff < 0) Qmin_eff = Qmin * Qmin / (Qmin + (Qmin - Qmin_eff)); }
    }
    fac = R(0.375) * (Qmax - Qmin);
    Qmed_max = Qmed + fac;
    Qmed_min = Qmed - fac;
    if (Qmed_max > Qmax_eff) Qmed_max = Qmax_eff;
    if (Qmed_min < Qmin_eff) Qmed_min = Qmin_eff;
    if (Qi < Qj) {
        if (QR < Qmin_eff) QR = Qmin_eff;
        if (QR > Qmed_max) QR = Qmed_max;
        if (QL > Qmax_eff) QL = Qmax_eff;
        if (QL < Qmed_min) QL = Qmed_min;
    } else {
        if (QR > Qmax_eff) QR = Qmax_eff;
        if (QR < Qmed_min) QR = Qmed_min;
        if (QL < Qmin_eff) QL = Qmin_eff;
        if (QL > Qmed_max) QL = Qmed_max;
    }
}

// Which estimates and fallbacks of the Riemann ladder a face took (bit per rung; a face that
// takes a rung several times, in its retries, counts once).  No counterpart in the reference:
// the force pass counts them on request, so that tests can see each rung run on the device.
enum RiemannRung : unsigned {
    kRungHllc = 1u << 0,          // first HLLC estimate accepted (get_wavespeeds_and_pressure_star, :496-512)
    kRungVacuum = 1u << 1,        // separating faster than sound: vacuum star state (:490-494)
    kRungRoe = 1u << 2,           // Roe-averaged wave speeds (:514-530)
    kRungPvrs = 1u << 3,          // primitive-variable estimate (:531-546)
    kRungKt = 1u << 4,            // Kurganov-Tadmor flux (Riemann_solver, :340-342)
    kRungExact = 1u << 5,         // exact solver (:352-355)
    kRungExactVacuum = 1u << 6,   // exact solver without an iterative solution: vacuum sample (:640-666)
    kRungUnphysical = 1u << 7,    // negative density or pressures (:284-290)
    kRungRetryFirstOrder = 1u << 8,  // first-order states at 1.4x the limiter (hydro_core_meshless_functions.h:351-380)
    kRungRetryZeroVel = 1u << 9,  // zero velocities at twice the limiter (:381-420)
    kRungFail = 1u << 10,         // no positive pressure anywhere (:421-425)
    kRungFace = 1u << 11,         // the face reached the Riemann solve
    kEntropicConsidered = 1u << 12,  // the entropic switch was evaluated (hydro_core_meshless_functions.h:515-518)
    kEntropicApplied = 1u << 13,  // and its energy flux replaced the Riemann one (:569-572)
};
constexpr int kRungCount = 14;

// get_wavespeeds_and_pressure_star, src/hydro/reimann.h:485-549; returns the estimate taken
template <typename R>
__device__ inline unsigned wavespeeds(const Gam<R> &G, const RState<R> &L, const RState<R> &Rs, const R *n,
                                      R vL, R vR, R csL, R csR, R hL, R hR, ROut<R> &o, R ptl)
{
    const R tiny = R(Tiny<R>::v);
    if ((vR - vL) > dmax(csL, csR)) {
        o.P_M = tiny; o.S_M = 0;
        return kRungVacuum;
    }
    R PT_L = L.p, PT_R = Rs.p;
    R S_L = dmin(vL, vR) - dmax(csL, csR);
    R S_R = dmax(vL, vR) + dmax(csL, csR);
    R wL = L.rho * (S_L - vL);
    R wR = Rs.rho * (S_R - vR);
    o.S_M = ((PT_R - PT_L) + wL * vL - wR * vR) / (wL - wR);
    o.P_M = (PT_L * wR - PT_R * wL + wL * wR * (vR - vL)) / (wR - wL);
    if (o.P_M <= tiny) { o.P_M = tiny; o.S_M = 0; }
    if (!((o.P_M <= 0) || (isnan(o.P_M)) || (o.P_M > ptl))) return kRungHllc;
    R sL = sqrt(L.rho), sR = sqrt(Rs.rho);
    R sinv = R(1) / (sL + sR);
    R vr0 = (sL * L.v0 + sR * Rs.v0) * sinv;
    R vr1 = (sL * L.v1 + sR * Rs.v1) * sinv;
    R vr2 = (sL * L.v2 + sR * Rs.v2) * sinv;
    R vline = vr0 * n[0] + vr1 * n[1] + vr2 * n[2];
    R h_roe = (sL * hL + sR * hR) * sinv;
    R cs_roe = sqrt(dmax(R(1.e-30), G.G9 * (h_roe - R(0.5) * (vr0 * vr0 + vr1 * vr1 + vr2 * vr2))));
    S_R = dmax(vR + csR, vline + cs_roe);
    S_L = dmin(vL - csL, vline - cs_roe);
    wR = Rs.rho * (S_R - vR);
    wL = -L.rho * (S_L - vL);
    o.S_M = ((wR * vR + wL * vL) + (PT_L - PT_R)) / (wR + wL);
    o.P_M = L.rho * (vL - S_L) * (vL - o.S_M) + PT_L;
    if (o.P_M <= tiny) { o.P_M = tiny; o.S_M = 0; }
    if (!((o.P_M <= 0) || (isnan(o.P_M)) || (o.P_M > ptl))) return kRungRoe;
    o.P_M = R(0.5) * ((PT_L + PT_R) + (vL - vR) * R(0.25) * (L.rho + Rs.rho) * (csL + csR));
    o.S_M = R(0.5) * (vR + vL) + R(2.0) * (PT_L - PT_R) / ((L.rho + Rs.rho) * (csL + csR));
    R Sp = dmax(dmax(fabs(vL - csL), fabs(vR - csR)), dmax(fabs(vL + csL), fabs(vR + csR)));
    if (o.S_M < -Sp) o.S_M = -Sp;
    if (o.S_M > Sp) o.S_M = Sp;
    if (o.P_M <= tiny) { o.P_M = tiny; o.S_M = 0; }
    return kRungPvrs;
}

// Riemann_solver_KurganovTadmor_PWK, src/hydro/reimann.h:445-477 (MFM branch)
template <typename R>
__device__ inline void kt_flux(const RState<R> &L, const RState<R> &Rs, const R *n, R vL, R vR, R csL, R csR,
                               R hL, R hR, ROut<R> &o)
{
    R dth = R(0.001) * R(0.5) * (L.rho + Rs.rho) * R(0.5) * (csL + csR);
    R d0 = Rs.rho * Rs.v0 - L.rho * L.v0, d1 = Rs.rho * Rs.v1 - L.rho * L.v1, d2 = Rs.rho * Rs.v2 - L.rho * L.v2;
    R dv2 = d0 * d0 + d1 * d1 + d2 * d2;
    R dvL = fabs(Rs.rho * vR - L.rho * vL);
    R alpha = dvL / sqrt(dth * dth + dv2);
    R S_L = alpha * csL; if (vL > 0) S_L += vL; else S_L -= vL;
    R S_R = alpha * csR; if (vR > 0) S_R += vR; else S_R -= vR;
    R S_M = dmax(S_L, S_R);
    o.S_M = dmax(csL + fabs(vL), csR + fabs(vR));
    R fl = L.rho * (vL + S_M);
    R fr = Rs.rho * (vR - S_M);
    R denom = R(1) / (L.rho * vL - Rs.rho * vR + S_M * (L.rho + Rs.rho));
    o.P_M = (fl * Rs.p - fr * L.p) * denom;
    R c = fl * fr * denom;
    o.f0 = (Rs.v0 - L.v0) * c + (o.P_M) * n[0];
    o.f1 = (Rs.v1 - L.v1) * c + (o.P_M) * n[1];
    o.f2 = (Rs.v2 - L.v2) * c + (o.P_M) * n[2];
    o.fp = (S_M * (fl * Rs.p + fr * L.p) + (hR - hL) * fl * fr) * denom;
}

// guess_for_pressure (:1028-1068) and iterative_Riemann_solver (:958-1026)
template <typename R>
__device__ inline R guess_p(const Gam<R> &G, const RState<R> &L, const RState<R> &Rs, R vL, R vR, R csL, R csR)
{
    R pv = R(0.5) * (L.p + Rs.p) - R(0.125) * (vR - vL) * (L.p + Rs.p) * (csL + csR);
    R pmin = dmin(L.p, Rs.p), pmax = dmax(L.p, Rs.p);
    if (pmin <= 0) return R(0.5) * (pmin + pmax);
    R q = pmax / pmin;
    if (q <= R(2.0) && (pmin <= pv && pv <= pmax)) return pv;
    if (pv < pmin) {
        R pnu = (csL + csR) - G.G7 * (vR - vL);
        R pde = csL / pow(L.p, G.G1) + csR / pow(Rs.p, G.G1);
        return pow(pnu / pde, G.G3);
    }
    R gel = sqrt((G.G5 / L.rho) / (G.G6 * L.p + pv));
    R ger = sqrt((G.G5 / Rs.rho) / (G.G6 * Rs.p + pv));
    R x = (gel * L.p + ger * Rs.p - (vR - vL)) / (gel + ger);
    if (x < pmin || x > pmax) x = pmin;
    return x;
}

template <typename R>
__device__ inline int iterative_riemann(const Gam<R> &G, const RState<R> &L, const RState<R> &Rs, R vL, R vR, R csL, R csR, ROut<R> &o)
{
    if constexpr (std::is_same<R, float>::value) {
        // FP32 form: the Newton iteration in FP64.  Near gamma = 1 the FP32 pressure function
        // (p^((gamma-1)/(2 gamma)) - 1) cancels to noise and the iteration would run to its
        // 1000-step limit without reaching the 1e-6 tolerance; in FP64 it converges in a few
        // steps (the rung is rare, its cost is not the pass's)
        const double g = (double)G.G0;
        const Gam<double> Gd{g, (g - 1.0) / (2.0 * g), (g + 1.0) / (2.0 * g), 2.0 * g / (g - 1.0), 2.0 / (g - 1.0), 2.0 / (g + 1.0),
                             (g - 1.0) / (g + 1.0), 0.5 * (g - 1.0), 1.0 / g, g - 1.0};
        const RState<double> Ld{(double)L.rho, (double)L.p, (double)L.v0, (double)L.v1, (double)L.v2};
        const RState<double> Rd{(double)Rs.rho, (double)Rs.p, (double)Rs.v0, (double)Rs.v1, (double)Rs.v2};
        ROut<double> od{0, 0, 0, 0, 0, 0};
        const int ok = iterative_riemann<double>(Gd, Ld, Rd, (double)vL, (double)vR, (double)csL, (double)csR, od);
        if (ok) {
            o.P_M = (R)od.P_M;
            o.S_M = (R)od.S_M;
        }
        return ok;
    }
    R dvel = vR - vL;
    R check = G.G4 * (csR + csL) - dvel;
    if (check < 0) return 0;
    R tol = R(100.0);
    R Pg = guess_p(G, L, Rs, vL, vR, csL, csR);
    R WL = 0, WR = 0, ZL, ZR;
    int it = 0;
    while ((tol > R(1.e-6)) && (it < 1000)) {
        R Pp = Pg;
        if (Pg > L.p) {
            R a0 = G.G5 / L.rho, a1 = G.G6 * L.p, a2 = sqrt(a0 / (Pg + a1));
            WL = (Pg - L.p) * a2;
            ZL = a2 * (R(1.0) - R(0.5) * (Pg - L.p) / (a1 + Pg));
        } else {
            R pr = Pg / L.p;
            WL = G.G4 * csL * (pow(pr, G.G1) - R(1));
            ZL = R(1) / (L.rho * csL) * pow(Pg / L.p, -G.G2);
        }
        if (Pg > Rs.p) {
            R a0 = G.G5 / Rs.rho, a1 = G.G6 * Rs.p, a2 = sqrt(a0 / (Pg + a1));
            WR = (Pg - Rs.p) * a2;
            ZR = a2 * (R(1.0) - R(0.5) * (Pg - Rs.p) / (a1 + Pg));
        } else {
            R pr = Pg / Rs.p;
            WR = G.G4 * csR * (pow(pr, G.G1) - R(1));
            ZR = R(1) / (Rs.rho * csR) * pow(pr, -G.G2);
        }
        if (it < 1000 / 2) Pg -= (WL + WR + dvel) / (ZL + ZR);
        else Pg -= R(0.5) * (WL + WR + dvel) / (ZL + ZR);
        if (Pg < R(0.1) * Pp) Pg = R(0.1) * Pp;
        tol = R(2.0) * fabs((Pg - Pp) / (Pg + Pp));
        it++;
    }
    if (it < 1000) {
        o.P_M = Pg;
        o.S_M = R(0.5) * (vL + vR) + R(0.5) * (WR - WL);
        return 1;
    }
    return 0;
}

// Riemann_solver_exact (:625-668), MFM: only P_M and S_M change
template <typename R>
__device__ inline unsigned exact_riemann(const Gam<R> &G, const RState<R> &L, const RState<R> &Rs, R vL, R vR, R csL, R csR, ROut<R> &o)
{
    if ((L.p == 0 && Rs.p == 0) || (L.rho == 0 && Rs.rho == 0)) { o.P_M = o.S_M = 0; return kRungExact | kRungExactVacuum; }
    if ((L.rho > 0) && (Rs.rho > 0)) {
        if (!iterative_riemann(G, L, Rs, vL, vR, csL, csR, o)) {
            R SL = vL + G.G4 * csL, SR = vR - G.G4 * csR;
            if (R(0.0) <= SL) { o.P_M = 0; o.S_M = vL + G.G4 * csL; }
            else if (R(0.0) >= SR) { o.P_M = 0; o.S_M = vR - G.G4 * csR; }
            else { o.P_M = 0; o.S_M = R(0.0); }
            return kRungExact | kRungExactVacuum;
        }
        return kRungExact;
    }
    if (L.rho > 0) { o.P_M = 0; o.S_M = vL + G.G4 * csL; }
    if (Rs.rho > 0) { o.P_M = 0; o.S_M = vR - G.G4 * csR; }
    return kRungExact | kRungExactVacuum;
}

// the sound speeds, enthalpies and normal velocities Riemann_solver derives from the face states
// (src/hydro/reimann.h:313-339, no EOS_GENERAL)
template <typename R> struct RPrep { R csL, csR, hL, hR, vL, vR; };
template <typename R>
__device__ __forceinline__ RPrep<R> riemann_prep(const Gam<R> &G, const RState<R> &L, const RState<R> &Rs, const R *n)
{
    RPrep<R> q;
    q.csL = sqrt(G.G0 * L.p / L.rho);
    q.csR = sqrt(G.G0 * Rs.p / Rs.rho);
    R uL = L.p / (G.G9 * L.rho);
    R uR = Rs.p / (G.G9 * Rs.rho);
    q.hL = L.p / L.rho + uL + R(0.5) * (L.v0 * L.v0 + L.v1 * L.v1 + L.v2 * L.v2);
    q.hR = Rs.p / Rs.rho + uR + R(0.5) * (Rs.v0 * Rs.v0 + Rs.v1 * Rs.v1 + Rs.v2 * Rs.v2);
    q.vL = L.v0 * n[0] + L.v1 * n[1] + L.v2 * n[2];
    q.vR = Rs.v0 * n[0] + Rs.v1 * n[1] + Rs.v2 * n[2];
    return q;
}

// Riemann_solver (:282-370) with HLLC_Riemann_solver (:382-396, MFM fluxes); returns the rungs taken
template <typename R>
__device__ inline unsigned riemann(const Gam<R> &G, const RState<R> &L, const RState<R> &Rs, const R *n, R ptl, ROut<R> &o)
{
    if ((L.p < 0 && Rs.p < 0) || (L.rho < 0) || (Rs.rho < 0)) { o.P_M = 0; return kRungUnphysical; }
    const RPrep<R> q = riemann_prep(G, L, Rs, n);
    unsigned rung = wavespeeds(G, L, Rs, n, q.vL, q.vR, q.csL, q.csR, q.hL, q.hR, o, ptl);
    o.f0 = o.P_M * n[0];
    o.f1 = o.P_M * n[1];
    o.f2 = o.P_M * n[2];
    o.fp = o.P_M * o.S_M;
    if ((o.P_M < 0) || (isnan(o.P_M))) {
        kt_flux(L, Rs, n, q.vL, q.vR, q.csL, q.csR, q.hL, q.hR, o);
        rung |= kRungKt;
    }
    if ((o.P_M < 0) || (isnan(o.P_M)) || (o.P_M > ptl)) rung |= exact_riemann(G, L, Rs, q.vL, q.vR, q.csL, q.csR, o);
    return rung;
}

// ---------------------------------------------------------------------------------------
// force pair: hydro_accumulate_neighbor (src/hydro/hydro_functions.h:57-546) with
// hydro_core_meshless_compute_pair (src/hydro/hydro_core_meshless_functions.h:49-616) and
// compute_finitevol_faces (src/hydro/compute_finitevol_faces_functions.h:48-168)
// ---------------------------------------------------------------------------------------
// per-particle record of the force pass (32 values: 256 B in FP64, 128 B in FP32), packed
// once per pass so that a neighbour costs a few wide loads instead of 32 scattered ones
// FP32 form: the value as rounded to FP32, so that no later addition fuses with the
// multiplication that made it (the FP32 kernels are built with FMA contraction; the kernel
// shapes then add the same rounded terms and give the same bits).  No effect on FP64.
template <typename R>
__device__ __forceinline__ R rounded(R v)
{
    if constexpr (std::is_same<R, float>::value) __asm__ volatile("" : "+v"(v));
    return v;
}

// FP32 form: a particle's position as its cell of a 1024^3 grid over its box (10 bits per axis,
// x in bits 20-29) and its FP32 offset from that cell's corner.  The FP64 form reads positions
// from the FP64 arrays (PosRec<double> is empty).
struct FinePos {
    uint32_t c;
    float o0, o1, o2;
};
template <typename R> struct PosRec {};
template <> struct PosRec<float> { FinePos fp; };

__device__ __forceinline__ FinePos fine_pos(double x, double y, double z, double box)
{
    const double cs = box * (1.0 / 1024.0);
    int c[3];
    float o[3];
    const double v[3] = {x, y, z};
    for (int a = 0; a < 3; a++) {
        int k = (int)floor(v[a] / cs);
        k = k < 0 ? 0 : (k > 1023 ? 1023 : k);
        c[a] = k;
        o[a] = (float)(v[a] - k * cs);
    }
    FinePos p;
    p.c = ((uint32_t)c[0] << 20) | ((uint32_t)c[1] << 10) | (uint32_t)c[2];
    p.o0 = o[0]; p.o1 = o[1]; p.o2 = o[2];
    return p;
}

// x_i - x_j, periodic nearest image, in FP32: the cell difference is exact (integers, wrapped
// over the 1024 cells of the box), scaled by the cell side, plus the offset difference.  Holds
// for any pair closer than half a box.
__device__ __forceinline__ void fine_dp(const FinePos &a, const FinePos &b, float cs, float &d0, float &d1, float &d2)
{
    int dc[3];
    for (int k = 0; k < 3; k++) {
        const int s = 20 - 10 * k;
        int d = (int)((a.c >> s) & 1023u) - (int)((b.c >> s) & 1023u);
        if (d > 512) d -= 1024;
        else if (d < -512) d += 1024;
        dc[k] = d;
    }
    d0 = (float)dc[0] * cs + (a.o0 - b.o0);
    d1 = (float)dc[1] * cs + (a.o1 - b.o1);
    d2 = (float)dc[2] * cs + (a.o2 - b.o2);
}

// h is FP64 in both forms (the FP32 form takes the kernel's edge factor from it, pair_kernel)
template <typename R> struct alignas(16) Cell : PosRec<R> {
    double h;
    R m, rho, p, cs, cn, fce, drk;
    R v0, v1, v2;
    R gr0, gr1, gr2, gp0, gp1, gp2;
    R gv[9];
    R t[6];
};

// dp = x_i - x_j (wrapped), r2 = |dp|^2 already known to satisfy the pair predicate.
// The Riemann solve of a face and its two retries (hydro_core_meshless_functions.h:351-427):
// the reconstructed states, then the first-order states with the face velocity at 1.4 times
// the pressure limit, then zero velocities at twice it.  false: no positive pressure.
template <typename T>
__device__ inline bool riemann_ladder(const Gam<T> &G, RState<T> L, RState<T> Rs, const T *n, T ptl, T pI, T pJ, T rhoI, T rhoJ,
                                      T vi0, T vi1, T vi2, T vj0, T vj1, T vj2, ROut<T> &o, unsigned &rung)
{
    o.P_M = o.S_M = o.f0 = o.f1 = o.f2 = o.fp = 0;
    rung |= riemann(G, L, Rs, n, ptl, o);
    if ((o.P_M < 0) || (isnan(o.P_M)) || (o.P_M > T(1.4) * ptl)) {
        Rs.p = pI; L.p = pJ; Rs.rho = rhoI; L.rho = rhoJ;
        Rs.v0 = vi0; Rs.v1 = vi1; Rs.v2 = vi2;
        L.v0 = vj0; L.v1 = vj1; L.v2 = vj2;
        rung |= kRungRetryFirstOrder | riemann(G, L, Rs, n, T(1.4) * ptl, o);
        if ((o.P_M < 0) || (isnan(o.P_M))) {
            Rs.p = pI; L.p = pJ; Rs.rho = rhoI; L.rho = rhoJ;
            Rs.v0 = Rs.v1 = Rs.v2 = 0; L.v0 = L.v1 = L.v2 = 0;
            rung |= kRungRetryZeroVel | riemann(G, L, Rs, n, T(2.0) * ptl, o);
            if ((o.P_M < 0) || (isnan(o.P_M))) { rung |= kRungFail; return false; }
        }
    }
    return true;
}

// The common case of the ladder alone: physical reconstructed states whose wave-speed estimate
// (HLLC, Roe or PVRS) gives a positive pressure within the limiter.  The ladder then takes no
// other rung, so the result is the ladder's, bit for bit.  false: the face needs the rest of the
// ladder (KT, the exact solver or a retry), which a kernel that calls this leaves to a second
// pass over its rows (the rare rungs' code then costs the common pass no registers).
template <typename T>
__device__ __forceinline__ bool riemann_common(const Gam<T> &G, const RState<T> &L, const RState<T> &Rs, const T *n, T ptl,
                                               ROut<T> &o, unsigned &rung)
{
    if ((L.p < 0 && Rs.p < 0) || (L.rho < 0) || (Rs.rho < 0)) return false;
    o.P_M = o.S_M = o.f0 = o.f1 = o.f2 = o.fp = 0;
    const RPrep<T> q = riemann_prep(G, L, Rs, n);
    const unsigned r = wavespeeds(G, L, Rs, n, q.vL, q.vR, q.csL, q.csR, q.hL, q.hR, o, ptl);
    if ((o.P_M < 0) || (isnan(o.P_M)) || (o.P_M > ptl)) return false;
    o.f0 = o.P_M * n[0];
    o.f1 = o.P_M * n[1];
    o.f2 = o.P_M * n[2];
    o.fp = o.P_M * o.S_M;
    rung |= r;
    return true;
}

// A row's force record with the fields of the reconstruction, the faces and the entropic switch
// in LDS (ColdRec, 26 values: 208 B per lane in FP64, three workgroups of 256 per CU) and the
// rest in registers: the record of i then costs a pair loop 12 registers instead of 64.
template <typename R> struct ColdRec {
    R m, p, cn, fce, drk, gr0, gr1, gr2, gp0, gp1, gp2;
    R gv[9];
    R t[6];
};
template <typename R> struct SplitCell {
    R h, rho, cs, v0, v1, v2;
    const R &m, &p, &cn, &fce, &drk, &gr0, &gr1, &gr2, &gp0, &gp1, &gp2;
    const R (&gv)[9];
    const R (&t)[6];
};

// returns 0 skipped, 1 fluxes produced (maybe zero), -1 no Riemann solution (stop 1234), 2 (COMMON
// only) the face needs a rare rung of the ladder; rung receives the RiemannRung bits of the face
// (callers that do not count them leave it unused).  COMMON: the Riemann solve takes only the
// ladder's common case (riemann_common).  IC: the record of i (Cell<R> or SplitCell<R>).
template <int KERN, typename R, bool COMMON = false, typename IC = Cell<R>>
__device__ inline int force_pair(const DevParams &P, const Gam<R> &G, const IC &I, const Cell<R> &J, R dp0, R dp1, R dp2, R r2,
                                 R &F0, R &F1, R &F2, R &Fp, R &vsig_ts, unsigned &rung)
{
    rung = 0;
    F0 = F1 = F2 = Fp = 0;
    vsig_ts = 0;
    if (J.m <= 0) return 0;
    if (J.rho <= 0) return 0;
    const double r2d = pair_r2d<R>(dp0, dp1, dp2, r2);
    if ((r2d >= I.h * I.h) && (r2d >= J.h * J.h)) return 0;
    if (r2 <= 0) return 0;
    const R r = sqrt(r2);
    const R rinv = R(1.0) / r;
    const R dv0 = I.v0 - J.v0, dv1 = I.v1 - J.v1, dv2 = I.v2 - J.v2;
    R vsig = I.cs + J.cs;
    R wk_i = 0, dwk_i = 0, wk_j = 0, dwk_j = 0;
    {
        const double rd = std::is_same<R, double>::value ? 0.0 : sqrt(r2d);
        pair_kernel<KERN, 0, R>(r, rd, I.h, wk_i, dwk_i);
        pair_kernel<KERN, 0, R>(r, rd, J.h, wk_j, dwk_j);
    }
    const R cnumcrit2 = R(kCnDanger) * R(kCnDanger) - I.cn * I.cn;
    const R vdotr2 = dp0 * dv0 + dp1 * dv1 + dp2 * dv2;
    if (vdotr2 < 0) vsig -= R(1.0) * vdotr2 * rinv;
    vsig_ts = vsig + fabs(I.cs - J.cs);
    const R V_i = I.m / I.rho;
    const R V_j = J.m / J.rho;
    const R leak = R(0.5) * (I.fce + J.fce);
    // faces
    R A0, A1, A2, An;
    {
        R wt_i = V_i, wt_j = V_j;
        if ((fabs(V_i - V_j) / dmin(V_i, V_j)) / R(3) > R(1.25))
            wt_i = wt_j = V_i * V_j * (wk_i + wk_j) / (V_i * wk_i + V_j * wk_j);
        R mi0 = I.t[0] * dp0 + I.t[3] * dp1 + I.t[5] * dp2;
        R mi1 = I.t[3] * dp0 + I.t[1] * dp1 + I.t[4] * dp2;
        R mi2 = I.t[5] * dp0 + I.t[4] * dp1 + I.t[2] * dp2;
        R mj0 = J.t[0] * dp0 + J.t[3] * dp1 + J.t[5] * dp2;
        R mj1 = J.t[3] * dp0 + J.t[1] * dp1 + J.t[4] * dp2;
        R mj2 = J.t[5] * dp0 + J.t[4] * dp1 + J.t[2] * dp2;
        R ci = wk_i * wt_i, cj = wk_j * wt_j;
        A0 = (ci * mi0 + cj * mj0) * R(1.0 * 1.0);
        A1 = (ci * mi1 + cj * mj1) * R(1.0 * 1.0);
        A2 = (ci * mi2 + cj * mj2) * R(1.0 * 1.0);
        An = A0 * A0 + A1 * A1 + A2 * A2;
        R fdp = A0 * dp0 + A1 * dp1 + A2 * dp2;
        if ((J.cn * J.cn > R(1.0e12) + cnumcrit2) || (fdp < 0)) {
            R fk = -(wt_i * V_i * dwk_i + wt_j * V_j * dwk_j) / r;
            fk *= R(1.0 * 1.0);
            A0 = dp0 * fk; A1 = dp1 * fk; A2 = dp2 * fk;
            An = fk * fk * r2;
        }
        An = sqrt(An);
    }
    R n[3];
    if constexpr (sizeof(R) != sizeof(double)) {
        // FP32: the area and the normal from the vector scaled by its largest component, so
        // that no square underflows or overflows whatever the unit system (the FP64 form keeps
        // the reference's sqrt of the sum of squares); a face with no area carries no flux
        const R amax = dmax(fabs(A0), dmax(fabs(A1), fabs(A2)));
        if (!(amax > 0)) return 1;
        const R s = R(1.0) / amax;
        const R a0 = A0 * s, a1 = A1 * s, a2 = A2 * s;
        const R len = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
        An = amax * len;
        const R inv = R(1.0) / len;
        n[0] = a0 * inv; n[1] = a1 * inv; n[2] = a2 * inv;
    } else {
        if (An == 0) return 1;
        R inv = R(1.0) / An;
        n[0] = A0 * inv; n[1] = A1 * inv; n[2] = A2 * inv;
    }
    R s_i = R(0.5) * r, s_j = R(-0.5) * r;
    s_i = R(0) - s_i;
    s_j = R(0) - s_j;
    R di[3], dj[3];
    di[0] = dp0 * rinv; di[1] = dp1 * rinv; di[2] = dp2 * rinv;
    for (int k = 0; k < 3; k++) { dj[k] = di[k] * s_j; di[k] *= s_i; }
    const R vf0 = rinv * ((-s_i) * J.v0 + s_j * I.v0);
    const R vf1 = rinv * ((-s_i) * J.v1 + s_j * I.v1);
    const R vf2 = rinv * ((-s_i) * J.v2 + s_j * I.v2);
    R fvi = I.v0 * n[0] + I.v1 * n[1] + I.v2 * n[2];
    R fvj = J.v0 * n[0] + J.v1 * n[1] + J.v2 * n[2];
    fvi /= R(1.0); fvj /= R(1.0);
    const R fadv = rinv * (-s_i * fvj + s_j * fvi);
    R v2a = 0;
    R vdr = vdotr2;
    vdr *= R(1.0) / (r * R(1.0));
    if (vdr < 0) v2a = vdr * vdr;
    R vdf = fvi - fvj;
    if (vdf < 0) v2a = dmax(v2a, vdf * vdf);
    int recon = 1;
    if (leak > R(1)) recon = 0;
    RState<R> L, Rs;
    reconstruct(I.rho, I.gr0, I.gr1, I.gr2, J.rho, J.gr0, J.gr1, J.gr2, di, dj, L.rho, Rs.rho, recon);
    reconstruct(I.p, I.gp0, I.gp1, I.gp2, J.p, J.gp0, J.gp1, J.gp2, di, dj, L.p, Rs.p, recon);
    reconstruct(I.v0 - vf0, I.gv[0], I.gv[1], I.gv[2], J.v0 - vf0, J.gv[0], J.gv[1], J.gv[2], di, dj, L.v0, Rs.v0, recon);
    reconstruct(I.v1 - vf1, I.gv[3], I.gv[4], I.gv[5], J.v1 - vf1, J.gv[3], J.gv[4], J.gv[5], di, dj, L.v1, Rs.v1, recon);
    reconstruct(I.v2 - vf2, I.gv[6], I.gv[7], I.gv[8], J.v2 - vf2, J.gv[6], J.gv[7], J.gv[8], di, dj, L.v2, Rs.v2, recon);
    const R pit = I.p + I.rho * v2a;
    const R pjt = J.p + J.rho * v2a;
    R ptl = R(1.1 * 1.0) * dmax(pit, pjt);
    if (recon == 0) ptl = dmax(ptl, dmax(dmax(I.p, J.p), R(2.) * dmax(I.rho, J.rho) * v2a));
    ROut<R> o;
    rung = kRungFace;
    if constexpr (COMMON) {
        if (!riemann_common<R>(G, L, Rs, n, ptl, o, rung)) return 2;
    } else {
        if (!riemann_ladder<R>(G, L, Rs, n, ptl, I.p, J.p, I.rho, J.rho, I.v0 - vf0, I.v1 - vf1, I.v2 - vf2, J.v0 - vf0, J.v1 - vf1,
                               J.v2 - vf2, o, rung))
            return -1;
    }
    if (!((o.P_M > 0) && (!isnan(o.P_M)))) return 1;
    o.fp += vf0 * o.f0 + vf1 * o.f1 + vf2 * o.f2;
    Fp = An * o.fp;
    F0 = An * o.f0; F1 = An * o.f1; F2 = An * o.f2;
    // face entropic switch (src/hydro/hydro_core_meshless_functions.h:515-573)
    {
        R smc = fabs(o.S_M) / dmin(I.cs, J.cs);
        if ((smc < R(0.5)) || (leak > R(1))) {
            rung |= kEntropicConsidered;
            int use = 1;
            R fpm = o.P_M * An;
            R pdvf = o.P_M * vdr / R(1.0);
            R pdvi = dwk_i * V_i * V_i * I.drk * pdvf;
            R pdvj = dwk_j * V_j * V_j * J.drk * pdvf;
            R du_old = fpm * (o.S_M + fadv);
            R du_new = R(0.5) * (pdvi - pdvj + fpm * (fvi + fvj));
            R cnum2 = J.cn * J.cn;
            R ratio = R(1.001);
            R toi = I.p / I.rho, toj = J.p / J.rho;
            if (use && smc > R(1.e-3) && cnum2 < cnumcrit2) {
                if (toi > ratio * toj) {
                    R dtoj = -du_old + fpm * fvj;
                    if (dtoj > 0) use = 0;
                    else if (dtoj < 0) { if (dtoj > -du_new + fpm * fvj) use = 0; }
                } else if (ratio * toi < toj) {
                    R dtoi = +du_old - fpm * fvi;
                    if (dtoi > 0) use = 0;
                    else if (dtoi < 0) { if (dtoi > +du_new - fpm * fvi) use = 0; }
                } else {
                    R dtoj = -du_old + fpm * fvj;
                    R dtoi = +du_old - fpm * fvi;
                    if (dtoj < 0) { if (dtoj > -du_new + fpm * fvj) use = 0; }
                    if (dtoi < 0) { if (dtoi > +du_new - fpm * fvi) use = 0; }
                }
            }
            if (cnum2 >= cnumcrit2) use = 1;
            if (use) {
                Fp += du_new - du_old;
                rung |= kEntropicApplied;
            }
        }
    }
    return 1;
}

// per-particle record of the gradient pass: the neighbour fields gradient_accumulate_neighbor reads
// the neighbour side of the gradient pass.  The differenced fields stay FP64 in both forms:
// a difference of two O(1) densities or pressures formed in FP32 loses the digits the gradient
// is made of (the mixed form then rounds the FP64 difference to FP32).
template <typename R> struct alignas(16) GradRec : PosRec<R> {
    R h, m;
    double rho, p, v0, v1, v2;
};

}  // namespace gz
