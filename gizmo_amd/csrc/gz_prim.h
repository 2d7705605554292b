// Device-wide sort, scan and reduce of the library, on rocPRIM directly.  Each call follows
// rocPRIM's two-phase form: with tmp == nullptr it only sets bytes to the temporary storage
// it needs; otherwise it launches on stream s.
#pragma once
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/functional.hpp>

#include <limits>

// stable LSD radix sort of (key, value) pairs on key bits [b0, b1)
template <typename K, typename V>
inline hipError_t gz_sort_pairs(void *tmp, size_t &bytes, const K *kin, K *kout, const V *vin, V *vout, int n, int b0, int b1,
                                hipStream_t s = 0)
{
    return rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, (unsigned)n, (unsigned)b0, (unsigned)b1, s);
}

// out[k] = in[0] + ... + in[k-1], out[0] = 0
template <typename T>
inline hipError_t gz_exclusive_sum(void *tmp, size_t &bytes, const T *in, T *out, int n, hipStream_t s = 0)
{
    return rocprim::exclusive_scan(tmp, bytes, in, out, T(0), (size_t)n, rocprim::plus<T>(), s);
}

// *out = max (min) of in[0..n-1]
template <typename T>
inline hipError_t gz_reduce_max(void *tmp, size_t &bytes, const T *in, T *out, int n, hipStream_t s = 0)
{
    return rocprim::reduce(tmp, bytes, in, out, std::numeric_limits<T>::lowest(), (size_t)n, rocprim::maximum<T>(), s);
}

template <typename T>
inline hipError_t gz_reduce_min(void *tmp, size_t &bytes, const T *in, T *out, int n, hipStream_t s = 0)
{
    return rocprim::reduce(tmp, bytes, in, out, std::numeric_limits<T>::max(), (size_t)n, rocprim::minimum<T>(), s);
}
