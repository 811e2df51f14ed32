// Device primitives shared by the detector and the solver: fill, exclusive
// scan and a stable least-significant-digit radix sort of 32-bit keys with a
// 32-bit payload (8-bit digits, wave64 ballot ranking in LDS).
#pragma once

#include "pa_common.hpp"

namespace pa {

template <class T> __global__ void k_fill(T* __restrict__ p, T v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Kernel fill instead of hipMemsetAsync so the same code can sit in a captured graph.
template <class T> inline void fill(T* p, T v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
    PA_LAUNCH_CHECK();
}

// Scratch for scan/sort; grown on demand.
struct PrimScratch {
    DBuf<uint32_t> scan_a, scan_b, scan_c;
    DBuf<uint32_t> hist, keys_alt, vals_alt, last;
};

// out[i] = sum_{j<i} in[j]; if total != nullptr, *total (device) = sum of all.
// in and out may alias.
void scan_exclusive(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, PrimScratch& s, hipStream_t st);

// Sorts (keys, vals) by the low `bits` bits of keys, stably.  Result in keys/vals.
void radix_sort_pairs(uint32_t* keys, uint32_t* vals, int64_t n, int bits, PrimScratch& s, hipStream_t st);

} // namespace pa
