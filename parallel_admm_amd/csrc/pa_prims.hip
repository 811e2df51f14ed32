// Scan and radix sort for gfx950 (wave64).
#include "pa_prims.hpp"

namespace pa {
namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// Block-wide exclusive scan of one value per thread (256 threads = 4 waves).
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh_wave, uint32_t* block_total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // inclusive scan inside the wave
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh_wave[wave] = x;
    __syncthreads();
    uint32_t base = 0;
    for (int w = 0; w < wave; w++) base += sh_wave[w];
    if (block_total) *block_total = sh_wave[0] + sh_wave[1] + sh_wave[2] + sh_wave[3];
    return base + x - v;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                          int64_t n, uint32_t* __restrict__ tile_sums) {
    __shared__ uint32_t sh_wave[4];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        sum += v[k];
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(sum, sh_wave, &tot);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ void k_scan_add(uint32_t* __restrict__ out, int64_t n, const uint32_t* __restrict__ tile_offsets) {
    const int64_t i = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
    const uint32_t add = tile_offsets[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int64_t e = i + (int64_t)k * SCAN_BLOCK;
        if (e < n) out[e] += add;
    }
}

__global__ void k_scan_total(const uint32_t* __restrict__ in, const uint32_t* __restrict__ out, int64_t n,
                             uint32_t* __restrict__ total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *total = n > 0 ? out[n - 1] + in[n - 1] : 0u;
}

// Histogram of one 8-bit digit per tile; hist is digit-major [256][ntiles].
constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;

__global__ void __launch_bounds__(RS_BLOCK) k_rs_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                     uint32_t* __restrict__ hist, int ntiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        int64_t e = base + (int64_t)k * RS_BLOCK + threadIdx.x;
        if (e < n) atomicAdd(&h[(keys[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: element order inside a tile is (round, wave, lane); the
// rank among equal digits comes from 8 wave ballots plus per-wave digit
// counts kept in LDS.
__global__ void __launch_bounds__(RS_BLOCK) k_rs_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                        uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int64_t n,
                                                        int shift, const uint32_t* __restrict__ offs, int ntiles) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    run[tid] = offs[(int64_t)tid * ntiles + blockIdx.x];
    wcnt[0][tid] = wcnt[1][tid] = wcnt[2][tid] = wcnt[3][tid] = 0;
    __syncthreads();
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int k = 0; k < RS_ITEMS; k++) {
        const int64_t e = base + (int64_t)k * RS_BLOCK + tid;
        const bool valid = e < n;
        uint32_t key = valid ? kin[e] : 0u;
        uint32_t val = valid ? vin[e] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            match &= bit ? m : ~m;
        }
        const uint32_t lrank = (uint32_t)__popcll(match & lt_mask);
        if (valid && lrank == 0) wcnt[wave][d] = (uint32_t)__popcll(match);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + lrank;
            for (int w = 0; w < wave; w++) pos += wcnt[w][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        run[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        wcnt[0][tid] = wcnt[1][tid] = wcnt[2][tid] = wcnt[3][tid] = 0;
        __syncthreads();
    }
}

void scan_rec(const uint32_t* in, uint32_t* out, int64_t n, PrimScratch& s, int depth, hipStream_t st) {
    if (n <= 0) return;
    const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    DBuf<uint32_t>& sums = depth == 0 ? s.scan_a : (depth == 1 ? s.scan_b : s.scan_c);
    PA_CHECK(depth < 3, PA_ERR_INVALID, "scan too large");
    sums.reserve((size_t)tiles + 1);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, st, in, out, n, sums.p);
    PA_LAUNCH_CHECK();
    if (tiles > 1) {
        scan_rec(sums.p, sums.p, tiles, s, depth + 1, st);
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, st, out, n, sums.p);
        PA_LAUNCH_CHECK();
    } else {
        // single tile: its offset is zero already
    }
}

} // namespace

void scan_exclusive(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, PrimScratch& s, hipStream_t st) {
    if (n <= 0) {
        if (total) fill<uint32_t>(total, 0u, 1, st);
        return;
    }
    if (total) {
        // The total needs the last input value, which `out` may overwrite when it aliases `in`.
        DBuf<uint32_t>& last = s.last;
        last.reserve(1);
        PA_HIP(hipMemcpyAsync(last.p, in + n - 1, sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        scan_rec(in, out, n, s, 0, st);
        hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, st, last.p, out + n - 1, (int64_t)1, total);
        PA_LAUNCH_CHECK();
    } else {
        scan_rec(in, out, n, s, 0, st);
    }
}

void radix_sort_pairs(uint32_t* keys, uint32_t* vals, int64_t n, int bits, PrimScratch& s, hipStream_t st) {
    if (n <= 1) return;
    const int ntiles = (int)((n + RS_TILE - 1) / RS_TILE);
    s.keys_alt.reserve((size_t)n);
    s.vals_alt.reserve((size_t)n);
    s.hist.reserve((size_t)ntiles * 256);
    DBuf<uint32_t>& hist = s.hist;
    uint32_t *ka = keys, *va = vals, *kb = s.keys_alt.p, *vb = s.vals_alt.p;
    int passes = (bits + 7) / 8;
    for (int pss = 0; pss < passes; pss++) {
        const int shift = 8 * pss;
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_BLOCK), 0, st, ka, n, shift, hist.p, ntiles);
        PA_LAUNCH_CHECK();
        scan_rec(hist.p, hist.p, (int64_t)ntiles * 256, s, 0, st);
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_BLOCK), 0, st, ka, va, kb, vb, n, shift, hist.p, ntiles);
        PA_LAUNCH_CHECK();
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        PA_HIP(hipMemcpyAsync(keys, ka, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
        PA_HIP(hipMemcpyAsync(vals, va, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    }
}

} // namespace pa
