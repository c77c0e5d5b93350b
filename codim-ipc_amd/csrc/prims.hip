// Device primitives: exclusive scan, stable LSD radix sort of (key, value) pairs and
// deterministic reductions.  Written for 64-wide wavefronts: within-wave ranking uses
// 64-bit ballots, block-level structures are 256 threads (four waves).
#include "common.hpp"

namespace cipc {

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = kBlock * SCAN_ITEMS;   // 2048 values per block

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}
// exclusive scan of one value per thread over a 256-thread block; returns block total
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t& excl, uint32_t* lds /*[4]*/) {
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan_u32(v);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    uint32_t base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint32_t t = lds[w];
        if (w < wid) base += t;
        total += t;
    }
    __syncthreads();
    excl = base + inc - v;
    return total;
}

__global__ void k_scan_block_sums(const uint32_t* __restrict__ in, long n, uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds[4];
    long base = (long)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        long idx = base + (long)i * kBlock + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    uint32_t excl;
    uint32_t tot = block_excl_scan_u32(s, excl, lds);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// single-block exclusive scan of up to any count (serial chunks per thread)
__global__ void k_scan_single(uint32_t* __restrict__ a, long n, uint32_t* __restrict__ total) {
    __shared__ uint32_t lds[4];
    long per = (n + kBlock - 1) / kBlock;
    long b = threadIdx.x * per, e = b + per < n ? b + per : n;
    uint32_t s = 0;
    for (long i = b; i < e; ++i) s += a[i];
    uint32_t excl;
    uint32_t tot = block_excl_scan_u32(s, excl, lds);
    uint32_t run = excl;
    for (long i = b; i < e; ++i) { uint32_t v = a[i]; a[i] = run; run += v; }
    if (threadIdx.x == 0 && total) *total = tot;
}

__global__ void k_scan_apply(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, long n,
                             const uint32_t* __restrict__ blockOff) {
    __shared__ uint32_t lds[4];
    long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        long idx = base + i;
        v[i] = idx < n ? in[idx] : 0;
        s += v[i];
    }
    uint32_t excl;
    block_excl_scan_u32(s, excl, lds);
    uint32_t run = excl + blockOff[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        long idx = base + i;
        if (idx < n) out[idx] = run;
        run += v[i];
    }
}

void scan_exclusive(const uint32_t* in, uint32_t* out, long n, uint32_t* total, PrimWork& w, hipStream_t s) {
    if (n <= 0) {
        if (total) CIPC_HIP(hipMemsetAsync(total, 0, sizeof(uint32_t), s));
        return;
    }
    int nb = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
    uint32_t* sums = w.scanBlocks.ensure(nb);
    hipLaunchKernelGGL(k_scan_block_sums, dim3(nb), dim3(kBlock), 0, s, in, n, sums);
    CIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kBlock), 0, s, sums, (long)nb, total);
    CIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kBlock), 0, s, in, out, n, sums);
    CIPC_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ radix sort
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = kBlock * RS_ITEMS;

__global__ void k_rs_hist(const uint64_t* __restrict__ keys, long n, int shift, uint32_t* __restrict__ hist, int nb) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    long base = (long)blockIdx.x * RS_TILE;
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        long idx = base + (long)i * kBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(long)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ void k_rs_scatter(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                             uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, long n, int shift,
                             const uint32_t* __restrict__ histScan, int nb) {
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    gbase[tid] = histScan[(long)tid * nb + blockIdx.x];
    run[tid] = 0;
    const uint64_t ltmask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    long base = (long)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; ++r) {
        long idx = base + (long)r * kBlock + tid;
        bool valid = idx < n;
        uint64_t key = valid ? kin[idx] : 0;
        uint32_t val = valid ? vin[idx] : 0;
        uint32_t dig = (uint32_t)((key >> shift) & 255u);
#pragma unroll
        for (int k = 0; k < 4; ++k) wcnt[k][tid] = 0;
        __syncthreads();
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t m = __ballot((dig >> b) & 1u);
            peers &= ((dig >> b) & 1u) ? m : ~m;
        }
        uint32_t rankInWave = (uint32_t)__popcll(peers & ltmask);
        if (valid && rankInWave == 0) wcnt[wid][dig] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = run[dig];
            for (int k = 0; k < wid; ++k) off += wcnt[k][dig];
            uint32_t dst = gbase[dig] + off + rankInWave;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
        run[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        __syncthreads();
    }
}

void radix_sort_pairs(uint64_t* keys, uint32_t* vals, long n, int bits, PrimWork& w, hipStream_t s) {
    if (n <= 1) return;
    int nb = (int)((n + RS_TILE - 1) / RS_TILE);
    uint64_t* k2 = w.keyAlt.ensure(n);
    uint32_t* v2 = w.valAlt.ensure(n);
    uint32_t* hist = w.hist.ensure((size_t)256 * nb);
    uint32_t* hs = w.histScan.ensure((size_t)256 * nb);
    uint64_t *ka = keys, *kb = k2;
    uint32_t *va = vals, *vb = v2;
    int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(kBlock), 0, s, ka, n, shift, hist, nb);
        CIPC_LAUNCH_CHECK();
        scan_exclusive(hist, hs, (long)256 * nb, nullptr, w, s);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(kBlock), 0, s, ka, va, kb, vb, n, shift, hs, nb);
        CIPC_LAUNCH_CHECK();
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        CIPC_HIP(hipMemcpyAsync(keys, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        CIPC_HIP(hipMemcpyAsync(vals, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
}

// ------------------------------------------------------------------ reductions
__global__ void k_chunk_sum(const double* __restrict__ in, long n, double* __restrict__ partial) {
    __shared__ double a[kBlock];
    long idx = (long)blockIdx.x * kBlock + threadIdx.x;
    a[threadIdx.x] = idx < n ? in[idx] : 0.0;
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) a[threadIdx.x] += a[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = a[0];
}
// Sequential sum in index order (the order the deterministic reductions define); the
// partials are staged through LDS by the whole workgroup so the single summing thread
// reads LDS instead of waiting on one global load per term.
__global__ void k_serial_sum(const double* __restrict__ partial, int n, double* __restrict__ out) {
    constexpr int kTile = 4096;
    __shared__ double buf[kTile];
    double t = 0;
    for (int base = 0; base < n; base += kTile) {
        const int m = n - base < kTile ? n - base : kTile;
        for (int i = threadIdx.x; i < m; i += blockDim.x) buf[i] = partial[base + i];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int i = 0; i < m; ++i) t += buf[i];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = t;
}
void chunk_sum(const double* in, long n, double* out, PrimWork& w, hipStream_t s) {
    int nb = n > 0 ? nblocks(n) : 0;
    double* part = w.redPartial.ensure(nb > 0 ? nb : 1);
    if (nb > 0) {
        hipLaunchKernelGGL(k_chunk_sum, dim3(nb), dim3(kBlock), 0, s, in, n, part);
        CIPC_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_serial_sum, dim3(1), dim3(kBlock), 0, s, part, nb, out);
    CIPC_LAUNCH_CHECK();
}
void sum_partials(const double* partials, int n, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_serial_sum, dim3(1), dim3(kBlock), 0, s, partials, n, out);
    CIPC_LAUNCH_CHECK();
}

}  // namespace cipc
