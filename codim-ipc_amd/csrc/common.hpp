// Shared host/device helpers for the C-IPC HIP library: error handling, grow-only
// device buffers, launch helpers and the device-primitive entry points (scan, radix
// sort, deterministic reductions) implemented in prims.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <stdexcept>
#include <vector>

namespace cipc {

struct HipError : std::runtime_error {
    int code;
    HipError(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};

#define CIPC_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw ::cipc::HipError(std::string(#call) + ": " + hipGetErrorString(e_), (int)e_); \
    } while (0)

#define CIPC_LAUNCH_CHECK() CIPC_HIP(hipGetLastError())

constexpr int kBlock = 256;
inline int nblocks(long n, int bs = kBlock) { return (int)((n + bs - 1) / bs); }
// XCD-aware item mapping for per-item kernels whose neighbouring items read neighbouring
// data: workgroups are dispatched to the 8 XCDs round robin, so workgroup b serves logical
// block (b % 8) * per + b / 8 and each XCD's L2 sees one contiguous eighth of the items.
// The grid must be xcd_grid(n) (a multiple of 8).  Only for uniform work per item: where
// the work clusters in space (CCD candidate walks in a contact region) the round-robin
// spread balances the XCDs better (measured: k_ccd_ee 14.9 ms round robin, 22.8 ms
// contiguous; k_query_pt 0.72 ms round robin, 0.65 ms contiguous).
inline int xcd_grid(long n, int bs = kBlock) { return 8 * ((nblocks(n, bs) + 7) / 8); }
__device__ __forceinline__ int xcd_item() {
    const int per = gridDim.x / 8;
    return ((int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8)) * (int)blockDim.x + (int)threadIdx.x;
}

// Grow-only device allocation; contents are not preserved on growth.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    T* ensure(size_t n) {
        if (n <= cap && p) return p;
        if (p) CIPC_HIP(hipFree(p));
        p = nullptr;
        size_t c = n < 64 ? 64 : n + n / 4;
        CIPC_HIP(hipMalloc(&p, c * sizeof(T)));
        cap = c;
        return p;
    }
    T* ensure_exact(size_t n) {
        if (n <= cap && p) return p;
        if (p) CIPC_HIP(hipFree(p));
        p = nullptr;
        size_t c = n < 1 ? 1 : n;
        CIPC_HIP(hipMalloc(&p, c * sizeof(T)));
        cap = c;
        return p;
    }
    operator T*() const { return p; }
};

template <class T>
inline void upload(DevBuf<T>& b, const T* h, size_t n, hipStream_t s) {
    b.ensure_exact(n);
    if (n) CIPC_HIP(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <class T>
inline void upload(DevBuf<T>& b, const std::vector<T>& h, hipStream_t s) { upload(b, h.data(), h.size(), s); }

// Scratch memory shared by the primitives.
struct PrimWork {
    DevBuf<uint32_t> scanBlocks, scanBlocks2, hist, histScan;
    DevBuf<uint64_t> keyAlt;
    DevBuf<uint32_t> valAlt;
    DevBuf<double> redPartial;
    DevBuf<uint32_t> u32tmp;
};

// exclusive prefix sum of n uint32 values; total (if not null) written to device memory
void scan_exclusive(const uint32_t* in, uint32_t* out, long n, uint32_t* total, PrimWork& w, hipStream_t s);
// stable LSD radix sort of (key, value) pairs on the low `bits` bits of the keys
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, long n, int bits, PrimWork& w, hipStream_t s);
// deterministic sums: chunks of 256 values folded by halving strides, chunk partials
// added in chunk order by one thread; result written to device memory `out`
void chunk_sum(const double* in, long n, double* out, PrimWork& w, hipStream_t s);
// sum of per-block partials already in device memory (fixed order)
void sum_partials(const double* partials, int n, double* out, hipStream_t s);

// a slot in an output list for every active lane of the wave: the active lanes are found
// by a ballot, the lowest adds their count to the counter, and each lane takes base + its
// rank among them (one atomic per wave; the slot order is arbitrary)
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* cnt) {
    const uint64_t act = __ballot(1);
    const int lane = (int)__lane_id();
    const int leader = __ffsll((unsigned long long)act) - 1;
    const uint32_t rank = (uint32_t)__popcll(act & ((1ull << lane) - 1ull));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(cnt, (uint32_t)__popcll(act));
    base = (uint32_t)__shfl((int)base, leader, 64);
    return base + rank;
}

// device reductions inside kernels
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// block-wide sum for blockDim = 256; result valid in thread 0.  Fixed association order.
__device__ __forceinline__ double block_sum_256(double v, double* lds /*[4]*/) {
    v = wave_sum(v);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) r = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    __syncthreads();
    return r;
}
__device__ __forceinline__ double block_min_256(double v, double* lds) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { double u = __shfl_xor(v, o, 64); v = u < v ? u : v; }
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) {
        r = lds[0];
        for (int i = 1; i < 4; ++i) r = lds[i] < r ? lds[i] : r;
    }
    __syncthreads();
    return r;
}

}  // namespace cipc
