// Small device utilities: index fill, bounding box.
#include "ctx.hpp"

namespace cipc {

__global__ void k_iota(uint32_t* __restrict__ a, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (uint32_t)i;
}
void fill_iota(uint32_t* a, long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_iota, dim3(nblocks(n)), dim3(kBlock), 0, s, a, n);
    CIPC_LAUNCH_CHECK();
}

__global__ void k_bbox_partial(const double* __restrict__ x, int nV, double* __restrict__ part) {
    __shared__ double lo[3][kBlock], hi[3][kBlock];
    int t = threadIdx.x;
    double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
    for (int v = blockIdx.x * blockDim.x + t; v < nV; v += gridDim.x * blockDim.x)
#pragma unroll
        for (int d = 0; d < 3; ++d) { double c = x[3 * v + d]; l[d] = c < l[d] ? c : l[d]; h[d] = c > h[d] ? c : h[d]; }
#pragma unroll
    for (int d = 0; d < 3; ++d) { lo[d][t] = l[d]; hi[d][t] = h[d]; }
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
        if (t < st)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                lo[d][t] = lo[d][t + st] < lo[d][t] ? lo[d][t + st] : lo[d][t];
                hi[d][t] = hi[d][t + st] > hi[d][t] ? hi[d][t + st] : hi[d][t];
            }
        __syncthreads();
    }
    if (t == 0)
        for (int d = 0; d < 3; ++d) { part[6 * blockIdx.x + d] = lo[d][0]; part[6 * blockIdx.x + 3 + d] = hi[d][0]; }
}
// min/max over nb partial records (lo xyz, hi xyz); min and max are exact, so the
// parallel order gives the serial result
__global__ void k_minmax6_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
    __shared__ double lo[3][kBlock], hi[3][kBlock];
    const int t = threadIdx.x;
    double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
    for (int b = t; b < nb; b += kBlock)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double a = part[6 * b + d], c = part[6 * b + 3 + d];
            l[d] = a < l[d] ? a : l[d];
            h[d] = c > h[d] ? c : h[d];
        }
#pragma unroll
    for (int d = 0; d < 3; ++d) { lo[d][t] = l[d]; hi[d][t] = h[d]; }
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
        if (t < st)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                lo[d][t] = lo[d][t + st] < lo[d][t] ? lo[d][t + st] : lo[d][t];
                hi[d][t] = hi[d][t + st] > hi[d][t] ? hi[d][t + st] : hi[d][t];
            }
        __syncthreads();
    }
    if (t == 0)
        for (int d = 0; d < 3; ++d) { out[d] = lo[d][0]; out[3 + d] = hi[d][0]; }
}
void minmax6_final(const double* part, int nb, double* out6, hipStream_t s) {
    hipLaunchKernelGGL(k_minmax6_final, dim3(1), dim3(kBlock), 0, s, part, nb, out6);
    CIPC_LAUNCH_CHECK();
}
void bbox_minmax(Ctx& c, const double* x, double* out6) {
    int nb = std::min(nblocks(c.nV), 1024);
    double* part = c.dscr2.ensure(6 * nb);
    hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(kBlock), 0, c.s, x, c.nV, part);
    CIPC_LAUNCH_CHECK();
    minmax6_final(part, nb, out6, c.s);
}

}  // namespace cipc
