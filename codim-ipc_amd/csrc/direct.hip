// Direct solve of the Newton system for small scenes: the reference factorises every
// Newton matrix with CHOLMOD (Math/DIRECT_SOLVER.h:14-66, stype -1) and reports a
// non-positive pivot as CHOLMOD_NOT_POSDEF, after which the step takes a gradient-descent
// direction (IMPLICIT_EULER.h:515-554).  Here: the assembled FP64 matrix (elastic BSR +
// contact BSR, Dirichlet rows already projected) is expanded into a dense row-major array
// padded to whole 64 x 64 tiles (identity on the padding) and factorised by a tiled
// right-looking Cholesky -- one workgroup per tile for the diagonal factor, the panel
// solves and the trailing updates -- then solved by tiled forward and backward
// substitution.  Dense, so limited to kDirectMaxDof unknowns (4.5 GiB of FP64 matrix); it
// is the fallback of the PCG solves for such scenes and a selectable solver.
#include "ctx.hpp"
#include <cmath>

namespace cipc {

constexpr int kDT = 64;          // tile edge
constexpr int kDirectBlock = 256;

__global__ void k_dense_zero(double* __restrict__ a, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = 0.0;
}
__global__ void k_dense_pad_identity(double* __restrict__ a, long ld, int n, int npad) {
    const int i = n + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npad) a[(long)i * ld + i] = 1.0;
}
// one thread per stored block: A[3 r + a][3 c + b] (+)= block; the BSR holds each (r, c)
// once, so a pattern's blocks never collide (the contact BSR is added after the elastic)
__global__ void k_dense_scatter(const int* __restrict__ rowOf, const int* __restrict__ col, const double* __restrict__ val, int nnzb,
                                double* __restrict__ a, long ld, int add) {
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= nnzb) return;
    const long r = rowOf[blk], c = col[blk];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double* dst = a + (3 * r + i) * ld + 3 * c + j;
            const double v = val[(long)(3 * i + j) * nnzb + blk];
            *dst = add ? *dst + v : v;
        }
}

// Cholesky of diagonal tile k in LDS (lower, in place); a non-positive pivot sets *bad and
// continues with pivot 1 so that the factorisation stays finite
__global__ __launch_bounds__(kDirectBlock) void k_potrf_tile(double* __restrict__ a, long ld, int k, int* __restrict__ bad) {
    __shared__ double T[kDT][kDT + 1];
    const long o = (long)k * kDT;
    for (int e = threadIdx.x; e < kDT * kDT; e += kDirectBlock) {
        const int i = e / kDT, j = e % kDT;
        T[i][j] = a[(o + i) * ld + o + j];
    }
    __syncthreads();
    for (int j = 0; j < kDT; ++j) {
        if (threadIdx.x == 0) {
            double d = T[j][j];
            if (!(d > 0)) {
                atomicOr(bad, 1);
                d = 1.0;
            }
            T[j][j] = sqrt(d);
        }
        __syncthreads();
        const double djj = T[j][j];
        for (int i = j + 1 + threadIdx.x; i < kDT; i += kDirectBlock) T[i][j] /= djj;
        __syncthreads();
        // trailing lower part of the tile
        const int m = kDT - j - 1;
        for (int e = threadIdx.x; e < m * m; e += kDirectBlock) {
            const int i = j + 1 + e / m, l = j + 1 + e % m;
            if (l <= i) T[i][l] -= T[i][j] * T[l][j];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < kDT * kDT; e += kDirectBlock) {
        const int i = e / kDT, j = e % kDT;
        a[(o + i) * ld + o + j] = j <= i ? T[i][j] : 0.0;
    }
}

// panel: tiles (i, k), i > k, become A_ik L_kk^-T (each row solved forward against L_kk)
__global__ __launch_bounds__(kDirectBlock) void k_trsm_tiles(double* __restrict__ a, long ld, int k) {
    __shared__ double L[kDT][kDT + 1];
    __shared__ double X[kDT][kDT + 1];
    const int i = k + 1 + blockIdx.x;
    const long ok = (long)k * kDT, oi = (long)i * kDT;
    for (int e = threadIdx.x; e < kDT * kDT; e += kDirectBlock) {
        const int r = e / kDT, c = e % kDT;
        L[r][c] = a[(ok + r) * ld + ok + c];
        X[r][c] = a[(oi + r) * ld + ok + c];
    }
    __syncthreads();
    // 4 threads per row would race on the running sum; one thread per row, 64 rows
    if (threadIdx.x < kDT) {
        const int r = threadIdx.x;
        for (int c = 0; c < kDT; ++c) {
            double s = X[r][c];
            for (int m = 0; m < c; ++m) s -= X[r][m] * L[c][m];
            X[r][c] = s / L[c][c];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kDT * kDT; e += kDirectBlock) {
        const int r = e / kDT, c = e % kDT;
        a[(oi + r) * ld + ok + c] = X[r][c];
    }
}

// trailing update of the lower tiles (i, j), k < j <= i: A_ij -= L_ik L_jk^T; block b
// enumerates the lower triangle of the trailing tile matrix row by row
__global__ __launch_bounds__(kDirectBlock) void k_syrk_tiles(double* __restrict__ a, long ld, int k, int nt) {
    __shared__ double P[kDT][kDT + 1];
    __shared__ double Q[kDT][kDT + 1];
    const int m = nt - k - 1;
    int b = blockIdx.x, ii = 0;
    while (b >= ii + 1) { b -= ii + 1; ++ii; }   // row ii of the trailing triangle holds ii + 1 tiles
    if (ii >= m) return;
    const int i = k + 1 + ii, j = k + 1 + b;
    const long ok = (long)k * kDT, oi = (long)i * kDT, oj = (long)j * kDT;
    for (int e = threadIdx.x; e < kDT * kDT; e += kDirectBlock) {
        const int r = e / kDT, c = e % kDT;
        P[r][c] = a[(oi + r) * ld + ok + c];
        Q[r][c] = a[(oj + r) * ld + ok + c];
    }
    __syncthreads();
    // 256 threads, 16 outputs each: thread t covers row t / 4 and columns (t % 4) + 4 s
    const int r = threadIdx.x / 4, c0 = threadIdx.x % 4;
    double acc[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) acc[s] = 0.0;
    for (int l = 0; l < kDT; ++l) {
        const double pv = P[r][l];
#pragma unroll
        for (int s = 0; s < 16; ++s) acc[s] += pv * Q[c0 + 4 * s][l];
    }
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int c = c0 + 4 * s;
        a[(oi + r) * ld + oj + c] -= acc[s];
    }
}

// forward substitution, tile k: y_k = L_kk^-1 b_k (b_k already holds the updates of the
// earlier tiles); backward: x_k = L_kk^-T y_k
__global__ void k_trsv_tile(const double* __restrict__ a, long ld, int k, double* __restrict__ y, int transposed) {
    __shared__ double L[kDT][kDT + 1];
    __shared__ double v[kDT];
    const long ok = (long)k * kDT;
    for (int e = threadIdx.x; e < kDT * kDT; e += blockDim.x) {
        const int r = e / kDT, c = e % kDT;
        L[r][c] = a[(ok + r) * ld + ok + c];
    }
    for (int e = threadIdx.x; e < kDT; e += blockDim.x) v[e] = y[ok + e];
    __syncthreads();
    if (threadIdx.x == 0) {
        if (!transposed) {
            for (int r = 0; r < kDT; ++r) {
                double s = v[r];
                for (int m = 0; m < r; ++m) s -= L[r][m] * v[m];
                v[r] = s / L[r][r];
            }
        } else {
            for (int r = kDT - 1; r >= 0; --r) {
                double s = v[r];
                for (int m = r + 1; m < kDT; ++m) s -= L[m][r] * v[m];
                v[r] = s / L[r][r];
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kDT; e += blockDim.x) y[ok + e] = v[e];
}
// forward: y_i -= L_ik y_k for tiles i > k (block per tile); backward: y_i -= L_ki^T y_k
// for tiles i < k
__global__ void k_gemv_tiles(const double* __restrict__ a, long ld, int k, double* __restrict__ y, int transposed) {
    __shared__ double yk[kDT];
    const int i = transposed ? blockIdx.x : k + 1 + blockIdx.x;
    const long ok = (long)k * kDT, oi = (long)i * kDT;
    for (int e = threadIdx.x; e < kDT; e += blockDim.x) yk[e] = y[ok + e];
    __syncthreads();
    if (threadIdx.x < kDT) {
        const int r = threadIdx.x;
        double s = 0;
        if (!transposed)
            for (int c = 0; c < kDT; ++c) s += a[(oi + r) * ld + ok + c] * yk[c];
        else
            for (int c = 0; c < kDT; ++c) s += a[(ok + c) * ld + oi + r] * yk[c];
        y[oi + r] -= s;
    }
}

int direct_solve(Ctx& c, const double* b, double* x) {
    const int n = 3 * c.nV;
    if (n > kDirectMaxDof) return 3;
    const int nt = (n + kDT - 1) / kDT, npad = nt * kDT;
    const long ld = npad;
    double* A = c.denseA.ensure_exact((size_t)npad * npad);
    hipLaunchKernelGGL(k_dense_zero, dim3(2048), dim3(kBlock), 0, c.s, A, (long)npad * npad);
    hipLaunchKernelGGL(k_dense_pad_identity, dim3(1), dim3(kDT), 0, c.s, A, ld, n, npad);
    CIPC_LAUNCH_CHECK();
    BSR* parts[2] = {&c.Ael, &c.Acon};
    for (int q = 0; q < 2; ++q) {
        BSR& M = *parts[q];
        if (M.nnzb == 0) continue;
        hipLaunchKernelGGL(k_dense_scatter, dim3(nblocks(M.nnzb)), dim3(kBlock), 0, c.s, M.rowOf.p, M.col.p, M.val.p, M.nnzb, A, ld, q);
        CIPC_LAUNCH_CHECK();
    }
    int* bad = c.denseBad.ensure(1);
    CIPC_HIP(hipMemsetAsync(bad, 0, sizeof(int), c.s));
    for (int k = 0; k < nt; ++k) {
        hipLaunchKernelGGL(k_potrf_tile, dim3(1), dim3(kDirectBlock), 0, c.s, A, ld, k, bad);
        const int m = nt - k - 1;
        if (m > 0) {
            hipLaunchKernelGGL(k_trsm_tiles, dim3(m), dim3(kDirectBlock), 0, c.s, A, ld, k);
            hipLaunchKernelGGL(k_syrk_tiles, dim3(m * (m + 1) / 2), dim3(kDirectBlock), 0, c.s, A, ld, k, nt);
        }
        CIPC_LAUNCH_CHECK();
    }
    double* y = c.denseY.ensure(npad);
    CIPC_HIP(hipMemsetAsync(y, 0, npad * sizeof(double), c.s));
    CIPC_HIP(hipMemcpyAsync(y, b, n * sizeof(double), hipMemcpyDeviceToDevice, c.s));
    for (int k = 0; k < nt; ++k) {
        hipLaunchKernelGGL(k_trsv_tile, dim3(1), dim3(kDT), 0, c.s, A, ld, k, y, 0);
        if (nt - k - 1 > 0) hipLaunchKernelGGL(k_gemv_tiles, dim3(nt - k - 1), dim3(kDT), 0, c.s, A, ld, k, y, 0);
    }
    for (int k = nt - 1; k >= 0; --k) {
        hipLaunchKernelGGL(k_trsv_tile, dim3(1), dim3(kDT), 0, c.s, A, ld, k, y, 1);
        if (k > 0) hipLaunchKernelGGL(k_gemv_tiles, dim3(k), dim3(kDT), 0, c.s, A, ld, k, y, 1);
    }
    CIPC_LAUNCH_CHECK();
    CIPC_HIP(hipMemcpyAsync(x, y, n * sizeof(double), hipMemcpyDeviceToDevice, c.s));
    CIPC_HIP(hipMemcpyAsync(c.hCnt + 48, bad, sizeof(int), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    c.directSolves++;
    return c.hCnt[48] ? 1 : 0;
}

}  // namespace cipc
