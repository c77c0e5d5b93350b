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

// Cholesky of diagonal tile k (lower, in place) by four waves: wave h keeps columns
// [16h, 16h+16) of every row in registers, so the trailing update of a column is 16
// register FMAs per lane against a broadcast copy of the column in LDS.  Per column two
// barriers: after the column's entries are published, and after its multipliers are
// (both LDS buffers alternate between columns, so no third is needed).  The entry (i, j)
// of the column being eliminated is published by the wave that holds it (picked from
// registers by a select tree on the bits of j); every lane derives the pivot and its
// row's multiplier from it with the same operations, so all lanes hold the same bits.
// A non-positive pivot sets *bad and continues with pivot 1 so that the factorisation
// stays finite.  With diag0, a pivot at or below tau x diag0 (the unknown's diagonal
// before elimination: a mode the matrix does not see) is dropped instead: a pivot at the
// unknown's own scale, zero column, dropped[] = 1.  With inv, the inverse of the factored
// tile is written there too.  Every sum runs in a fixed order: trailing updates in
// column order, substitution sums in index order.
static_assert(kDirectBlock == 4 * kDT, "the tile factor runs four waves of one lane per row");
__global__ __launch_bounds__(kDirectBlock) void k_potrf_tile(double* __restrict__ a, long ld, int k, int* __restrict__ bad,
                                                             const double* __restrict__ diag0, int* __restrict__ dropped, double tau,
                                                             double* __restrict__ inv) {
    __shared__ double T[kDT][kDT + 1];
    __shared__ double colj[2][kDT];
    __shared__ double cur[2][kDT];
    __shared__ double dg0[kDT];
    const int tid = threadIdx.x, i = tid & 63, h = tid >> 6;
    const long o = (long)k * kDT;
    if (h == 0) dg0[i] = diag0 ? diag0[o + i] : 0.0;
    for (int r = h; r < kDT; r += 4) T[r][i] = a[(o + r) * ld + o + i];
    __syncthreads();
    double row[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) row[q] = T[i][16 * h + q];
    for (int j = 0; j < kDT; ++j) {
        const int b = j & 1;
        if (h == (j >> 4)) {
            double t4[4];
            const bool s8 = j & 8, s4 = j & 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) t4[q] = s8 ? (s4 ? row[q + 12] : row[q + 8]) : (s4 ? row[q + 4] : row[q]);
            const bool s2 = j & 2, s1 = j & 1;
            cur[b][i] = s2 ? (s1 ? t4[3] : t4[2]) : (s1 ? t4[1] : t4[0]);
        }
        __syncthreads();
        double d = cur[b][j];
        const double rij = cur[b][i];
        int dr = 0;
        bool neg = false;
        if (diag0) {
            const double d0 = dg0[j];
            if (!(d > tau * d0)) {
                dr = 1;
                d = d0 > 0 ? d0 : 1.0;
            }
        } else if (!(d > 0)) {
            neg = true;
            d = 1.0;
        }
        if (tid == j) {
            if (diag0) dropped[o + j] = dr;
            else if (neg) atomicOr(bad, 1);
        }
        const double pv = sqrt(d);
        double lij;
        if (i == j) lij = pv;
        else if (i > j) lij = dr ? 0.0 : rij / pv;
        else lij = 0.0;
        if (h == 0) {
            T[i][j] = lij;
            colj[b][i] = i > j ? lij : 0.0;
        }
        __syncthreads();
        if (i > j) {
#pragma unroll
            for (int q = 0; q < 16; ++q) row[q] = fma(-lij, colj[b][16 * h + q], row[q]);
        }
    }
    __syncthreads();
    for (int r = h; r < kDT; r += 4) a[(o + r) * ld + o + i] = T[r][i];
    if (inv) {
        // column i of the inverse, rows [16h, 16h+16) in wave h: X[r] = (delta_ri -
        // sum_{m<r} L[r][m] X[m]) / L[r][r] with the sum taken in m order, quarter by
        // quarter: wave s finishes its rows, then the later waves add their terms
        // from them
        __shared__ double X[kDT][kDT + 1];
        double sacc[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) sacc[q] = 16 * h + q == i ? 1.0 : 0.0;
        for (int st = 0; st < 4; ++st) {
            if (h == st) {
                double xq[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int r = 16 * st + q;
#pragma unroll
                    for (int m = 0; m < q; ++m) sacc[q] = fma(-T[r][16 * st + m], xq[m], sacc[q]);
                    xq[q] = r < i ? 0.0 : sacc[q] / T[r][r];
                    X[r][i] = xq[q];
                }
            }
            __syncthreads();
            if (h > st) {
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int r = 16 * h + q;
#pragma unroll
                    for (int m = 0; m < 16; ++m) sacc[q] = fma(-T[r][16 * st + m], X[16 * st + m][i], sacc[q]);
                }
            }
        }
        double* dst = inv + (long)k * kDT * kDT;
        for (int r = h; r < kDT; r += 4) dst[r * kDT + i] = X[r][i];
    }
}

// panel: tiles (i, k), i > k, become A_ik L_kk^-T (each row solved forward against L_kk);
// columns of dropped pivots are zero
__global__ __launch_bounds__(kDirectBlock) void k_trsm_tiles(double* __restrict__ a, long ld, int k, const int* __restrict__ dropped) {
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
    // one thread per row (the running sums of a row are sequential)
    if (threadIdx.x < kDT) {
        const int r = threadIdx.x;
        for (int c = 0; c < kDT; ++c) {
            double s = X[r][c];
            for (int m = 0; m < c; ++m) s -= X[r][m] * L[c][m];
            X[r][c] = (dropped && dropped[ok + c]) ? 0.0 : s / L[c][c];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kDT * kDT; e += kDirectBlock) {
        const int r = e / kDT, c = e % kDT;
        a[(oi + r) * ld + ok + c] = X[r][c];
    }
}

// C (64 x 64 tile, row-major, leading dimension ldc) -= P Q^T  (BT) or P Q  (!BT), P and Q
// 64 x 64 tiles (ldp, ldq); overwrite: C = P Q instead.  Both operands are staged in LDS
// (Pt[l][r] = P[r][l]; Qt[l][c] = Q[c][l] or Q[l][c]) so that a thread's four rows and four
// columns are consecutive words; each thread owns a 4 x 4 block of C.  All reads of P and
// Q precede the barrier, so C may be the tile Q.
template <bool BT>
__device__ __forceinline__ void tile_update(double* __restrict__ C, long ldc, const double* P, long ldp, const double* Q, long ldq,
                                            double (*Pt)[kDT + 2], double (*Qt)[kDT + 2], bool overwrite,
                                            const int* colMask = nullptr) {
    // colMask (BT only): output columns c with colMask[c] != 0 come out zero
    for (int e = threadIdx.x; e < kDT * kDT; e += kDirectBlock) {
        const int r = e / kDT, l = e % kDT;
        Pt[l][r] = P[r * ldp + l];
        if (BT) Qt[l][r] = (colMask && colMask[r]) ? 0.0 : Q[r * ldq + l];
        else Qt[r][l] = Q[r * ldq + l];
    }
    __syncthreads();
    const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
#pragma unroll 4
    for (int l = 0; l < kDT; ++l) {
        double pv[4], qv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            pv[i] = Pt[l][4 * tr + i];
            qv[i] = Qt[l][4 * tc + i];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = fma(pv[i], qv[j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double* dst = C + (4 * tr + i) * ldc + 4 * tc + j;
            *dst = overwrite ? acc[i][j] : *dst - acc[i][j];
        }
}

// trailing update of the lower tiles (i, j), k < j <= i: A_ij -= L_ik L_jk^T; block b
// enumerates the lower triangle of the trailing tile matrix row by row
__global__ __launch_bounds__(kDirectBlock) void k_syrk_tiles(double* __restrict__ a, long ld, int k, int nt) {
    __shared__ double Pt[kDT][kDT + 2];
    __shared__ double Qt[kDT][kDT + 2];
    const int m = nt - k - 1;
    int b = blockIdx.x, ii = 0;
    while (b >= ii + 1) { b -= ii + 1; ++ii; }   // row ii of the trailing triangle holds ii + 1 tiles
    if (ii >= m) return;
    const int i = k + 1 + ii, j = k + 1 + b;
    const long ok = (long)k * kDT, oi = (long)i * kDT, oj = (long)j * kDT;
    tile_update<true>(a + oi * ld + oj, ld, a + oi * ld + ok, ld, a + oj * ld + ok, ld, Pt, Qt, false);
}

// panel by product with the inverted diagonal tile: A_ik <- A_ik inv(L_kk)^T, columns of
// dropped pivots zero
__global__ __launch_bounds__(kDirectBlock) void k_panel_tiles(double* __restrict__ a, long ld, int k, const double* __restrict__ inv,
                                                             const int* __restrict__ dropped) {
    __shared__ double Pt[kDT][kDT + 2];
    __shared__ double Qt[kDT][kDT + 2];
    const int i = k + 1 + blockIdx.x;
    const long ok = (long)k * kDT, oi = (long)i * kDT;
    double* Aik = a + oi * ld + ok;
    tile_update<true>(Aik, ld, Aik, ld, inv + (long)k * kDT * kDT, kDT, Pt, Qt, true, dropped ? dropped + ok : nullptr);
}

void dense_cholesky(Ctx& c, double* A, int npad, int* bad, double* tileInv, const double* diag0, int* dropped, double tau) {
    const int nt = npad / kDT;
    const long ld = npad;
    for (int k = 0; k < nt; ++k) {
        // the diagonal tile's inverse (tileInv) turns the panel into a product; without it
        // the panel rows are solved by substitution (the direct solve's exact path)
        hipLaunchKernelGGL(k_potrf_tile, dim3(1), dim3(kDirectBlock), 0, c.s, A, ld, k, bad, diag0, dropped, tau, tileInv);
        const int m = nt - k - 1;
        if (m > 0) {
            if (tileInv) hipLaunchKernelGGL(k_panel_tiles, dim3(m), dim3(kDirectBlock), 0, c.s, A, ld, k, (const double*)tileInv,
                                            (const int*)dropped);
            else hipLaunchKernelGGL(k_trsm_tiles, dim3(m), dim3(kDirectBlock), 0, c.s, A, ld, k, (const int*)dropped);
            hipLaunchKernelGGL(k_syrk_tiles, dim3(m * (m + 1) / 2), dim3(kDirectBlock), 0, c.s, A, ld, k, nt);
        }
        CIPC_LAUNCH_CHECK();
    }
}

// L^-1 by blocked forward substitution on the identity, right-looking over tile rows k:
// row k:  W_kj = inv(L_kk) W_kj  (j <= k; W_kj holds the updated right-hand side)
// update: W_ij -= L_ik W_kj      (i > k, j <= k)
__global__ __launch_bounds__(kDirectBlock) void k_inv_row(const double* __restrict__ inv, double* __restrict__ W, long ld, int k) {
    __shared__ double Pt[kDT][kDT + 2];
    __shared__ double Qt[kDT][kDT + 2];
    const long ok = (long)k * kDT, oj = (long)blockIdx.x * kDT;   // tile column j = blockIdx.x <= k
    double* Wkj = W + ok * ld + oj;
    tile_update<false>(Wkj, ld, inv + (long)k * kDT * kDT, kDT, Wkj, ld, Pt, Qt, true);
}
__global__ __launch_bounds__(kDirectBlock) void k_inv_update(const double* __restrict__ L, double* __restrict__ W, long ld, int k) {
    __shared__ double Pt[kDT][kDT + 2];
    __shared__ double Qt[kDT][kDT + 2];
    const int j = blockIdx.x % (k + 1), i = k + 1 + blockIdx.x / (k + 1);
    const long ok = (long)k * kDT, oi = (long)i * kDT, oj = (long)j * kDT;
    tile_update<false>(W + oi * ld + oj, ld, L + oi * ld + ok, ld, W + ok * ld + oj, ld, Pt, Qt, false);
}
__global__ void k_dense_identity(double* __restrict__ a, long ld, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[(long)i * ld + i] = 1.0;
}

// W = L^-1; tileInv holds the inverses of L's diagonal tiles (dense_cholesky with tileInv)
void dense_lower_inverse(Ctx& c, const double* L, double* W, int npad, const double* tileInv) {
    const int nt = npad / kDT;
    const long ld = npad;
    hipLaunchKernelGGL(k_dense_zero, dim3(2048), dim3(kBlock), 0, c.s, W, (long)npad * npad);
    hipLaunchKernelGGL(k_dense_identity, dim3(nblocks(npad)), dim3(kBlock), 0, c.s, W, ld, npad);
    CIPC_LAUNCH_CHECK();
    for (int k = 0; k < nt; ++k) {
        hipLaunchKernelGGL(k_inv_row, dim3(k + 1), dim3(kDirectBlock), 0, c.s, tileInv, W, ld, k);
        if (k + 1 < nt) hipLaunchKernelGGL(k_inv_update, dim3((nt - k - 1) * (k + 1)), dim3(kDirectBlock), 0, c.s, L, W, ld, k);
        CIPC_LAUNCH_CHECK();
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
    dense_cholesky(c, A, npad, bad, nullptr);
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
