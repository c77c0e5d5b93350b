// Newton linear solve: block-Jacobi preconditioned conjugate gradient on the
// assembled Hessian H = A_elastic (static BSR pattern, mass and Dirichlet projection
// included) + A_contact (per-Newton-iteration BSR).  Stands in for the reference's
// direct solve (Math/DIRECT_SOLVER.h:14-66, CHOLMOD); the relative-residual stopping
// rule ||r|| <= tol ||b|| follows the reference's iterative-solver precedent
// (AMGCL_SOLVER.h:53-65, rel-tol 1e-5).
//
// Three forms share one set of kernels (templates on the matrix value type M and the
// vector type V):
//   fp64   M = V = double;
//   mixed  M = float (FP32 copies of the Hessian blocks, half the SpMV bytes), V = double
//          (Krylov vectors, SpMV accumulation), inside FP64 iterative refinement;
//   fp32   M = V = float (blocks, vectors and SpMV accumulation in FP32, dot products
//          summed in FP64), inside FP64 iterative refinement.
// The refinement (x += d, r = b - H x with the FP64 blocks) repeats the inner solve on
// the FP64 residual until ||r|| <= tol ||b|| holds in FP64.
//
// One CG iteration is two kernels and no atomics:
//   k_pcg_spmv   beta from the previous update's partial sums, p = z + beta p_old formed
//                on the fly, q = A p, p written, partial sums of p.q per workgroup;
//   k_pcg_update x += alpha p, r -= alpha q, z = Dinv r, partial sums of r.z and r.r.
// Each kernel starts by summing the previous kernel's per-workgroup partials itself
// (every workgroup in the same fixed order, so all hold the identical scalar); kernel
// boundaries on the one stream order those writes before the reads, so no fence or
// ticket counter is needed.  The SpMV runs a fixed grid (Ctx::spmvGrid workgroups) so the
// update kernel has few partials to sum.
//
// Matrix blocks are stored structure-of-arrays (entry k of all blocks contiguous);
// LANES lanes share a block row and stride over its blocks, so the lanes of a wave read
// consecutive addresses.  Rows are mapped XCD-aware: workgroup b runs on XCD b % 8 and
// XCD x owns one contiguous eighth of the rows, so gathered neighbour values stay in
// that XCD's L2.
//
// The host launches iterations in chunks and polls an on-device done flag between
// chunks; kernels launched after convergence return at once.
#include "ctx.hpp"
#include "learn.hpp"
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <hip/hip_ext.h>
#include <cmath>

namespace cipc {

// solver scalars (sc[]) and flags
enum { S_RZ0 = 0, S_RZ1, S_BB, S_RR, S_TOL2 };   // S_TOL2: the squared relative target the SpMV tests
enum { F_DONE = 0, F_ITERS, F_FAIL };

constexpr int kSpmvBlock = 256;
constexpr int kUpdBlock = 256;
constexpr int kXcd = 8;
// with profiling on, every kProfEvery-th SpMV launch is bracketed by timing events (events
// on every launch slow the solve by about 10%), starting at launch kProfFirst of each
// chunk: the chunk's first launch follows the host's poll of the done flag with the GPU
// idle, which is not representative
constexpr int kProfEvery = 8, kProfFirst = 3;

// Sum over the workgroup; every thread receives the same value (butterfly shuffles
// give bit-identical results in all lanes, waves are added in a fixed order).
template <int BS>
__device__ __forceinline__ double block_allsum(double v, double* lds) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) r += lds[w];
    __syncthreads();
    return r;
}

// rows per XCD, rounded up to whole groups of `rowsPerGroup`
inline int rows_per_xcd(int nrow, int rowsPerGroup) {
    int r = (nrow + kXcd - 1) / kXcd;
    return (r + rowsPerGroup - 1) / rowsPerGroup * rowsPerGroup;
}

// One block row's share of lane `lane`: s += A_row,c (z_c + beta p_c) over the row's
// blocks [beg, end).  The lane that meets the diagonal block keeps z_row + beta p_row
// in `own` (every row stores its diagonal block), so the row's new direction needs no
// separate load.
// UNR = 2: each pass takes the lane's next two blocks (b and b + LANES) and issues all of
// their loads before any product, doubling the loads in flight per lane; a missing second
// block re-reads the first and contributes zero.
template <class M, class V, int LANES, bool WITHP, int UNR = 1>
__device__ __forceinline__ void bsr_row(int beg, int end, const int* __restrict__ col, const M* __restrict__ val, long nnzb, int row,
                                        const V* __restrict__ z, const V* __restrict__ p, V beta, V* s, V* own, bool& hasOwn) {
    static_assert(UNR == 1 || UNR == 2, "");
    for (int b = beg; b < end; b += UNR * LANES) {
        int bu[UNR];
        bool on[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            on[u] = u == 0 || b + u * LANES < end;
            bu[u] = on[u] ? b + u * LANES : b;
        }
        int c[UNR];
        V a[UNR][9], x[UNR][3];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            c[u] = col[bu[u]];
#pragma unroll
            for (int k = 0; k < 9; ++k) a[u][k] = (V)val[k * nnzb + bu[u]];
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
#pragma unroll
            for (int k = 0; k < 3; ++k) x[u][k] = z[3 * c[u] + k];
            if (WITHP) {
                V pv[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) pv[k] = p[3 * c[u] + k];
#pragma unroll
                for (int k = 0; k < 3; ++k) x[u][k] += beta * pv[k];
            }
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if (on[u] && c[u] == row) {
                own[0] = x[u][0]; own[1] = x[u][1]; own[2] = x[u][2];
                hasOwn = true;
            }
            if (!on[u]) x[u][0] = x[u][1] = x[u][2] = 0;
            s[0] += a[u][0] * x[u][0] + a[u][1] * x[u][1] + a[u][2] * x[u][2];
            s[1] += a[u][3] * x[u][0] + a[u][4] * x[u][1] + a[u][5] * x[u][2];
            s[2] += a[u][6] * x[u][0] + a[u][7] * x[u][1] + a[u][8] * x[u][2];
        }
    }
}

// One block row of E + K in a single lane loop: positions j < nE of the row walk the
// elastic blocks [eb, ee), the rest the contact blocks [cb, ce) (each lane picks its
// pattern's arrays per position), so the lanes of a row fill up across both patterns and
// no second dependent row-bound load follows the elastic pass.  The diagonal block that
// fills `own` is the elastic one (contact rows also hold diagonal blocks).
template <class M, class V, int LANES, bool WITHP, int UNR>
__device__ __forceinline__ void bsr_row_ek(int lane, int eb, int ee, int cb, int ce, const int* __restrict__ ecol,
                                           const M* __restrict__ ev, long enz, const int* __restrict__ ccol, const M* __restrict__ cv,
                                           long cnz, int row, const V* __restrict__ z, const V* __restrict__ p, V beta, V* s, V* own,
                                           bool& hasOwn) {
    static_assert(UNR == 1 || UNR == 2, "");
    const int nE = ee - eb, nT = nE + (ce - cb);
    for (int j = lane; j < nT; j += UNR * LANES) {
        int c[UNR];
        bool on[UNR], el[UNR];
        V a[UNR][9], x[UNR][3];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int ju = u == 0 || j + u * LANES < nT ? j + u * LANES : j;
            on[u] = u == 0 || j + u * LANES < nT;
            el[u] = ju < nE;
            const int b = el[u] ? eb + ju : cb + (ju - nE);
            const int* cp = el[u] ? ecol : ccol;
            const M* vp = el[u] ? ev : cv;
            const long nz = el[u] ? enz : cnz;
            c[u] = cp[b];
#pragma unroll
            for (int k = 0; k < 9; ++k) a[u][k] = (V)vp[k * nz + b];
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
#pragma unroll
            for (int k = 0; k < 3; ++k) x[u][k] = z[3 * c[u] + k];
            if (WITHP) {
                V pv[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) pv[k] = p[3 * c[u] + k];
#pragma unroll
                for (int k = 0; k < 3; ++k) x[u][k] += beta * pv[k];
            }
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if (on[u] && el[u] && c[u] == row) {
                own[0] = x[u][0]; own[1] = x[u][1]; own[2] = x[u][2];
                hasOwn = true;
            }
            if (!on[u]) x[u][0] = x[u][1] = x[u][2] = 0;
            s[0] += a[u][0] * x[u][0] + a[u][1] * x[u][1] + a[u][2] * x[u][2];
            s[1] += a[u][3] * x[u][0] + a[u][4] * x[u][1] + a[u][5] * x[u][2];
            s[2] += a[u][6] * x[u][0] + a[u][7] * x[u][1] + a[u][8] * x[u][2];
        }
    }
}

// Iteration `it`: convergence of the residual left by the previous update is decided
// first; then p_new = z + beta p_old (p_new = z at it = 0, WITHP = false), q = A p_new,
// partial p.q.  Row bounds of the next row group are loaded one group ahead.
template <class M, class V, int LANES, int BS, bool WITHP, int UNR = 1>
__global__ __launch_bounds__(BS) void k_pcg_spmv(const int* __restrict__ itDev, int itArg, int maxIt, const int* __restrict__ erp,
                                                 const int* __restrict__ ecol, const M* __restrict__ ev, int enz,
                                                 const int* __restrict__ crp, const int* __restrict__ ccol, const M* __restrict__ cv,
                                                 int cnz, int nrow, int rowsPerXcd, const V* __restrict__ z,
                                                 const V* __restrict__ pold, V* __restrict__ pnew, V* __restrict__ q,
                                                 const double* __restrict__ partRZ, int nRZ, const double* __restrict__ partRR, int nRR,
                                                 double* __restrict__ partS, double* __restrict__ sc, int* __restrict__ flags) {
    __shared__ double lds[BS / 64];
    constexpr int RPG = BS / LANES;   // rows per group
    const int xb = blockIdx.x % kXcd, jb = blockIdx.x / kXcd, nj = gridDim.x / kXcd;
    const int lane = threadIdx.x % LANES;
    const int rowBase = xb * rowsPerXcd + (int)threadIdx.x / LANES;
    // first group's row bounds (elastic and contact), issued before the done flag and the
    // scalar reduction are read
    int g = jb;
    int row = rowBase + g * RPG;
    bool live = g * RPG < rowsPerXcd && row < nrow;
    int beg = live ? erp[row] : 0, end = live ? erp[row + 1] : 0;
    int cbeg = live && cnz > 0 ? crp[row] : 0, cend = live && cnz > 0 ? crp[row + 1] : 0;
    if (flags[F_DONE]) return;
    const int it = itDev ? *itDev + itArg : itArg;   // a captured chunk reads its base iteration

    double arz = 0, arr = 0;
#pragma unroll 4
    for (int i = threadIdx.x; i < nRZ; i += BS) arz += partRZ[i];
#pragma unroll 4
    for (int i = threadIdx.x; i < nRR; i += BS) arr += partRR[i];
    const double rz = block_allsum<BS>(arz, lds);
    const double rr = block_allsum<BS>(arr, lds);
    const double bb = it == 0 ? rr : sc[S_BB];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (rr <= sc[S_TOL2] * bb || it >= maxIt) {
        if (lead) {
            if (it == 0) sc[S_BB] = rr;
            sc[S_RR] = rr;
            flags[F_ITERS] = it;
            flags[F_DONE] = 1;
        }
        return;
    }
    const V beta = WITHP ? (V)(rz / sc[S_RZ0 + ((it - 1) & 1)]) : (V)0;

    double pq = 0;
    for (; g * RPG < rowsPerXcd; g += nj) {
        const int gn = g + nj;
        const int rowN = rowBase + gn * RPG;
        const bool liveN = gn * RPG < rowsPerXcd && rowN < nrow;
        const int begN = liveN ? erp[rowN] : 0, endN = liveN ? erp[rowN + 1] : 0;
        const int cbegN = liveN && cnz > 0 ? crp[rowN] : 0, cendN = liveN && cnz > 0 ? crp[rowN + 1] : 0;
        V s[3] = {0, 0, 0}, own[3] = {0, 0, 0};
        bool hasOwn = false;
        if (row < nrow)
            bsr_row_ek<M, V, LANES, WITHP, UNR>(lane, beg, end, cbeg, cend, ecol, ev, enz, ccol, cv, cnz, row, z, pold, beta, s, own,
                                                hasOwn);
#pragma unroll
        for (int o = LANES / 2; o > 0; o >>= 1)
#pragma unroll
            for (int k = 0; k < 3; ++k) s[k] += __shfl_xor(s[k], o, LANES);
        if (row < nrow && hasOwn) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                pnew[3 * row + k] = own[k];
                q[3 * row + k] = s[k];
                pq += (double)own[k] * (double)s[k];
            }
        }
        row = rowN;
        beg = begN;
        end = endN;
        cbeg = cbegN;
        cend = cendN;
    }
    const double bpq = block_allsum<BS>(pq, lds);
    if (threadIdx.x == 0) partS[blockIdx.x] = bpq;
    if (lead) {
        sc[S_RZ0 + (it & 1)] = rz;
        if (it == 0) sc[S_BB] = rr;
    }
}

// the SpMV instance for a launch shape: LANES lanes per block row, UNR blocks per lane pass
template <class M, class V>
static auto spmv_kernel(int lanes, int unr, bool withP) -> decltype(&k_pcg_spmv<M, V, 16, kSpmvBlock, true, 1>) {
#define CIPC_SPMV_PICK(L, U) \
    if (lanes == L && unr == U) return withP ? k_pcg_spmv<M, V, L, kSpmvBlock, true, U> : k_pcg_spmv<M, V, L, kSpmvBlock, false, U>;
    CIPC_SPMV_PICK(8, 1)
    CIPC_SPMV_PICK(8, 2)
    CIPC_SPMV_PICK(4, 1)
    CIPC_SPMV_PICK(4, 2)
    CIPC_SPMV_PICK(16, 2)
    CIPC_SPMV_PICK(16, 1)
#undef CIPC_SPMV_PICK
    return nullptr;
}

// Iteration `it`: alpha = r.z / p.q, x += alpha p, r -= alpha q, z = Dinv r, partials
// of r.z and r.r (written as partU[block], partU[nU + block]).
template <class V, int BS, bool JAC>
__global__ __launch_bounds__(BS) void k_pcg_update(const int* __restrict__ itDev, int itArg, int nrow, int rowsPerXcd, V* __restrict__ x,
                                                   V* __restrict__ r, V* __restrict__ z, const V* __restrict__ p, const V* __restrict__ q,
                                                   const V* __restrict__ dinv, const double* __restrict__ partS, int nS,
                                                   double* __restrict__ partU, double* __restrict__ sc, int* __restrict__ flags) {
    if (flags[F_DONE]) return;
    const int it = itDev ? *itDev + itArg : itArg;
    __shared__ double lds[BS / 64];
    double a = 0;
    for (int i = threadIdx.x; i < nS; i += BS) a += partS[i];
    const int row = (int)(blockIdx.x % kXcd) * rowsPerXcd + (int)(blockIdx.x / kXcd) * BS + (int)threadIdx.x;
    const bool valid = row < nrow;
    V pv[3] = {0, 0, 0}, qv[3] = {0, 0, 0}, xv[3] = {0, 0, 0}, rv[3] = {0, 0, 0}, D[9];
    if (valid) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            pv[k] = p[3 * row + k]; qv[k] = q[3 * row + k];
            xv[k] = x[3 * row + k]; rv[k] = r[3 * row + k];
        }
        if (JAC)
#pragma unroll
            for (int k = 0; k < 9; ++k) D[k] = dinv[9l * row + k];
    }
    const double pq = block_allsum<BS>(a, lds);
    if (!(pq > 0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            flags[F_FAIL] = 1;
            flags[F_ITERS] = it;
            flags[F_DONE] = 1;
        }
        return;
    }
    const V alpha = (V)(sc[S_RZ0 + (it & 1)] / pq);
    double rz = 0, rr = 0;
    if (valid) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            x[3 * row + k] = xv[k] + alpha * pv[k];
            rv[k] = rv[k] - alpha * qv[k];
            r[3 * row + k] = rv[k];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (JAC) {
                const V zk = D[3 * k] * rv[0] + D[3 * k + 1] * rv[1] + D[3 * k + 2] * rv[2];
                z[3 * row + k] = zk;
                rz += (double)rv[k] * (double)zk;
            }
            rr += (double)rv[k] * (double)rv[k];
        }
    }
    const double brz = block_allsum<BS>(rz, lds);
    const double brr = block_allsum<BS>(rr, lds);
    if (threadIdx.x == 0) {
        partU[blockIdx.x] = brz;
        partU[gridDim.x + blockIdx.x] = brr;
    }
}

// Learned form, iteration `it`, fused with the first factor product: alpha = r.z / p.q,
// r' = r - alpha q (written to the other r buffer: rows are read by their neighbours
// while the owner updates them), x += alpha p, y = F r' over the rows of the lower BF16
// factor F, partial sums of r'.r'.  Every row of F stores its diagonal block; the lane
// that meets it updates the row's r and x.  Fixed grid, rows mapped XCD-aware.
template <int LANES, int BS, class V>
__global__ __launch_bounds__(BS) void k_pcg_update_lower(const int* __restrict__ itDev, int itArg, int nrow, int rowsPerXcd,
                                                         const int* __restrict__ rp,
                                                         const int* __restrict__ col, const uint32_t* __restrict__ val, int nnz,
                                                         V* __restrict__ x, const V* __restrict__ p, const V* __restrict__ q,
                                                         const V* __restrict__ rOld, V* __restrict__ rNew, V* __restrict__ y,
                                                         const double* __restrict__ partS, int nS, double* __restrict__ partRR,
                                                         double* __restrict__ sc, int* __restrict__ flags) {
    if (flags[F_DONE]) return;
    const int it = itDev ? *itDev + itArg : itArg;
    __shared__ double lds[BS / 64];
    constexpr int RPG = BS / LANES;
    const int xb = blockIdx.x % kXcd, jb = blockIdx.x / kXcd, nj = gridDim.x / kXcd;
    const int lane = threadIdx.x % LANES;
    const int rowBase = xb * rowsPerXcd + (int)threadIdx.x / LANES;
    int g = jb;
    int row = rowBase + g * RPG;
    bool live = g * RPG < rowsPerXcd && row < nrow;
    int beg = live ? rp[row] : 0, end = live ? rp[row + 1] : 0;
    double a = 0;
    for (int i = threadIdx.x; i < nS; i += BS) a += partS[i];
    const double pq = block_allsum<BS>(a, lds);
    if (!(pq > 0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            flags[F_FAIL] = 1;
            flags[F_ITERS] = it;
            flags[F_DONE] = 1;
        }
        return;
    }
    const V alpha = (V)(sc[S_RZ0 + (it & 1)] / pq);
    double rr = 0;
    for (; g * RPG < rowsPerXcd; g += nj) {
        const int gn = g + nj, rowN = rowBase + gn * RPG;
        const bool liveN = gn * RPG < rowsPerXcd && rowN < nrow;
        const int begN = liveN ? rp[rowN] : 0, endN = liveN ? rp[rowN + 1] : 0;
        V acc[3] = {0, 0, 0};
        for (int b = beg + lane; b < end; b += LANES) {
            const int c = col[b];
            V rv[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) rv[k] = rOld[3 * c + k] - alpha * q[3 * c + k];
            V f[9];
            lp_unpack9<V>(val, nnz, b, f);
            acc[0] += f[0] * rv[0] + f[1] * rv[1] + f[2] * rv[2];
            acc[1] += f[3] * rv[0] + f[4] * rv[1] + f[5] * rv[2];
            acc[2] += f[6] * rv[0] + f[7] * rv[1] + f[8] * rv[2];
            if (c == row) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    rNew[3 * row + k] = rv[k];
                    x[3 * row + k] += alpha * p[3 * row + k];
                    rr += (double)rv[k] * (double)rv[k];
                }
            }
        }
#pragma unroll
        for (int o = LANES / 2; o > 0; o >>= 1)
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[k] += __shfl_xor(acc[k], o, LANES);
        if (row < nrow && lane == 0) {
#pragma unroll
            for (int k = 0; k < 3; ++k) y[3 * row + k] = acc[k];
        }
        row = rowN;
        beg = begN;
        end = endN;
    }
    const double brr = block_allsum<BS>(rr, lds);
    if (threadIdx.x == 0) partRR[blockIdx.x] = brr;
}

// r = b, x = 0, z = Dinv r; partials of r.z and r.r in the update kernel's layout;
// the solver flags are cleared.
template <class V, int BS>
__global__ __launch_bounds__(BS) void k_pcg_init(int nrow, int rowsPerXcd, const V* __restrict__ b, V* __restrict__ x, V* __restrict__ r,
                                                 V* __restrict__ z, const V* __restrict__ dinv, double* __restrict__ partU,
                                                 int* __restrict__ flags) {
    __shared__ double lds[BS / 64];
    const int row = (int)(blockIdx.x % kXcd) * rowsPerXcd + (int)(blockIdx.x / kXcd) * BS + (int)threadIdx.x;
    double rz = 0, rr = 0;
    if (row < nrow) {
        V rv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            rv[k] = b[3 * row + k];
            r[3 * row + k] = rv[k];
            x[3 * row + k] = 0;
            rr += (double)rv[k] * (double)rv[k];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const V zk = dinv[9l * row + 3 * k] * rv[0] + dinv[9l * row + 3 * k + 1] * rv[1] + dinv[9l * row + 3 * k + 2] * rv[2];
            z[3 * row + k] = zk;
            rz += (double)rv[k] * (double)zk;
        }
    }
    const double brz = block_allsum<BS>(rz, lds);
    const double brr = block_allsum<BS>(rr, lds);
    if (threadIdx.x == 0) {
        partU[blockIdx.x] = brz;
        partU[gridDim.x + blockIdx.x] = brr;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        flags[F_DONE] = 0;
        flags[F_ITERS] = 0;
        flags[F_FAIL] = 0;
    }
}

// diagonal 3x3 blocks of A_el + A_con (FP64), inverted (symmetric adjugate), stored as V
template <class V>
__global__ void k_dinv(int nV, const int* __restrict__ rpe, const int* __restrict__ ce, const double* __restrict__ ve, int nze,
                       const int* __restrict__ rpc, const int* __restrict__ cc, const double* __restrict__ vc, int nzc,
                       V* __restrict__ dinv) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = rpe[v]; b < rpe[v + 1]; ++b)
        if (ce[b] == v) {
#pragma unroll
            for (int k = 0; k < 9; ++k) a[k] += ve[(long)k * nze + b];
        }
    if (nzc > 0)
        for (int b = rpc[v]; b < rpc[v + 1]; ++b)
            if (cc[b] == v) {
#pragma unroll
                for (int k = 0; k < 9; ++k) a[k] += vc[(long)k * nzc + b];
            }
    double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double r[9];
    if (det > 0 && isfinite(det)) {
        double id = 1.0 / det;
        r[0] = c00 * id; r[1] = (a[2] * a[7] - a[1] * a[8]) * id; r[2] = (a[1] * a[5] - a[2] * a[4]) * id;
        r[3] = c01 * id; r[4] = (a[0] * a[8] - a[2] * a[6]) * id; r[5] = (a[2] * a[3] - a[0] * a[5]) * id;
        r[6] = c02 * id; r[7] = (a[1] * a[6] - a[0] * a[7]) * id; r[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) r[k] = 0;
        r[0] = a[0] > 0 ? 1.0 / a[0] : 1.0; r[4] = a[4] > 0 ? 1.0 / a[4] : 1.0; r[8] = a[8] > 0 ? 1.0 / a[8] : 1.0;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) dinv[9l * v + k] = (V)r[k];
}

// ---- mixed-precision helpers
__global__ void k_f64_to_f32(long n, const double* __restrict__ a, float* __restrict__ b) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
__global__ void k_acc_f32(long n, const float* __restrict__ d, double* __restrict__ x) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += (double)d[i];
}
// r = b - (A_el + A_con) x in FP64, rows of LANES lanes; per-workgroup partial of r.r
template <int LANES>
__global__ __launch_bounds__(kBlock) void k_residual(const int* __restrict__ erp, const int* __restrict__ ecol, const double* __restrict__ ev,
                                                     int enz, const int* __restrict__ crp, const int* __restrict__ ccol,
                                                     const double* __restrict__ cv, int cnz, int nrow, const double* __restrict__ b,
                                                     const double* __restrict__ x, double* __restrict__ r, double* __restrict__ part) {
    __shared__ double lds[kBlock / 64];
    const long gid = (long)blockIdx.x * kBlock + threadIdx.x;
    const int row = (int)(gid / LANES), lane = (int)(gid % LANES);
    double s[3] = {0, 0, 0};
    if (row < nrow) {
        double own[3];
        bool hasOwn = false;
        bsr_row<double, double, LANES, false>(erp[row] + lane, erp[row + 1], ecol, ev, enz, row, x, x, 0.0, s, own, hasOwn);
        if (cnz > 0) bsr_row<double, double, LANES, false>(crp[row] + lane, crp[row + 1], ccol, cv, cnz, -1, x, x, 0.0, s, own, hasOwn);
    }
#pragma unroll
    for (int o = LANES / 2; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) s[k] += __shfl_xor(s[k], o, LANES);
    double rr = 0;
    if (row < nrow && lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double rk = b[3 * row + k] - s[k];
            r[3 * row + k] = rk;
            rr += rk * rk;
        }
    }
    const double br = block_allsum<kBlock>(rr, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = br;
}

// device iteration counter of the captured chunks: set before the first replay, advanced
// by the chunk's last node
__global__ void k_it_set(int* itDev, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *itDev = v;
}
__global__ void k_it_advance(int* itDev, int by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *itDev += by;
}

// ---- one preconditioned CG solve in precision V; returns 0 converged, 1 not SPD
// (p^T A p <= 0), 2 iteration limit
template <class M, class V>
static int pcg_run(Ctx& c, const M* valE, const M* valK, const V* b, V* xsol, V* r, V* z, V* pbuf, V* q, const V* dinv, double tol, int maxIter,
                   int* iters, double* relres, bool learned = false) {
    const int nV = c.nV, n = 3 * nV;
    BSR& E = c.Ael;
    BSR& K = c.Acon;
    double* sc = c.scal.p + 32;               // solver scalars
    int* flags = (int*)(c.counters.p + 32);   // solver flags
    Ctx::Tuner& T = c.tune;
    // the autotuner's SpMV phase changes the SpMV shape from chunk to chunk (all shapes
    // compute the same q and p.q partials, so the iteration is unchanged)
    const bool spmvTuning = T.pcg && T.pcgPhase == 1;
    int nSmax = c.spmvGrid;
    if (spmvTuning)
        for (const auto& k : T.spmvCand) nSmax = std::max(nSmax, k[0]);
    const int rpxU = rows_per_xcd(nV, kUpdBlock), nU = kXcd * (rpxU / kUpdBlock);
    const int nP = lp_apply_partials(c);
    double* partS = c.dscr.ensure((size_t)nSmax + 2l * nU + 2l * nP);
    double* partU = partS + nSmax;
    double* partP = partU + 2l * nU;
    double* partF = partP + nP;               // r.r partials of the fused learned update
    // r.z partials: the update kernel's (block Jacobi) or the preconditioner product's
    const double* partRZ = learned ? partP : partU;
    const int nRZ = learned ? nP : nU;
    // two-level correction of the learned preconditioner (coarse.hip)
    const bool coarse = learned && c.coarseOn && c.cg.ready;
    V* ltmp = nullptr;
    V* rb[2] = {r, r};                        // learned form: r ping-pongs between r and a second buffer

    hipLaunchKernelGGL((k_pcg_init<V, kUpdBlock>), dim3(nU), dim3(kUpdBlock), 0, c.s, nV, rpxU, b, xsol, r, z, dinv, partU, flags);
    CIPC_LAUNCH_CHECK();
    if (learned) {
        static_assert(sizeof(double) >= sizeof(V), "");
        ltmp = (V*)c.lpTmp.ensure((size_t)2 * n);
        rb[1] = ltmp + n;
        if (coarse) coarse_restrict_solve<V>(c, (const V*)r, nullptr);
        lp_apply<V>(c, (const V*)r, z, ltmp, partP, coarse);
    }

    const int chunk = 16;
    int* hf = (int*)(c.hCnt + 32);
    // sampled SpMV timing: for the caller's profile and for the autotuner's SpMV phase
    // (the caller's profile samples the form's own block type: FP64-block retries of a
    // reduced-precision form are not its SpMV)
    const bool profThis = c.profiling && (sizeof(M) == sizeof(double)) == (c.profPrecision == 0);
    const bool sampling = profThis || spmvTuning;
    if (sampling && c.evPool.size() < 2 * chunk) {
        for (size_t i = c.evPool.size(); i < 2 * chunk; ++i) {
            hipEvent_t e;
            CIPC_HIP(hipEventCreate(&e));
            c.evPool.push_back(e);
        }
    }
    const double tol2 = tol * tol;
    // the target is read by the SpMV from device memory, so captured chunks serve every
    // tolerance (refinement sweeps change it)
    c.hScal[36] = tol2;
    CIPC_HIP(hipMemcpyAsync(sc + S_TOL2, c.hScal + 36, sizeof(double), hipMemcpyHostToDevice, c.s));
    const double mb = (double)sizeof(M), vb = (double)sizeof(V);
    const double spmvBytes = (9 * mb + 4) * ((double)E.nnzb + K.nnzb) + 4.0 * (nV + 1) * (K.nnzb > 0 ? 2 : 1) + 4 * 3 * vb * nV;
    int* itDev = (int*)(c.counters.p + 24);   // base iteration of a captured chunk
    // one iteration's launches: iteration `it` (itDev == nullptr) or base + k (captured)
    auto launch_iteration = [&](int it, int k, const int* itd, int nS, int sLanes, int sUnroll, int rpxS, bool timed) {
        const V* pold = pbuf + (size_t)((it + 1) & 1) * n;
        V* pnew = pbuf + (size_t)(it & 1) * n;
        const int itArg = itd ? k : it;
        auto kern = spmv_kernel<M, V>(sLanes, sUnroll, it > 0);
        const double* partRR = learned && it > 0 ? partF : partU + nU;
        const int nRR = learned && it > 0 ? nP : nU;
        if (timed)
            hipExtLaunchKernelGGL(kern, dim3(nS), dim3(kSpmvBlock), 0, c.s, c.evPool[2 * k], c.evPool[2 * k + 1], 0, itd, itArg, maxIter,
                                  E.rowptr.p, E.col.p, valE, E.nnzb, K.rowptr.p, K.col.p, valK, K.nnzb, nV, rpxS, (const V*)z, pold, pnew,
                                  q, partRZ, nRZ, partRR, nRR, partS, sc, flags);
        else
            hipLaunchKernelGGL(kern, dim3(nS), dim3(kSpmvBlock), 0, c.s, itd, itArg, maxIter, E.rowptr.p, E.col.p, valE, E.nnzb,
                               K.rowptr.p, K.col.p, valK, K.nnzb, nV, rpxS, (const V*)z, pold, pnew, q, partRZ, nRZ, partRR, nRR, partS,
                               sc, flags);
        if (!learned) {
            hipLaunchKernelGGL((k_pcg_update<V, kUpdBlock, true>), dim3(nU), dim3(kUpdBlock), 0, c.s, itd, itArg, nV, rpxU, xsol, r, z,
                               (const V*)pnew, (const V*)q, dinv, (const double*)partS, nS, partU, sc, flags);
        } else {
            LearnedPrecond& lp = *c.lp;
            V* rNew = rb[(it + 1) & 1];
            auto fused = c.lpLanes == 2    ? k_pcg_update_lower<2, kSpmvBlock, V>
                         : c.lpLanes == 4  ? k_pcg_update_lower<4, kSpmvBlock, V>
                         : c.lpLanes == 16 ? k_pcg_update_lower<16, kSpmvBlock, V>
                                           : k_pcg_update_lower<8, kSpmvBlock, V>;
            const int rpxL = rows_per_xcd(nV, kSpmvBlock / c.lpLanes);
            hipLaunchKernelGGL(fused, dim3(nP), dim3(kSpmvBlock), 0, c.s, itd, itArg, nV, rpxL, (const int*)lp.gLrp.p, (const int*)lp.gLcol.p,
                               (const uint32_t*)lp.FL.p, lp.nLow, xsol, (const V*)pnew, (const V*)q, (const V*)rb[it & 1], rNew, ltmp,
                               (const double*)partS, nS, partF, sc, flags);
            if (coarse) coarse_restrict_solve<V>(c, (const V*)rNew, flags);
            lp_apply_upper<V>(c, (const V*)ltmp, z, (const V*)rNew, partP, flags, coarse);
        }
    };
    int it = 0;
    bool itDevSet = false;
    for (;;) {
        int launched = 0;
        int cand = -1, nS = c.spmvGrid, sLanes = c.spmvLanes, sUnroll = c.spmvUnroll;
        if (spmvTuning) {
            cand = (int)(T.spmvTurn++ % (long)T.spmvCand.size());
            nS = T.spmvCand[cand][0]; sLanes = T.spmvCand[cand][1]; sUnroll = T.spmvCand[cand][2];
        }
        const int rpxS = rows_per_xcd(nV, kSpmvBlock / sLanes);
        // a whole chunk after the first, with no launch timed, replays one captured graph of
        // its 16 iterations (launch arguments identical from chunk to chunk: the iteration
        // number comes from the device counter, the buffer parities repeat every 16)
        const bool asGraph = c.useGraphs && !sampling && it > 0 && it % chunk == 0 && it + chunk - 1 <= maxIter;
        if (asGraph) {
            // (the target tolerance is read from sc[S_TOL2] at run time, not captured)
            std::vector<uintptr_t> key = {(uintptr_t)(void*)spmv_kernel<M, V>(sLanes, sUnroll, true), (uintptr_t)nS, (uintptr_t)rpxS,
                                          (uintptr_t)maxIter,
                                          (uintptr_t)E.rowptr.p, (uintptr_t)E.col.p, (uintptr_t)valE, (uintptr_t)E.nnzb,
                                          (uintptr_t)K.rowptr.p, (uintptr_t)K.col.p, (uintptr_t)valK, (uintptr_t)K.nnzb, (uintptr_t)nV,
                                          (uintptr_t)z, (uintptr_t)pbuf, (uintptr_t)q, (uintptr_t)partRZ, (uintptr_t)nRZ,
                                          (uintptr_t)partF, (uintptr_t)partU, (uintptr_t)nU, (uintptr_t)nP, (uintptr_t)partS,
                                          (uintptr_t)sc, (uintptr_t)flags, (uintptr_t)xsol, (uintptr_t)r, (uintptr_t)dinv,
                                          (uintptr_t)learned, (uintptr_t)ltmp, (uintptr_t)rb[1], (uintptr_t)partP, (uintptr_t)sizeof(M),
                                          (uintptr_t)sizeof(V)};
            if (learned) {
                LearnedPrecond& lp = *c.lp;
                const std::vector<uintptr_t> lk = lp_graph_key(c);
                key.insert(key.end(), lk.begin(), lk.end());
            }
            if (coarse) {
                const CoarseSpace& g = c.cg;
                key.insert(key.end(), {(uintptr_t)g.nAgg, (uintptr_t)g.npad, (uintptr_t)g.agg.p, (uintptr_t)g.aggPtr.p,
                                       (uintptr_t)g.aggVerts.p, (uintptr_t)g.d.p, (uintptr_t)g.rc.p, (uintptr_t)g.wc.p, (uintptr_t)g.zc.p,
                                       (uintptr_t)g.W32.p, (uintptr_t)g.WT32.p, (uintptr_t)g.W16.p, (uintptr_t)g.WT16.p,
                                       (uintptr_t)g.bf16});
            } else {
                key.push_back(0);
            }
            if (!c.pcgGraph.exec || c.pcgGraph.key != key) {
                c.pcgGraph.reset();
                hipGraph_t g = nullptr;
                CIPC_HIP(hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal));
                try {
                    for (int k = 0; k < chunk; ++k) launch_iteration(it + k, k, itDev, nS, sLanes, sUnroll, rpxS, false);
                    hipLaunchKernelGGL(k_it_advance, dim3(1), dim3(64), 0, c.s, itDev, chunk);
                } catch (...) {
                    (void)hipStreamEndCapture(c.s, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                CIPC_HIP(hipStreamEndCapture(c.s, &g));
                hipGraphExec_t ex = nullptr;
                const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                CIPC_HIP(ie);
                c.pcgGraph.exec = ex;
                c.pcgGraph.key = key;
                c.pcgGraphCaptures++;
            }
            if (!itDevSet) {
                hipLaunchKernelGGL(k_it_set, dim3(1), dim3(64), 0, c.s, itDev, it);
                itDevSet = true;
            }
            CIPC_HIP(hipGraphLaunch(c.pcgGraph.exec, c.s));
            c.pcgGraphLaunches++;
            it += chunk;
            launched = chunk;
        } else {
            itDevSet = false;
            for (int k = 0; k < chunk && it <= maxIter; ++k, ++it, ++launched)
                launch_iteration(it, k, nullptr, nS, sLanes, sUnroll, rpxS, sampling && k % kProfEvery == kProfFirst);
        }
        CIPC_LAUNCH_CHECK();
        CIPC_HIP(hipMemcpyAsync(hf, flags, 3 * sizeof(int), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        if (sampling) {
            // SpMV launches of iterations before the one that detected convergence did
            // the full work; the detecting launch and later ones return early
            const int done = hf[F_DONE] ? hf[F_ITERS] : it;
            const int base = it - launched;
            for (int k = kProfFirst; k < launched && base + k < done; k += kProfEvery) {
                float ms = 0;
                CIPC_HIP(hipEventElapsedTime(&ms, c.evPool[2 * k], c.evPool[2 * k + 1]));
                if (profThis) {
                    c.profMs += ms;
                    c.profLaunches++;
                    c.profBytes += spmvBytes;
                    c.profFlops += 18.0 * ((double)E.nnzb + K.nnzb) + 8.0 * n;
                }
                if (cand >= 0) {
                    T.spmvMs[cand] += ms;
                    T.spmvN[cand]++;
                }
            }
        }
        if (hf[F_DONE] || it > maxIter) break;
    }
    CIPC_HIP(hipMemcpyAsync(c.hScal + 32, sc, 4 * sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    if (iters) *iters = hf[F_ITERS];
    const double bb = c.hScal[32 + S_BB], rr = c.hScal[32 + S_RR];
    if (relres) *relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    if (hf[F_FAIL]) return 1;
    if (!hf[F_DONE] || rr > tol2 * bb) return 2;
    return 0;
}

// ||b - H x||^2 in FP64 (r written)
static double residual_norm2(Ctx& c, const double* b, const double* x, double* r) {
    BSR& E = c.Ael;
    BSR& K = c.Acon;
    const int nV = c.nV;
    const int nb = nblocks(4l * nV);
    double* part = c.dscr2.ensure((size_t)nb + 1);
    hipLaunchKernelGGL(k_residual<4>, dim3(nb), dim3(kBlock), 0, c.s, E.rowptr.p, E.col.p, E.val.p, E.nnzb, K.rowptr.p, K.col.p, K.val.p, K.nnzb,
                       nV, b, x, r, part);
    CIPC_LAUNCH_CHECK();
    sum_partials(part, nb, c.scal.p + 40, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hScal + 40, c.scal.p + 40, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return c.hScal[40];
}

__global__ void k_acc_f64(long n, const double* __restrict__ d, double* __restrict__ x) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

// FP64 iterative refinement: x = 0, r = b; repeat { solve A_M d = r in precision V to a
// relative tolerance chosen from the remaining gap; x += d; r = b - A x in FP64 } until
// ||r|| <= tol ||b||.  rb is the precision-V copy of r (unused when V = double).
template <class M, class V>
static int refine(Ctx& c, const M* vE, const M* vK, const V* dinv, const double* b, double* xsol, double tol, int maxIter, int* iters,
                  double* relres, V* wr, V* wz, V* wp, V* wq, V* d, V* rb, bool learned) {
    // with the coarse level a stalled refinement returns at once (2): the caller redoes the
    // solve with FP64 blocks instead of spending the rest of the budget
    const bool stallExit = learned && c.coarseOn && c.cg.ready;
    const long n = 3l * c.nV;
    double* r = c.rres.ensure(n);
    CIPC_HIP(hipMemsetAsync(xsol, 0, n * sizeof(double), c.s));
    const double bnorm = std::sqrt(residual_norm2(c, b, xsol, r));
    if (relres) *relres = 0;
    if (iters) *iters = 0;
    if (bnorm == 0) return 0;
    double rnorm = bnorm;
    int total = 0;
    constexpr int kMaxRefine = 8;
    // FP32 Krylov recurrences stall near 1e-6 relative; FP64 ones do not
    const double floorTol = sizeof(V) == sizeof(float) ? 1e-6 : 1e-12;
    // aim just below the target: the inner residual drifts from the true FP64 residual by
    // the rounding of the FP32 blocks, and pushing the inner solve further costs more
    // iterations than the short second sweep it would save (measured on the 320k stack:
    // the drift is additive -- the FP32 blocks' error times the solution -- so an inner
    // target below it only over-solves the sweep)
    constexpr double kInnerMargin = 0.9;
    for (int k = 0; k < kMaxRefine; ++k) {
        const double inner = std::max(floorTol, std::min(0.5, kInnerMargin * tol * bnorm / rnorm));
        const V* rhs;
        if constexpr (sizeof(V) == sizeof(double)) {
            rhs = (const V*)r;
        } else {
            hipLaunchKernelGGL(k_f64_to_f32, dim3(nblocks(n)), dim3(kBlock), 0, c.s, n, (const double*)r, rb);
            CIPC_LAUNCH_CHECK();
            rhs = rb;
        }
        int it = 0;
        double rr = 0;
        int rc = pcg_run<M, V>(c, vE, vK, rhs, d, wr, wz, wp, wq, dinv, inner, std::max(1, maxIter - total), &it, &rr, learned);
        total += it;
        c.lastRefinements = k + 1;
        // A breakdown (p.q <= 0) of the reduced-precision inner solve keeps its iterate up
        // to the last good iteration: the FP32 blocks' rounding can make a direction that H
        // barely resists indefinite, which says nothing about H itself.  The FP64
        // refinement goes on from the true residual while the sweeps make progress.
        if constexpr (sizeof(V) == sizeof(double))
            hipLaunchKernelGGL(k_acc_f64, dim3(nblocks(n)), dim3(kBlock), 0, c.s, n, (const double*)d, xsol);
        else
            hipLaunchKernelGGL(k_acc_f32, dim3(nblocks(n)), dim3(kBlock), 0, c.s, n, (const float*)d, xsol);
        CIPC_LAUNCH_CHECK();
        const double rprev = rnorm;
        rnorm = std::sqrt(residual_norm2(c, b, xsol, r));
        if (relres) *relres = rnorm / bnorm;
        if (iters) *iters = total;
        if (rnorm <= tol * bnorm) return 0;
        if (rc == 1 && !(rnorm < 0.9 * rprev)) return 1;
        // a sweep after the first that gains less than 4x: the FP32 blocks' rounding holds
        // the refinement back
        if (stallExit && k >= 1 && !(rnorm < 0.25 * rprev)) return 2;
        if (total >= maxIter) return 2;
    }
    return 2;
}

static int solve_form(Ctx& c, const double* b, double* xsol, double tol, int maxIter, int* iters, double* relres, bool learned) {
    const int nV = c.nV;
    const long n = 3l * nV;
    BSR& E = c.Ael;
    BSR& K = c.Acon;
    if (K.nnzb == 0) { K.rowptr.ensure(nV + 1); }
    c.lastRefinements = 0;
    if (c.P.precision == 0) {
        double* dinv = c.pdinv.ensure(9l * nV);
        hipLaunchKernelGGL(k_dinv<double>, dim3(nblocks(nV)), dim3(kBlock), 0, c.s, nV, E.rowptr.p, E.col.p, E.val.p, E.nnzb, K.rowptr.p,
                           K.col.p, K.val.p, K.nnzb, dinv);
        CIPC_LAUNCH_CHECK();
        return pcg_run<double, double>(c, E.val.p, K.val.p, b, xsol, c.pr.ensure(n), c.pz.ensure(n), c.pp.ensure(2 * n), c.pq.ensure(n), dinv, tol,
                                       maxIter, iters, relres, learned);
    }
    // mixed forms: FP32 copies of the blocks; FP64 iterative refinement around the solve
    float* vE = c.valE32.ensure(9l * E.nnzb + 1);
    float* vK = c.valK32.ensure(9l * K.nnzb + 1);
    hipLaunchKernelGGL(k_f64_to_f32, dim3(nblocks(9l * E.nnzb)), dim3(kBlock), 0, c.s, 9l * E.nnzb, E.val.p, vE);
    if (K.nnzb > 0) hipLaunchKernelGGL(k_f64_to_f32, dim3(nblocks(9l * K.nnzb)), dim3(kBlock), 0, c.s, 9l * K.nnzb, K.val.p, vK);
    CIPC_LAUNCH_CHECK();
    if (c.P.precision == 1) {
        double* dinv = c.pdinv.ensure(9l * nV);
        hipLaunchKernelGGL(k_dinv<double>, dim3(nblocks(nV)), dim3(kBlock), 0, c.s, nV, E.rowptr.p, E.col.p, E.val.p, E.nnzb, K.rowptr.p,
                           K.col.p, K.val.p, K.nnzb, dinv);
        CIPC_LAUNCH_CHECK();
        return refine<float, double>(c, vE, vK, dinv, b, xsol, tol, maxIter, iters, relres, c.pr.ensure(n), c.pz.ensure(n),
                                     c.pp.ensure(2 * n), c.pq.ensure(n), c.xacc.ensure(n), nullptr, learned);
    }
    float* dinv = c.fdinv.ensure(9l * nV);
    hipLaunchKernelGGL(k_dinv<float>, dim3(nblocks(nV)), dim3(kBlock), 0, c.s, nV, E.rowptr.p, E.col.p, E.val.p, E.nnzb, K.rowptr.p, K.col.p,
                       K.val.p, K.nnzb, dinv);
    CIPC_LAUNCH_CHECK();
    return refine<float, float>(c, vE, vK, dinv, b, xsol, tol, maxIter, iters, relres, c.fr.ensure(n), c.fz.ensure(n), c.fp.ensure(2 * n),
                                c.fq.ensure(n), c.fx.ensure(n), c.fb.ensure(n), learned);
}

static double norm2(Ctx& c, const double* a, long n);

// The learned preconditioner runs with an iteration budget; when it does not converge
// within it (or meets a non-positive curvature) the solve goes on with block Jacobi:
// from the learned iterate when the budget ran out and that iterate's residual is below
// half of ||b|| (the correction is solved to tol ||b|| / ||r0||, so the stopping rule
// ||b - H x|| <= tol ||b|| is unchanged), else from zero.
static int pcg_solve_impl(Ctx& c, const double* b, double* xsol, double tol, int maxIter, int* iters, double* relres) {
    c.lastPrecondFallback = 0;
    const bool learned = c.precond == 1 && !c.lpOffThisStep;
    c.lastFallbackContinued = 0;
    if (!learned) {
        int it = 0;
        const int rc = solve_form(c, b, xsol, tol, maxIter, &it, relres, false);
        if (!c.solveWarm) c.lastJacobiIters = it;   // the learned budget's cap counts from-zero solves only
        if (iters) *iters = it;
        return rc;
    }
    if (c.lpStepStart || c.lpSinceForward >= c.lpEvery || !lp_refresh(c)) {
        lp_prepare(c);
        c.lpSinceForward = 0;
        c.lpStepStart = false;
    }
    if (c.coarseOn) {
        CoarseSpace& g = c.cg;
        if (g.stepStart || ++g.sinceSetup >= g.every || !g.ready) {
            coarse_prepare(c, c.solveX ? c.solveX : c.x.p);
            g.sinceSetup = 0;
            g.stepStart = false;
        }
    }
    c.lpSinceForward++;
    c.precondSolves++;
    int it = 0;
    double rr = 0;
    int budget = std::min(maxIter, c.lpBudget);
    // (with the coarse level the learned solve is far stronger than block Jacobi: no cap)
    if (c.lastJacobiIters > 0 && !(c.coarseOn && c.cg.ready))
        budget = std::min(budget, std::max(200, c.lastJacobiIters + c.lastJacobiIters / 4));
    int rc;
    // FP64-block attempts with the coarse level get four times the budget: block Jacobi,
    // the only thing left after them, is far slower on the systems that reach them
    const int budget64 = std::min(maxIter, 4 * budget);
    const bool ranFp64 = c.fp64ThisStep && c.coarseOn && c.cg.ready && c.P.precision != 0;
    if (ranFp64) {
        // hard systems come in runs: after a retry the step's remaining solves start in FP64
        const int saved = c.P.precision;
        c.P.precision = 0;
        try {
            rc = solve_form(c, b, xsol, tol, budget64, &it, &rr, true);
        } catch (...) {
            c.P.precision = saved;
            throw;
        }
        c.P.precision = saved;
    } else {
        rc = solve_form(c, b, xsol, tol, budget, &it, &rr, true);
    }
    c.lastLearnedRc = rc;
    c.lastLearnedIters = it;
    if (rc == 0) {
        if (iters) *iters = it;
        if (relres) *relres = rr;
        return 0;
    }
    // With the coarse level, a reduced-precision solve that fails is retried with FP64
    // blocks and the same preconditioner, from the learned iterate: on the hardest contact
    // systems the FP32 blocks' rounding (about 6e-8 relative) stalls the refinement, while
    // the two-level preconditioner itself still converges in a few hundred iterations.
    if (!ranFp64 && c.coarseOn && c.cg.ready && c.P.precision != 0) {
        const long n = 3l * c.nV;
        double* xl = c.fallbackX.ensure(n);
        double* r0 = c.fallbackR.ensure(n);
        CIPC_HIP(hipMemcpyAsync(xl, xsol, n * sizeof(double), hipMemcpyDeviceToDevice, c.s));
        const double bn2 = norm2(c, b, n);
        const double rn2 = residual_norm2(c, b, xl, r0);
        const bool from = rn2 > 0 && rn2 < bn2;
        const int saved = c.P.precision;
        c.P.precision = 0;
        int it3 = 0;
        double rr3 = 0;
        int rc3;
        try {
            rc3 = from ? solve_form(c, r0, xsol, tol * std::sqrt(bn2 / rn2), budget64, &it3, &rr3, true)
                       : solve_form(c, b, xsol, tol, budget64, &it3, &rr3, true);
        } catch (...) {
            c.P.precision = saved;
            throw;
        }
        c.P.precision = saved;
        c.fp64Retries++;
        c.fp64ThisStep = true;
        if (from) {
            hipLaunchKernelGGL(k_acc_f64, dim3(nblocks(n)), dim3(kBlock), 0, c.s, n, (const double*)xl, xsol);
            CIPC_LAUNCH_CHECK();
            rr3 *= std::sqrt(rn2 / bn2);
        }
        it += it3;
        if (rc3 == 0) {
            if (iters) *iters = it;
            if (relres) *relres = rr3;
            return 0;
        }
        rc = rc3;
    }
    c.lastPrecondFallback = 1;
    c.precondFallbacks++;
    // the rest of the step stays on block Jacobi, unless the coarse level is on: then the
    // hard solve is an exception and the next ones are learned again
    if (!(c.coarseOn && c.cg.ready)) c.lpOffThisStep = true;
    int it2 = 0;
    if (rc == 2) {
        const long n = 3l * c.nV;
        double* xl = c.fallbackX.ensure(n);
        double* r0 = c.fallbackR.ensure(n);
        CIPC_HIP(hipMemcpyAsync(xl, xsol, n * sizeof(double), hipMemcpyDeviceToDevice, c.s));
        const double bn2 = norm2(c, b, n);
        const double rn2 = residual_norm2(c, b, xl, r0);
        if (rn2 > 0 && rn2 < 0.25 * bn2) {
            double rr2 = 0;
            rc = solve_form(c, r0, xsol, tol * std::sqrt(bn2 / rn2), maxIter, &it2, &rr2, false);
            hipLaunchKernelGGL(k_acc_f64, dim3(nblocks(n)), dim3(kBlock), 0, c.s, n, (const double*)xl, xsol);
            CIPC_LAUNCH_CHECK();
            if (relres) *relres = rr2 * std::sqrt(rn2 / bn2);
            // lastJacobiIters (the learned budget's cap) keeps counting from-zero solves only
            c.lastFallbackContinued = 1;
            if (iters) *iters = it + it2;
            return rc;
        }
    }
    rc = solve_form(c, b, xsol, tol, maxIter, &it2, relres, false);
    if (!c.solveWarm) c.lastJacobiIters = it2;
    if (iters) *iters = it + it2;
    return rc;
}

// Autotuner of the PCG launch shapes.  Phase 1 (SpMV grid x {16 lanes x 1 block, 8 x 1,
// 8 x 2 blocks per lane pass}; 16 x 2 was the slowest shape in every measurement): the Newton solves switch the SpMV shape every chunk of 16 iterations, in
// turn, and time it on sampled launches, so every shape is timed on the same matrices;
// once each has kTuneSamples samples the fastest mean is kept.  Phase 2 (learned
// preconditioner only: factor-product grid x lanes per row): each Newton solve uses the
// next candidate; every candidate runs twice (the list is walked twice, so slow drift of
// the matrix between solves affects all alike), costed by the solve's wall time per PCG
// iteration (stream drained before and after), and the lower of its two costs counts.
// When the number of stored Hessian blocks later moves more than 5% from the tuned
// matrix's (contacts forming or releasing), the tuning reruns, at most kRetunes times.
constexpr int kRetunes = 2;
constexpr long kTuneSamples = 32;

static void tune_phase(Ctx& c, int phase) {
    Ctx::Tuner& T = c.tune;
    T.pcgCand.clear();
    if (phase == 1) {
        T.spmvCand.clear();
        for (int sg : {1024, 2048})
            for (auto ls : {std::array<int, 2>{16, 1}, {8, 1}, {8, 2}}) T.spmvCand.push_back({sg, ls[0], ls[1]});
        T.spmvMs.assign(T.spmvCand.size(), 0.0);
        T.spmvN.assign(T.spmvCand.size(), 0);
        T.spmvTurn = 0;
    } else {
        for (int pass = 0; pass < 2; ++pass)
            for (int lg : {1024, 2048})
                for (int ln : {2, 4, 8}) T.pcgCand.push_back({lg, ln});
        T.pcgCost.assign(T.pcgCand.size(), 0.0);
        T.pcgIdx = 0;
    }
    T.pcgPhase = phase;
    T.pcg = true;
}

static void tune_done(Ctx& c, long nnzb) {
    Ctx::Tuner& T = c.tune;
    T.pcg = false;
    T.tunedNnzb = nnzb;
    if (T.pcgPhase == 1 && c.precond == 1) tune_phase(c, 2);
}

static int pcg_solve_tuned(Ctx& c, const double* b, double* xsol, double tol, int maxIter, int* iters, double* relres) {
    Ctx::Tuner& T = c.tune;
    const long nnzb = (long)c.Ael.nnzb + c.Acon.nnzb;
    if (!T.pcg && T.enabled && T.retunes > 0 && std::fabs((double)(nnzb - T.tunedNnzb)) > 0.05 * (double)T.tunedNnzb) {
        T.retunes--;
        tune_phase(c, 1);
    }
    if (!T.pcg) return pcg_solve_impl(c, b, xsol, tol, maxIter, iters, relres);
    if (T.pcgPhase == 1) {
        const int rc = pcg_solve_impl(c, b, xsol, tol, maxIter, iters, relres);
        long nMin = T.spmvN[0];
        for (long v : T.spmvN) nMin = std::min(nMin, v);
        if (nMin >= kTuneSamples) {
            int best = 0;
            for (int i = 1; i < (int)T.spmvCand.size(); ++i)
                if (T.spmvMs[i] / (double)T.spmvN[i] < T.spmvMs[best] / (double)T.spmvN[best]) best = i;
            c.spmvGrid = T.spmvCand[best][0]; c.spmvLanes = T.spmvCand[best][1]; c.spmvUnroll = T.spmvCand[best][2];
            tune_done(c, nnzb);
        }
        return rc;
    }
    c.lpGrid = T.pcgCand[T.pcgIdx][0];
    c.lpLanes = T.pcgCand[T.pcgIdx][1];
    CIPC_HIP(hipStreamSynchronize(c.s));
    const auto t0 = std::chrono::steady_clock::now();
    int it = 0;
    const int rc = pcg_solve_impl(c, b, xsol, tol, maxIter, &it, relres);
    CIPC_HIP(hipStreamSynchronize(c.s));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (iters) *iters = it;
    T.pcgCost[T.pcgIdx] = it > 0 ? ms / it : 1e300;
    if (++T.pcgIdx == (int)T.pcgCand.size()) {
        const int half = (int)T.pcgCand.size() / 2;
        int best = 0;
        for (int i = 1; i < half; ++i)
            if (std::min(T.pcgCost[i], T.pcgCost[i + half]) < std::min(T.pcgCost[best], T.pcgCost[best + half])) best = i;
        c.lpGrid = T.pcgCand[best][0];
        c.lpLanes = T.pcgCand[best][1];
        tune_done(c, nnzb);
    }
    return rc;
}

__global__ void k_sq_partial(long n, const double* __restrict__ a, double* __restrict__ part) {
    __shared__ double lds[kBlock / 64];
    double v = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) v += a[i] * a[i];
    const double s = block_allsum<kBlock>(v, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
static double norm2(Ctx& c, const double* a, long n) {
    const int nb = std::min(nblocks(n), 1024);
    double* part = c.dscr2.ensure((size_t)nb + 1);
    hipLaunchKernelGGL(k_sq_partial, dim3(nb), dim3(kBlock), 0, c.s, n, a, part);
    CIPC_LAUNCH_CHECK();
    sum_partials(part, nb, c.scal.p + 41, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hScal + 41, c.scal.p + 41, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return c.hScal[41];
}

// With an initial guess x0: r0 = b - H x0 in FP64; when ||r0|| < ||b|| / 2 the solve runs
// on H d = r0 to the tolerance that gives ||b - H (x0 + d)|| <= tol ||b||, and x = x0 + d;
// otherwise x0 is ignored.  The stopping rule is the same as from a zero start.
static int pcg_solve_from(Ctx& c, const double* b, double* xsol, double tol, int maxIter, int* iters, double* relres, const double* x0) {
    if (!x0) return pcg_solve_tuned(c, b, xsol, tol, maxIter, iters, relres);
    const long n = 3l * c.nV;
    if (c.Acon.nnzb == 0) c.Acon.rowptr.ensure(c.nV + 1);
    double* r0 = c.warmR.ensure(n);
    const double bn2 = norm2(c, b, n);
    const double rn2 = residual_norm2(c, b, x0, r0);
    c.lastWarmRatio = bn2 > 0 ? std::sqrt(rn2 / bn2) : 1.0;
    if (!(rn2 < 0.25 * bn2)) return pcg_solve_tuned(c, b, xsol, tol, maxIter, iters, relres);
    double rr = 0;
    c.solveWarm = true;   // a correction solve: its iteration count is not a from-zero count
    int rc;
    try {
        rc = pcg_solve_tuned(c, r0, xsol, tol * std::sqrt(bn2 / rn2), maxIter, iters, &rr);
    } catch (...) {
        c.solveWarm = false;
        throw;
    }
    c.solveWarm = false;
    hipLaunchKernelGGL(k_acc_f64, dim3(nblocks(n)), dim3(kBlock), 0, c.s, n, x0, xsol);
    CIPC_LAUNCH_CHECK();
    if (relres) *relres = rr * std::sqrt(rn2 / bn2);
    return rc;
}

// dense direct solve and its FP64 relative residual ||b - H x|| / ||b||
static int direct_checked(Ctx& c, const double* b, double* xsol, double* relres) {
    const long n = 3l * c.nV;
    const int rc = direct_solve(c, b, xsol);
    if (rc == 0 && relres) {
        const double bn2 = norm2(c, b, n);
        const double rn2 = residual_norm2(c, b, xsol, c.warmR.ensure(n));
        *relres = bn2 > 0 ? std::sqrt(rn2 / bn2) : 0.0;
    }
    return rc;
}

// The reference factorises every Newton matrix (Solve_Direct, DIRECT_SOLVER.h:14-66).
// With linSolver 1 the dense direct solve does the same here; with linSolver 0 the PCG
// solves, and the direct solve takes over a solve that reached maxIter when the system
// has at most kDirectMaxDof unknowns.  A non-positive pivot returns 1, as the reference's
// CHOLMOD_NOT_POSDEF, and the step turns to the gradient direction.
int pcg_solve(Ctx& c, const double* b, double* xsol, double tol, int maxIter, int* iters, double* relres, const double* x0) {
    const long n = 3l * c.nV;
    if (c.Acon.nnzb == 0) c.Acon.rowptr.ensure(c.nV + 1);
    if (c.P.linSolver == 1) {
        if (iters) *iters = 0;
        return direct_checked(c, b, xsol, relres);
    }
    int it = 0;
    const int rc = pcg_solve_from(c, b, xsol, tol, maxIter, &it, relres, x0);
    if (iters) *iters = it;
    if (rc != 2 || n > kDirectMaxDof) return rc;
    return direct_checked(c, b, xsol, relres);
}

void autotune_start(Ctx& c, int what) {
    Ctx::Tuner& T = c.tune;
    if (what & 1) {
        T.enabled = true;
        T.retunes = kRetunes;
        tune_phase(c, 1);
    }
    if (what & 2) {
        T.cellCand = {0.5, 0.7, 1.0, 1.4, 2.0, 0.5, 0.7, 1.0, 1.4, 2.0};
        T.cellCost.assign(T.cellCand.size(), 0.0);
        T.cellIdx = 0;
        T.cell = true;
    }
}

}  // namespace cipc
