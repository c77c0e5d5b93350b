// Newton linear solve: block-Jacobi preconditioned conjugate gradient on the
// assembled Hessian H = A_elastic (static BSR pattern, mass and Dirichlet projection
// included) + A_contact (per-Newton-iteration BSR).  Stands in for the reference's
// direct solve (Math/DIRECT_SOLVER.h:14-66, CHOLMOD); the relative-residual stopping
// rule ||r|| <= tol ||b|| follows the reference's iterative-solver precedent
// (AMGCL_SOLVER.h:53-65, rel-tol 1e-5).
//
// One iteration is two kernels and no atomics:
//   k_pcg_spmv   q = A (z + beta p_old) for every block row, p_new written, partial
//                sums of p_new . q per workgroup;
//   k_pcg_update x += alpha p, r -= alpha q, z = Dinv r, partial sums of r.z, r.r.
// Each kernel starts by summing the previous kernel's per-workgroup partials itself
// (every workgroup, same fixed order, so every workgroup holds the identical scalar);
// kernel boundaries on the one stream order the partials' writes before those reads,
// so no fence or ticket counter is needed.  The SpMV keeps A z and A p_old in separate
// accumulators so the scalar reduction's latency overlaps the matrix stream, and beta
// is applied when the row is finished.
//
// Matrix blocks are stored structure-of-arrays (entry k of all blocks contiguous);
// LANES lanes share a block row and stride over its blocks, so the lanes of a wave
// read consecutive addresses.  Rows are mapped XCD-aware: workgroup b runs on XCD
// b % 8 and XCD x owns one contiguous eighth of the rows, so the gathered neighbour
// values stay in that XCD's L2.
//
// The host launches iterations in chunks and polls an on-device done flag between
// chunks; kernels launched after convergence return at once.
#include "ctx.hpp"
#include <hip/hip_ext.h>
#include <cmath>

namespace cipc {

// solver scalars (sc[]) and flags
enum { S_RZ0 = 0, S_RZ1, S_BB, S_RR };
enum { F_DONE = 0, F_ITERS, F_FAIL };

constexpr int kSpmvLanes = 8;
constexpr int kSpmvBlock = 512;
constexpr int kUpdBlock = 256;
constexpr int kXcd = 8;

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

// rows per XCD, rounded up to whole workgroups of `rowsPerBlock`
inline int rows_per_xcd(int nrow, int rowsPerBlock) {
    int r = (nrow + kXcd - 1) / kXcd;
    return (r + rowsPerBlock - 1) / rowsPerBlock * rowsPerBlock;
}
__device__ __forceinline__ int xcd_row_base(int rowsPerXcd, int rowsPerBlock) {
    return (int)(blockIdx.x % kXcd) * rowsPerXcd + (int)(blockIdx.x / kXcd) * rowsPerBlock;
}

template <int LANES>
__device__ __forceinline__ void bsr_row_acc(const int* __restrict__ rp, const int* __restrict__ col, const double* __restrict__ val,
                                            long nnzb, int row, int lane, const double* __restrict__ z,
                                            const double* __restrict__ p, bool withP, double* az, double* ap) {
    const int e = rp[row + 1];
    for (int b = rp[row] + lane; b < e; b += LANES) {
        const int c = col[b];
        double a[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) a[k] = val[k * nnzb + b];
        const double z0 = z[3 * c], z1 = z[3 * c + 1], z2 = z[3 * c + 2];
        az[0] += a[0] * z0 + a[1] * z1 + a[2] * z2;
        az[1] += a[3] * z0 + a[4] * z1 + a[5] * z2;
        az[2] += a[6] * z0 + a[7] * z1 + a[8] * z2;
        if (withP) {
            const double p0 = p[3 * c], p1 = p[3 * c + 1], p2 = p[3 * c + 2];
            ap[0] += a[0] * p0 + a[1] * p1 + a[2] * p2;
            ap[1] += a[3] * p0 + a[4] * p1 + a[5] * p2;
            ap[2] += a[6] * p0 + a[7] * p1 + a[8] * p2;
        }
    }
}

// Iteration `it`: p_new = z + beta p_old (p_new = z at it = 0), q = A p_new, partial
// p_new . q.  Convergence of the residual left by the previous update is decided here.
template <int LANES, int BS>
__global__ __launch_bounds__(BS) void k_pcg_spmv(int it, int maxIt, double tol2, const int* __restrict__ erp,
                                                 const int* __restrict__ ecol, const double* __restrict__ ev, int enz,
                                                 const int* __restrict__ crp, const int* __restrict__ ccol,
                                                 const double* __restrict__ cv, int cnz, int nrow, int rowsPerXcd,
                                                 const double* __restrict__ z, const double* __restrict__ pold,
                                                 double* __restrict__ pnew, double* __restrict__ q,
                                                 const double* __restrict__ partU, int nU, double* __restrict__ partS,
                                                 double* __restrict__ sc, int* __restrict__ flags) {
    if (flags[F_DONE]) return;
    __shared__ double lds[BS / 64];
    double arz = 0, arr = 0;
    for (int i = threadIdx.x; i < nU; i += BS) { arz += partU[i]; arr += partU[nU + i]; }

    constexpr int RPB = BS / LANES;
    const int row = xcd_row_base(rowsPerXcd, RPB) + (int)threadIdx.x / LANES;
    const int lane = (int)threadIdx.x % LANES;
    const bool withP = it > 0;
    double az[3] = {0, 0, 0}, ap[3] = {0, 0, 0};
    if (row < nrow) {
        bsr_row_acc<LANES>(erp, ecol, ev, enz, row, lane, z, pold, withP, az, ap);
        if (cnz > 0) bsr_row_acc<LANES>(crp, ccol, cv, cnz, row, lane, z, pold, withP, az, ap);
    }
#pragma unroll
    for (int o = LANES / 2; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            az[k] += __shfl_xor(az[k], o, LANES);
            ap[k] += __shfl_xor(ap[k], o, LANES);
        }

    const double rz = block_allsum<BS>(arz, lds);
    const double rr = block_allsum<BS>(arr, lds);
    const double bb = it == 0 ? rr : sc[S_BB];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (rr <= tol2 * bb || it >= maxIt) {
        if (lead) {
            if (it == 0) sc[S_BB] = rr;
            sc[S_RR] = rr;
            flags[F_ITERS] = it;
            flags[F_DONE] = 1;
        }
        return;
    }
    const double beta = withP ? rz / sc[S_RZ0 + ((it - 1) & 1)] : 0.0;
    double pq = 0;
    if (row < nrow && lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double zk = z[3 * row + k];
            const double pk = withP ? zk + beta * pold[3 * row + k] : zk;
            const double qk = withP ? az[k] + beta * ap[k] : az[k];
            pnew[3 * row + k] = pk;
            q[3 * row + k] = qk;
            pq += pk * qk;
        }
    }
    const double bpq = block_allsum<BS>(pq, lds);
    if (threadIdx.x == 0) partS[blockIdx.x] = bpq;
    if (lead) {
        sc[S_RZ0 + (it & 1)] = rz;
        if (it == 0) sc[S_BB] = rr;
    }
}

// Iteration `it`: alpha = r.z / p.q, x += alpha p, r -= alpha q, z = Dinv r, partials
// of r.z and r.r (written as partU[block], partU[nU + block]).
template <int BS>
__global__ __launch_bounds__(BS) void k_pcg_update(int it, int nrow, int rowsPerXcd, double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ z, const double* __restrict__ p,
                                                   const double* __restrict__ q, const double* __restrict__ dinv,
                                                   const double* __restrict__ partS, int nS, double* __restrict__ partU,
                                                   double* __restrict__ sc, int* __restrict__ flags) {
    if (flags[F_DONE]) return;
    __shared__ double lds[BS / 64];
    double a = 0;
    for (int i = threadIdx.x; i < nS; i += BS) a += partS[i];
    const int row = xcd_row_base(rowsPerXcd, BS) + (int)threadIdx.x;
    const bool valid = row < nrow;
    double pv[3] = {0, 0, 0}, qv[3] = {0, 0, 0}, xv[3] = {0, 0, 0}, rv[3] = {0, 0, 0}, D[9];
    if (valid) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            pv[k] = p[3 * row + k]; qv[k] = q[3 * row + k];
            xv[k] = x[3 * row + k]; rv[k] = r[3 * row + k];
        }
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
    const double alpha = sc[S_RZ0 + (it & 1)] / pq;
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
            const double zk = D[3 * k] * rv[0] + D[3 * k + 1] * rv[1] + D[3 * k + 2] * rv[2];
            z[3 * row + k] = zk;
            rz += rv[k] * zk;
            rr += rv[k] * rv[k];
        }
    }
    const double brz = block_allsum<BS>(rz, lds);
    const double brr = block_allsum<BS>(rr, lds);
    if (threadIdx.x == 0) {
        partU[blockIdx.x] = brz;
        partU[gridDim.x + blockIdx.x] = brr;
    }
}

// r = b, x = 0, z = Dinv r; partials of r.z and r.r in the update kernel's layout;
// the solver flags are cleared.
template <int BS>
__global__ __launch_bounds__(BS) void k_pcg_init(int nrow, int rowsPerXcd, const double* __restrict__ b, double* __restrict__ x,
                                                 double* __restrict__ r, double* __restrict__ z, const double* __restrict__ dinv,
                                                 double* __restrict__ partU, int* __restrict__ flags) {
    __shared__ double lds[BS / 64];
    const int row = xcd_row_base(rowsPerXcd, BS) + (int)threadIdx.x;
    double rz = 0, rr = 0;
    if (row < nrow) {
        double rv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            rv[k] = b[3 * row + k];
            r[3 * row + k] = rv[k];
            x[3 * row + k] = 0;
            rr += rv[k] * rv[k];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double zk = dinv[9l * row + 3 * k] * rv[0] + dinv[9l * row + 3 * k + 1] * rv[1] + dinv[9l * row + 3 * k + 2] * rv[2];
            z[3 * row + k] = zk;
            rz += rv[k] * zk;
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

// diagonal 3x3 blocks of A_el + A_con, inverted (symmetric adjugate)
__global__ void k_dinv(int nV, const int* __restrict__ rpe, const int* __restrict__ ce, const double* __restrict__ ve, int nze,
                       const int* __restrict__ rpc, const int* __restrict__ cc, const double* __restrict__ vc, int nzc,
                       double* __restrict__ dinv) {
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
    for (int k = 0; k < 9; ++k) dinv[9 * v + k] = r[k];
}

// returns 0 converged, 1 not SPD (p^T A p <= 0), 2 iteration limit
int pcg_solve(Ctx& c, const double* b, double* xsol, double tol, int maxIter, int* iters, double* relres) {
    const int nV = c.nV, n = 3 * nV;
    BSR& E = c.Ael;
    BSR& K = c.Acon;
    if (K.nnzb == 0) { K.rowptr.ensure(nV + 1); }
    double* r = c.pr.ensure(n);
    double* z = c.pz.ensure(n);
    double* pbuf = c.pp.ensure(2l * n);       // p_old / p_new, alternating by iteration
    double* q = c.pq.ensure(n);
    double* dinv = c.pdinv.ensure(9l * nV);
    double* sc = c.scal.p + 32;               // solver scalars
    int* flags = (int*)(c.counters.p + 32);   // solver flags

    constexpr int RPB = kSpmvBlock / kSpmvLanes;
    const int rpxS = rows_per_xcd(nV, RPB), nS = kXcd * (rpxS / RPB);
    const int rpxU = rows_per_xcd(nV, kUpdBlock), nU = kXcd * (rpxU / kUpdBlock);
    double* partS = c.dscr.ensure((size_t)nS + 2l * nU);
    double* partU = partS + nS;

    hipLaunchKernelGGL(k_dinv, dim3(nblocks(nV)), dim3(kBlock), 0, c.s, nV, E.rowptr.p, E.col.p, E.val.p, E.nnzb, K.rowptr.p, K.col.p,
                       K.val.p, K.nnzb, dinv);
    CIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_pcg_init<kUpdBlock>, dim3(nU), dim3(kUpdBlock), 0, c.s, nV, rpxU, b, xsol, r, z, dinv, partU, flags);
    CIPC_LAUNCH_CHECK();

    const int chunk = 16;
    int* hf = (int*)(c.hCnt + 32);
    if (c.profiling && c.evPool.size() < 2 * chunk) {
        for (size_t i = c.evPool.size(); i < 2 * chunk; ++i) {
            hipEvent_t e;
            CIPC_HIP(hipEventCreate(&e));
            c.evPool.push_back(e);
        }
    }
    const double tol2 = tol * tol;
    const double spmvBytes = 76.0 * E.nnzb + 4.0 * (nV + 1) + (K.nnzb > 0 ? 76.0 * K.nnzb + 4.0 * (nV + 1) : 0.0);
    int it = 0;
    for (;;) {
        int launched = 0;
        for (int k = 0; k < chunk && it <= maxIter; ++k, ++it, ++launched) {
            const double* pold = pbuf + (size_t)((it + 1) & 1) * n;
            double* pnew = pbuf + (size_t)(it & 1) * n;
            if (c.profiling)
                hipExtLaunchKernelGGL(k_pcg_spmv<kSpmvLanes, kSpmvBlock>, dim3(nS), dim3(kSpmvBlock), 0, c.s, c.evPool[2 * k],
                                      c.evPool[2 * k + 1], 0, it, maxIter, tol2, E.rowptr.p, E.col.p, E.val.p, E.nnzb, K.rowptr.p, K.col.p,
                                      K.val.p, K.nnzb, nV, rpxS, z, pold, pnew, q, partU, nU, partS, sc, flags);
            else
                hipLaunchKernelGGL((k_pcg_spmv<kSpmvLanes, kSpmvBlock>), dim3(nS), dim3(kSpmvBlock), 0, c.s, it, maxIter, tol2, E.rowptr.p,
                                   E.col.p, E.val.p, E.nnzb, K.rowptr.p, K.col.p, K.val.p, K.nnzb, nV, rpxS, z, pold, pnew, q, partU, nU,
                                   partS, sc, flags);
            hipLaunchKernelGGL(k_pcg_update<kUpdBlock>, dim3(nU), dim3(kUpdBlock), 0, c.s, it, nV, rpxU, xsol, r, z, (const double*)pnew, q,
                               dinv, partS, nS, partU, sc, flags);
        }
        CIPC_LAUNCH_CHECK();
        CIPC_HIP(hipMemcpyAsync(hf, flags, 3 * sizeof(int), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        if (c.profiling) {
            // SpMV launches of iterations before the one that detected convergence did
            // the full work; the detecting launch and later ones return early
            const int done = hf[F_DONE] ? hf[F_ITERS] : it;
            const int base = it - launched;
            for (int k = 0; k < launched && base + k < done; ++k) {
                float ms = 0;
                CIPC_HIP(hipEventElapsedTime(&ms, c.evPool[2 * k], c.evPool[2 * k + 1]));
                c.profMs += ms;
                c.profLaunches++;
                c.profBytes += spmvBytes + 96.0 * nV;
                c.profFlops += 18.0 * (E.nnzb + K.nnzb) + 8.0 * n;
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

}  // namespace cipc
