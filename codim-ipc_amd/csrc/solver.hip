// Newton linear solve: block-Jacobi preconditioned conjugate gradient on the
// assembled Hessian H = A_elastic (static BSR pattern, mass and Dirichlet projection
// included) + A_contact (per-Newton-iteration BSR).  Stands in for the reference's
// direct solve (Math/DIRECT_SOLVER.h:14-66, CHOLMOD); the relative-residual stopping
// rule ||r|| <= tol ||b|| follows the reference's iterative-solver precedent
// (AMGCL_SOLVER.h:53-65, rel-tol 1e-5).
//
// SpMV: a group of LANES lanes per block row, each lane strides over the row's
// blocks; blocks are stored structure-of-arrays (entry k of all blocks contiguous)
// so the lanes of a wave read consecutive addresses; the group reduces with
// cross-lane shuffles.  Scalars (alpha, beta, residual) never leave the device
// inside an iteration: dot products finish in the last block to arrive (ticket
// counter, partials summed in a fixed order), and every kernel returns early once
// the on-device "done" flag is set, so the host only polls every few iterations.
#include "ctx.hpp"
#include <cmath>

namespace cipc {

enum { S_RZ = 0, S_PQ, S_RZNEW, S_RR, S_BETA, S_BNORM, S_TOL, S_ALPHA };
enum { F_DONE = 0, F_ITERS, F_TICKET, F_FAIL, F_MAXIT };

template <int LANES>
__global__ void k_spmv(const int* __restrict__ rowptr, const int* __restrict__ col, const double* __restrict__ val, int nnzb, int nrow,
                       const double* __restrict__ p, double* __restrict__ q, int accumulate, const int* __restrict__ flags) {
    if (flags && flags[F_DONE]) return;
    long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int row = (int)(gid / LANES), lane = (int)(gid % LANES);
    double s0 = 0, s1 = 0, s2 = 0;
    if (row < nrow) {
        int e = rowptr[row + 1];
        for (int b = rowptr[row] + lane; b < e; b += LANES) {
            int c = col[b];
            double x0 = p[3 * c], x1 = p[3 * c + 1], x2 = p[3 * c + 2];
            s0 += val[b] * x0 + val[(long)nnzb + b] * x1 + val[2l * nnzb + b] * x2;
            s1 += val[3l * nnzb + b] * x0 + val[4l * nnzb + b] * x1 + val[5l * nnzb + b] * x2;
            s2 += val[6l * nnzb + b] * x0 + val[7l * nnzb + b] * x1 + val[8l * nnzb + b] * x2;
        }
    }
#pragma unroll
    for (int o = LANES / 2; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, LANES);
        s1 += __shfl_xor(s1, o, LANES);
        s2 += __shfl_xor(s2, o, LANES);
    }
    if (row < nrow && lane == 0) {
        if (accumulate) { q[3 * row] += s0; q[3 * row + 1] += s1; q[3 * row + 2] += s2; }
        else { q[3 * row] = s0; q[3 * row + 1] = s1; q[3 * row + 2] = s2; }
    }
}

// finish a block reduction of up to two values; the last block to arrive sums the
// per-block partials in a fixed order and calls fin(total0, total1)
template <class Fin>
__device__ __forceinline__ void last_block_finish(double v0, double v1, double* part, int* flags, Fin fin) {
    __shared__ double lds[4];
    __shared__ int amLast;
    double b0 = block_sum_256(v0, lds);
    double b1 = block_sum_256(v1, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = b0;
        part[2 * blockIdx.x + 1] = b1;
        __threadfence();
        int t = atomicAdd(&flags[F_TICKET], 1);
        amLast = (t == (int)gridDim.x - 1);
    }
    __syncthreads();
    if (!amLast) return;
    __threadfence();
    double a0 = 0, a1 = 0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) {
        // agent-scope atomic loads: the partials come from other CUs and must not be
        // served from this CU's (non-coherent) vector L1
        a0 += __hip_atomic_load(&part[2 * i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a1 += __hip_atomic_load(&part[2 * i + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    double t0 = block_sum_256(a0, lds);
    double t1 = block_sum_256(a1, lds);
    if (threadIdx.x == 0) {
        flags[F_TICKET] = 0;
        fin(t0, t1);
    }
}

// row-wise: q += A_con p (over all rows), then pq = p . q
__global__ void k_spmv_con_dot(const int* __restrict__ rowptr, const int* __restrict__ col, const double* __restrict__ val, int nnzb,
                               int nrow, const double* __restrict__ p, double* __restrict__ q, double* __restrict__ sc,
                               double* __restrict__ part, int* __restrict__ flags) {
    if (flags[F_DONE]) return;
    constexpr int LANES = 4;
    long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int row = (int)(gid / LANES), lane = (int)(gid % LANES);
    double s0 = 0, s1 = 0, s2 = 0;
    if (row < nrow && nnzb > 0) {
        int e = rowptr[row + 1];
        for (int b = rowptr[row] + lane; b < e; b += LANES) {
            int c = col[b];
            double x0 = p[3 * c], x1 = p[3 * c + 1], x2 = p[3 * c + 2];
            s0 += val[b] * x0 + val[(long)nnzb + b] * x1 + val[2l * nnzb + b] * x2;
            s1 += val[3l * nnzb + b] * x0 + val[4l * nnzb + b] * x1 + val[5l * nnzb + b] * x2;
            s2 += val[6l * nnzb + b] * x0 + val[7l * nnzb + b] * x1 + val[8l * nnzb + b] * x2;
        }
    }
#pragma unroll
    for (int o = LANES / 2; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, LANES);
        s1 += __shfl_xor(s1, o, LANES);
        s2 += __shfl_xor(s2, o, LANES);
    }
    double d = 0;
    if (row < nrow && lane == 0) {
        double q0 = q[3 * row] + s0, q1 = q[3 * row + 1] + s1, q2 = q[3 * row + 2] + s2;
        q[3 * row] = q0; q[3 * row + 1] = q1; q[3 * row + 2] = q2;
        d = p[3 * row] * q0 + p[3 * row + 1] * q1 + p[3 * row + 2] * q2;
    }
    last_block_finish(d, 0.0, part, flags, [&](double t0, double) { sc[S_PQ] = t0; });
}

__global__ void k_pcg_update(int nV, double* __restrict__ x, double* __restrict__ r, double* __restrict__ z, const double* __restrict__ p,
                             const double* __restrict__ q, const double* __restrict__ dinv, double* __restrict__ sc,
                             double* __restrict__ part, int* __restrict__ flags) {
    if (flags[F_DONE]) return;
    double pq = sc[S_PQ];
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0, rr = 0;
    if (pq > 0) {
        double alpha = sc[S_RZ] / pq;
        if (v < nV) {
            double rv[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                x[3 * v + k] += alpha * p[3 * v + k];
                rv[k] = r[3 * v + k] - alpha * q[3 * v + k];
                r[3 * v + k] = rv[k];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double zk = dinv[9 * v + 3 * k] * rv[0] + dinv[9 * v + 3 * k + 1] * rv[1] + dinv[9 * v + 3 * k + 2] * rv[2];
                z[3 * v + k] = zk;
                rz += rv[k] * zk;
                rr += rv[k] * rv[k];
            }
        }
    }
    last_block_finish(rz, rr, part, flags, [&](double t0, double t1) {
        int it = flags[F_ITERS] + 1;
        flags[F_ITERS] = it;
        if (!(pq > 0)) { flags[F_FAIL] = 1; flags[F_DONE] = 1; return; }
        sc[S_BETA] = t0 / sc[S_RZ];
        sc[S_RZ] = t0;
        sc[S_RR] = t1;
        if (sqrt(t1) <= sc[S_TOL] * sc[S_BNORM] || it >= flags[F_MAXIT]) flags[F_DONE] = 1;
    });
}
__global__ void k_pcg_p(int n, double* __restrict__ p, const double* __restrict__ z, const double* __restrict__ sc, const int* __restrict__ flags) {
    if (flags[F_DONE]) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + sc[S_BETA] * p[i];
}

// r = b, x = 0, z = Dinv r, p = z; partial sums of r.z and b.b
__global__ void k_pcg_init(int nV, const double* __restrict__ b, double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                           double* __restrict__ p, const double* __restrict__ dinv, double* __restrict__ sc, double* __restrict__ part,
                           int* __restrict__ flags) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0, bb = 0;
    if (v < nV) {
        double rv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) { rv[k] = b[3 * v + k]; r[3 * v + k] = rv[k]; x[3 * v + k] = 0; bb += rv[k] * rv[k]; }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double zk = dinv[9 * v + 3 * k] * rv[0] + dinv[9 * v + 3 * k + 1] * rv[1] + dinv[9 * v + 3 * k + 2] * rv[2];
            z[3 * v + k] = zk; p[3 * v + k] = zk;
            rz += rv[k] * zk;
        }
    }
    last_block_finish(rz, bb, part, flags, [&](double t0, double t1) {
        sc[S_RZ] = t0;
        sc[S_RR] = t1;
        sc[S_BNORM] = sqrt(t1);
        flags[F_ITERS] = 0;
        flags[F_FAIL] = 0;
        flags[F_DONE] = (t1 == 0.0) ? 1 : 0;
    });
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

void spmv_full(Ctx& c, const double* p, double* q) {
    BSR& E = c.Ael;
    BSR& K = c.Acon;
    hipLaunchKernelGGL(k_spmv<8>, dim3(nblocks(8l * E.nrow)), dim3(kBlock), 0, c.s, E.rowptr.p, E.col.p, E.val.p, E.nnzb, E.nrow, p, q, 0,
                       (const int*)nullptr);
    CIPC_LAUNCH_CHECK();
    if (K.nnzb > 0) {
        hipLaunchKernelGGL(k_spmv<4>, dim3(nblocks(4l * K.nrow)), dim3(kBlock), 0, c.s, K.rowptr.p, K.col.p, K.val.p, K.nnzb, K.nrow, p, q,
                           1, (const int*)nullptr);
        CIPC_LAUNCH_CHECK();
    }
}

// returns 0 converged, 1 not SPD (p^T A p <= 0), 2 iteration limit
int pcg_solve(Ctx& c, const double* b, double* xsol, double tol, int maxIter, int* iters, double* relres) {
    const int nV = c.nV, n = 3 * nV;
    BSR& E = c.Ael;
    BSR& K = c.Acon;
    if (K.nnzb == 0) { K.rowptr.ensure(nV + 1); }
    double* r = c.pr.ensure(n);
    double* z = c.pz.ensure(n);
    double* p = c.pp.ensure(n);
    double* q = c.pq.ensure(n);
    double* dinv = c.pdinv.ensure(9l * nV);
    double* sc = c.scal.p + 32;               // 16 solver scalars
    int* flags = (int*)(c.counters.p + 32);   // 16 solver flags
    int nbv = nblocks(nV);
    int nbr = nblocks(4l * nV);
    double* part = c.dscr.ensure(2l * std::max(nbv, nbr) + 2);
    hipLaunchKernelGGL(k_dinv, dim3(nbv), dim3(kBlock), 0, c.s, nV, E.rowptr.p, E.col.p, E.val.p, E.nnzb, K.rowptr.p, K.col.p, K.val.p, K.nnzb,
                       dinv);
    CIPC_LAUNCH_CHECK();
    int hflags[5] = {0, 0, 0, 0, maxIter};
    CIPC_HIP(hipMemcpyAsync(flags, hflags, sizeof(hflags), hipMemcpyHostToDevice, c.s));
    double htol = tol;
    CIPC_HIP(hipMemcpyAsync(sc + S_TOL, &htol, sizeof(double), hipMemcpyHostToDevice, c.s));
    hipLaunchKernelGGL(k_pcg_init, dim3(nbv), dim3(kBlock), 0, c.s, nV, b, xsol, r, z, p, dinv, sc, part, flags);
    CIPC_LAUNCH_CHECK();
    const int chunk = 16;
    int* hf = (int*)(c.hCnt + 32);
    for (;;) {
        for (int k = 0; k < chunk; ++k) {
            hipLaunchKernelGGL(k_spmv<8>, dim3(nblocks(8l * nV)), dim3(kBlock), 0, c.s, E.rowptr.p, E.col.p, E.val.p, E.nnzb, nV, p, q, 0,
                               (const int*)flags);
            hipLaunchKernelGGL(k_spmv_con_dot, dim3(nbr), dim3(kBlock), 0, c.s, K.rowptr.p, K.col.p, K.val.p, K.nnzb, nV, p, q, sc, part, flags);
            hipLaunchKernelGGL(k_pcg_update, dim3(nbv), dim3(kBlock), 0, c.s, nV, xsol, r, z, p, q, dinv, sc, part, flags);
            hipLaunchKernelGGL(k_pcg_p, dim3(nblocks(n)), dim3(kBlock), 0, c.s, n, p, z, sc, flags);
        }
        CIPC_LAUNCH_CHECK();
        CIPC_HIP(hipMemcpyAsync(hf, flags, 4 * sizeof(int), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        if (hf[F_DONE]) break;
    }
    CIPC_HIP(hipMemcpyAsync(c.hScal + 32, sc, 8 * sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    if (iters) *iters = hf[F_ITERS];
    double bn = c.hScal[32 + S_BNORM];
    if (relres) *relres = bn > 0 ? std::sqrt(c.hScal[32 + S_RR]) / bn : 0.0;
    if (hf[F_FAIL]) return 1;
    if (bn > 0 && std::sqrt(c.hScal[32 + S_RR]) > tol * bn) return 2;
    return 0;
}

}  // namespace cipc
