// Learned factorised-approximate-inverse preconditioner: graph, features, network
// forward/backward, Kaporin loss, Adam, and the preconditioner product used by the PCG.
// See learn.hpp for the model.  No counterpart in the reference (it factorises with
// CHOLMOD, Math/DIRECT_SOLVER.h:14-66); the block-Jacobi PCG stays the fallback.
//
// All reductions are deterministic: per-node sums walk the node's edges in CSR order,
// weight gradients are summed per workgroup over a fixed row range and the workgroup
// partials are added in index order.
#include "learn.hpp"
#include <cstdlib>
#include <algorithm>
#include <cmath>
#include <random>

namespace cipc {

LpLayout::LpLayout() {
    int o = 0;
    auto take = [&](int n) { int r = o; o += n; return r; };
    we = take(kLpNodeIn * kLpHidden); be = take(kLpHidden);
    wx = take(kLpEdgeIn * kLpHidden); bx = take(kLpHidden);
    for (int l = 0; l < kLpLayers; ++l) {
        wm[l] = take(3 * kLpHidden * kLpHidden); bm[l] = take(kLpHidden);
        wu[l] = take(2 * kLpHidden * kLpHidden); bu[l] = take(kLpHidden);
    }
    wo = take(3 * kLpHidden * 9); bo = take(9);
    wd = take(kLpHidden * 6); bd = take(6);
    total = o;
}

// ------------------------------------------------------------------ small helpers
__device__ __forceinline__ float bf16_round(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
}
__global__ void k_lp_quantize(int n, const float* __restrict__ w, float* __restrict__ q) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = bf16_round(w[i]);
}
__global__ void k_lp_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, float lr,
                          float b1, float b2, float c1, float c2, float eps) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1 - b1) * gi, vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
__global__ void k_lp_zero(long n, float* __restrict__ a) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.f;
}

// ------------------------------------------------------------------ features
// S_i = diag(H_ii)^-1/2 (H = elastic + contact), node features: scaled diagonal block (9),
// mean of m_i S^2, log(1 + contact blocks in the row), Dirichlet flag
__global__ void k_lp_node_features(int nV, const int* __restrict__ diagBlk, const double* __restrict__ ev, int enz, const int* __restrict__ crp,
                                   const int* __restrict__ ccol, const double* __restrict__ cv, int cnz, const double* __restrict__ mass,
                                   const uint8_t* __restrict__ dbc, double* __restrict__ S, float* __restrict__ nodeF) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nV) return;
    double D[9];
    const int bd = diagBlk[i];
#pragma unroll
    for (int k = 0; k < 9; ++k) D[k] = ev[(long)k * enz + bd];
    int ncon = 0;
    if (cnz > 0) {
        ncon = crp[i + 1] - crp[i];
        for (int b = crp[i]; b < crp[i + 1]; ++b)
            if (ccol[b] == i)
#pragma unroll
                for (int k = 0; k < 9; ++k) D[k] += cv[(long)k * cnz + b];
    }
    double s[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s[k] = D[4 * k] > 0 ? 1.0 / sqrt(D[4 * k]) : 1.0;
        S[3 * i + k] = s[k];
    }
    float* f = nodeF + (long)kLpNodeIn * i;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) f[3 * a + b] = (float)(s[a] * D[3 * a + b] * s[b]);
    f[9] = (float)(mass[i] * (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) / 3.0);
    f[10] = logf(1.0f + (float)ncon);
    f[11] = dbc[i] ? 1.f : 0.f;
}
// edge features: S_i (H_el + H_con)_ij S_j; the contact block of (i, j), if any, is found
// by binary search in row i of the contact matrix (columns sorted)
__global__ void k_lp_edge_features(int nE, const int* __restrict__ eBlock, const int* __restrict__ eSrc, const int* __restrict__ eDst,
                                   const double* __restrict__ ev, int enz, const int* __restrict__ crp, const int* __restrict__ ccol,
                                   const double* __restrict__ cv, int cnz, const double* __restrict__ S, float* __restrict__ edgeF) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const int b = eBlock[e], i = eSrc[e], j = eDst[e];
    double a[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = ev[(long)k * enz + b];
    if (cnz > 0) {
        int lo = crp[i], hi = crp[i + 1];
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ccol[mid] < j) lo = mid + 1;
            else hi = mid;
        }
        if (lo < crp[i + 1] && ccol[lo] == j)
#pragma unroll
            for (int k = 0; k < 9; ++k) a[k] += cv[(long)k * cnz + lo];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) edgeF[(long)kLpEdgeIn * e + 3 * r + c] = (float)(S[3 * i + r] * a[3 * r + c] * S[3 * j + c]);
}

// ------------------------------------------------------------------ dense layers
// out[r][o] = act(b[o] + sum_seg x_seg[idx_seg(r)] . W[seg rows][o]) (+ resid[r][o]);
// W row-major [Ktot][NOUT].  16 rows per workgroup; weights and the rows' inputs in LDS.
struct LpSeg {
    const float* x;
    const int* idx;   // input row of output row r (nullptr: r)
    int k;            // features per row (row stride)
};
constexpr int kLpRows = 16;
constexpr int kLpMaxK = 3 * kLpHidden;

template <int NOUT>
__global__ __launch_bounds__(256) void k_lp_dense(int nrows, LpSeg s0, LpSeg s1, LpSeg s2, const float* __restrict__ W,
                                                  const float* __restrict__ bias, int relu, const float* __restrict__ resid,
                                                  float* __restrict__ out) {
    __shared__ float Wl[kLpMaxK * NOUT];
    __shared__ float Xl[kLpRows][kLpMaxK + 1];
    const int K = s0.k + s1.k + s2.k;
    for (int t = threadIdx.x; t < K * NOUT; t += 256) Wl[t] = W[t];
    const int r0 = blockIdx.x * kLpRows;
    for (int t = threadIdx.x; t < kLpRows * K; t += 256) {
        const int rr = t / K, k = t % K, r = r0 + rr;
        float v = 0.f;
        if (r < nrows) {
            if (k < s0.k) v = s0.x[(long)(s0.idx ? s0.idx[r] : r) * s0.k + k];
            else if (k < s0.k + s1.k) v = s1.x[(long)(s1.idx ? s1.idx[r] : r) * s1.k + (k - s0.k)];
            else v = s2.x[(long)(s2.idx ? s2.idx[r] : r) * s2.k + (k - s0.k - s1.k)];
        }
        Xl[rr][k] = v;
    }
    __syncthreads();
    const int rr = threadIdx.x / 16, ol = threadIdx.x % 16, r = r0 + rr;
    constexpr int Q = (NOUT + 15) / 16;
    float acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = (ol + 16 * q < NOUT) ? bias[ol + 16 * q] : 0.f;
    for (int k = 0; k < K; ++k) {
        const float x = Xl[rr][k];
#pragma unroll
        for (int q = 0; q < Q; ++q)
            if (ol + 16 * q < NOUT) acc[q] += x * Wl[k * NOUT + ol + 16 * q];
    }
    if (r >= nrows) return;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int o = ol + 16 * q;
        if (o >= NOUT) continue;
        float y = acc[q];
        if (relu) y = y > 0.f ? y : 0.f;
        if (resid) y += resid[(long)r * NOUT + o];
        out[(long)r * NOUT + o] = y;
    }
}

// dX[r][k] (+)= sum_o dY[r][o] W[(koff + k)][o] for k < KOUT (input gradient of one segment)
template <int NOUT>
// (addIn may alias dX: each element is read and written by the same thread)
__global__ __launch_bounds__(256) void k_lp_dense_t(int nrows, const float* __restrict__ dY, const float* __restrict__ W, int koff, int KOUT,
                                                    const float* addIn, float* dX) {
    __shared__ float Wl[kLpHidden * NOUT];
    __shared__ float Yl[kLpRows][NOUT + 1];
    for (int t = threadIdx.x; t < KOUT * NOUT; t += 256) Wl[t] = W[(long)koff * NOUT + t];
    const int r0 = blockIdx.x * kLpRows;
    for (int t = threadIdx.x; t < kLpRows * NOUT; t += 256) {
        const int rr = t / NOUT, o = t % NOUT, r = r0 + rr;
        Yl[rr][o] = r < nrows ? dY[(long)r * NOUT + o] : 0.f;
    }
    __syncthreads();
    const int rr = threadIdx.x / 16, kl = threadIdx.x % 16, r = r0 + rr;
    if (r >= nrows) return;
    for (int k = kl; k < KOUT; k += 16) {
        float acc = 0.f;
        for (int o = 0; o < NOUT; ++o) acc += Yl[rr][o] * Wl[k * NOUT + o];
        if (addIn) acc += addIn[(long)r * KOUT + k];
        dX[(long)r * KOUT + k] = acc;
    }
}

// weight gradient partials: workgroup g sums rows [g*chunk, (g+1)*chunk) of
// x_seg[idx(r)]^T dY[r] (and dY for the bias) into part[g][Ktot*NOUT + NOUT]
template <int NOUT>
__global__ __launch_bounds__(256) void k_lp_wgrad(int nrows, int chunk, LpSeg s0, LpSeg s1, LpSeg s2, const float* __restrict__ dY,
                                                  float* __restrict__ part) {
    __shared__ float Xl[kLpRows][kLpMaxK + 1];
    __shared__ float Yl[kLpRows][NOUT + 1];
    const int K = s0.k + s1.k + s2.k, P = K * NOUT + NOUT;
    constexpr int MAXP = (kLpMaxK * NOUT + NOUT + 255) / 256;
    float acc[MAXP];
#pragma unroll
    for (int j = 0; j < MAXP; ++j) acc[j] = 0.f;
    const int rbeg = blockIdx.x * chunk, rend = min(nrows, rbeg + chunk);
    for (int r0 = rbeg; r0 < rend; r0 += kLpRows) {
        __syncthreads();
        for (int t = threadIdx.x; t < kLpRows * K; t += 256) {
            const int rr = t / K, k = t % K, r = r0 + rr;
            float v = 0.f;
            if (r < rend) {
                if (k < s0.k) v = s0.x[(long)(s0.idx ? s0.idx[r] : r) * s0.k + k];
                else if (k < s0.k + s1.k) v = s1.x[(long)(s1.idx ? s1.idx[r] : r) * s1.k + (k - s0.k)];
                else v = s2.x[(long)(s2.idx ? s2.idx[r] : r) * s2.k + (k - s0.k - s1.k)];
            }
            Xl[rr][k] = v;
        }
        for (int t = threadIdx.x; t < kLpRows * NOUT; t += 256) {
            const int rr = t / NOUT, o = t % NOUT, r = r0 + rr;
            Yl[rr][o] = r < rend ? dY[(long)r * NOUT + o] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < MAXP; ++j) {
            const int pidx = threadIdx.x + 256 * j;
            if (pidx >= P) break;
            float s = 0.f;
            if (pidx < K * NOUT) {
                const int k = pidx / NOUT, o = pidx % NOUT;
                for (int rr = 0; rr < kLpRows; ++rr) s += Xl[rr][k] * Yl[rr][o];
            } else {
                const int o = pidx - K * NOUT;
                for (int rr = 0; rr < kLpRows; ++rr) s += Yl[rr][o];
            }
            acc[j] += s;
        }
    }
#pragma unroll
    for (int j = 0; j < MAXP; ++j) {
        const int pidx = threadIdx.x + 256 * j;
        if (pidx < P) part[(long)blockIdx.x * P + pidx] = acc[j];
    }
}
// grad[w] += sum_g part[g][w] (weights), grad[b] += sum_g part[g][K*NOUT + o] (bias)
__global__ void k_lp_wgrad_sum(int ngroups, int P, int KN, const float* __restrict__ part, float* __restrict__ gw, float* __restrict__ gb) {
    int pidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (pidx >= P) return;
    float s = 0.f;
    for (int g = 0; g < ngroups; ++g) s += part[(long)g * P + pidx];
    if (pidx < KN) gw[pidx] += s;
    else gb[pidx - KN] += s;
}

// agg[i][f] = sum over node i's edges of m[e][f]
__global__ void k_lp_agg(int nV, const int* __restrict__ erp, const float* __restrict__ m, float* __restrict__ agg) {
    const int i = blockIdx.x * 4 + threadIdx.x / kLpHidden, f = threadIdx.x % kLpHidden;
    if (i >= nV) return;
    float s = 0.f;
    for (int e = erp[i]; e < erp[i + 1]; ++e) s += m[(long)e * kLpHidden + f];
    agg[(long)i * kLpHidden + f] = s;
}
// dM[e] = (de[e] + dagg[src(e)]) masked by m[e] > 0
__global__ void k_lp_msg_grad(long n, const int* __restrict__ eSrc, const float* __restrict__ de, const float* __restrict__ dagg,
                              const float* __restrict__ m, float* __restrict__ dM) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long e = t / kLpHidden;
    const int f = (int)(t % kLpHidden);
    const float g = de[t] + dagg[(long)eSrc[e] * kLpHidden + f];
    dM[t] = m[t] > 0.f ? g : 0.f;
}
// d = d masked by y > 0
__global__ void k_lp_relu_mask(long n, const float* __restrict__ y, const float* __restrict__ d, float* __restrict__ out) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = y[t] > 0.f ? d[t] : 0.f;
}
// dh[i] += sum_{e in row i} dSrc[e] + dDst[eT[e]]
__global__ void k_lp_gather_node(int nV, const int* __restrict__ erp, const int* __restrict__ eT, const float* __restrict__ dSrc,
                                 const float* __restrict__ dDst, float* __restrict__ dh) {
    const int i = blockIdx.x * 4 + threadIdx.x / kLpHidden, f = threadIdx.x % kLpHidden;
    if (i >= nV) return;
    float s = 0.f;
    for (int e = erp[i]; e < erp[i + 1]; ++e) s += dSrc[(long)e * kLpHidden + f] + dDst[(long)eT[e] * kLpHidden + f];
    dh[(long)i * kLpHidden + f] += s;
}
// decoder gradients of the lower edges back to nodes: dh[i] += sum_{lower l, src(l)=i} dS[l] + sum_{lower l, dst(l)=i} dD[l]
__global__ void k_lp_gather_lower(int nV, const int* __restrict__ sOff, const int* __restrict__ sIdx, const int* __restrict__ dOff,
                                  const int* __restrict__ dIdx, const float* __restrict__ dS, const float* __restrict__ dD, float* __restrict__ dh) {
    const int i = blockIdx.x * 4 + threadIdx.x / kLpHidden, f = threadIdx.x % kLpHidden;
    if (i >= nV) return;
    float s = 0.f;
    for (int k = sOff[i]; k < sOff[i + 1]; ++k) s += dS[(long)sIdx[k] * kLpHidden + f];
    for (int k = dOff[i]; k < dOff[i + 1]; ++k) s += dD[(long)dIdx[k] * kLpHidden + f];
    dh[(long)i * kLpHidden + f] += s;
}
// de[lEdge[l]] = dE[l] for lower edges (de cleared before)
__global__ void k_lp_scatter_lower(int nL, const int* __restrict__ lEdge, const float* __restrict__ dE, float* __restrict__ de) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nL * kLpHidden) return;
    const long l = t / kLpHidden;
    const int f = (int)(t % kLpHidden);
    de[(long)lEdge[l] * kLpHidden + f] = dE[t];
}


// ------------------------------------------------------------------ MFMA inference layers
// Inference form of the dense layers on the matrix cores: inputs and weights in BF16,
// FP32 accumulation (v_mfma_f32_16x16x16_bf16).  A wave computes 16 rows x 16 NT
// outputs: for every 16-wide k step it feeds one 16x16 A tile (lane l: row l % 16,
// k = 4 (l / 16) .. +3, from the gathered input row) and NT 16x16 B tiles (lane l:
// column 16 j + l % 16, the same 4 k's) held transposed in LDS; the accumulators hold
// rows 4 (l / 16) + t, column 16 j + l % 16.
typedef short lp_v4i16 __attribute__((ext_vector_type(4)));
typedef float lp_v4f32 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// Message layer (192 -> 64, rows gathered through the edge lists): one 16-row tile per
// wave; eNext (optional): eNext = eIn + out, the edge-state update fused into the epilogue
template <int K>
__global__ __launch_bounds__(256) void k_lp_dense_mfma(int nrows, LpSeg s0, LpSeg s1, LpSeg s2, const float* __restrict__ W,
                                                       const float* __restrict__ bias, int relu, float* __restrict__ out,
                                                       const float* __restrict__ eIn, float* __restrict__ eNext) {
    static_assert(K % 64 == 0 && K <= kLpMaxK, "segments are 64 wide");
    constexpr int PITCH = K + 4;   // shorts; keeps 8-byte alignment of every 4-k group
    __shared__ short Wt[kLpHidden * PITCH];
    for (int t = threadIdx.x; t < K * kLpHidden; t += 256) {
        const int k = t / kLpHidden, n = t % kLpHidden;
        Wt[n * PITCH + k] = bf16_bits(W[t]);
    }
    __syncthreads();
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int i = lane % 16, kg = lane / 16;
    const int rbase = blockIdx.x * 64 + wave * 16;
    const int r = rbase + i;
    const bool valid = r < nrows;
    long rowIn[3] = {0, 0, 0};
    if (valid) {
        rowIn[0] = s0.idx ? s0.idx[r] : r;
        rowIn[1] = s1.idx ? s1.idx[r] : r;
        rowIn[2] = s2.idx ? s2.idx[r] : r;
    }
    lp_v4f32 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = (lp_v4f32){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < K / 16; ++ks) {
        const int k0 = ks * 16 + 4 * kg;
        const int seg = k0 / kLpHidden, kk = k0 % kLpHidden;
        const float* src = seg == 0 ? s0.x : (seg == 1 ? s1.x : s2.x);
        float4 xv = make_float4(0.f, 0.f, 0.f, 0.f);
        if (valid) xv = *(const float4*)(src + rowIn[seg] * kLpHidden + kk);
        const lp_v4i16 a = {bf16_bits(xv.x), bf16_bits(xv.y), bf16_bits(xv.z), bf16_bits(xv.w)};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const lp_v4i16 b = *(const lp_v4i16*)&Wt[(16 * j + i) * PITCH + k0];
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int row = rbase + 4 * kg + t, col = 16 * j + i;
            if (row < nrows) {
                float y = acc[j][t] + bias[col];
                if (relu) y = y > 0.f ? y : 0.f;
                out[(long)row * kLpHidden + col] = y;
                if (eNext) eNext[(long)row * kLpHidden + col] = eIn[(long)row * kLpHidden + col] + y;
            }
        }
}


// General inference layer on the matrix cores (the node-level layers and the edge encoder):
//   out[r][o] = act(b[o] + sum_seg x_seg[idx_seg(r)] . W[seg rows][o]) (+ resid[r][o]),
//   o < nout <= 16 NT; optionally eNext = eIn + out.
// The K axis is the concatenation of the segments, each padded with zeros to a multiple
// of 16.  The weights are packed once per call into BF16 [16 NT][KP] (transposed,
// zero-padded; k_lp_wt_pack) and staged in LDS; each wave then walks 16-row tiles
// (grid-stride), so a workgroup stages them once for many tiles.
__global__ void k_lp_wt_pack(const float* __restrict__ W, int k0, int k1, int k2, int nout, int NT, int KP, short* __restrict__ Wt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 16 * NT * KP) return;
    const int n = t / KP, kp = t % KP;
    const int P0 = (k0 + 15) / 16 * 16, P1 = P0 + (k1 + 15) / 16 * 16;
    int seg = 0, kk = kp, base = 0;
    if (kp >= P1) { seg = 2; kk = kp - P1; base = k0 + k1; }
    else if (kp >= P0) { seg = 1; kk = kp - P0; base = k0; }
    const int ks = seg == 0 ? k0 : (seg == 1 ? k1 : k2);
    Wt[t] = (n < nout && kk < ks) ? bf16_bits(W[(long)(base + kk) * nout + n]) : (short)0;
}

template <int NT, int KP>
__global__ __launch_bounds__(256) void k_lp_mfma(int nrows, LpSeg s0, LpSeg s1, LpSeg s2, const short* __restrict__ WtG,
                                                 const float* __restrict__ bias, int nout, int relu, const float* __restrict__ resid,
                                                 float* __restrict__ out, const float* __restrict__ eIn, float* __restrict__ eNext) {
    static_assert(KP % 16 == 0 && KP <= kLpMaxK, "");
    static_assert(NT == 1 || NT * 16 == kLpHidden, "64-wide outputs or one 16-wide tile");
    constexpr int PITCH = KP + 4;   // shorts; keeps 8-byte alignment of every 4-k group
    __shared__ short Wt[16 * NT * PITCH];
    for (int t = threadIdx.x; t < 16 * NT * KP; t += 256) Wt[(t / KP) * PITCH + t % KP] = WtG[t];
    __syncthreads();
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int i = lane % 16, kg = lane / 16;
    const int P0 = (s0.k + 15) / 16 * 16, P1 = P0 + (s1.k + 15) / 16 * 16;
    const int nTiles = (nrows + 15) / 16;
    for (int tile = blockIdx.x * 4 + wave; tile < nTiles; tile += gridDim.x * 4) {
        const int rbase = tile * 16;
        const int r = rbase + i;
        const bool valid = r < nrows;
        long rowIn[3] = {0, 0, 0};
        if (valid) {
            rowIn[0] = s0.idx ? s0.idx[r] : r;
            rowIn[1] = s1.idx ? s1.idx[r] : r;
            rowIn[2] = s2.idx ? s2.idx[r] : r;
        }
        // the tile's A operand (this lane's 4 k of every 16-wide step), all loads first
        lp_v4i16 a[KP / 16];
        if (KP >= 64) {
            // wide layers: every segment is 64 wide (one base address per segment)
#pragma unroll
            for (int sg = 0; sg < KP / 64; ++sg) {
                const LpSeg& S = sg == 0 ? s0 : (sg == 1 ? s1 : s2);
                const float* src = S.x + rowIn[sg] * kLpHidden + 4 * kg;
                float4 xv[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) xv[q] = valid ? *(const float4*)(src + 16 * q) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    a[4 * sg + q] = (lp_v4i16){bf16_bits(xv[q].x), bf16_bits(xv[q].y), bf16_bits(xv[q].z), bf16_bits(xv[q].w)};
            }
        } else
#pragma unroll
        for (int ks = 0; ks < KP / 16; ++ks) {
            const int kp0 = ks * 16 + 4 * kg;
            const int seg = kp0 < P0 ? 0 : (kp0 < P1 ? 1 : 2);
            const int kk = kp0 - (seg == 0 ? 0 : (seg == 1 ? P0 : P1));
            const LpSeg& S = seg == 0 ? s0 : (seg == 1 ? s1 : s2);
            const float* src = S.x + rowIn[seg] * S.k;
            float4 xv = make_float4(0.f, 0.f, 0.f, 0.f);
            if (valid) {
                if ((S.k & 3) == 0 && kk + 4 <= S.k) {
                    xv = *(const float4*)(src + kk);
                } else {
                    if (kk < S.k) xv.x = src[kk];
                    if (kk + 1 < S.k) xv.y = src[kk + 1];
                    if (kk + 2 < S.k) xv.z = src[kk + 2];
                    if (kk + 3 < S.k) xv.w = src[kk + 3];
                }
            }
            a[ks] = (lp_v4i16){bf16_bits(xv.x), bf16_bits(xv.y), bf16_bits(xv.z), bf16_bits(xv.w)};
        }
        lp_v4f32 acc[NT];
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[j] = (lp_v4f32){0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
        for (int ks = 0; ks < KP / 16; ++ks) {
            const int kp0 = ks * 16 + 4 * kg;
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const lp_v4i16 bv = *(const lp_v4i16*)&Wt[(16 * j + i) * PITCH + kp0];
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ks], bv, acc[j], 0, 0, 0);
            }
        }
#pragma unroll
        for (int j = 0; j < NT; ++j)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = rbase + 4 * kg + t, col = 16 * j + i;
                if (row < nrows && col < nout) {
                    float y = acc[j][t] + bias[col];
                    if (relu) y = y > 0.f ? y : 0.f;
                    if (resid) y += resid[(long)row * nout + col];
                    out[(long)row * nout + col] = y;
                    if (eNext) eNext[(long)row * nout + col] = eIn[(long)row * nout + col] + y;
                }
            }
    }
}

// ------------------------------------------------------------------ G assembly / products
// BF16 (round to nearest even) of a finite float, as the high half of its bit pattern
__device__ __forceinline__ uint32_t lp_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
// 9 BF16 values of a 3x3 block packed in five 32-bit words (entries 2w, 2w+1 in word w;
// entry 8 alone in word 4), stored SoA [5][n]
__device__ __forceinline__ void lp_store_packed(uint32_t* __restrict__ dst, long n, long pos, const uint32_t* h) {
#pragma unroll
    for (int w = 0; w < 4; ++w) dst[w * n + pos] = h[2 * w] | (h[2 * w + 1] << 16);
    dst[4 * n + pos] = h[8];
}

// Writes G (FP32, SoA over the elastic pattern: the training / inspection form) and the
// two compact BF16 block CSRs of F = G S used by the PCG: FL = F (lower incl. diagonal)
// and FU = F^T, so that M^-1 r = S G^T G S r = FU (FL r).  F's column scaling by
// S = diag(H)^-1/2 removes the scaling vectors from both products; the same rounded F
// sits on both sides, so M^-1 stays symmetric positive definite.
__global__ void k_lp_assemble_G(int nL, int nV, int nnzb, const int* __restrict__ lEdge, const int* __restrict__ eBlock,
                                const int* __restrict__ blkT, const int* __restrict__ diagBlk, const int* __restrict__ bcol,
                                const double* __restrict__ S, const float* __restrict__ Gl, const float* __restrict__ raw,
                                float* __restrict__ G, int nLow, const int* __restrict__ posLE, const int* __restrict__ posLD,
                                const int* __restrict__ posUE, const int* __restrict__ posUD, uint32_t* __restrict__ FL,
                                uint32_t* __restrict__ FU) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    float g[9];
    int col, pl, pu;
    if (t < nL) {
        const int b = eBlock[lEdge[t]], tb = blkT[b];
        pl = posLE[t]; pu = posUE[t]; col = bcol[b];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float v = Gl[9l * t + 3 * a + c];
                g[3 * a + c] = v;
                G[(long)(3 * a + c) * nnzb + b] = v;
                G[(long)(3 * c + a) * nnzb + tb] = v;
            }
    } else if (t < nL + nV) {
        const int i = t - nL, b = diagBlk[i];
        pl = posLD[i]; pu = posUD[i]; col = i;
        const float* r = raw + 6l * i;
        g[0] = expf(r[0]); g[1] = 0.f; g[2] = 0.f; g[3] = r[3]; g[4] = expf(r[1]); g[5] = 0.f; g[6] = r[4]; g[7] = r[5]; g[8] = expf(r[2]);
#pragma unroll
        for (int k = 0; k < 9; ++k) G[(long)k * nnzb + b] = g[k];
    } else {
        return;
    }
    uint32_t hl[9], hu[9];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint32_t h = lp_bf16(g[3 * a + c] * (float)S[3 * col + c]);
            hl[3 * a + c] = h;
            hu[3 * c + a] = h;
        }
    lp_store_packed(FL, nLow, pl, hl);
    lp_store_packed(FU, nLow, pu, hu);
}

// y_i = sum_b F_b x_col(b) over a compact block CSR of BF16 blocks packed five words per
// block (SoA [5][nnz]); FP32 or FP64 arithmetic by V.  When r != nullptr, per-workgroup
// partial sums of r.y.  Fixed grid, rows mapped XCD-aware, next row group's bounds
// loaded ahead.
template <int LANES, int BS, class V>
__global__ __launch_bounds__(BS) void k_lp_spmv_bf16(int nrow, int rowsPerXcd, const int* __restrict__ rp, const int* __restrict__ col,
                                                     const uint32_t* __restrict__ val, int nnz, const V* __restrict__ x,
                                                     V* __restrict__ y, const V* __restrict__ r, double* __restrict__ part,
                                                     const int* __restrict__ done, const int* __restrict__ cgAgg,
                                                     const double* __restrict__ cgD, const double* __restrict__ cgZ) {
    if (done && *done) return;   // the PCG has converged: later iterations of the chunk do nothing
    __shared__ double lds[BS / 64];
    constexpr int RPG = BS / LANES;
    const int xb = blockIdx.x % 8, jb = blockIdx.x / 8, nj = gridDim.x / 8;
    const int lane = threadIdx.x % LANES;
    const int rowBase = xb * rowsPerXcd + (int)threadIdx.x / LANES;
    int g = jb;
    int row = rowBase + g * RPG;
    bool live = g * RPG < rowsPerXcd && row < nrow;
    int beg = live ? rp[row] : 0, end = live ? rp[row + 1] : 0;
    double dot = 0;
    for (; g * RPG < rowsPerXcd; g += nj) {
        const int gn = g + nj, rowN = rowBase + gn * RPG;
        const bool liveN = gn * RPG < rowsPerXcd && rowN < nrow;
        const int begN = liveN ? rp[rowN] : 0, endN = liveN ? rp[rowN + 1] : 0;
        V acc[3] = {0, 0, 0};
        for (int b = beg + lane; b < end; b += LANES) {
            const int c = col[b];
            const V x0 = x[3 * c], x1 = x[3 * c + 1], x2 = x[3 * c + 2];
            uint32_t w[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) w[k] = val[(long)k * nnz + b];
            V a[9];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                a[2 * k] = (V)__uint_as_float(w[k] << 16);
                a[2 * k + 1] = (V)__uint_as_float(w[k] & 0xFFFF0000u);
            }
            a[8] = (V)__uint_as_float(w[4] << 16);
            acc[0] += a[0] * x0 + a[1] * x1 + a[2] * x2;
            acc[1] += a[3] * x0 + a[4] * x1 + a[5] * x2;
            acc[2] += a[6] * x0 + a[7] * x1 + a[8] * x2;
        }
#pragma unroll
        for (int o = LANES / 2; o > 0; o >>= 1)
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[k] += __shfl_xor(acc[k], o, LANES);
        if (row < nrow && lane == 0) {
            // two-level correction: y_i += T_i zc = t_a + w_a x d_i (rows of the coarse space)
            const int ca = cgZ ? cgAgg[row] : -1;
            if (ca >= 0) {
                const double* zc = cgZ + 6 * ca;
                const double d0 = cgD[3 * row], d1 = cgD[3 * row + 1], d2 = cgD[3 * row + 2];
                acc[0] += (V)(zc[0] + zc[4] * d2 - zc[5] * d1);
                acc[1] += (V)(zc[1] + zc[5] * d0 - zc[3] * d2);
                acc[2] += (V)(zc[2] + zc[3] * d1 - zc[4] * d0);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                y[3 * row + k] = acc[k];
                if (r) dot += (double)r[3 * row + k] * (double)acc[k];
            }
        }
        row = rowN;
        beg = begN;
        end = endN;
    }
    if (part) {
        double v = wave_sum(dot);
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0;
            for (int w = 0; w < BS / 64; ++w) t += lds[w];
            part[blockIdx.x] = t;
        }
    }
}

// z_j = s_j . sum_{blocks b of row j, col >= j} T_b y_col, T_b = G_b at upper positions
// (they hold transposed lower blocks) and G_b^T on the diagonal.  Grid-stride over row
// groups so that the optional per-workgroup partial sums of r.z stay few.
__global__ void k_lp_apply_upper(int nV, const int* __restrict__ rp, const int* __restrict__ col, const float* __restrict__ G, int nnzb,
                                 const double* __restrict__ s, const double* __restrict__ y, double* __restrict__ z,
                                 const double* __restrict__ r, double* __restrict__ part) {
    constexpr int LANES = 16;
    __shared__ double lds[4];
    const int lane = threadIdx.x % LANES;
    double rz = 0;
    for (int g = blockIdx.x; g * (kBlock / LANES) < nV; g += gridDim.x) {
        const int row = g * (kBlock / LANES) + (int)threadIdx.x / LANES;
        double acc[3] = {0, 0, 0};
        if (row < nV) {
            for (int b = rp[row] + lane; b < rp[row + 1]; b += LANES) {
                const int c = col[b];
                if (c < row) continue;
                const double y0 = y[3 * c], y1 = y[3 * c + 1], y2 = y[3 * c + 2];
                float gv[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) gv[k] = G[(long)k * nnzb + b];
                if (c == row) {
                    acc[0] += (double)gv[0] * y0 + (double)gv[3] * y1 + (double)gv[6] * y2;
                    acc[1] += (double)gv[1] * y0 + (double)gv[4] * y1 + (double)gv[7] * y2;
                    acc[2] += (double)gv[2] * y0 + (double)gv[5] * y1 + (double)gv[8] * y2;
                } else {
                    acc[0] += (double)gv[0] * y0 + (double)gv[1] * y1 + (double)gv[2] * y2;
                    acc[1] += (double)gv[3] * y0 + (double)gv[4] * y1 + (double)gv[5] * y2;
                    acc[2] += (double)gv[6] * y0 + (double)gv[7] * y1 + (double)gv[8] * y2;
                }
            }
        }
#pragma unroll
        for (int o = LANES / 2; o > 0; o >>= 1)
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[k] += __shfl_xor(acc[k], o, LANES);
        if (row < nV && lane == 0) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double zk = s ? s[3 * row + k] * acc[k] : acc[k];
                z[3 * row + k] = zk;
                if (r) rz += r[3 * row + k] * zk;
            }
        }
    }
    if (part) {
        const double b = block_sum_256(rz, lds);
        if (threadIdx.x == 0) part[blockIdx.x] = b;
    }
}

// v = S H S u  over the elastic and contact blocks; partial u.v per workgroup
__global__ void k_lp_hhat(int nV, const int* __restrict__ erp, const int* __restrict__ ecol, const double* __restrict__ ev, int enz,
                          const int* __restrict__ crp, const int* __restrict__ ccol, const double* __restrict__ cv, int cnz,
                          const double* __restrict__ s, const double* __restrict__ u, double* __restrict__ v, double* __restrict__ part) {
    constexpr int LANES = 16;
    __shared__ double lds[4];
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int row = (int)(gid / LANES), lane = (int)(gid % LANES);
    double acc[3] = {0, 0, 0};
    if (row < nV) {
        for (int pass = 0; pass < 2; ++pass) {
            const int* rp = pass ? crp : erp;
            const int* cl = pass ? ccol : ecol;
            const double* vv = pass ? cv : ev;
            const long nz = pass ? cnz : enz;
            if (pass && cnz == 0) break;
            for (int b = rp[row] + lane; b < rp[row + 1]; b += LANES) {
                const int c = cl[b];
                const double x0 = s[3 * c] * u[3 * c], x1 = s[3 * c + 1] * u[3 * c + 1], x2 = s[3 * c + 2] * u[3 * c + 2];
                acc[0] += vv[b] * x0 + vv[nz + b] * x1 + vv[2 * nz + b] * x2;
                acc[1] += vv[3 * nz + b] * x0 + vv[4 * nz + b] * x1 + vv[5 * nz + b] * x2;
                acc[2] += vv[6 * nz + b] * x0 + vv[7 * nz + b] * x1 + vv[8 * nz + b] * x2;
            }
        }
    }
#pragma unroll
    for (int o = LANES / 2; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] += __shfl_xor(acc[k], o, LANES);
    double uv = 0;
    if (row < nV && lane == 0)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double vk = s[3 * row + k] * acc[k];
            v[3 * row + k] = vk;
            uv += u[3 * row + k] * vk;
        }
    const double b = block_sum_256(uv, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// Rademacher probe vectors (splitmix64 of seed, probe, entry)
__global__ void k_lp_probe(long n, uint64_t seed, double* __restrict__ w) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    w[i] = (z >> 63) ? 1.0 : -1.0;
}

// dL/dG on the lower pattern: c * sum_p w_p[i] v_p[j]^T for lower edges (i > j) and the
// diagonal; the diagonal part is mapped to the raw decoder outputs (exp diagonal, log-det
// term -2/n per diagonal output)
__global__ void k_lp_grad_G(int nL, int nV, int np, long nvec, const int* __restrict__ lEdge, const int* __restrict__ eSrc,
                            const int* __restrict__ eDst, const double* __restrict__ w, const double* __restrict__ v, double c,
                            const float* __restrict__ raw, double logdetScale, float* __restrict__ dGl, float* __restrict__ draw) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int i, j;
    if (t < nL) {
        const int e = lEdge[t];
        i = eSrc[e];
        j = eDst[e];
    } else if (t < nL + nV) {
        i = j = t - nL;
    } else {
        return;
    }
    double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = 0; p < np; ++p)
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) g[3 * a + b] += w[p * nvec + 3 * i + a] * v[p * nvec + 3 * j + b];
    if (t < nL) {
#pragma unroll
        for (int k = 0; k < 9; ++k) dGl[9l * t + k] = (float)(c * g[k]);
    } else {
        const float* r = raw + 6l * i;
        float* d = draw + 6l * i;
#pragma unroll
        for (int k = 0; k < 3; ++k) d[k] = (float)(c * g[4 * k] * exp((double)r[k]) - logdetScale);
        d[3] = (float)(c * g[3]);
        d[4] = (float)(c * g[6]);
        d[5] = (float)(c * g[7]);
    }
}
__global__ void k_lp_sum_rawdiag(int nV, const float* __restrict__ raw, double* __restrict__ part) {
    __shared__ double lds[4];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double s = i < nV ? (double)raw[6l * i] + raw[6l * i + 1] + raw[6l * i + 2] : 0.0;
    const double b = block_sum_256(s, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// ------------------------------------------------------------------ host side
static LearnedPrecond& LP(Ctx& c) {
    if (!c.lp) c.lp.reset(new LearnedPrecond());
    return *c.lp;
}
static double host_sum(Ctx& c, const double* part, int n) {
    sum_partials(part, n, c.scal.p + 44, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hScal + 44, c.scal.p + 44, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return c.hScal[44];
}

void lp_build_graph(Ctx& c) {
    LearnedPrecond& lp = LP(c);
    BSR& E = c.Ael;
    const int nV = E.nrow, nnzb = E.nnzb;
    std::vector<int> rp(nV + 1), col(nnzb);
    CIPC_HIP(hipMemcpyAsync(rp.data(), E.rowptr.p, sizeof(int) * (nV + 1), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipMemcpyAsync(col.data(), E.col.p, sizeof(int) * nnzb, hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    std::vector<int> diagBlk(nV, -1), blkT(nnzb, -1), blkEdge(nnzb, -1), eBlock, eSrc, eDst, erp(nV + 1, 0);
    for (int i = 0; i < nV; ++i) {
        erp[i] = (int)eBlock.size();
        for (int b = rp[i]; b < rp[i + 1]; ++b) {
            if (col[b] == i) { diagBlk[i] = b; continue; }
            blkEdge[b] = (int)eBlock.size();
            eBlock.push_back(b);
            eSrc.push_back(i);
            eDst.push_back(col[b]);
        }
    }
    erp[nV] = (int)eBlock.size();
    for (int i = 0; i < nV; ++i)
        if (diagBlk[i] < 0) throw HipError("learned preconditioner: a row has no diagonal block", -1);
    for (int i = 0; i < nV; ++i)
        for (int b = rp[i]; b < rp[i + 1]; ++b) {
            const int j = col[b];
            const int* lo = col.data() + rp[j];
            const int* hi = col.data() + rp[j + 1];
            const int* f = std::lower_bound(lo, hi, i);
            if (f == hi || *f != i) throw HipError("learned preconditioner: Hessian pattern is not symmetric", -1);
            blkT[b] = (int)(f - col.data());
        }
    const int nE = (int)eBlock.size();
    std::vector<int> eT(nE), lEdge;
    for (int e = 0; e < nE; ++e) eT[e] = blkEdge[blkT[eBlock[e]]];
    for (int e = 0; e < nE; ++e)
        if (eDst[e] < eSrc[e]) lEdge.push_back(e);
    const int nL = (int)lEdge.size();
    // per node: lower edges with src = i, lower edges with dst = i
    std::vector<int> sOff(nV + 1, 0), dOff(nV + 1, 0), sIdx(nL), dIdx(nL);
    for (int l = 0; l < nL; ++l) { sOff[eSrc[lEdge[l]] + 1]++; dOff[eDst[lEdge[l]] + 1]++; }
    for (int i = 0; i < nV; ++i) { sOff[i + 1] += sOff[i]; dOff[i + 1] += dOff[i]; }
    {
        std::vector<int> fs(sOff.begin(), sOff.end() - 1), fd(dOff.begin(), dOff.end() - 1);
        for (int l = 0; l < nL; ++l) { sIdx[fs[eSrc[lEdge[l]]]++] = l; dIdx[fd[eDst[lEdge[l]]]++] = l; }
    }
    upload(lp.eBlock, eBlock, c.s); upload(lp.eSrc, eSrc, c.s); upload(lp.eDst, eDst, c.s); upload(lp.erp, erp, c.s);
    upload(lp.eT, eT, c.s); upload(lp.lEdge, lEdge, c.s); upload(lp.blkT, blkT, c.s); upload(lp.diagBlk, diagBlk, c.s);
    upload(lp.lSrcOff, sOff, c.s); upload(lp.lSrcIdx, sIdx, c.s); upload(lp.lDstOff, dOff, c.s); upload(lp.lDstIdx, dIdx, c.s);
    // lower-edge gather indices for the decoder: src node, dst node, edge
    std::vector<int> lS(nL), lD(nL);
    for (int l = 0; l < nL; ++l) { lS[l] = eSrc[lEdge[l]]; lD[l] = eDst[lEdge[l]]; }
    upload(lp.lowSrc, lS, c.s); upload(lp.lowDst, lD, c.s);
    // compact lower (G) and upper (G^T) block CSRs over the same rows, columns ascending
    std::vector<int> lowIdxOfEdge(nE, -1);
    for (int l = 0; l < nL; ++l) lowIdxOfEdge[lEdge[l]] = l;
    std::vector<int> Lrp(nV + 1, 0), Lcol, Urp(nV + 1, 0), Ucol, pLE(nL), pLD(nV), pUE(nL), pUD(nV);
    for (int i = 0; i < nV; ++i) {
        Lrp[i] = (int)Lcol.size();
        for (int b = rp[i]; b < rp[i + 1] && col[b] <= i; ++b) {
            if (col[b] == i) pLD[i] = (int)Lcol.size();
            else pLE[lowIdxOfEdge[blkEdge[b]]] = (int)Lcol.size();
            Lcol.push_back(col[b]);
        }
        Urp[i] = (int)Ucol.size();
        for (int b = rp[i]; b < rp[i + 1]; ++b) {
            if (col[b] < i) continue;
            if (col[b] == i) pUD[i] = (int)Ucol.size();
            else pUE[lowIdxOfEdge[eT[blkEdge[b]]]] = (int)Ucol.size();   // (i, j>i) holds G_ji^T
            Ucol.push_back(col[b]);
        }
    }
    Lrp[nV] = (int)Lcol.size();
    Urp[nV] = (int)Ucol.size();
    if (Lcol.size() != Ucol.size()) throw HipError("learned preconditioner: lower/upper pattern mismatch", -1);
    lp.nLow = (int)Lcol.size();
    upload(lp.gLrp, Lrp, c.s); upload(lp.gLcol, Lcol, c.s); upload(lp.gUrp, Urp, c.s); upload(lp.gUcol, Ucol, c.s);
    upload(lp.posLE, pLE, c.s); upload(lp.posLD, pLD, c.s); upload(lp.posUE, pUE, c.s); upload(lp.posUD, pUD, c.s);
    lp.FL.ensure_exact(5 * (size_t)lp.nLow);
    lp.FU.ensure_exact(5 * (size_t)lp.nLow);
    CIPC_HIP(hipStreamSynchronize(c.s));
    lp.nV = nV; lp.nnzb = nnzb; lp.nE = nE; lp.nL = nL;
    const size_t H = kLpHidden;
    lp.S.ensure_exact(3 * (size_t)nV);
    lp.nodeF.ensure_exact((size_t)kLpNodeIn * nV);
    lp.edgeF.ensure_exact((size_t)kLpEdgeIn * std::max(nE, 1));
    for (int l = 0; l <= kLpLayers; ++l) { lp.h[l].ensure_exact(H * nV); lp.e[l].ensure_exact(H * std::max(nE, 1)); }
    for (int l = 0; l < kLpLayers; ++l) {
        lp.m[l].ensure_exact(H * std::max(nE, 1));
        lp.u[l].ensure_exact(H * nV);
        lp.agg[l].ensure_exact(H * nV);
    }
    lp.Gl.ensure_exact(9 * (size_t)std::max(nL, 1));
    lp.raw.ensure_exact(6 * (size_t)nV);
    lp.Gblk.ensure_exact(9 * (size_t)nnzb);
    lp.wpartial.ensure_exact((size_t)256 * (kLpMaxK * kLpHidden + kLpHidden));
    lp.graphReady = true;
    lp.forwardReady = false;
}

void lp_init_weights(Ctx& c, uint64_t seed) {
    LearnedPrecond& lp = LP(c);
    const LpLayout& L = lp.L;
    std::vector<float> w(L.total, 0.f);
    std::mt19937_64 g(seed);
    auto fill = [&](int off, int fanIn, int fanOut) {
        // He-uniform for ReLU layers
        const float a = std::sqrt(6.0f / (float)fanIn);
        std::uniform_real_distribution<float> U(-a, a);
        for (int k = 0; k < fanIn * fanOut; ++k) w[off + k] = U(g);
    };
    fill(L.we, kLpNodeIn, kLpHidden);
    fill(L.wx, kLpEdgeIn, kLpHidden);
    for (int l = 0; l < kLpLayers; ++l) {
        fill(L.wm[l], 3 * kLpHidden, kLpHidden);
        fill(L.wu[l], 2 * kLpHidden, kLpHidden);
        // residual branches start small so the stack starts near identity
        for (int k = 0; k < 2 * kLpHidden * kLpHidden; ++k) w[L.wu[l] + k] *= 0.1f;
    }
    // decoder at zero: G = I (point Jacobi) before training
    set_weights(c, w.data(), L.total);
}

int lp_num_params() { return LpLayout().total; }

void lp_factor(Ctx& c, bool useMfma, float* G, int n) {
    LearnedPrecond& lp = LP(c);
    const bool mf = lp.mfma;
    lp.mfma = useMfma;
    lp_prepare(c);
    lp.mfma = mf;
    if (n != 9 * lp.nnzb) throw HipError("learned preconditioner: factor buffer size", -1);
    CIPC_HIP(hipMemcpyAsync(G, lp.Gblk.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
}

void set_weights(Ctx& c, const float* w, int n) {
    LearnedPrecond& lp = LP(c);
    if (n != lp.L.total) throw HipError("learned preconditioner: wrong parameter count", -1);
    upload(lp.W, w, (size_t)n, c.s);
    lp.Wq.ensure_exact(n);
    lp.gW.ensure_exact(n);
    lp.mW.ensure_exact(n);
    lp.vW.ensure_exact(n);
    hipLaunchKernelGGL(k_lp_zero, dim3(nblocks(n)), dim3(kBlock), 0, c.s, (long)n, lp.mW.p);
    hipLaunchKernelGGL(k_lp_zero, dim3(nblocks(n)), dim3(kBlock), 0, c.s, (long)n, lp.vW.p);
    hipLaunchKernelGGL(k_lp_quantize, dim3(nblocks(n)), dim3(kBlock), 0, c.s, n, (const float*)lp.W.p, lp.Wq.p);
    CIPC_LAUNCH_CHECK();
    CIPC_HIP(hipStreamSynchronize(c.s));
    lp.adamT = 0;
    lp.weightsReady = true;
    lp.forwardReady = false;
    lp.forwardReady = false;
}
void get_weights(Ctx& c, float* w, int n) {
    LearnedPrecond& lp = LP(c);
    if (!lp.weightsReady || n != lp.L.total) throw HipError("learned preconditioner: no weights / wrong count", -1);
    CIPC_HIP(hipMemcpyAsync(w, lp.W.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
}

void get_adam(Ctx& c, float* m, float* v, int* t, int n) {
    LearnedPrecond& lp = LP(c);
    if (!lp.weightsReady || n != lp.L.total) throw HipError("learned preconditioner: no weights / wrong count", -1);
    CIPC_HIP(hipMemcpyAsync(m, lp.mW.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipMemcpyAsync(v, lp.vW.p, sizeof(float) * n, hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    *t = lp.adamT;
}
void set_adam(Ctx& c, const float* m, const float* v, int t, int n) {
    LearnedPrecond& lp = LP(c);
    if (!lp.weightsReady || n != lp.L.total) throw HipError("learned preconditioner: no weights / wrong count", -1);
    CIPC_HIP(hipMemcpyAsync(lp.mW.p, m, sizeof(float) * n, hipMemcpyHostToDevice, c.s));
    CIPC_HIP(hipMemcpyAsync(lp.vW.p, v, sizeof(float) * n, hipMemcpyHostToDevice, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    lp.adamT = t;
}

static const LpSeg kNoSeg = {nullptr, nullptr, 0};

template <int NOUT>
static void dense(Ctx& c, int nrows, LpSeg s0, LpSeg s1, LpSeg s2, const float* W, const float* b, int relu, const float* resid, float* out) {
    if (nrows <= 0) return;
    hipLaunchKernelGGL(k_lp_dense<NOUT>, dim3((nrows + kLpRows - 1) / kLpRows), dim3(256), 0, c.s, nrows, s0, s1, s2, W, b, relu, resid, out);
    CIPC_LAUNCH_CHECK();
}
template <int NOUT>
static void dense_t(Ctx& c, int nrows, const float* dY, const float* W, int koff, const float* addIn, float* dX) {
    if (nrows <= 0) return;
    hipLaunchKernelGGL(k_lp_dense_t<NOUT>, dim3((nrows + kLpRows - 1) / kLpRows), dim3(256), 0, c.s, nrows, dY, W, koff, kLpHidden, addIn, dX);
    CIPC_LAUNCH_CHECK();
}
template <int NOUT>
static void wgrad(Ctx& c, int nrows, LpSeg s0, LpSeg s1, LpSeg s2, const float* dY, float* gw, float* gb) {
    if (nrows <= 0) return;
    LearnedPrecond& lp = LP(c);
    const int K = s0.k + s1.k + s2.k, P = K * NOUT + NOUT;
    const int groups = std::min(256, std::max(1, (nrows + 255) / 256));
    int chunk = (nrows + groups - 1) / groups;
    chunk = (chunk + kLpRows - 1) / kLpRows * kLpRows;
    float* part = lp.wpartial.ensure((size_t)groups * P);
    hipLaunchKernelGGL(k_lp_wgrad<NOUT>, dim3(groups), dim3(256), 0, c.s, nrows, chunk, s0, s1, s2, dY, part);
    hipLaunchKernelGGL(k_lp_wgrad_sum, dim3(nblocks(P)), dim3(kBlock), 0, c.s, groups, P, K * NOUT, (const float*)part, gw, gb);
    CIPC_LAUNCH_CHECK();
}
__global__ void k_lp_add(long n, const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ out) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] + b[i];
}
static void add(Ctx& c, long n, const float* a, const float* b, float* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_lp_add, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.s, n, a, b, out);
    CIPC_LAUNCH_CHECK();
}

// workgroups of the general MFMA layers (waves walk 16-row tiles grid-stride)
constexpr int kLpMfmaGrid = 8192;

static void mfma_dense(Ctx& c, int nrows, LpSeg s0, LpSeg s1, LpSeg s2, const float* W, const float* b, int nout, int relu,
                       const float* resid, float* out, const float* eIn = nullptr, float* eNext = nullptr) {
    if (nrows <= 0) return;
    LearnedPrecond& lp = LP(c);
    const int KP = (s0.k + 15) / 16 * 16 + (s1.k + 15) / 16 * 16 + (s2.k + 15) / 16 * 16;
    const int NT = (nout + 15) / 16;
    if (KP > kLpMaxK || !(NT == 1 || NT == 4)) throw HipError("learned preconditioner: MFMA layer shape", -1);
    short* wt = lp.wt16.ensure_exact((size_t)64 * kLpMaxK);
    const int nw = 16 * NT * KP;
    hipLaunchKernelGGL(k_lp_wt_pack, dim3(nblocks(nw)), dim3(kBlock), 0, c.s, W, s0.k, s1.k, s2.k, nout, NT, KP, wt);
    const int grid = std::min(kLpMfmaGrid, ((nrows + 15) / 16 + 3) / 4);
    auto kern = NT == 4 ? (KP == 16 ? k_lp_mfma<4, 16> : KP == 64 ? k_lp_mfma<4, 64> : k_lp_mfma<4, 128>)
                        : (KP == 16 ? k_lp_mfma<1, 16> : KP == 64 ? k_lp_mfma<1, 64> : k_lp_mfma<1, 128>);
    const bool wide = (s0.k == 64) && (s1.k == 0 || s1.k == 64) && s2.k == 0;
    if (!(KP == 16 || (wide && (KP == 64 || KP == 128)))) throw HipError("learned preconditioner: MFMA layer shape", -1);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, c.s, nrows, s0, s1, s2, (const short*)wt, b, nout, relu, resid, out, eIn, eNext);
    CIPC_LAUNCH_CHECK();
}

static void features(Ctx& c) {
    LearnedPrecond& lp = LP(c);
    BSR& E = c.Ael;
    BSR& K = c.Acon;
    if (K.nnzb == 0) K.rowptr.ensure(c.nV + 1);
    hipLaunchKernelGGL(k_lp_node_features, dim3(nblocks(lp.nV)), dim3(kBlock), 0, c.s, lp.nV, (const int*)lp.diagBlk.p, (const double*)E.val.p,
                       E.nnzb, (const int*)K.rowptr.p, (const int*)K.col.p, (const double*)K.val.p, K.nnzb, (const double*)c.mass.p,
                       proj_mask(c), lp.S.p, lp.nodeF.p);
    if (lp.nE > 0)
        hipLaunchKernelGGL(k_lp_edge_features, dim3(nblocks(lp.nE)), dim3(kBlock), 0, c.s, lp.nE, (const int*)lp.eBlock.p, (const int*)lp.eSrc.p,
                           (const int*)lp.eDst.p, (const double*)E.val.p, E.nnzb, (const int*)K.rowptr.p, (const int*)K.col.p,
                           (const double*)K.val.p, K.nnzb, (const double*)lp.S.p, lp.edgeF.p);
    CIPC_LAUNCH_CHECK();
}

// G, FL, FU from the decoder outputs Gl, raw and the scaling S
static void lp_assemble(Ctx& c) {
    LearnedPrecond& lp = LP(c);
    const int nV = lp.nV, nL = lp.nL;
    hipLaunchKernelGGL(k_lp_assemble_G, dim3(nblocks(nL + nV)), dim3(kBlock), 0, c.s, nL, nV, lp.nnzb, (const int*)lp.lEdge.p,
                       (const int*)lp.eBlock.p, (const int*)lp.blkT.p, (const int*)lp.diagBlk.p, (const int*)c.Ael.col.p,
                       (const double*)lp.S.p, (const float*)lp.Gl.p, (const float*)lp.raw.p, lp.Gblk.p, lp.nLow,
                       (const int*)lp.posLE.p, (const int*)lp.posLD.p, (const int*)lp.posUE.p, (const int*)lp.posUD.p, lp.FL.p,
                       lp.FU.p);
    CIPC_LAUNCH_CHECK();
    lp.forwardReady = true;
}

// inference form: the encoders, message layers, node updates and the diagonal decoder on
// the matrix cores (BF16 inputs and weights, FP32 accumulation and FP32 activations); the
// off-diagonal decoder (nine outputs per lower edge) in FP32
static void lp_forward_mfma(Ctx& c) {
    LearnedPrecond& lp = LP(c);
    const LpLayout& L = lp.L;
    const float* Wq = lp.Wq.p;
    const int nV = lp.nV, nE = lp.nE, nL = lp.nL;
    mfma_dense(c, nV, {lp.nodeF.p, nullptr, kLpNodeIn}, kNoSeg, kNoSeg, Wq + L.we, Wq + L.be, kLpHidden, 1, nullptr, lp.h[0].p);
    mfma_dense(c, nE, {lp.edgeF.p, nullptr, kLpEdgeIn}, kNoSeg, kNoSeg, Wq + L.wx, Wq + L.bx, kLpHidden, 1, nullptr, lp.e[0].p);
    for (int l = 0; l < kLpLayers; ++l) {
        const LpSeg ms0 = {lp.h[l].p, lp.eSrc.p, kLpHidden}, ms1 = {lp.h[l].p, lp.eDst.p, kLpHidden}, ms2 = {lp.e[l].p, nullptr, kLpHidden};
        if (nE > 0) {
            hipLaunchKernelGGL(k_lp_dense_mfma<3 * kLpHidden>, dim3((nE + 63) / 64), dim3(256), 0, c.s, nE, ms0, ms1, ms2, Wq + L.wm[l],
                               Wq + L.bm[l], 1, lp.m[l].p, (const float*)lp.e[l].p, lp.e[l + 1].p);
            CIPC_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_lp_agg, dim3((nV + 3) / 4), dim3(256), 0, c.s, nV, (const int*)lp.erp.p, (const float*)lp.m[l].p, lp.agg[l].p);
        CIPC_LAUNCH_CHECK();
        // h_{l+1} = h_l + relu(Wu [h_l, agg] + bu)
        mfma_dense(c, nV, {lp.h[l].p, nullptr, kLpHidden}, {lp.agg[l].p, nullptr, kLpHidden}, kNoSeg, Wq + L.wu[l], Wq + L.bu[l], kLpHidden,
                   1, lp.h[l].p, lp.h[l + 1].p);
    }
    const float* hL = lp.h[kLpLayers].p;
    dense<9>(c, nL, {hL, lp.lowSrc.p, kLpHidden}, {hL, lp.lowDst.p, kLpHidden}, {lp.e[kLpLayers].p, lp.lEdge.p, kLpHidden}, Wq + L.wo,
             Wq + L.bo, 0, nullptr, lp.Gl.p);
    mfma_dense(c, nV, {hL, nullptr, kLpHidden}, kNoSeg, kNoSeg, Wq + L.wd, Wq + L.bd, 6, 0, nullptr, lp.raw.p);
}

void lp_forward(Ctx& c) {
    LearnedPrecond& lp = LP(c);
    const LpLayout& L = lp.L;
    const float* Wq = lp.Wq.p;
    const int nV = lp.nV, nE = lp.nE, nL = lp.nL;
    const long H = kLpHidden;
    if (lp.mfma) {
        lp_forward_mfma(c);
        lp_assemble(c);
        return;
    }
    dense<kLpHidden>(c, nV, {lp.nodeF.p, nullptr, kLpNodeIn}, kNoSeg, kNoSeg, Wq + L.we, Wq + L.be, 1, nullptr, lp.h[0].p);
    dense<kLpHidden>(c, nE, {lp.edgeF.p, nullptr, kLpEdgeIn}, kNoSeg, kNoSeg, Wq + L.wx, Wq + L.bx, 1, nullptr, lp.e[0].p);
    for (int l = 0; l < kLpLayers; ++l) {
        const LpSeg ms0 = {lp.h[l].p, lp.eSrc.p, kLpHidden}, ms1 = {lp.h[l].p, lp.eDst.p, kLpHidden}, ms2 = {lp.e[l].p, nullptr, kLpHidden};
        dense<kLpHidden>(c, nE, ms0, ms1, ms2, Wq + L.wm[l], Wq + L.bm[l], 1, nullptr, lp.m[l].p);
        add(c, H * nE, lp.e[l].p, lp.m[l].p, lp.e[l + 1].p);
        hipLaunchKernelGGL(k_lp_agg, dim3((nV + 3) / 4), dim3(256), 0, c.s, nV, (const int*)lp.erp.p, (const float*)lp.m[l].p, lp.agg[l].p);
        CIPC_LAUNCH_CHECK();
        dense<kLpHidden>(c, nV, {lp.h[l].p, nullptr, kLpHidden}, {lp.agg[l].p, nullptr, kLpHidden}, kNoSeg, Wq + L.wu[l], Wq + L.bu[l], 1,
                         nullptr, lp.u[l].p);
        add(c, H * nV, lp.h[l].p, lp.u[l].p, lp.h[l + 1].p);
    }
    const float* hL = lp.h[kLpLayers].p;
    dense<9>(c, nL, {hL, lp.lowSrc.p, kLpHidden}, {hL, lp.lowDst.p, kLpHidden}, {lp.e[kLpLayers].p, lp.lEdge.p, kLpHidden}, Wq + L.wo,
             Wq + L.bo, 0, nullptr, lp.Gl.p);
    dense<6>(c, nV, {hL, nullptr, kLpHidden}, kNoSeg, kNoSeg, Wq + L.wd, Wq + L.bd, 0, nullptr, lp.raw.p);
    lp_assemble(c);
}

bool lp_refresh(Ctx& c) {
    LearnedPrecond& lp = LP(c);
    if (!lp.forwardReady || !lp.graphReady || lp.nV != c.nV || lp.nnzb != c.Ael.nnzb) return false;
    BSR& E = c.Ael;
    BSR& K = c.Acon;
    if (K.nnzb == 0) K.rowptr.ensure(c.nV + 1);
    const int nL = lp.nL, nV = lp.nV;
    hipLaunchKernelGGL(k_lp_node_features, dim3(nblocks(nV)), dim3(kBlock), 0, c.s, nV, (const int*)lp.diagBlk.p, (const double*)E.val.p,
                       E.nnzb, (const int*)K.rowptr.p, (const int*)K.col.p, (const double*)K.val.p, K.nnzb, (const double*)c.mass.p,
                       proj_mask(c), lp.S.p, lp.nodeF.p);
    hipLaunchKernelGGL(k_lp_assemble_G, dim3(nblocks(nL + nV)), dim3(kBlock), 0, c.s, nL, nV, lp.nnzb, (const int*)lp.lEdge.p,
                       (const int*)lp.eBlock.p, (const int*)lp.blkT.p, (const int*)lp.diagBlk.p, (const int*)c.Ael.col.p,
                       (const double*)lp.S.p, (const float*)lp.Gl.p, (const float*)lp.raw.p, lp.Gblk.p, lp.nLow,
                       (const int*)lp.posLE.p, (const int*)lp.posLD.p, (const int*)lp.posUE.p, (const int*)lp.posUD.p, lp.FL.p,
                       lp.FU.p);
    CIPC_LAUNCH_CHECK();
    return true;
}

void lp_prepare(Ctx& c) {
    LearnedPrecond& lp = LP(c);
    if (!lp.weightsReady) throw HipError("learned preconditioner: no weights loaded", -1);
    if (!lp.graphReady || lp.nV != c.nV || lp.nnzb != c.Ael.nnzb) lp_build_graph(c);
    features(c);
    lp_forward(c);
}

constexpr int kLpSpmvBlock = 256;
int lp_apply_partials(const Ctx& c) { return c.lpGrid; }

template <int LANES, class V>
static void lp_spmv_go(Ctx& c, const int* rp, const int* col, const uint32_t* val, int nnz, const V* x, V* y, const V* r, double* part,
                       const int* done, bool coarse) {
    constexpr int RPG = kLpSpmvBlock / LANES;
    const int nV = c.lp->nV;
    int rpx = (nV + 7) / 8;
    rpx = (rpx + RPG - 1) / RPG * RPG;
    const CoarseSpace& g = c.cg;
    hipLaunchKernelGGL((k_lp_spmv_bf16<LANES, kLpSpmvBlock, V>), dim3(c.lpGrid), dim3(kLpSpmvBlock), 0, c.s, nV, rpx, rp, col, val, nnz,
                       x, y, r, part, done, coarse ? (const int*)g.agg.p : nullptr, coarse ? (const double*)g.d.p : nullptr,
                       coarse ? (const double*)g.zc.p : nullptr);
}
template <class V>
static void lp_spmv(Ctx& c, const int* rp, const int* col, const uint32_t* val, int nnz, const V* x, V* y, const V* r, double* part,
                    const int* done = nullptr, bool coarse = false) {
    switch (c.lpLanes) {
        case 2: lp_spmv_go<2, V>(c, rp, col, val, nnz, x, y, r, part, done, coarse); break;
        case 4: lp_spmv_go<4, V>(c, rp, col, val, nnz, x, y, r, part, done, coarse); break;
        case 16: lp_spmv_go<16, V>(c, rp, col, val, nnz, x, y, r, part, done, coarse); break;
        default: lp_spmv_go<8, V>(c, rp, col, val, nnz, x, y, r, part, done, coarse); break;
    }
}

std::vector<uintptr_t> lp_graph_key(const Ctx& c) {
    const LearnedPrecond& lp = *c.lp;
    return {(uintptr_t)c.lpGrid, (uintptr_t)c.lpLanes, (uintptr_t)lp.nV, (uintptr_t)lp.nLow, (uintptr_t)lp.gLrp.p, (uintptr_t)lp.gLcol.p,
            (uintptr_t)lp.FL.p, (uintptr_t)lp.gUrp.p, (uintptr_t)lp.gUcol.p, (uintptr_t)lp.FU.p};
}

template <class V>
int lp_apply(Ctx& c, const V* r, V* z, V* tmp, double* part, bool coarse) {
    LearnedPrecond& lp = LP(c);
    // tmp = F r;  z = F^T tmp (+ P zc, the coarse correction of r already in c.cg.zc)
    // (+ partial sums of r.z), F = G S
    lp_spmv<V>(c, (const int*)lp.gLrp.p, (const int*)lp.gLcol.p, (const uint32_t*)lp.FL.p, lp.nLow, r, tmp, (const V*)nullptr, nullptr);
    lp_spmv<V>(c, (const int*)lp.gUrp.p, (const int*)lp.gUcol.p, (const uint32_t*)lp.FU.p, lp.nLow, (const V*)tmp, z, r, part, nullptr, coarse);
    CIPC_LAUNCH_CHECK();
    return c.lpGrid;
}
template int lp_apply<double>(Ctx&, const double*, double*, double*, double*, bool);
template int lp_apply<float>(Ctx&, const float*, float*, float*, double*, bool);

template <class V>
void lp_apply_upper(Ctx& c, const V* y, V* z, const V* r, double* part, const int* done, bool coarse) {
    LearnedPrecond& lp = LP(c);
    lp_spmv<V>(c, (const int*)lp.gUrp.p, (const int*)lp.gUcol.p, (const uint32_t*)lp.FU.p, lp.nLow, y, z, r, part, done, coarse);
    CIPC_LAUNCH_CHECK();
}
template void lp_apply_upper<double>(Ctx&, const double*, double*, const double*, double*, const int*, bool);
template void lp_apply_upper<float>(Ctx&, const float*, float*, const float*, double*, const int*, bool);

double lp_loss_grad(Ctx& c, int nProbes, uint64_t seed, float* gradOut) {
    LearnedPrecond& lp = LP(c);
    const bool mf = lp.mfma;
    lp.mfma = false;   // training differentiates the FP32 forward
    lp_prepare(c);
    lp.mfma = mf;
    const LpLayout& L = lp.L;
    BSR& E = c.Ael;
    BSR& K = c.Acon;
    const int nV = lp.nV, nE = lp.nE, nL = lp.nL;
    const long n3 = 3l * nV, H = kLpHidden;
    const int nb = nblocks(16l * nV);   // one row group of 16 rows per workgroup (k_lp_hhat, k_lp_apply_upper)
    double* w = lp.probes.ensure((size_t)nProbes * n3);
    double* u = lp.pu.ensure((size_t)nProbes * n3);
    double* v = lp.pv.ensure((size_t)nProbes * n3);
    double* part = lp.lossPart.ensure((size_t)nProbes * nb + nb);
    for (int p = 0; p < nProbes; ++p) {
        hipLaunchKernelGGL(k_lp_probe, dim3(nblocks(n3)), dim3(kBlock), 0, c.s, n3, seed * 0x100000001B3ull + (uint64_t)p * 0x632BE59BD9B4E019ull,
                           w + p * n3);
        hipLaunchKernelGGL(k_lp_apply_upper, dim3(nb), dim3(kBlock), 0, c.s, nV, (const int*)E.rowptr.p, (const int*)E.col.p,
                           (const float*)lp.Gblk.p, lp.nnzb, (const double*)nullptr, (const double*)(w + p * n3), u + p * n3,
                           (const double*)nullptr, (double*)nullptr);
        hipLaunchKernelGGL(k_lp_hhat, dim3(nb), dim3(kBlock), 0, c.s, nV, (const int*)E.rowptr.p, (const int*)E.col.p, (const double*)E.val.p,
                           E.nnzb, (const int*)K.rowptr.p, (const int*)K.col.p, (const double*)K.val.p, K.nnzb, (const double*)lp.S.p,
                           (const double*)(u + p * n3), v + p * n3, part + (long)p * nb);
        CIPC_LAUNCH_CHECK();
    }
    const double trace = host_sum(c, part, nProbes * nb) / nProbes;
    hipLaunchKernelGGL(k_lp_sum_rawdiag, dim3(nblocks(nV)), dim3(kBlock), 0, c.s, nV, (const float*)lp.raw.p, part);
    CIPC_LAUNCH_CHECK();
    const double sumRaw = host_sum(c, part, nblocks(nV));
    const double nn = (double)n3;
    const double loss = std::log(trace / nn) - 2.0 / nn * sumRaw;
    lp.lastLoss = loss;
    lp.lastTrace = trace;

    // ---- backward
    float* gW = lp.gW.p;
    hipLaunchKernelGGL(k_lp_zero, dim3(nblocks(L.total)), dim3(kBlock), 0, c.s, (long)L.total, gW);
    const long nEmax = std::max<long>(std::max<long>(nE, nL), nV);
    float* dGl = lp.dGl.ensure(9 * (size_t)std::max(nL, 1));
    float* draw = lp.draw.ensure(6 * (size_t)nV);
    float* dh = lp.dh.ensure(H * nV);
    float* dagg = lp.dagg.ensure(H * nV);
    float* de = lp.de.ensure(H * std::max(nE, 1));
    float* de2 = lp.de2.ensure(H * std::max(nE, 1));
    float* t1 = lp.dtmp.ensure(H * nEmax);
    float* t2 = lp.dtmp2.ensure(H * nEmax);
    float* t3 = lp.dtmp3.ensure(H * nEmax);
    hipLaunchKernelGGL(k_lp_grad_G, dim3(nblocks(nL + nV)), dim3(kBlock), 0, c.s, nL, nV, nProbes, n3, (const int*)lp.lEdge.p,
                       (const int*)lp.eSrc.p, (const int*)lp.eDst.p, (const double*)w, (const double*)v, 2.0 / (nProbes * trace),
                       (const float*)lp.raw.p, 2.0 / nn, dGl, draw);
    CIPC_LAUNCH_CHECK();
    const float* Wq = lp.Wq.p;
    const float* hL = lp.h[kLpLayers].p;
    // decoder, lower blocks
    wgrad<9>(c, nL, {hL, lp.lowSrc.p, kLpHidden}, {hL, lp.lowDst.p, kLpHidden}, {lp.e[kLpLayers].p, lp.lEdge.p, kLpHidden}, dGl, gW + L.wo,
             gW + L.bo);
    dense_t<9>(c, nL, dGl, Wq + L.wo, 0, nullptr, t1);
    dense_t<9>(c, nL, dGl, Wq + L.wo, kLpHidden, nullptr, t2);
    dense_t<9>(c, nL, dGl, Wq + L.wo, 2 * kLpHidden, nullptr, t3);
    hipLaunchKernelGGL(k_lp_zero, dim3(nblocks(H * nV)), dim3(kBlock), 0, c.s, H * nV, dh);
    if (nE > 0) hipLaunchKernelGGL(k_lp_zero, dim3((unsigned)((H * nE + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.s, H * nE, de);
    if (nL > 0) {
        hipLaunchKernelGGL(k_lp_scatter_lower, dim3((unsigned)((H * nL + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.s, nL, (const int*)lp.lEdge.p,
                           (const float*)t3, de);
        hipLaunchKernelGGL(k_lp_gather_lower, dim3((nV + 3) / 4), dim3(256), 0, c.s, nV, (const int*)lp.lSrcOff.p, (const int*)lp.lSrcIdx.p,
                           (const int*)lp.lDstOff.p, (const int*)lp.lDstIdx.p, (const float*)t1, (const float*)t2, dh);
    }
    CIPC_LAUNCH_CHECK();
    // decoder, diagonal blocks
    wgrad<6>(c, nV, {hL, nullptr, kLpHidden}, kNoSeg, kNoSeg, draw, gW + L.wd, gW + L.bd);
    dense_t<6>(c, nV, draw, Wq + L.wd, 0, dh, dh);
    // message-passing layers in reverse
    for (int l = kLpLayers - 1; l >= 0; --l) {
        // h[l+1] = h[l] + u[l],  u[l] = relu(Wu [h[l], agg[l]] + bu)
        hipLaunchKernelGGL(k_lp_relu_mask, dim3(nblocks(H * nV)), dim3(kBlock), 0, c.s, H * nV, (const float*)lp.u[l].p, (const float*)dh, t1);
        CIPC_LAUNCH_CHECK();
        wgrad<kLpHidden>(c, nV, {lp.h[l].p, nullptr, kLpHidden}, {lp.agg[l].p, nullptr, kLpHidden}, kNoSeg, t1, gW + L.wu[l], gW + L.bu[l]);
        dense_t<kLpHidden>(c, nV, t1, Wq + L.wu[l], kLpHidden, nullptr, dagg);
        dense_t<kLpHidden>(c, nV, t1, Wq + L.wu[l], 0, dh, dh);
        if (nE == 0) continue;
        // e[l+1] = e[l] + m[l],  m[l] = relu(Wm [h[l]_src, h[l]_dst, e[l]] + bm),  agg[l] = row sums of m[l]
        hipLaunchKernelGGL(k_lp_msg_grad, dim3((unsigned)((H * nE + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.s, H * nE, (const int*)lp.eSrc.p,
                           (const float*)de, (const float*)dagg, (const float*)lp.m[l].p, t2);
        CIPC_LAUNCH_CHECK();
        wgrad<kLpHidden>(c, nE, {lp.h[l].p, lp.eSrc.p, kLpHidden}, {lp.h[l].p, lp.eDst.p, kLpHidden}, {lp.e[l].p, nullptr, kLpHidden}, t2,
                         gW + L.wm[l], gW + L.bm[l]);
        dense_t<kLpHidden>(c, nE, t2, Wq + L.wm[l], 0, nullptr, t3);
        dense_t<kLpHidden>(c, nE, t2, Wq + L.wm[l], kLpHidden, nullptr, de2);
        dense_t<kLpHidden>(c, nE, t2, Wq + L.wm[l], 2 * kLpHidden, de, de);
        hipLaunchKernelGGL(k_lp_gather_node, dim3((nV + 3) / 4), dim3(256), 0, c.s, nV, (const int*)lp.erp.p, (const int*)lp.eT.p,
                           (const float*)t3, (const float*)de2, dh);
        CIPC_LAUNCH_CHECK();
    }
    // encoders
    hipLaunchKernelGGL(k_lp_relu_mask, dim3(nblocks(H * nV)), dim3(kBlock), 0, c.s, H * nV, (const float*)lp.h[0].p, (const float*)dh, t1);
    CIPC_LAUNCH_CHECK();
    wgrad<kLpHidden>(c, nV, {lp.nodeF.p, nullptr, kLpNodeIn}, kNoSeg, kNoSeg, t1, gW + L.we, gW + L.be);
    if (nE > 0) {
        hipLaunchKernelGGL(k_lp_relu_mask, dim3((unsigned)((H * nE + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.s, H * nE, (const float*)lp.e[0].p,
                           (const float*)de, t2);
        CIPC_LAUNCH_CHECK();
        wgrad<kLpHidden>(c, nE, {lp.edgeF.p, nullptr, kLpEdgeIn}, kNoSeg, kNoSeg, t2, gW + L.wx, gW + L.bx);
    }
    if (gradOut) {
        CIPC_HIP(hipMemcpyAsync(gradOut, gW, sizeof(float) * L.total, hipMemcpyDeviceToHost, c.s));
    }
    CIPC_HIP(hipStreamSynchronize(c.s));
    return loss;
}

double lp_train_step(Ctx& c, int nProbes, uint64_t seed, double lr) {
    LearnedPrecond& lp = LP(c);
    const double loss = lp_loss_grad(c, nProbes, seed, nullptr);
    const int n = lp.L.total;
    lp.adamT++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)lp.adamT), c2 = 1.f - std::pow(b2, (float)lp.adamT);
    hipLaunchKernelGGL(k_lp_adam, dim3(nblocks(n)), dim3(kBlock), 0, c.s, n, lp.W.p, (const float*)lp.gW.p, lp.mW.p, lp.vW.p, (float)lr, b1, b2, c1,
                       c2, 1e-8f);
    hipLaunchKernelGGL(k_lp_quantize, dim3(nblocks(n)), dim3(kBlock), 0, c.s, n, (const float*)lp.W.p, lp.Wq.p);
    CIPC_LAUNCH_CHECK();
    lp.forwardReady = false;
    return loss;
}

}  // namespace cipc
