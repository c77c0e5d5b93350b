// Intersection-free step size: FEM/IPC.h:1879-2244 (Compute_Intersection_Free_StepSize)
// with the CCD spatial hash of Grid/SPATIAL_HASH.h:432-622 and the additive CCD of
// Math/Distance/CCD.h:188-395.
//
// The reference's candidate pairs are those whose swept voxel ranges overlap on its
// CCD voxel grid (voxel = mean boundary-edge length at x, origin = min corner of the
// swept boundary nodes minus offset/2); that grid, the step clamp
// (alpha /= alpha * mean|p| / voxel when > 1) and the overlap test are reproduced
// exactly, so the candidate set, and with it the step size, matches the reference
// rule.  Candidates are enumerated through cells of f x f x f voxels (f grown until the
// cell table fits), then filtered by the exact voxel-range overlap.
//
// Each candidate runs the additive CCD in FP64 with the running per-primitive step as
// its cut-off, as the reference does; every loop stops after at most ACCD_MAX_ITERS
// advances (touching stencils would otherwise never end) with the conservative time
// accumulated so far.  The global step is a min-reduction over positive doubles
// through their bit patterns (atomicMin on uint64).
#include "ctx.hpp"
#include "dev_math.hpp"
#include <algorithm>
#include <cmath>

namespace cipc {

static constexpr long ACCD_MAX_ITERS = 1000000;

#pragma clang fp contract(off)

__device__ __forceinline__ bool pt_ccd_broadphase(d3 p, d3 t0, d3 t1, d3 t2, d3 dp, d3 dt0, d3 dt1, d3 dt2, double dist) {
    d3 pe = p + dp, q0 = t0 + dt0, q1 = t1 + dt1, q2 = t2 + dt2;
    double maxp, minp, mx, mn;
    maxp = dmax(p.x, pe.x); minp = dmin(p.x, pe.x);
    mx = dmax(dmax(dmax(dmax(dmax(t0.x, t1.x), t2.x), q0.x), q1.x), q2.x);
    mn = dmin(dmin(dmin(dmin(dmin(t0.x, t1.x), t2.x), q0.x), q1.x), q2.x);
    if (minp - mx > dist || mn - maxp > dist) return false;
    maxp = dmax(p.y, pe.y); minp = dmin(p.y, pe.y);
    mx = dmax(dmax(dmax(dmax(dmax(t0.y, t1.y), t2.y), q0.y), q1.y), q2.y);
    mn = dmin(dmin(dmin(dmin(dmin(t0.y, t1.y), t2.y), q0.y), q1.y), q2.y);
    if (minp - mx > dist || mn - maxp > dist) return false;
    maxp = dmax(p.z, pe.z); minp = dmin(p.z, pe.z);
    mx = dmax(dmax(dmax(dmax(dmax(t0.z, t1.z), t2.z), q0.z), q1.z), q2.z);
    mn = dmin(dmin(dmin(dmin(dmin(t0.z, t1.z), t2.z), q0.z), q1.z), q2.z);
    if (minp - mx > dist || mn - maxp > dist) return false;
    return true;
}
__device__ __forceinline__ bool ee_ccd_broadphase(d3 a0, d3 a1, d3 b0, d3 b1, d3 da0, d3 da1, d3 db0, d3 db1, double dist) {
    d3 A2 = a0 + da0, A3 = a1 + da1, B2 = b0 + db0, B3 = b1 + db1;
    double mxa, mna, mxb, mnb;
    mxa = dmax(dmax(dmax(a0.x, a1.x), A2.x), A3.x); mna = dmin(dmin(dmin(a0.x, a1.x), A2.x), A3.x);
    mxb = dmax(dmax(dmax(b0.x, b1.x), B2.x), B3.x); mnb = dmin(dmin(dmin(b0.x, b1.x), B2.x), B3.x);
    if (mna - mxb > dist || mnb - mxa > dist) return false;
    mxa = dmax(dmax(dmax(a0.y, a1.y), A2.y), A3.y); mna = dmin(dmin(dmin(a0.y, a1.y), A2.y), A3.y);
    mxb = dmax(dmax(dmax(b0.y, b1.y), B2.y), B3.y); mnb = dmin(dmin(dmin(b0.y, b1.y), B2.y), B3.y);
    if (mna - mxb > dist || mnb - mxa > dist) return false;
    mxa = dmax(dmax(dmax(a0.z, a1.z), A2.z), A3.z); mna = dmin(dmin(dmin(a0.z, a1.z), A2.z), A3.z);
    mxb = dmax(dmax(dmax(b0.z, b1.z), B2.z), B3.z); mnb = dmin(dmin(dmin(b0.z, b1.z), B2.z), B3.z);
    if (mna - mxb > dist || mnb - mxa > dist) return false;
    return true;
}

// CCD.h:280-328
__device__ bool pt_accd(d3 p, d3 t0, d3 t1, d3 t2, d3 dp, d3 dt0, d3 dt1, d3 dt2, double eta, double th, double& toc) {
    d3 mov = {(dt0.x + dt1.x + dt2.x + dp.x) / 4, (dt0.y + dt1.y + dt2.y + dp.y) / 4, (dt0.z + dt1.z + dt2.z + dp.z) / 4};
    dt0 = dt0 - mov; dt1 = dt1 - mov; dt2 = dt2 - mov; dp = dp - mov;
    double m2 = dmax(dmax(sq(dt0), sq(dt1)), sq(dt2));
    double maxDispMag = sqrt(sq(dp)) + sqrt(m2);
    if (maxDispMag == 0) return false;
    double dist2_cur = dist2_pt_unclassified(p, t0, t1, t2);
    double dist_cur = sqrt(dist2_cur);
    double gap = eta * (dist2_cur - th * th) / (dist_cur + th);
    double toc_prev = toc;
    double t = 0;
    for (long it = 0;; ++it) {
        if (it >= ACCD_MAX_ITERS) { toc = t; return true; }
        double lb = (1 - eta) * (dist2_cur - th * th) / ((dist_cur + th) * maxDispMag);
        p = p + lb * dp; t0 = t0 + lb * dt0; t1 = t1 + lb * dt1; t2 = t2 + lb * dt2;
        dist2_cur = dist2_pt_unclassified(p, t0, t1, t2);
        dist_cur = sqrt(dist2_cur);
        if (t != 0 && ((dist2_cur - th * th) / (dist_cur + th) < gap)) break;
        t += lb;
        if (t > toc_prev) { toc = t; return false; }
    }
    toc = t;
    return true;
}
// CCD.h:331-395
__device__ bool ee_accd(d3 ea0, d3 ea1, d3 eb0, d3 eb1, d3 dea0, d3 dea1, d3 deb0, d3 deb1, double eta, double th, double& toc) {
    d3 mov = {(dea0.x + dea1.x + deb0.x + deb1.x) / 4, (dea0.y + dea1.y + deb0.y + deb1.y) / 4, (dea0.z + dea1.z + deb0.z + deb1.z) / 4};
    dea0 = dea0 - mov; dea1 = dea1 - mov; deb0 = deb0 - mov; deb1 = deb1 - mov;
    double maxDispMag = sqrt(dmax(sq(dea0), sq(dea1))) + sqrt(dmax(sq(deb0), sq(deb1)));
    if (maxDispMag == 0) return false;
    double th2 = th * th;
    double dist2_cur = dist2_ee_unclassified(ea0, ea1, eb0, eb1);
    double dFunc = dist2_cur - th2;
    if (dFunc <= 0) {
        double l[4] = {sq(ea0 - eb0), sq(ea0 - eb1), sq(ea1 - eb0), sq(ea1 - eb1)};
        double m = l[0];
        for (int i = 1; i < 4; ++i) if (l[i] < m) m = l[i];
        dist2_cur = m; dFunc = dist2_cur - th2;
    }
    double dist_cur = sqrt(dist2_cur);
    double gap = eta * dFunc / (dist_cur + th);
    double toc_prev = toc;
    double t = 0;
    for (long it = 0;; ++it) {
        if (it >= ACCD_MAX_ITERS) { toc = t; return true; }
        double lb = (1 - eta) * dFunc / ((dist_cur + th) * maxDispMag);
        ea0 = ea0 + lb * dea0; ea1 = ea1 + lb * dea1; eb0 = eb0 + lb * deb0; eb1 = eb1 + lb * deb1;
        dist2_cur = dist2_ee_unclassified(ea0, ea1, eb0, eb1);
        dFunc = dist2_cur - th2;
        if (dFunc <= 0) {
            double l[4] = {sq(ea0 - eb0), sq(ea0 - eb1), sq(ea1 - eb0), sq(ea1 - eb1)};
            double m = l[0];
            for (int i = 1; i < 4; ++i) if (l[i] < m) m = l[i];
            dist2_cur = m; dFunc = dist2_cur - th2;
        }
        dist_cur = sqrt(dist2_cur);
        if (t != 0 && (dFunc / (dist_cur + th) < gap)) break;
        t += lb;
        if (t > toc_prev) { toc = t; return false; }
    }
    toc = t;
    return true;
}

// CCD.h:237-277
__device__ __forceinline__ bool pe_ccd_broadphase(d3 p, d3 e0, d3 e1, d3 dp, d3 de0, d3 de1, double dist) {
    d3 pe = p + dp, q0 = e0 + de0, q1 = e1 + de1;
    double maxp, minp, mx, mn;
    maxp = dmax(p.x, pe.x); minp = dmin(p.x, pe.x);
    mx = dmax(dmax(dmax(e0.x, e1.x), q0.x), q1.x); mn = dmin(dmin(dmin(e0.x, e1.x), q0.x), q1.x);
    if (minp - mx > dist || mn - maxp > dist) return false;
    maxp = dmax(p.y, pe.y); minp = dmin(p.y, pe.y);
    mx = dmax(dmax(dmax(e0.y, e1.y), q0.y), q1.y); mn = dmin(dmin(dmin(e0.y, e1.y), q0.y), q1.y);
    if (minp - mx > dist || mn - maxp > dist) return false;
    maxp = dmax(p.z, pe.z); minp = dmin(p.z, pe.z);
    mx = dmax(dmax(dmax(e0.z, e1.z), q0.z), q1.z); mn = dmin(dmin(dmin(e0.z, e1.z), q0.z), q1.z);
    if (minp - mx > dist || mn - maxp > dist) return false;
    return true;
}
__device__ __forceinline__ bool pp_ccd_broadphase(d3 p0, d3 p1, d3 dp0, d3 dp1, double dist) {
    d3 q0 = p0 + dp0, q1 = p1 + dp1;
    if (dmin(p0.x, q0.x) - dmax(p1.x, q1.x) > dist || dmin(p1.x, q1.x) - dmax(p0.x, q0.x) > dist) return false;
    if (dmin(p0.y, q0.y) - dmax(p1.y, q1.y) > dist || dmin(p1.y, q1.y) - dmax(p0.y, q0.y) > dist) return false;
    if (dmin(p0.z, q0.z) - dmax(p1.z, q1.z) > dist || dmin(p1.z, q1.z) - dmax(p0.z, q0.z) > dist) return false;
    return true;
}
// CCD.h:397-441
__device__ bool pe_accd(d3 p, d3 e0, d3 e1, d3 dp, d3 de0, d3 de1, double eta, double th, double& toc) {
    d3 mov = {(dp.x + de0.x + de1.x) / 3, (dp.y + de0.y + de1.y) / 3, (dp.z + de0.z + de1.z) / 3};
    de0 = de0 - mov; de1 = de1 - mov; dp = dp - mov;
    double maxDispMag = sqrt(sq(dp)) + sqrt(dmax(sq(de0), sq(de1)));
    if (maxDispMag == 0) return false;
    double dist2_cur = dist2_pe_unclassified(p, e0, e1);
    double dist_cur = sqrt(dist2_cur);
    double gap = eta * (dist2_cur - th * th) / (dist_cur + th);
    double toc_prev = toc;
    double t = 0;
    for (long it = 0;; ++it) {
        if (it >= ACCD_MAX_ITERS) { toc = t; return true; }
        double lb = (1 - eta) * (dist2_cur - th * th) / ((dist_cur + th) * maxDispMag);
        p = p + lb * dp; e0 = e0 + lb * de0; e1 = e1 + lb * de1;
        dist2_cur = dist2_pe_unclassified(p, e0, e1);
        dist_cur = sqrt(dist2_cur);
        if (t != 0 && (dist2_cur - th * th) / (dist_cur + th) < gap) break;
        t += lb;
        if (t > toc_prev) { toc = t; return false; }
    }
    toc = t;
    return true;
}
// CCD.h:443-483
__device__ bool pp_accd(d3 p0, d3 p1, d3 dp0, d3 dp1, double eta, double th, double& toc) {
    d3 mov = {(dp0.x + dp1.x) / 2, (dp0.y + dp1.y) / 2, (dp0.z + dp1.z) / 2};
    dp1 = dp1 - mov; dp0 = dp0 - mov;
    double maxDispMag = sqrt(sq(dp0)) + sqrt(sq(dp1));
    if (maxDispMag == 0) return false;
    double dist2_cur = dist2_pp(p0, p1);
    double dist_cur = sqrt(dist2_cur);
    double gap = eta * (dist2_cur - th * th) / (dist_cur + th);
    double toc_prev = toc;
    double t = 0;
    for (long it = 0;; ++it) {
        if (it >= ACCD_MAX_ITERS) { toc = t; return true; }
        double lb = (1 - eta) * (dist2_cur - th * th) / ((dist_cur + th) * maxDispMag);
        p0 = p0 + lb * dp0; p1 = p1 + lb * dp1;
        dist2_cur = dist2_pp(p0, p1);
        dist_cur = sqrt(dist2_cur);
        if (t != 0 && (dist2_cur - th * th) / (dist_cur + th) < gap) break;
        t += lb;
        if (t > toc_prev) { toc = t; return false; }
    }
    toc = t;
    return true;
}

// ------------------------------------------------------------------ clamp inputs
__global__ void k_edge_len(const double* __restrict__ x, const int* __restrict__ bEdge, int nBE, double* __restrict__ len) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nBE) return;
    len[e] = sqrt(sq(ld3(x, bEdge[e]) - ld3(x, bEdge[nBE + e])));
}
__global__ void k_pabs(const double* __restrict__ p, const int* __restrict__ bNode, int nBN, double* __restrict__ out) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nBN) return;
    int v = bNode[s];
    out[s] = (fabs(p[3 * v]) + fabs(p[3 * v + 1])) + fabs(p[3 * v + 2]);
}
// swept bounding box of the boundary nodes (x and x + alpha p)
__global__ void k_swept_partial(const double* __restrict__ x, const double* __restrict__ p, double alpha, const int* __restrict__ bNode,
                                int nBN, double* __restrict__ part) {
    __shared__ double lo[3][kBlock], hi[3][kBlock];
    int t = threadIdx.x;
    double l[3] = {1e300, 1e300, 1e300}, h[3] = {-1e300, -1e300, -1e300};
    for (int s = blockIdx.x * blockDim.x + t; s < nBN; s += gridDim.x * blockDim.x) {
        int v = bNode[s];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double a = x[3 * v + d], b = x[3 * v + d] + alpha * p[3 * v + d];
            double mn = dmin(a, b), mx = dmax(a, b);
            l[d] = mn < l[d] ? mn : l[d]; h[d] = mx > h[d] ? mx : h[d];
        }
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
        for (int d = 0; d < 3; ++d) { part[6 * blockIdx.x + d] = lo[d][0]; part[6 * blockIdx.x + 3 + d] = hi[d][0]; }
}
struct VoxGrid { double lb[3]; double inv; double halfXi; double alpha; int f; int edim[3]; };

// per boundary node: voxel index range of its swept, offset-padded box (SPATIAL_HASH.h:546-553)
__global__ void k_node_vox(const double* __restrict__ x, const double* __restrict__ p, const int* __restrict__ bNode, int nBN, VoxGrid G,
                           int* __restrict__ vlo, int* __restrict__ vhi, int nV) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nBN) return;
    int v = bNode[s];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double a = x[3 * v + d], b = x[3 * v + d] + G.alpha * p[3 * v + d];
        double mn = dmin(a, b) - G.halfXi, mx = dmax(a, b) + G.halfXi;
        vlo[d * nV + v] = (int)floor((mn - G.lb[d]) * G.inv);
        vhi[d * nV + v] = (int)floor((mx - G.lb[d]) * G.inv);
    }
}
#pragma clang fp contract(on)

__device__ __forceinline__ int ecell(int vox, int f, int dim) {
    int c = vox >= 0 ? vox / f : 0;
    return c < dim ? c : dim - 1;
}
// primitive voxel range = union of its nodes' ranges; count of enumeration cells
// (0 for primitives covering more than kLargeCells cells, which go to the large list)
__global__ void k_prim_vox(const int* __restrict__ verts, int n, int nv, const int* __restrict__ vlo, const int* __restrict__ vhi, int nV,
                           VoxGrid G, int* __restrict__ plo, int* __restrict__ phi, int4* __restrict__ box, uint32_t* __restrict__ cnt,
                           uint8_t* __restrict__ lgFlag, int* __restrict__ lgList, uint32_t* __restrict__ lgCount) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t c = 1;
    int blo[3], bhi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int lo = 1 << 30, hi = -(1 << 30);
        for (int k = 0; k < nv; ++k) {
            int v = verts[(long)k * n + i];
            lo = min(lo, vlo[d * nV + v]);
            hi = max(hi, vhi[d * nV + v]);
        }
        plo[(long)d * n + i] = lo;
        phi[(long)d * n + i] = hi;
        blo[d] = lo;
        bhi[d] = hi;
        c *= (uint64_t)(ecell(hi, G.f, G.edim[d]) - ecell(lo, G.f, G.edim[d]) + 1);
    }
    // the same box as one 32-byte record (lo xyz, hi x | hi yz) for the candidate loops
    box[2l * i] = make_int4(blo[0], blo[1], blo[2], bhi[0]);
    box[2l * i + 1] = make_int4(bhi[1], bhi[2], 0, 0);
    const bool large = c > kLargeCells;
    lgFlag[i] = large ? 1 : 0;
    cnt[i] = large ? 0u : (uint32_t)c;
    if (large) lgList[atomicAdd(lgCount, 1u)] = i;
}
__global__ void k_prim_vox_emit(int n, const int* __restrict__ plo, const int* __restrict__ phi, VoxGrid G, const uint8_t* __restrict__ lgFlag,
                                const uint32_t* __restrict__ off, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || lgFlag[i]) return;
    int a[3], b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) { a[d] = ecell(plo[(long)d * n + i], G.f, G.edim[d]); b[d] = ecell(phi[(long)d * n + i], G.f, G.edim[d]); }
    uint32_t o = off[i];
    for (int iz = a[2]; iz <= b[2]; ++iz)
        for (int iy = a[1]; iy <= b[1]; ++iy)
            for (int ix = a[0]; ix <= b[0]; ++ix) {
                keys[o] = (uint64_t)ix + (uint64_t)G.edim[0] * ((uint64_t)iy + (uint64_t)G.edim[1] * (uint64_t)iz);
                vals[o] = (uint32_t)i;
                ++o;
            }
}
__global__ void k_cell_bounds2(const uint64_t* __restrict__ keys, long n, uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) start[k] = (uint32_t)i;
    if (i == n - 1 || keys[i + 1] != k) end[k] = (uint32_t)(i + 1);
}

__device__ __forceinline__ void atomic_min_pos(unsigned long long* a, double v) {
    atomicMin(a, (unsigned long long)__double_as_longlong(v));
}

// voxel box of primitive i from its 32-byte record: lo = (a.x, a.y, a.z), hi = (a.w, b.x, b.y)
struct VBox {
    int4 a, b;
};
__device__ __forceinline__ VBox ld_box(const int4* __restrict__ box, int i) { return {box[2l * i], box[2l * i + 1]}; }
__device__ __forceinline__ bool box_overlap(const VBox& A, const VBox& B) {
    return !(A.a.x > B.a.w || B.a.x > A.a.w || A.a.y > B.b.x || B.a.y > A.b.x || A.a.z > B.b.y || B.a.z > A.b.y);
}

// point vI against triangle t (IPC.h:2120-2164): exact voxel-range overlap (the
// triangle's box record), shared vertex and Dirichlet filters, swept-box broad phase,
// additive CCD with the running step `largest` as cut-off
__device__ __forceinline__ void pt_ccd_candidate_box(const double* __restrict__ x, const double* __restrict__ p, const int* __restrict__ tri,
                                                     int nT, const uint8_t* __restrict__ dbc, const VBox& T, int vI, const int* nlo,
                                                     const int* nhi, d3 P, d3 dP, int t, double xi, double& largest,
                                                     uint32_t* __restrict__ zero) {
    if (nlo[0] > T.a.w || T.a.x > nhi[0] || nlo[1] > T.b.x || T.a.y > nhi[1] || nlo[2] > T.b.y || T.a.z > nhi[2]) return;
    int a = tri[t], b = tri[nT + t], c = tri[2 * nT + t];
    if (vI == a || vI == b || vI == c) return;
    if (dbc[vI] && dbc[a] && dbc[b] && dbc[c]) return;
    d3 t0 = ld3(x, a), t1 = ld3(x, b), t2 = ld3(x, c);
    d3 d0 = ld3(p, a), d1 = ld3(p, b), d2 = ld3(p, c);
    if (!pt_ccd_broadphase(P, t0, t1, t2, dP, d0, d1, d2, xi)) return;
    double la = largest;
    if (pt_accd(P, t0, t1, t2, dP, d0, d1, d2, 0.1, xi, la)) {
        if (largest > la) largest = la;
        if (la == 0) atomicOr(zero, 1u);
    }
}
// edge pair eA < eB (IPC.h:2166-2242) with both box records loaded; edge A first in
// the additive CCD
__device__ __forceinline__ void ee_ccd_candidate_box(const double* __restrict__ x, const double* __restrict__ p, const int* __restrict__ bEdge,
                                                     int nBE, const uint8_t* __restrict__ dbc, const VBox& A, const VBox& B, int eA, int eB,
                                                     double xi, double& largest, uint32_t* __restrict__ zero) {
    if (!box_overlap(A, B)) return;
    int a0 = bEdge[eA], a1 = bEdge[nBE + eA], b0 = bEdge[eB], b1 = bEdge[nBE + eB];
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
    if (dbc[a0] && dbc[a1] && dbc[b0] && dbc[b1]) return;
    d3 A0 = ld3(x, a0), A1 = ld3(x, a1), dA0 = ld3(p, a0), dA1 = ld3(p, a1);
    d3 B0 = ld3(x, b0), B1 = ld3(x, b1), dB0 = ld3(p, b0), dB1 = ld3(p, b1);
    if (!ee_ccd_broadphase(A0, A1, B0, B1, dA0, dA1, dB0, dB1, xi)) return;
    double la = largest;
    if (ee_accd(A0, A1, B0, B1, dA0, dA1, dB0, dB1, 0.1, xi, la)) {
        if (largest > la) largest = la;
        if (la == 0) atomicOr(zero, 1u);
    }
}
__global__ void k_ccd_pt(const double* __restrict__ x, const double* __restrict__ p, const int* __restrict__ bNode, int nBN,
                         const int* __restrict__ tri, int nT, const uint8_t* __restrict__ dbc, const int* __restrict__ vlo,
                         const int* __restrict__ vhi, int nV, const int4* __restrict__ tbox, VoxGrid G,
                         const uint32_t* __restrict__ start, const uint32_t* __restrict__ end, const int4* __restrict__ crec,
                         const int* __restrict__ lgList, const uint32_t* __restrict__ lgCount, double alpha0, double xi,
                         unsigned long long* __restrict__ gmin, uint32_t* __restrict__ zero) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nBN) return;
    int vI = bNode[s];
    int nlo[3], nhi[3], qlo[3], qhi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        nlo[d] = vlo[d * nV + vI]; nhi[d] = vhi[d * nV + vI];
        qlo[d] = ecell(nlo[d], G.f, G.edim[d]); qhi[d] = ecell(nhi[d], G.f, G.edim[d]);
    }
    d3 P = ld3(x, vI), dP = ld3(p, vI);
    double largest = alpha0;
    for (int iz = qlo[2]; iz <= qhi[2]; ++iz)
        for (int iy = qlo[1]; iy <= qhi[1]; ++iy)
            for (int ix = qlo[0]; ix <= qhi[0]; ++ix) {
                uint64_t cell = (uint64_t)ix + (uint64_t)G.edim[0] * ((uint64_t)iy + (uint64_t)G.edim[1] * (uint64_t)iz);
                uint32_t st = start[cell], en = end[cell];
                for (uint32_t k = st; k < en; ++k) {
                    const VBox T = ld_box(crec, (int)k);   // the cell's records are contiguous
                    const int t = T.b.z;
                    if (ix != max(qlo[0], ecell(T.a.x, G.f, G.edim[0])) || iy != max(qlo[1], ecell(T.a.y, G.f, G.edim[1])) ||
                        iz != max(qlo[2], ecell(T.a.z, G.f, G.edim[2])))
                        continue;
                    pt_ccd_candidate_box(x, p, tri, nT, dbc, T, vI, nlo, nhi, P, dP, t, xi, largest, zero);
                }
            }
    const uint32_t nl = *lgCount;
    for (uint32_t k = 0; k < nl; ++k) {
        const int t = lgList[k];
        pt_ccd_candidate_box(x, p, tri, nT, dbc, ld_box(tbox, t), vI, nlo, nhi, P, dP, t, xi, largest, zero);
    }
    if (largest < alpha0) atomic_min_pos(gmin, largest);
}

// pair coverage as in k_query_ee: grid-listed edges take grid-listed partners of
// larger index plus every large edge; large edges take large partners of larger index
__global__ void k_ccd_ee(const double* __restrict__ x, const double* __restrict__ p, const int* __restrict__ bEdge, int nBE,
                         const uint8_t* __restrict__ dbc, const int4* __restrict__ ebox, VoxGrid G,
                         const uint32_t* __restrict__ start, const uint32_t* __restrict__ end, const int4* __restrict__ crec,
                         const uint8_t* __restrict__ lgFlag, const int* __restrict__ lgList, const uint32_t* __restrict__ lgCount,
                         double alpha0, double xi, unsigned long long* __restrict__ gmin, uint32_t* __restrict__ zero) {
    int eI = blockIdx.x * blockDim.x + threadIdx.x;
    if (eI >= nBE) return;
    double largest = alpha0;
    const uint32_t nl = *lgCount;
    const VBox A = ld_box(ebox, eI);
    if (lgFlag[eI]) {
        for (uint32_t k = 0; k < nl; ++k) {
            int eJ = lgList[k];
            if (eJ > eI) ee_ccd_candidate_box(x, p, bEdge, nBE, dbc, A, ld_box(ebox, eJ), eI, eJ, xi, largest, zero);
        }
    } else {
        const int qlo[3] = {ecell(A.a.x, G.f, G.edim[0]), ecell(A.a.y, G.f, G.edim[1]), ecell(A.a.z, G.f, G.edim[2])};
        const int qhi[3] = {ecell(A.a.w, G.f, G.edim[0]), ecell(A.b.x, G.f, G.edim[1]), ecell(A.b.y, G.f, G.edim[2])};
        for (int iz = qlo[2]; iz <= qhi[2]; ++iz)
            for (int iy = qlo[1]; iy <= qhi[1]; ++iy)
                for (int ix = qlo[0]; ix <= qhi[0]; ++ix) {
                    uint64_t cell = (uint64_t)ix + (uint64_t)G.edim[0] * ((uint64_t)iy + (uint64_t)G.edim[1] * (uint64_t)iz);
                    uint32_t st = start[cell], en = end[cell];
                    for (uint32_t k = st; k < en; ++k) {
                        const VBox B = ld_box(crec, (int)k);   // the cell's records are contiguous
                        const int eJ = B.b.z;
                        if (eJ <= eI) continue;
                        if (ix != max(qlo[0], ecell(B.a.x, G.f, G.edim[0])) || iy != max(qlo[1], ecell(B.a.y, G.f, G.edim[1])) ||
                            iz != max(qlo[2], ecell(B.a.z, G.f, G.edim[2])))
                            continue;
                        ee_ccd_candidate_box(x, p, bEdge, nBE, dbc, A, B, eI, eJ, xi, largest, zero);
                    }
                }
        for (uint32_t k = 0; k < nl; ++k) {
            int eJ = lgList[k];
            const VBox B = ld_box(ebox, eJ);
            if (eJ > eI) ee_ccd_candidate_box(x, p, bEdge, nBE, dbc, A, B, eI, eJ, xi, largest, zero);
            else ee_ccd_candidate_box(x, p, bEdge, nBE, dbc, B, A, eJ, eI, xi, largest, zero);
        }
    }
    if (largest < alpha0) atomic_min_pos(gmin, largest);
}

// particle (boundary-node slot s >= codimBN1) against rod segments, IPC.h:2095-2133,
// and particle pairs, IPC.h:2135-2162 (a pair is run with the lower slot first, as the
// reference does); same voxel-overlap candidate rule and cell coverage as k_ccd_ee
__device__ __forceinline__ void pe_ccd_candidate(const double* __restrict__ x, const double* __restrict__ p, const int* __restrict__ bEdge,
                                                 int nBE, const uint8_t* __restrict__ dbc, const int* __restrict__ elo,
                                                 const int* __restrict__ ehi, int vI, const int* nlo, const int* nhi, d3 P, d3 dP, int e,
                                                 double xi, double& largest) {
    if (nlo[0] > ehi[e] || elo[e] > nhi[0] || nlo[1] > ehi[nBE + e] || elo[nBE + e] > nhi[1] || nlo[2] > ehi[2l * nBE + e] ||
        elo[2l * nBE + e] > nhi[2])
        return;
    int a = bEdge[e], b = bEdge[nBE + e];
    if (vI == a || vI == b) return;
    if (dbc[vI] && dbc[a] && dbc[b]) return;
    d3 e0 = ld3(x, a), e1 = ld3(x, b), de0 = ld3(p, a), de1 = ld3(p, b);
    if (!pe_ccd_broadphase(P, e0, e1, dP, de0, de1, xi)) return;
    double la = largest;
    if (pe_accd(P, e0, e1, dP, de0, de1, 0.1, xi, la)) {
        if (largest > la) largest = la;
    }
}
__device__ __forceinline__ void pp_ccd_candidate(const double* __restrict__ x, const double* __restrict__ p, const uint8_t* __restrict__ dbc,
                                                 const int* __restrict__ vlo, const int* __restrict__ vhi, int nV, int vA, int vB,
                                                 double xi, double& largest) {
    if (vlo[vA] > vhi[vB] || vlo[vB] > vhi[vA] || vlo[nV + vA] > vhi[nV + vB] || vlo[nV + vB] > vhi[nV + vA] ||
        vlo[2 * nV + vA] > vhi[2 * nV + vB] || vlo[2 * nV + vB] > vhi[2 * nV + vA])
        return;
    if (dbc[vA] && dbc[vB]) return;
    d3 A = ld3(x, vA), B = ld3(x, vB), dA = ld3(p, vA), dB = ld3(p, vB);
    if (!pp_ccd_broadphase(A, B, dA, dB, xi)) return;
    double la = largest;
    if (pp_accd(A, B, dA, dB, 0.1, xi, la)) {
        if (largest > la) largest = la;
    }
}
__global__ void k_ccd_codim(const double* __restrict__ x, const double* __restrict__ p, const int* __restrict__ bNode, int nBN,
                            int codimBN1, const int* __restrict__ bEdge, int nBE, int nRod, const uint8_t* __restrict__ dbc,
                            const int* __restrict__ vlo, const int* __restrict__ vhi, int nV, const int* __restrict__ elo,
                            const int* __restrict__ ehi, VoxGrid G, const uint32_t* __restrict__ estart, const uint32_t* __restrict__ eend,
                            const uint32_t* __restrict__ evals, const int* __restrict__ lgListE, const uint32_t* __restrict__ lgCountE,
                            const int* __restrict__ plo, const uint32_t* __restrict__ pstart, const uint32_t* __restrict__ pend,
                            const uint32_t* __restrict__ pvals, const uint8_t* __restrict__ lgFlagP, const int* __restrict__ lgListP,
                            const uint32_t* __restrict__ lgCountP, double alpha0, double xi, unsigned long long* __restrict__ gmin) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int s = codimBN1 + j;
    if (s >= nBN) return;
    const int nPart = nBN - codimBN1, firstRod = nBE - nRod;
    int vI = bNode[s];
    int nlo[3], nhi[3], qlo[3], qhi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        nlo[d] = vlo[d * nV + vI]; nhi[d] = vhi[d * nV + vI];
        qlo[d] = ecell(nlo[d], G.f, G.edim[d]); qhi[d] = ecell(nhi[d], G.f, G.edim[d]);
    }
    d3 P = ld3(x, vI), dP = ld3(p, vI);
    double largest = alpha0;
    const bool large = lgFlagP[j] != 0;
    if (!large) {
        for (int iz = qlo[2]; iz <= qhi[2]; ++iz)
            for (int iy = qlo[1]; iy <= qhi[1]; ++iy)
                for (int ix = qlo[0]; ix <= qhi[0]; ++ix) {
                    uint64_t cell = (uint64_t)ix + (uint64_t)G.edim[0] * ((uint64_t)iy + (uint64_t)G.edim[1] * (uint64_t)iz);
                    for (uint32_t k = estart[cell], en = eend[cell]; k < en; ++k) {
                        int e = (int)evals[k];
                        if (e < firstRod) continue;
                        if (ix != max(qlo[0], ecell(elo[e], G.f, G.edim[0])) || iy != max(qlo[1], ecell(elo[nBE + e], G.f, G.edim[1])) ||
                            iz != max(qlo[2], ecell(elo[2l * nBE + e], G.f, G.edim[2])))
                            continue;
                        pe_ccd_candidate(x, p, bEdge, nBE, dbc, elo, ehi, vI, nlo, nhi, P, dP, e, xi, largest);
                    }
                    for (uint32_t k = pstart[cell], en = pend[cell]; k < en; ++k) {
                        int jJ = (int)pvals[k];
                        if (jJ <= j) continue;
                        if (ix != max(qlo[0], ecell(plo[jJ], G.f, G.edim[0])) || iy != max(qlo[1], ecell(plo[nPart + jJ], G.f, G.edim[1])) ||
                            iz != max(qlo[2], ecell(plo[2 * nPart + jJ], G.f, G.edim[2])))
                            continue;
                        pp_ccd_candidate(x, p, dbc, vlo, vhi, nV, vI, bNode[codimBN1 + jJ], xi, largest);
                    }
                }
    } else {
        // a large particle meets the grid-listed rod segments only through the large-edge
        // rule below; test every rod segment directly
        for (int e = firstRod; e < nBE; ++e) pe_ccd_candidate(x, p, bEdge, nBE, dbc, elo, ehi, vI, nlo, nhi, P, dP, e, xi, largest);
    }
    if (!large) {
        const uint32_t nle = *lgCountE;
        for (uint32_t k = 0; k < nle; ++k) {
            int e = lgListE[k];
            if (e >= firstRod) pe_ccd_candidate(x, p, bEdge, nBE, dbc, elo, ehi, vI, nlo, nhi, P, dP, e, xi, largest);
        }
    }
    const uint32_t nlp = *lgCountP;
    for (uint32_t k = 0; k < nlp; ++k) {
        int jJ = lgListP[k];
        if (jJ > j) pp_ccd_candidate(x, p, dbc, vlo, vhi, nV, vI, bNode[codimBN1 + jJ], xi, largest);
        else if (!large) pp_ccd_candidate(x, p, dbc, vlo, vhi, nV, bNode[codimBN1 + jJ], vI, xi, largest);
    }
    if (largest < alpha0) atomic_min_pos(gmin, largest);
}

static int bits_for2(uint64_t n) {
    int b = 1;
    while (b < 64 && (1ull << b) < n) ++b;
    return b;
}

// primitive boxes in cell order: record k = box of the k-th (cell, primitive) entry, its
// spare third word of the second half holding the primitive, so a query walks a cell's
// candidates through contiguous memory instead of gathering one box per candidate
__global__ void k_cell_records(const uint32_t* __restrict__ vals, long total, const int4* __restrict__ box, int4* __restrict__ rec) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const int v = (int)vals[k];
    int4 hi = box[2l * v + 1];
    hi.z = v;
    rec[2 * k] = box[2l * v];
    rec[2 * k + 1] = hi;
}

static long build_vox_cells(Ctx& c, const int* verts, int n, int nv, const int* vlo, const int* vhi, const VoxGrid& G, DevBuf<int>& plo,
                            DevBuf<int>& phi, DevBuf<int>& box, DevBuf<uint64_t>& keys, DevBuf<uint32_t>& vals, DevBuf<uint32_t>& start,
                            DevBuf<uint32_t>& end, DevBuf<uint8_t>& lgFlag, DevBuf<int>& lgList, uint32_t* lgCount) {
    long nCells = (long)G.edim[0] * G.edim[1] * G.edim[2];
    start.ensure(nCells);
    end.ensure(nCells);
    CIPC_HIP(hipMemsetAsync(start.p, 0, nCells * sizeof(uint32_t), c.s));
    CIPC_HIP(hipMemsetAsync(end.p, 0, nCells * sizeof(uint32_t), c.s));
    CIPC_HIP(hipMemsetAsync(lgCount, 0, sizeof(uint32_t), c.s));
    lgFlag.ensure((size_t)n + 1);
    lgList.ensure((size_t)n + 1);
    if (n == 0) return 0;
    plo.ensure(3l * n);
    phi.ensure(3l * n);
    box.ensure(8l * n);
    uint32_t* cnt = c.u32a.ensure(n + 1);
    uint32_t* off = c.u32b.ensure(n + 1);
    hipLaunchKernelGGL(k_prim_vox, dim3(nblocks(n)), dim3(kBlock), 0, c.s, verts, n, nv, vlo, vhi, c.nV, G, plo.p, phi.p, (int4*)box.p, cnt, lgFlag.p,
                       lgList.p, lgCount);
    CIPC_LAUNCH_CHECK();
    scan_exclusive(cnt, off, n, c.counters.p + 9, c.pw, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hCnt + 9, c.counters.p + 9, sizeof(uint32_t), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    long total = c.hCnt[9];
    if (total == 0) return 0;
    keys.ensure(total);
    vals.ensure(total);
    hipLaunchKernelGGL(k_prim_vox_emit, dim3(nblocks(n)), dim3(kBlock), 0, c.s, n, plo.p, phi.p, G, lgFlag.p, off, keys.p, vals.p);
    CIPC_LAUNCH_CHECK();
    radix_sort_pairs(keys.p, vals.p, total, bits_for2((uint64_t)nCells), c.pw, c.s);
    hipLaunchKernelGGL(k_cell_bounds2, dim3(nblocks(total)), dim3(kBlock), 0, c.s, keys.p, total, start.p, end.p);
    CIPC_LAUNCH_CHECK();
    return total;
}

double ccd_stepsize(Ctx& c, const double* x, const double* p, double alpha, double* clamped, int* zeroToi) {
    if (zeroToi) *zeroToi = 0;
    const double xi = c.P.xi;
    // step clamp, SPATIAL_HASH.h:454-482 (deterministic chunked sums)
    double voxelSize = 1.0;
    {
        double* len = c.dscr2.ensure(std::max(c.nBE, c.nBN) + 1);
        if (c.nBE) {
            hipLaunchKernelGGL(k_edge_len, dim3(nblocks(c.nBE)), dim3(kBlock), 0, c.s, x, c.bEdge.p, c.nBE, len);
            CIPC_LAUNCH_CHECK();
            chunk_sum(len, c.nBE, c.scal.p + 8, c.pw, c.s);
        }
        double* pa = c.dscr.ensure(c.nBN + 1);
        hipLaunchKernelGGL(k_pabs, dim3(nblocks(c.nBN)), dim3(kBlock), 0, c.s, p, c.bNode.p, c.nBN, pa);
        CIPC_LAUNCH_CHECK();
        chunk_sum(pa, c.nBN, c.scal.p + 9, c.pw, c.s);
        CIPC_HIP(hipMemcpyAsync(c.hScal + 8, c.scal.p + 8, 2 * sizeof(double), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        if (c.nBE) voxelSize *= c.hScal[8] / c.nBE;
        double pSize = c.hScal[9] / (c.nBN * 3);
        double spanSize = alpha * pSize / voxelSize;
        if (spanSize > 1) alpha /= spanSize;
    }
    if (clamped) *clamped = alpha;
    // voxel grid, SPATIAL_HASH.h:484-544
    VoxGrid G;
    {
        int nb = std::min(nblocks(c.nBN), 1024);
        double* part = c.dscr2.ensure(6l * nb);
        hipLaunchKernelGGL(k_swept_partial, dim3(nb), dim3(kBlock), 0, c.s, x, p, alpha, c.bNode.p, c.nBN, part);
        CIPC_LAUNCH_CHECK();
        minmax6_final(part, nb, c.scal.p + 10, c.s);
        CIPC_HIP(hipMemcpyAsync(c.hScal + 10, c.scal.p + 10, 6 * sizeof(double), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        double lb[3], range[3];
        for (int d = 0; d < 3; ++d) {
            lb[d] = c.hScal[10 + d] - xi / 2;
            double rt = c.hScal[13 + d] + xi / 2;
            range[d] = rt - lb[d];
        }
        double inv = 1.0 / voxelSize;
        double vcl = 1;
        for (int d = 0; d < 3; ++d) vcl *= std::max(1.0, std::ceil(range[d] * inv));
        if (vcl > 1e9) { voxelSize *= std::pow(vcl / 1.0e9, 1.0 / 3); inv = 1.0 / voxelSize; }
        for (int d = 0; d < 3; ++d) G.lb[d] = lb[d];
        G.inv = inv;
        G.halfXi = xi / 2;
        G.alpha = alpha;
        long vc[3];
        for (int d = 0; d < 3; ++d) vc[d] = std::max(1L, (long)std::ceil(range[d] * inv)) + 1;
        G.f = 1;
        for (;;) {
            double cells = 1;
            for (int d = 0; d < 3; ++d) cells *= (double)((vc[d] + G.f - 1) / G.f);
            if (cells <= (double)(1 << 23)) break;
            G.f++;
        }
        for (int d = 0; d < 3; ++d) G.edim[d] = (int)((vc[d] + G.f - 1) / G.f);
    }
    int* vlo = c.i32a.ensure(3l * c.nV);
    int* vhi = c.i32b.ensure(3l * c.nV);
    hipLaunchKernelGGL(k_node_vox, dim3(nblocks(c.nBN)), dim3(kBlock), 0, c.s, x, p, c.bNode.p, c.nBN, G, vlo, vhi, c.nV);
    CIPC_LAUNCH_CHECK();
    static_assert(sizeof(unsigned long long) == sizeof(double), "");
    unsigned long long* gmin = (unsigned long long*)(c.scal.p + 24);
    double a0 = alpha;
    CIPC_HIP(hipMemcpyAsync(gmin, &a0, sizeof(double), hipMemcpyHostToDevice, c.s));
    CIPC_HIP(hipMemsetAsync(c.counters.p + 13, 0, sizeof(uint32_t), c.s));
    // point-triangle
    uint32_t* lgCountT = c.counters.p + 18;
    uint32_t* lgCountE = c.counters.p + 19;
    const long totT = build_vox_cells(c, c.btri.p, c.nBT, 3, vlo, vhi, G, c.ccdTlo, c.ccdThi, c.ccdTbox, c.k64a, c.u32c, c.cellStart, c.cellEnd,
                                      c.lgFlagT, c.lgListT, lgCountT);
    int4* recT = (int4*)c.ccdTrec.ensure(8 * (size_t)std::max(totT, 1L));
    if (totT > 0) {
        hipLaunchKernelGGL(k_cell_records, dim3(nblocks(totT)), dim3(kBlock), 0, c.s, c.u32c.p, totT, (const int4*)c.ccdTbox.p, recT);
        CIPC_LAUNCH_CHECK();
    }
    if (c.nBN && c.nBT) {
        hipLaunchKernelGGL(k_ccd_pt, dim3(nblocks(c.nBN)), dim3(kBlock), 0, c.s, x, p, c.bNode.p, c.nBN, c.btri.p, c.nBT, c.dbc.p, vlo, vhi, c.nV,
                           (const int4*)c.ccdTbox.p, G, c.cellStart.p, c.cellEnd.p, (const int4*)recT, c.lgListT.p, lgCountT, alpha, xi, gmin,
                           c.counters.p + 13);
        CIPC_LAUNCH_CHECK();
    }
    // edge-edge
    const long totE = build_vox_cells(c, c.bEdge.p, c.nBE, 2, vlo, vhi, G, c.ccdElo, c.ccdEhi, c.ccdEbox, c.k64b, c.u32d, c.ecellStart,
                                      c.ecellEnd, c.lgFlagE, c.lgListE, lgCountE);
    int4* recE = (int4*)c.ccdErec.ensure(8 * (size_t)std::max(totE, 1L));
    if (totE > 0) {
        hipLaunchKernelGGL(k_cell_records, dim3(nblocks(totE)), dim3(kBlock), 0, c.s, c.u32d.p, totE, (const int4*)c.ccdEbox.p, recE);
        CIPC_LAUNCH_CHECK();
    }
    if (c.nBE) {
        hipLaunchKernelGGL(k_ccd_ee, dim3(nblocks(c.nBE)), dim3(kBlock), 0, c.s, x, p, c.bEdge.p, c.nBE, c.dbc.p, (const int4*)c.ccdEbox.p, G,
                           c.ecellStart.p, c.ecellEnd.p, (const int4*)recE, c.lgFlagE.p, c.lgListE.p, lgCountE, alpha, xi, gmin,
                           c.counters.p + 13);
        CIPC_LAUNCH_CHECK();
    }
    // particles against rod segments and particle pairs
    if (c.nPart) {
        uint32_t* lgCountP = c.counters.p + 21;
        build_vox_cells(c, c.partSlot.p, c.nPart, 1, vlo, vhi, G, c.ccdPlo, c.ccdPhi, c.ccdPbox, c.pKeys, c.pVals, c.pcellStart, c.pcellEnd,
                        c.lgFlagP, c.lgListP, lgCountP);
        hipLaunchKernelGGL(k_ccd_codim, dim3(nblocks(c.nPart)), dim3(kBlock), 0, c.s, x, p, c.bNode.p, c.nBN, c.codimBN1, c.bEdge.p, c.nBE,
                           c.nRod, c.dbc.p, vlo, vhi, c.nV, c.ccdElo.p, c.ccdEhi.p, G, c.ecellStart.p, c.ecellEnd.p, c.u32d.p,
                           c.lgListE.p, lgCountE, c.ccdPlo.p, c.pcellStart.p, c.pcellEnd.p, c.pVals.p, c.lgFlagP.p, c.lgListP.p,
                           lgCountP, alpha, xi, gmin);
        CIPC_LAUNCH_CHECK();
    }
    CIPC_HIP(hipMemcpyAsync(c.hScal + 24, gmin, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipMemcpyAsync(c.hCnt + 13, c.counters.p + 13, sizeof(uint32_t), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    if (zeroToi && c.hCnt[13]) *zeroToi = 1;
    return c.hScal[24];
}

}  // namespace cipc
