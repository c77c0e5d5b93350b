// Contact kernels: spatial-grid broad phase, point-triangle and edge-edge narrow
// phase with compaction, PP/PE deduplication with multiplicity, barrier energy /
// gradient / Hessian (per-stencil PSD projection), minimum distance.
//
// Reference: FEM/IPC.h:19-740 (Compute_Constraint_Set), :742-941, :943-1256,
// :1258-1731 (Compute_Barrier{,_Gradient,_Hessian}), :2246-2388 (Compute_Min_Dist2).
//
// Broad phase: a uniform grid (cell = mean rest boundary-edge length, grown so the
// grid has at most 2^23 cells) holds each triangle / edge in every cell its AABB
// touches; entries are sorted by cell (stable radix sort) and each cell's run is
// found through dense start/end tables.  A query visits the cells of its box and
// accepts a candidate only in the first cell common to both boxes, so every pair is
// tested once; indices are clamped into the grid, which keeps the search complete
// for points outside it.  The active set is defined by the reference's predicates
// (AABB filter with d_hat + offset, region type, d^2 < (d_hat + offset)^2), so it does
// not depend on the grid.
//
// Output order (deterministic): point-triangle by (boundary node, triangle),
// edge-edge family by (edge, edge), then deduplicated PP/PE entries in the
// lexicographic int4 order of the reference's std::map.
#include "ctx.hpp"
#include "dev_math.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>

namespace cipc {

__device__ __forceinline__ int cell_of(double v, double lb, double inv, int dim) {
    double f = floor((v - lb) * inv);
    if (!(f >= 0.0)) return 0;
    if (f >= (double)(dim - 1)) return dim - 1;
    return (int)f;
}

// ------------------------------------------------------------------ grid build
// kind 0: triangles (tri SoA), kind 1: boundary edges (bEdge SoA).  Primitives whose
// box covers more than kLargeCells cells get count 0 and go to the large list.
__global__ void k_prim_range(const double* __restrict__ x, const int* __restrict__ verts, int n, int nv, GridParams G,
                             int* __restrict__ lo, uint32_t* __restrict__ count, uint8_t* __restrict__ lgFlag,
                             int* __restrict__ lgList, uint32_t* __restrict__ lgCount) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int k = 0; k < nv; ++k) {
        int v = verts[(long)k * n + i];
#pragma unroll
        for (int d = 0; d < 3; ++d) { double c = x[3 * v + d]; mn[d] = c < mn[d] ? c : mn[d]; mx[d] = c > mx[d] ? c : mx[d]; }
    }
    uint64_t cnt = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int a = cell_of(mn[d], G.lb[d], G.inv, G.dim[d]), b = cell_of(mx[d], G.lb[d], G.inv, G.dim[d]);
        lo[(long)d * n + i] = a;
        cnt *= (uint64_t)(b - a + 1);
    }
    const bool large = cnt > kLargeCells;
    lgFlag[i] = large ? 1 : 0;
    count[i] = large ? 0u : (uint32_t)cnt;
    if (large) lgList[atomicAdd(lgCount, 1u)] = i;
}
__global__ void k_prim_emit(const double* __restrict__ x, const int* __restrict__ verts, int n, int nv, GridParams G,
                            const uint8_t* __restrict__ lgFlag, const uint32_t* __restrict__ off, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || lgFlag[i]) return;
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int k = 0; k < nv; ++k) {
        int v = verts[(long)k * n + i];
#pragma unroll
        for (int d = 0; d < 3; ++d) { double c = x[3 * v + d]; mn[d] = c < mn[d] ? c : mn[d]; mx[d] = c > mx[d] ? c : mx[d]; }
    }
    int a[3], b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) { a[d] = cell_of(mn[d], G.lb[d], G.inv, G.dim[d]); b[d] = cell_of(mx[d], G.lb[d], G.inv, G.dim[d]); }
    uint32_t o = off[i];
    for (int iz = a[2]; iz <= b[2]; ++iz)
        for (int iy = a[1]; iy <= b[1]; ++iy)
            for (int ix = a[0]; ix <= b[0]; ++ix) {
                keys[o] = (uint64_t)ix + (uint64_t)G.dim[0] * ((uint64_t)iy + (uint64_t)G.dim[1] * (uint64_t)iz);
                vals[o] = (uint32_t)i;
                ++o;
            }
}
__global__ void k_cell_bounds(const uint64_t* __restrict__ keys, long n, uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) start[k] = (uint32_t)i;
    if (i == n - 1 || keys[i + 1] != k) end[k] = (uint32_t)(i + 1);
}

static int bits_for(uint64_t n) {
    int b = 1;
    while (b < 64 && (1ull << b) < n) ++b;
    return b;
}

// builds the sorted cell lists of one primitive kind (large primitives listed apart,
// their count left in device memory at lgCount); returns number of cell entries
static long build_cells(Ctx& c, const double* x, const int* verts, int n, int nv, const GridParams& G, DevBuf<int>& lo,
                        DevBuf<uint64_t>& keys, DevBuf<uint32_t>& vals, DevBuf<uint32_t>& start, DevBuf<uint32_t>& end,
                        DevBuf<uint8_t>& lgFlag, DevBuf<int>& lgList, uint32_t* lgCount) {
    long nCells = (long)G.dim[0] * G.dim[1] * G.dim[2];
    start.ensure(nCells);
    end.ensure(nCells);
    CIPC_HIP(hipMemsetAsync(start.p, 0, nCells * sizeof(uint32_t), c.s));
    CIPC_HIP(hipMemsetAsync(end.p, 0, nCells * sizeof(uint32_t), c.s));
    CIPC_HIP(hipMemsetAsync(lgCount, 0, sizeof(uint32_t), c.s));
    lgFlag.ensure((size_t)n + 1);
    lgList.ensure((size_t)n + 1);
    if (n == 0) return 0;
    lo.ensure((size_t)3 * n);
    uint32_t* cnt = c.u32a.ensure(n + 1);
    uint32_t* off = c.u32b.ensure(n + 1);
    hipLaunchKernelGGL(k_prim_range, dim3(nblocks(n)), dim3(kBlock), 0, c.s, x, verts, n, nv, G, lo.p, cnt, lgFlag.p, lgList.p, lgCount);
    CIPC_LAUNCH_CHECK();
    scan_exclusive(cnt, off, n, c.counters.p + 8, c.pw, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hCnt + 8, c.counters.p + 8, sizeof(uint32_t), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    long total = c.hCnt[8];
    if (total == 0) return 0;
    keys.ensure(total);
    vals.ensure(total);
    hipLaunchKernelGGL(k_prim_emit, dim3(nblocks(n)), dim3(kBlock), 0, c.s, x, verts, n, nv, G, lgFlag.p, off, keys.p, vals.p);
    CIPC_LAUNCH_CHECK();
    radix_sort_pairs(keys.p, vals.p, total, bits_for((uint64_t)nCells), c.pw, c.s);
    hipLaunchKernelGGL(k_cell_bounds, dim3(nblocks(total)), dim3(kBlock), 0, c.s, keys.p, total, start.p, end.p);
    CIPC_LAUNCH_CHECK();
    return total;
}

// ------------------------------------------------------------------ narrow phase
struct CSOut {
    uint64_t* key; int* c4; uint32_t* cnt; uint32_t cap;   // kept entries
    uint64_t* dkey; uint32_t* dcnt; uint32_t dcap;          // PP/PE entries to deduplicate
};

__device__ __forceinline__ uint64_t dup_key(int c0, int c1, int c2, int nV) {
    // lexicographic order of (c0 < 0, c1 >= 0, c2 >= -1) as signed ints
    uint64_t a = (uint64_t)(nV - 1 - (-c0 - 1)), b = (uint64_t)c1, d = (uint64_t)(c2 + 1);
    return (a * (uint64_t)nV + b) * (uint64_t)(nV + 1) + d;
}
// Output slots are reserved once per wavefront: the lanes emitting at this point (whatever
// the divergence) are found by a ballot, the lowest of them adds their count to the
// counter, and each lane takes base + its rank among them (mbcnt).  The slot order is
// arbitrary; the set is sorted by key afterwards, so the result is deterministic.
__device__ __forceinline__ void emit_kept(const CSOut& o, uint64_t key, int c0, int c1, int c2, int c3) {
    const uint32_t i = wave_reserve(o.cnt);
    if (i < o.cap) {
        o.key[i] = key;
        o.c4[4 * (long)i] = c0; o.c4[4 * (long)i + 1] = c1; o.c4[4 * (long)i + 2] = c2; o.c4[4 * (long)i + 3] = c3;
    }
}
__device__ __forceinline__ void emit_dup(const CSOut& o, uint64_t key) {
    const uint32_t i = wave_reserve(o.dcnt);
    if (i < o.dcap) o.dkey[i] = key;
}

#pragma clang fp contract(off)
__device__ __forceinline__ bool pt_cd_broadphase(d3 p, d3 t0, d3 t1, d3 t2, double dist) {
    double mx, mn;
    mx = dmax(dmax(t0.x, t1.x), t2.x); mn = dmin(dmin(t0.x, t1.x), t2.x);
    if (p.x - mx > dist || mn - p.x > dist) return false;
    mx = dmax(dmax(t0.y, t1.y), t2.y); mn = dmin(dmin(t0.y, t1.y), t2.y);
    if (p.y - mx > dist || mn - p.y > dist) return false;
    mx = dmax(dmax(t0.z, t1.z), t2.z); mn = dmin(dmin(t0.z, t1.z), t2.z);
    if (p.z - mx > dist || mn - p.z > dist) return false;
    return true;
}
__device__ __forceinline__ bool ee_cd_broadphase(d3 a0, d3 a1, d3 b0, d3 b1, double dist) {
    if (dmin(a0.x, a1.x) - dmax(b0.x, b1.x) > dist || dmin(b0.x, b1.x) - dmax(a0.x, a1.x) > dist) return false;
    if (dmin(a0.y, a1.y) - dmax(b0.y, b1.y) > dist || dmin(b0.y, b1.y) - dmax(a0.y, a1.y) > dist) return false;
    if (dmin(a0.z, a1.z) - dmax(b0.z, b1.z) > dist || dmin(b0.z, b1.z) - dmax(a0.z, a1.z) > dist) return false;
    return true;
}

// point vI (boundary-node slot svI) against triangle t: IPC.h:57-140
__device__ __forceinline__ void pt_candidate(const CSOut& o, const double* __restrict__ x, const int* __restrict__ tri, int nT,
                                             const uint8_t* __restrict__ dbc, int svI, int vI, d3 p, int t, double dHat,
                                             double dHat2, int nV) {
    int a = tri[t], b = tri[nT + t], cc = tri[2 * nT + t];
    if (vI == a || vI == b || vI == cc) return;
    if (dbc[vI] && dbc[a] && dbc[b] && dbc[cc]) return;
    d3 t0 = ld3(x, a), t1 = ld3(x, b), t2 = ld3(x, cc);
    if (!pt_cd_broadphase(p, t0, t1, t2, dHat)) return;
    int ty = pt_type(p, t0, t1, t2);
    double d;
    int c1 = -1, c2 = -1;
    switch (ty) {
        case 0: d = dist2_pp(p, t0); c1 = a; break;
        case 1: d = dist2_pp(p, t1); c1 = b; break;
        case 2: d = dist2_pp(p, t2); c1 = cc; break;
        case 3: d = dist2_pe(p, t0, t1); c1 = a; c2 = b; break;
        case 4: d = dist2_pe(p, t1, t2); c1 = b; c2 = cc; break;
        case 5: d = dist2_pe(p, t2, t0); c1 = cc; c2 = a; break;
        default: d = dist2_pt(p, t0, t1, t2); break;
    }
    if (!(d < dHat2)) return;
    if (ty == 6) emit_kept(o, (uint64_t)svI * (uint64_t)nT + (uint64_t)t, -vI - 1, a, b, cc);
    else emit_dup(o, dup_key(-vI - 1, c1, c2, nV));
}

// edge pair eA < eB: IPC.h:142-330 (edge A supplies a0/a1 of the encoding)
__device__ __forceinline__ void ee_candidate(const CSOut& o, const double* __restrict__ x, const double* __restrict__ X0,
                                             const int* __restrict__ bEdge, int nBE, const uint8_t* __restrict__ dbc, int eA, int a0,
                                             int a1, d3 ea0, d3 ea1, d3 ra0, d3 ra1, int eB, double dHat, double dHat2, int nV) {
    int b0 = bEdge[eB], b1 = bEdge[nBE + eB];
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
    if (dbc[a0] && dbc[a1] && dbc[b0] && dbc[b1]) return;
    d3 eb0 = ld3(x, b0), eb1 = ld3(x, b1);
    if (!ee_cd_broadphase(ea0, ea1, eb0, eb1, dHat)) return;
    double crossN = ee_cross_norm2(ea0, ea1, eb0, eb1);
    double eps = ee_mollifier_threshold(ra0, ra1, ld3(X0, b0), ld3(X0, b1));
    bool moll = crossN < eps;
    double d;
    int c[4];
    switch (ee_type(ea0, ea1, eb0, eb1)) {
        case 0: d = dist2_pp(ea0, eb0);
            if (moll) { c[0] = a0; c[1] = b0; c[2] = -a1 - 1; c[3] = -b1 - 1; } else { c[0] = -a0 - 1; c[1] = b0; c[2] = -1; c[3] = -1; } break;
        case 1: d = dist2_pp(ea0, eb1);
            if (moll) { c[0] = a0; c[1] = b1; c[2] = -a1 - 1; c[3] = -b0 - 1; } else { c[0] = -a0 - 1; c[1] = b1; c[2] = -1; c[3] = -1; } break;
        case 2: d = dist2_pe(ea0, eb0, eb1);
            if (moll) { c[0] = a0; c[1] = b0; c[2] = b1; c[3] = -a1 - 1; } else { c[0] = -a0 - 1; c[1] = b0; c[2] = b1; c[3] = -1; } break;
        case 3: d = dist2_pp(ea1, eb0);
            if (moll) { c[0] = a1; c[1] = b0; c[2] = -a0 - 1; c[3] = -b1 - 1; } else { c[0] = -a1 - 1; c[1] = b0; c[2] = -1; c[3] = -1; } break;
        case 4: d = dist2_pp(ea1, eb1);
            if (moll) { c[0] = a1; c[1] = b1; c[2] = -a0 - 1; c[3] = -b0 - 1; } else { c[0] = -a1 - 1; c[1] = b1; c[2] = -1; c[3] = -1; } break;
        case 5: d = dist2_pe(ea1, eb0, eb1);
            if (moll) { c[0] = a1; c[1] = b0; c[2] = b1; c[3] = -a0 - 1; } else { c[0] = -a1 - 1; c[1] = b0; c[2] = b1; c[3] = -1; } break;
        case 6: d = dist2_pe(eb0, ea0, ea1);
            if (moll) { c[0] = b0; c[1] = a0; c[2] = a1; c[3] = -b1 - 1; } else { c[0] = -b0 - 1; c[1] = a0; c[2] = a1; c[3] = -1; } break;
        case 7: d = dist2_pe(eb1, ea0, ea1);
            if (moll) { c[0] = b1; c[1] = a0; c[2] = a1; c[3] = -b0 - 1; } else { c[0] = -b1 - 1; c[1] = a0; c[2] = a1; c[3] = -1; } break;
        default: d = dist2_ee(ea0, ea1, eb0, eb1);
            if (moll) { c[0] = a0; c[1] = a1; c[2] = -b0 - 1; c[3] = b1; } else { c[0] = a0; c[1] = a1; c[2] = b0; c[3] = b1; } break;
    }
    if (!(d < dHat2)) return;
    if (c[0] >= 0) emit_kept(o, (uint64_t)eA * (uint64_t)nBE + (uint64_t)eB, c[0], c[1], c[2], c[3]);
    else emit_dup(o, dup_key(c[0], c[1], c[2], nV));
}

__global__ void k_query_pt(const double* __restrict__ x, const int* __restrict__ bNode, int nBN, const int* __restrict__ tri,
                           int nT, const uint8_t* __restrict__ dbc, GridParams G, const uint32_t* __restrict__ start,
                           const uint32_t* __restrict__ end, const uint32_t* __restrict__ vals, const int* __restrict__ triLo,
                           const int* __restrict__ lgList, const uint32_t* __restrict__ lgCount, double dHat, double dHat2,
                           int nV, CSOut o) {
    int svI = xcd_item();
    if (svI >= nBN) return;
    int vI = bNode[svI];
    d3 p = ld3(x, vI);
    int qlo[3], qhi[3];
    double pc[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) { qlo[d] = cell_of(pc[d] - dHat, G.lb[d], G.inv, G.dim[d]); qhi[d] = cell_of(pc[d] + dHat, G.lb[d], G.inv, G.dim[d]); }
    for (int iz = qlo[2]; iz <= qhi[2]; ++iz)
        for (int iy = qlo[1]; iy <= qhi[1]; ++iy)
            for (int ix = qlo[0]; ix <= qhi[0]; ++ix) {
                uint64_t cell = (uint64_t)ix + (uint64_t)G.dim[0] * ((uint64_t)iy + (uint64_t)G.dim[1] * (uint64_t)iz);
                uint32_t s = start[cell], e = end[cell];
                for (uint32_t k = s; k < e; ++k) {
                    int t = (int)vals[k];
                    int l0 = triLo[t], l1 = triLo[nT + t], l2 = triLo[2 * nT + t];
                    if (ix != max(qlo[0], l0) || iy != max(qlo[1], l1) || iz != max(qlo[2], l2)) continue;
                    pt_candidate(o, x, tri, nT, dbc, svI, vI, p, t, dHat, dHat2, nV);
                }
            }
    const uint32_t nl = *lgCount;
    for (uint32_t k = 0; k < nl; ++k) pt_candidate(o, x, tri, nT, dbc, svI, vI, p, lgList[k], dHat, dHat2, nV);
}

// Every unordered edge pair is tested once: a grid-listed edge finds grid-listed
// edges of larger index in the first common cell and every large edge from the
// list; a large edge only tests large edges of larger index.
__global__ void k_query_ee(const double* __restrict__ x, const double* __restrict__ X0, const int* __restrict__ bEdge, int nBE,
                           const uint8_t* __restrict__ dbc, GridParams G, const uint32_t* __restrict__ start,
                           const uint32_t* __restrict__ end, const uint32_t* __restrict__ vals, const int* __restrict__ edgeLo,
                           const uint8_t* __restrict__ lgFlag, const int* __restrict__ lgList, const uint32_t* __restrict__ lgCount,
                           double dHat, double dHat2, int nV, CSOut o) {
    int eI = blockIdx.x * blockDim.x + threadIdx.x;
    if (eI >= nBE) return;
    int a0 = bEdge[eI], a1 = bEdge[nBE + eI];
    d3 ea0 = ld3(x, a0), ea1 = ld3(x, a1);
    d3 ra0 = ld3(X0, a0), ra1 = ld3(X0, a1);
    const uint32_t nl = *lgCount;
    if (lgFlag[eI]) {
        for (uint32_t k = 0; k < nl; ++k) {
            int eJ = lgList[k];
            if (eJ > eI) ee_candidate(o, x, X0, bEdge, nBE, dbc, eI, a0, a1, ea0, ea1, ra0, ra1, eJ, dHat, dHat2, nV);
        }
        return;
    }
    int qlo[3], qhi[3];
    double mn[3] = {dmin(ea0.x, ea1.x), dmin(ea0.y, ea1.y), dmin(ea0.z, ea1.z)};
    double mx[3] = {dmax(ea0.x, ea1.x), dmax(ea0.y, ea1.y), dmax(ea0.z, ea1.z)};
#pragma unroll
    for (int d = 0; d < 3; ++d) { qlo[d] = cell_of(mn[d] - dHat, G.lb[d], G.inv, G.dim[d]); qhi[d] = cell_of(mx[d] + dHat, G.lb[d], G.inv, G.dim[d]); }
    for (int iz = qlo[2]; iz <= qhi[2]; ++iz)
        for (int iy = qlo[1]; iy <= qhi[1]; ++iy)
            for (int ix = qlo[0]; ix <= qhi[0]; ++ix) {
                uint64_t cell = (uint64_t)ix + (uint64_t)G.dim[0] * ((uint64_t)iy + (uint64_t)G.dim[1] * (uint64_t)iz);
                uint32_t s = start[cell], e = end[cell];
                for (uint32_t k = s; k < e; ++k) {
                    int eJ = (int)vals[k];
                    if (eJ <= eI) continue;
                    int l0 = edgeLo[eJ], l1 = edgeLo[nBE + eJ], l2 = edgeLo[2 * nBE + eJ];
                    if (ix != max(qlo[0], l0) || iy != max(qlo[1], l1) || iz != max(qlo[2], l2)) continue;
                    ee_candidate(o, x, X0, bEdge, nBE, dbc, eI, a0, a1, ea0, ea1, ra0, ra1, eJ, dHat, dHat2, nV);
                }
            }
    for (uint32_t k = 0; k < nl; ++k) {
        int eJ = lgList[k];
        if (eJ > eI) {
            ee_candidate(o, x, X0, bEdge, nBE, dbc, eI, a0, a1, ea0, ea1, ra0, ra1, eJ, dHat, dHat2, nV);
        } else {
            int b0 = bEdge[eJ], b1 = bEdge[nBE + eJ];
            ee_candidate(o, x, X0, bEdge, nBE, dbc, eJ, b0, b1, ld3(x, b0), ld3(x, b1), ld3(X0, b0), ld3(X0, b1), eI, dHat, dHat2, nV);
        }
    }
}
// rod nodes and particles (boundary-node slots >= codimBN0) against rod segments (the
// last nRod boundary edges), IPC.h:271-317; particles (slots >= codimBN1) against the
// particles after them, IPC.h:319-349.  All results are PP/PE entries for the
// multiplicity merge.
__device__ __forceinline__ bool pe_cd_broadphase(d3 p, d3 e0, d3 e1, double dist) {
    if (p.x - dmax(e0.x, e1.x) > dist || dmin(e0.x, e1.x) - p.x > dist) return false;
    if (p.y - dmax(e0.y, e1.y) > dist || dmin(e0.y, e1.y) - p.y > dist) return false;
    if (p.z - dmax(e0.z, e1.z) > dist || dmin(e0.z, e1.z) - p.z > dist) return false;
    return true;
}
__device__ __forceinline__ void pe_codim_candidate(const CSOut& o, const double* __restrict__ x, const int* __restrict__ bEdge, int nBE,
                                                   const uint8_t* __restrict__ dbc, int vI, d3 p, int eJ, double dHat, double dHat2,
                                                   int nV) {
    int a = bEdge[eJ], b = bEdge[nBE + eJ];
    if (vI == a || vI == b) return;
    if (dbc[vI] && dbc[a] && dbc[b]) return;
    d3 e0 = ld3(x, a), e1 = ld3(x, b);
    if (!pe_cd_broadphase(p, e0, e1, dHat)) return;
    double d;
    int c1, c2 = -1;
    switch (pe_type(p, e0, e1)) {
        case 0: d = dist2_pp(p, e0); c1 = a; break;
        case 1: d = dist2_pp(p, e1); c1 = b; break;
        default: d = dist2_pe(p, e0, e1); c1 = a; c2 = b; break;
    }
    if (d < dHat2) emit_dup(o, dup_key(-vI - 1, c1, c2, nV));
}
__global__ void k_query_codim(const double* __restrict__ x, const int* __restrict__ bNode, int nBN, int codimBN0, int codimBN1,
                              const int* __restrict__ bEdge, int nBE, int nRod, const uint8_t* __restrict__ dbc, GridParams G,
                              const uint32_t* __restrict__ estart, const uint32_t* __restrict__ eend, const uint32_t* __restrict__ evals,
                              const int* __restrict__ edgeLo, const int* __restrict__ lgListE, const uint32_t* __restrict__ lgCountE,
                              const uint32_t* __restrict__ pstart, const uint32_t* __restrict__ pend, const uint32_t* __restrict__ pvals,
                              double dHat, double dHat2, int nV, CSOut o) {
    int s = codimBN0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nBN) return;
    int vI = bNode[s];
    d3 p = ld3(x, vI);
    int qlo[3], qhi[3];
    double pc[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) { qlo[d] = cell_of(pc[d] - dHat, G.lb[d], G.inv, G.dim[d]); qhi[d] = cell_of(pc[d] + dHat, G.lb[d], G.inv, G.dim[d]); }
    const int firstRod = nBE - nRod;
    for (int iz = qlo[2]; iz <= qhi[2]; ++iz)
        for (int iy = qlo[1]; iy <= qhi[1]; ++iy)
            for (int ix = qlo[0]; ix <= qhi[0]; ++ix) {
                uint64_t cell = (uint64_t)ix + (uint64_t)G.dim[0] * ((uint64_t)iy + (uint64_t)G.dim[1] * (uint64_t)iz);
                for (uint32_t k = estart[cell], e = eend[cell]; k < e; ++k) {
                    int eJ = (int)evals[k];
                    if (eJ < firstRod) continue;
                    if (ix != max(qlo[0], edgeLo[eJ]) || iy != max(qlo[1], edgeLo[nBE + eJ]) || iz != max(qlo[2], edgeLo[2 * nBE + eJ])) continue;
                    pe_codim_candidate(o, x, bEdge, nBE, dbc, vI, p, eJ, dHat, dHat2, nV);
                }
            }
    const uint32_t nl = *lgCountE;
    for (uint32_t k = 0; k < nl; ++k) {
        int eJ = lgListE[k];
        if (eJ >= firstRod) pe_codim_candidate(o, x, bEdge, nBE, dbc, vI, p, eJ, dHat, dHat2, nV);
    }
    if (s < codimBN1) return;
    // a point lies in exactly one cell, so each particle pair is met once
    for (int iz = qlo[2]; iz <= qhi[2]; ++iz)
        for (int iy = qlo[1]; iy <= qhi[1]; ++iy)
            for (int ix = qlo[0]; ix <= qhi[0]; ++ix) {
                uint64_t cell = (uint64_t)ix + (uint64_t)G.dim[0] * ((uint64_t)iy + (uint64_t)G.dim[1] * (uint64_t)iz);
                for (uint32_t k = pstart[cell], e = pend[cell]; k < e; ++k) {
                    int sJ = codimBN1 + (int)pvals[k];
                    if (sJ <= s) continue;
                    int vJ = bNode[sJ];
                    if (dbc[vI] && dbc[vJ]) continue;
                    double d = dist2_pp(p, ld3(x, vJ));
                    if (d < dHat2) emit_dup(o, dup_key(-vI - 1, vJ, -1, nV));
                }
            }
}
#pragma clang fp contract(on)

// gather sorted kept entries into the SoA constraint array
__global__ void k_gather_c4(const uint32_t* __restrict__ idx, const int* __restrict__ c4, int n, int* __restrict__ C, int nC,
                            int offset) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long s = idx[i];
#pragma unroll
    for (int k = 0; k < 4; ++k) C[(long)k * nC + offset + i] = c4[4 * s + k];
}
__global__ void k_run_flags(const uint64_t* __restrict__ key, int n, uint32_t* __restrict__ head) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || key[i] != key[i - 1]) ? 1u : 0u;
}
__global__ void k_run_emit(const uint64_t* __restrict__ key, int n, const uint32_t* __restrict__ pos, int* __restrict__ C, int nC,
                           int offset, int nV) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!(i == 0 || key[i] != key[i - 1])) return;
    int j = i + 1;
    while (j < n && key[j] == key[i]) ++j;
    uint64_t k = key[i];
    uint64_t d = k % (uint64_t)(nV + 1), r = k / (uint64_t)(nV + 1);
    uint64_t b = r % (uint64_t)nV, a = r / (uint64_t)nV;
    int c0 = -((int)(nV - 1 - (int)a)) - 1;
    long o = offset + pos[i];
    C[o] = c0;
    C[(long)nC + o] = (int)b;
    C[2l * nC + o] = (int)d - 1;
    C[3l * nC + o] = -(j - i);
}

static void constraint_set_impl(Ctx& c, const double* x, DevBuf<int>& out, int& nOut) {
    const double dHat = std::sqrt(c.dHat2) + c.P.xi;     // IPC.h:53-54
    const double dHat2 = dHat * dHat;
    // grid over the current bounding box of all vertices
    {
        double* bb = c.scal.p + 16;
        bbox_minmax(c, x, bb);
        CIPC_HIP(hipMemcpyAsync(c.hScal + 16, bb, 6 * sizeof(double), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
    }
    GridParams G;
    double cs = c.cellSize * c.cellMult;
    double rng[3];
    for (int d = 0; d < 3; ++d) { G.lb[d] = c.hScal[16 + d]; rng[d] = c.hScal[19 + d] - c.hScal[16 + d]; }
    for (;;) {
        double nc = 1;
        for (int d = 0; d < 3; ++d) nc *= std::floor(rng[d] / cs) + 1;
        if (nc <= (double)(1 << 23)) break;
        cs *= 1.25;
    }
    G.inv = 1.0 / cs;
    for (int d = 0; d < 3; ++d) G.dim[d] = (int)std::floor(rng[d] / cs) + 1;

    uint32_t* lgCountT = c.counters.p + 16;
    uint32_t* lgCountE = c.counters.p + 17;
    build_cells(c, x, c.btri.p, c.nBT, 3, G, c.triLo, c.k64a, c.u32c, c.cellStart, c.cellEnd, c.lgFlagT, c.lgListT, lgCountT);
    build_cells(c, x, c.bEdge.p, c.nBE, 2, G, c.edgeLo, c.k64b, c.u32d, c.ecellStart, c.ecellEnd, c.lgFlagE, c.lgListE, lgCountE);
    if (c.nPart)
        build_cells(c, x, c.partSlot.p, c.nPart, 1, G, c.pLo, c.pKeys, c.pVals, c.pcellStart, c.pcellEnd, c.lgFlagP, c.lgListP,
                    c.counters.p + 20);

    // narrow phase with capacity retry
    CSOut o;
    uint32_t cap = c.csCap;
    for (;;) {
        o.cap = cap; o.dcap = cap;
        o.key = c.csKey.ensure(2 * (size_t)cap);
        o.c4 = c.csC4.ensure(8 * (size_t)cap);
        o.dkey = c.csDKey.ensure((size_t)cap);
        o.cnt = c.counters.p;
        o.dcnt = c.counters.p + 1;
        CIPC_HIP(hipMemsetAsync(c.counters.p, 0, 4 * sizeof(uint32_t), c.s));
        // point-triangle kept entries fill [0, cap), edge-edge kept entries [cap, 2 cap)
        hipLaunchKernelGGL(k_query_pt, dim3(xcd_grid(c.nBN)), dim3(kBlock), 0, c.s, x, c.bNode.p, c.nBN, c.btri.p, c.nBT, c.dbc.p, G,
                           c.cellStart.p, c.cellEnd.p, c.u32c.p, c.triLo.p, c.lgListT.p, lgCountT, dHat, dHat2, c.nV, o);
        CIPC_LAUNCH_CHECK();
        CSOut oEE = o;
        oEE.key = o.key + cap; oEE.c4 = o.c4 + 4 * (size_t)cap; oEE.cnt = c.counters.p + 2;
        hipLaunchKernelGGL(k_query_ee, dim3(nblocks(c.nBE)), dim3(kBlock), 0, c.s, x, c.X0.p, c.bEdge.p, c.nBE, c.dbc.p, G,
                           c.ecellStart.p, c.ecellEnd.p, c.u32d.p, c.edgeLo.p, c.lgFlagE.p, c.lgListE.p, lgCountE, dHat, dHat2, c.nV,
                           oEE);
        CIPC_LAUNCH_CHECK();
        if (c.nBN > c.codimBN0) {
            hipLaunchKernelGGL(k_query_codim, dim3(nblocks(c.nBN - c.codimBN0)), dim3(kBlock), 0, c.s, x, c.bNode.p, c.nBN, c.codimBN0,
                               c.codimBN1, c.bEdge.p, c.nBE, c.nRod, c.dbc.p, G, c.ecellStart.p, c.ecellEnd.p, c.u32d.p, c.edgeLo.p,
                               c.lgListE.p, lgCountE, c.nPart ? c.pcellStart.p : c.ecellStart.p, c.nPart ? c.pcellEnd.p : c.ecellEnd.p,
                               c.nPart ? c.pVals.p : c.u32d.p, dHat, dHat2, c.nV, o);
            CIPC_LAUNCH_CHECK();
        }
        CIPC_HIP(hipMemcpyAsync(c.hCnt, c.counters.p, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, c.s));
        CIPC_HIP(hipStreamSynchronize(c.s));
        uint32_t nPT = c.hCnt[0], nDup = c.hCnt[1], nEE = c.hCnt[2];
        uint32_t need = std::max(std::max(nPT, nDup), nEE);
        if (need > cap) {
            while (cap < need) cap *= 2;
            c.csCap = cap;
            continue;
        }
        // deduplicate PP/PE entries: sort, run heads, scan
        uint32_t nUnique = 0;
        uint32_t* dpos = c.csPos.ensure((size_t)std::max<uint32_t>(nDup, 1) + 1);
        uint32_t* si = c.csIdx.ensure((size_t)std::max<uint32_t>(std::max(std::max(nPT, nEE), nDup), 1));
        if (nDup) {
            fill_iota(si, nDup, c.s);
            radix_sort_pairs(o.dkey, si, nDup, bits_for((uint64_t)c.nV * c.nV * (c.nV + 1)), c.pw, c.s);
            uint32_t* h = c.csHead.ensure(nDup);
            hipLaunchKernelGGL(k_run_flags, dim3(nblocks(nDup)), dim3(kBlock), 0, c.s, o.dkey, (int)nDup, h);
            CIPC_LAUNCH_CHECK();
            scan_exclusive(h, dpos, nDup, c.counters.p + 4, c.pw, c.s);
            CIPC_HIP(hipMemcpyAsync(c.hCnt + 4, c.counters.p + 4, sizeof(uint32_t), hipMemcpyDeviceToHost, c.s));
            CIPC_HIP(hipStreamSynchronize(c.s));
            nUnique = c.hCnt[4];
        }
        int nC = (int)(nPT + nEE + nUnique);
        out.ensure((size_t)4 * std::max(nC, 1));
        if (nPT) {
            fill_iota(si, nPT, c.s);
            radix_sort_pairs(o.key, si, nPT, bits_for((uint64_t)c.nBN * c.nBT), c.pw, c.s);
            hipLaunchKernelGGL(k_gather_c4, dim3(nblocks(nPT)), dim3(kBlock), 0, c.s, si, o.c4, (int)nPT, out.p, nC, 0);
            CIPC_LAUNCH_CHECK();
        }
        if (nEE) {
            fill_iota(si, nEE, c.s);
            radix_sort_pairs(oEE.key, si, nEE, bits_for((uint64_t)c.nBE * c.nBE), c.pw, c.s);
            hipLaunchKernelGGL(k_gather_c4, dim3(nblocks(nEE)), dim3(kBlock), 0, c.s, si, oEE.c4, (int)nEE, out.p, nC, (int)nPT);
            CIPC_LAUNCH_CHECK();
        }
        if (nDup) {
            hipLaunchKernelGGL(k_run_emit, dim3(nblocks(nDup)), dim3(kBlock), 0, c.s, o.dkey, (int)nDup, dpos, out.p, nC,
                               (int)(nPT + nEE), c.nV);
            CIPC_LAUNCH_CHECK();
        }
        nOut = nC;
        return;
    }
}

// Autotuner of the grid cell size: while it runs, each call uses the next candidate
// multiplier of the mean rest edge length and its wall time is recorded (stream drained
// before and after; two rounds over the candidates); then the fastest multiplier is
// kept.  The constraint set itself does not depend on the cell size.
void constraint_set(Ctx& c, const double* x, DevBuf<int>& out, int& nOut) {
    Ctx::Tuner& T = c.tune;
    if (!T.cell) return constraint_set_impl(c, x, out, nOut);
    c.cellMult = T.cellCand[T.cellIdx];
    CIPC_HIP(hipStreamSynchronize(c.s));
    const auto t0 = std::chrono::steady_clock::now();
    constraint_set_impl(c, x, out, nOut);
    CIPC_HIP(hipStreamSynchronize(c.s));
    T.cellCost[T.cellIdx] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (++T.cellIdx == (int)T.cellCand.size()) {
        // sum the rounds per multiplier
        const int m = (int)T.cellCand.size() / 2;
        int best = 0;
        for (int i = 1; i < m; ++i)
            if (T.cellCost[i] + T.cellCost[i + m] < T.cellCost[best] + T.cellCost[best + m]) best = i;
        c.cellMult = T.cellCand[best];
        T.cell = false;
    }
}

// ------------------------------------------------------------------ per-constraint evaluation
struct LocalC {
    int nv;
    int vid[4];
    double d2;
};

// decode a constraint (reference encoding) into its 4-vertex stencil; for mollified
// entries the stencil is the two edges (edgeVInd of IPC.h:831, 845, 859)
// entry k of constraint i at C[k * sk + i * si]: the library's SoA layout (sk = nC, si =
// 1) or int4 rows (sk = 1, si = 4)
__device__ __forceinline__ void decode(const int* __restrict__ C, long sk, long si, int i, int* cc, int& kind, int* vid, int& nv,
                                       double& mult) {
#pragma unroll
    for (int k = 0; k < 4; ++k) cc[k] = C[(long)k * sk + (long)i * si];
    mult = 1;
    if (cc[0] >= 0) {
        nv = 4;
        if (cc[2] >= 0 && cc[3] >= 0) { kind = 0; vid[0] = cc[0]; vid[1] = cc[1]; vid[2] = cc[2]; vid[3] = cc[3]; }                  // EE
        else if (cc[3] >= 0) { kind = 1; vid[0] = cc[0]; vid[1] = cc[1]; vid[2] = -cc[2] - 1; vid[3] = cc[3]; }                      // mollified EE
        else if (cc[2] >= 0) { kind = 2; vid[0] = cc[0]; vid[1] = -cc[3] - 1; vid[2] = cc[1]; vid[3] = cc[2]; }                      // mollified PE
        else { kind = 3; vid[0] = cc[0]; vid[1] = -cc[2] - 1; vid[2] = cc[1]; vid[3] = -cc[3] - 1; }                                 // mollified PP
    } else {
        vid[0] = -cc[0] - 1;
        if (cc[3] >= 0) { kind = 4; nv = 4; vid[1] = cc[1]; vid[2] = cc[2]; vid[3] = cc[3]; }                                        // PT
        else if (cc[2] >= 0) { kind = 5; nv = 3; vid[1] = cc[1]; vid[2] = cc[2]; vid[3] = -1; mult = -cc[3]; }                       // PE
        else { kind = 6; nv = 2; vid[1] = cc[1]; vid[2] = -1; vid[3] = -1; mult = -cc[3]; }                                          // PP
    }
}
__device__ __forceinline__ double con_dist2(const double* __restrict__ x, int kind, const int* vid) {
#pragma clang fp contract(off)
    switch (kind) {
        case 0: case 1: return dist2_ee(ld3(x, vid[0]), ld3(x, vid[1]), ld3(x, vid[2]), ld3(x, vid[3]));
        case 2: return dist2_pe(ld3(x, vid[0]), ld3(x, vid[2]), ld3(x, vid[3]));
        case 3: return dist2_pp(ld3(x, vid[0]), ld3(x, vid[2]));
        case 4: return dist2_pt(ld3(x, vid[0]), ld3(x, vid[1]), ld3(x, vid[2]), ld3(x, vid[3]));
        case 5: return dist2_pe(ld3(x, vid[0]), ld3(x, vid[1]), ld3(x, vid[2]));
        default: return dist2_pp(ld3(x, vid[0]), ld3(x, vid[1]));
    }
}

// ROWS: C is int4 rows checked against [0, nV) after decoding (device-resident input the
// host never saw; an entry out of range contributes nothing and sets bit 2 of *bad),
// else the library's own SoA set
template <bool ROWS>
__global__ void k_barrier_energy(const double* __restrict__ x, const double* __restrict__ X0, const int* __restrict__ C, int nC,
                                 double kappa, double xi2, double shat, double* __restrict__ partial, uint32_t* __restrict__ bad,
                                 int nV) {
    __shared__ double lds[4];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    if (i < nC) {
        int cc[4], vid[4], kind, nv;
        double mult;
        decode(C, ROWS ? 1 : nC, ROWS ? 4 : 1, i, cc, kind, vid, nv, mult);
        bool inRange = true;
        if (ROWS) {
#pragma unroll
            for (int k = 0; k < 4; ++k) inRange = inRange && cc[k] >= -nV && cc[k] < nV && (k >= nv || (vid[k] >= 0 && vid[k] < nV));
        }
        double d2 = inRange ? con_dist2(x, kind, vid) : 1.0;
        double s = d2 - xi2;
        if (!inRange) {
            atomicOr(bad, 2u);
        } else if (s <= 0) {
            atomicOr(bad, 1u);
        } else {
            double b = barrier_e(s, shat, kappa);
            if (kind >= 1 && kind <= 3) {
                double eps = ee_mollifier_threshold(ld3(X0, vid[0]), ld3(X0, vid[1]), ld3(X0, vid[2]), ld3(X0, vid[3]));
                double m = ee_mollifier(ld3(x, vid[0]), ld3(x, vid[1]), ld3(x, vid[2]), ld3(x, vid[3]), eps, nullptr, nullptr);
                e = b * m;
            } else {
                e = (cc[3] < -1) ? b * mult : b;
            }
        }
    }
    double t = block_sum_256(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// distance gradient / Hessian of a constraint embedded into its 4-vertex stencil
__device__ __forceinline__ void con_dist_derivs(const double* __restrict__ x, int kind, const int* vid, double* gd, double* Hd) {
#pragma unroll
    for (int i = 0; i < 12; ++i) gd[i] = 0;
    if (Hd) {
#pragma unroll
        for (int i = 0; i < 144; ++i) Hd[i] = 0;
    }
    switch (kind) {
        case 0: case 1:
            dist2_ee_gh(ld3(x, vid[0]), ld3(x, vid[1]), ld3(x, vid[2]), ld3(x, vid[3]), gd, Hd);
            break;
        case 4:
            dist2_pt_gh(ld3(x, vid[0]), ld3(x, vid[1]), ld3(x, vid[2]), ld3(x, vid[3]), gd, Hd);
            break;
        case 2: {   // PE over stencil slots (0, 2, 3)
            double g9[9], H9[81];
            dist2_pe_gh(ld3(x, vid[0]), ld3(x, vid[2]), ld3(x, vid[3]), g9, Hd ? H9 : nullptr);
            const int sl[3] = {0, 2, 3};
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    gd[sl[a] * 3 + k] = g9[a * 3 + k];
                    if (Hd) {
#pragma unroll
                        for (int b = 0; b < 3; ++b)
#pragma unroll
                            for (int l = 0; l < 3; ++l) Hd[(sl[a] * 3 + k) * 12 + sl[b] * 3 + l] = H9[(a * 3 + k) * 9 + b * 3 + l];
                    }
                }
            break;
        }
        case 3: {   // PP over stencil slots (0, 2)
            double g6[6], H6[36];
            dist2_pp_gh(ld3(x, vid[0]), ld3(x, vid[2]), g6, Hd ? H6 : nullptr);
            const int sl[2] = {0, 2};
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    gd[sl[a] * 3 + k] = g6[a * 3 + k];
                    if (Hd) {
#pragma unroll
                        for (int b = 0; b < 2; ++b)
#pragma unroll
                            for (int l = 0; l < 3; ++l) Hd[(sl[a] * 3 + k) * 12 + sl[b] * 3 + l] = H6[(a * 3 + k) * 6 + b * 3 + l];
                    }
                }
            break;
        }
        case 5: {   // PE, native 3-vertex layout
            double H9[81];
            dist2_pe_gh(ld3(x, vid[0]), ld3(x, vid[1]), ld3(x, vid[2]), gd, Hd ? H9 : nullptr);
            if (Hd) {
#pragma unroll
                for (int i = 0; i < 9; ++i)
#pragma unroll
                    for (int j = 0; j < 9; ++j) Hd[(i / 3 * 3 + i % 3) * 12 + j] = H9[i * 9 + j];
            }
            break;
        }
        default: {  // PP
            double H6[36];
            dist2_pp_gh(ld3(x, vid[0]), ld3(x, vid[1]), gd, Hd ? H6 : nullptr);
            if (Hd) {
#pragma unroll
                for (int i = 0; i < 6; ++i)
#pragma unroll
                    for (int j = 0; j < 6; ++j) Hd[i * 12 + j] = H6[i * 6 + j];
            }
            break;
        }
    }
}

// local gradients: SoA [3][4 nC] plus the vertex of each slot (-1 when unused or
// when the barrier is not evaluated)
__global__ void k_barrier_grad_local(const double* __restrict__ x, const double* __restrict__ X0, const int* __restrict__ C, int nC,
                                     double kappa, double xi2, double shat, double* __restrict__ gl, int* __restrict__ gv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC) return;
    int cc[4], vid[4], kind, nv;
    double mult;
    decode(C, nC, 1, i, cc, kind, vid, nv, mult);
    double d2 = con_dist2(x, kind, vid);
    double s = d2 - xi2;
    long ns = 4l * nC;
    if (!(s > 0)) {
#pragma unroll
        for (int a = 0; a < 4; ++a) gv[4l * i + a] = -1;
        return;
    }
    double gd[12];
    con_dist_derivs(x, kind, vid, gd, nullptr);
    double bg = barrier_g(s, shat, kappa);
    double gout[12];
    if (kind >= 1 && kind <= 3) {
        double eps = ee_mollifier_threshold(ld3(X0, vid[0]), ld3(X0, vid[1]), ld3(X0, vid[2]), ld3(X0, vid[3]));
        double ge[12];
        double m = ee_mollifier(ld3(x, vid[0]), ld3(x, vid[1]), ld3(x, vid[2]), ld3(x, vid[3]), eps, ge, nullptr);
        double b = barrier_e(s, shat, kappa);
#pragma unroll
        for (int k = 0; k < 12; ++k) gout[k] = m * bg * gd[k] + b * ge[k];
    } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) gout[k] = mult * bg * gd[k];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        bool used = a < nv;
        gv[4l * i + a] = used ? vid[a] : -1;
#pragma unroll
        for (int k = 0; k < 3; ++k) gl[k * ns + 4l * i + a] = used ? gout[a * 3 + k] : 0.0;
    }
}

// Stage 1 of the barrier Hessian: each constraint's local Hessian (12 x 12 slot layout,
// leading 3 nv x 3 nv block used) into Hc[144 i], and nvv[i] = nv (0: not evaluated)
__global__ void __launch_bounds__(64) k_barrier_hess_build(const double* __restrict__ x, const double* __restrict__ X0,
                                                           const int* __restrict__ C, int nC, double kappa, double xi2, double shat,
                                                           double* __restrict__ Hc, int* __restrict__ nvv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC) return;
    int cc[4], vid[4], kind, nv;
    double mult;
    decode(C, nC, 1, i, cc, kind, vid, nv, mult);
    double d2 = con_dist2(x, kind, vid);
    double s = d2 - xi2;
    double* H = Hc + 144l * i;
    if (!(s > 0)) {
        nvv[i] = 0;
        return;
    }
    double gd[12], Hd[144];
    con_dist_derivs(x, kind, vid, gd, Hd);
    double bg = barrier_g(s, shat, kappa), bh = barrier_h(s, shat, kappa);
    if (kind >= 1 && kind <= 3) {
        double eps = ee_mollifier_threshold(ld3(X0, vid[0]), ld3(X0, vid[1]), ld3(X0, vid[2]), ld3(X0, vid[3]));
        double ge[12], He[144];
        double m = ee_mollifier(ld3(x, vid[0]), ld3(x, vid[1]), ld3(x, vid[2]), ld3(x, vid[3]), eps, ge, He);
        double b = barrier_e(s, shat, kappa);
#pragma unroll
        for (int r = 0; r < 12; ++r)
#pragma unroll
            for (int q = 0; q < 12; ++q)
                H[r * 12 + q] = bg * gd[r] * ge[q] + bg * ge[r] * gd[q] + b * He[r * 12 + q] + m * bh * gd[r] * gd[q] + m * bg * Hd[r * 12 + q];
        nvv[i] = 4;
    } else {
        const int n3 = 3 * nv;
#pragma unroll
        for (int r = 0; r < 12; ++r)
#pragma unroll
            for (int q = 0; q < 12; ++q)
                if (r < n3 && q < n3) H[r * 12 + q] = mult * (bh * gd[r] * gd[q] + bg * Hd[r * 12 + q]);
        nvv[i] = nv;
    }
}

// Stage 2: PSD projection of each local Hessian (translation-reduced Jacobi in registers,
// one constraint per thread) written as SoA [9][ns] blocks (ns >= 16 nC); block (a, b) of
// constraint i at slot 16 i + 4 a + b, zero outside the constraint's nv x nv blocks.
__global__ void __launch_bounds__(64) k_barrier_hess_proj(const double* __restrict__ Hc, const int* __restrict__ nvv, int nC,
                                                          double* __restrict__ hl, long ns) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nC) return;
    const long base = 16l * i;
    const int nv = nvv[i];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (a >= nv || b >= nv)
#pragma unroll
                for (int e = 0; e < 9; ++e) hl[(long)e * ns + base + 4 * a + b] = 0.0;
    auto store = [&](int r, int c, double v) { hl[(long)((r % 3) * 3 + c % 3) * ns + base + 4 * (r / 3) + c / 3] = v; };
    const double* H = Hc + 144l * i;
    if (nv == 4) make_pd_translation<4>(H, 12, store);
    else if (nv == 3) make_pd_translation<3>(H, 12, store);
    else if (nv == 2) make_pd_translation<2>(H, 12, store);
}

__global__ void k_con_dist2(const double* __restrict__ x, const int* __restrict__ C, int nC, double* __restrict__ d2,
                            double* __restrict__ partial) {
    __shared__ double lds[4];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 1e300;
    if (i < nC) {
        int cc[4], vid[4], kind, nv;
        double mult;
        decode(C, nC, 1, i, cc, kind, vid, nv, mult);
        v = con_dist2(x, kind, vid);
        if (d2) d2[i] = v;
    }
    double m = block_min_256(v, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}
__global__ void k_min_partials(const double* __restrict__ partial, int n, double* __restrict__ out) {
    if (threadIdx.x == 0) {
        double m = 1e300;
        for (int i = 0; i < n; ++i) m = partial[i] < m ? partial[i] : m;
        *out = m;
    }
}

// ------------------------------------------------------------------ host wrappers
double barrier_energy(Ctx& c, const double* x, const int* C, int nC, double kappa, int* bad) {
    if (nC == 0) { if (bad) *bad = 0; return 0.0; }
    const double xi2 = c.P.xi * c.P.xi, shat = c.dHat2 + 2 * std::sqrt(c.dHat2) * c.P.xi;   // IPC.h:756-757
    int nb = nblocks(nC);
    double* part = c.red.ensure(nb);
    CIPC_HIP(hipMemsetAsync(c.counters.p + 12, 0, sizeof(uint32_t), c.s));
    hipLaunchKernelGGL(k_barrier_energy<false>, dim3(nb), dim3(kBlock), 0, c.s, x, c.X0.p, C, nC, kappa, xi2, shat, part, c.counters.p + 12,
                       c.nV);
    CIPC_LAUNCH_CHECK();
    sum_partials(part, nb, c.scal.p + 1, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hScal + 1, c.scal.p + 1, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipMemcpyAsync(c.hCnt + 12, c.counters.p + 12, sizeof(uint32_t), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    if (bad) *bad = c.hCnt[12] ? 1 : 0;
    return c.hScal[1];
}

__global__ void k_clear_u32(uint32_t* __restrict__ a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0u;
}
// The same energy over int4 constraint rows (C[4 i + k], the reference's layout) with every
// buffer supplied by the caller: per-block partials part[nblocks(nC)], the sum *out and
// the validity bits *bad (bit 1: a distance at or below the offset, bit 2: an entry
// outside the vertex range), all device memory.  *bad is cleared by a fill kernel.
// Nothing is allocated and nothing is synchronised; the launches go to c.s.
void barrier_energy_rows(Ctx& c, const double* x, const int* Crows, int nC, double kappa, double* part, double* out, uint32_t* bad) {
    const double xi2 = c.P.xi * c.P.xi, shat = c.dHat2 + 2 * std::sqrt(c.dHat2) * c.P.xi;   // IPC.h:756-757
    hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, c.s, bad, 1);
    CIPC_LAUNCH_CHECK();
    if (nC == 0) {
        hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, c.s, reinterpret_cast<uint32_t*>(out), 2);   // +0.0
        CIPC_LAUNCH_CHECK();
        return;
    }
    const int nb = nblocks(nC);
    hipLaunchKernelGGL(k_barrier_energy<true>, dim3(nb), dim3(kBlock), 0, c.s, x, c.X0.p, Crows, nC, kappa, xi2, shat, part, bad, c.nV);
    CIPC_LAUNCH_CHECK();
    sum_partials(part, nb, out, c.s);
}

__global__ void k_vert_bounds(const uint64_t* __restrict__ key, long n, int nV, uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = key[i];
    if (k >= (uint64_t)nV) return;
    if (i == 0 || key[i - 1] != k) start[k] = (uint32_t)i;
    if (i == n - 1 || key[i + 1] != k) end[k] = (uint32_t)(i + 1);
}
// slots of vertices flagged in `skip` (rows the caller projects out) are not gathered: a
// static obstacle's vertex (a ground corner) can sit in 10^5 stencils, which one thread
// would otherwise sum serially
__global__ void k_grad_keys(const int* __restrict__ gv, long n, int nV, const uint8_t* __restrict__ skip, uint64_t* __restrict__ key) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = gv[i];
    key[i] = (v < 0 || (skip && skip[v])) ? (uint64_t)nV : (uint64_t)v;
}
__global__ void k_grad_gather(const double* __restrict__ gl, long ns, const uint32_t* __restrict__ slot, const uint32_t* __restrict__ start,
                              const uint32_t* __restrict__ end, int nV, double* __restrict__ g) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    uint32_t s = start[v], e = end[v];
    double a = 0, b = 0, d = 0;
    for (uint32_t k = s; k < e; ++k) {
        long q = slot[k];
        a += gl[q]; b += gl[ns + q]; d += gl[2 * ns + q];
    }
    g[3 * v] += a; g[3 * v + 1] += b; g[3 * v + 2] += d;
}

// g += barrier gradient (deterministic: slots sorted by vertex, summed in slot order)
void barrier_gradient(Ctx& c, const double* x, const int* C, int nC, double kappa, double* g, const uint8_t* skip) {
    if (nC == 0) return;
    const double xi2 = c.P.xi * c.P.xi, shat = c.dHat2 + 2 * std::sqrt(c.dHat2) * c.P.xi;
    long ns = 4l * nC;
    double* gl = c.conGLocal.ensure(3 * ns);
    int* gv = c.conGVert.ensure(ns);
    hipLaunchKernelGGL(k_barrier_grad_local, dim3(nblocks(nC)), dim3(kBlock), 0, c.s, x, c.X0.p, C, nC, kappa, xi2, shat, gl, gv);
    CIPC_LAUNCH_CHECK();
    scatter_local_gradient(c, gl, gv, ns, g, skip);
}

// g += local gradient slots (SoA [3][ns], vertex gv per slot, -1 unused), summed per
// vertex in slot order (radix sort by vertex, stable); vertices flagged in skip keep g
void scatter_local_gradient(Ctx& c, const double* gl, const int* gv, long ns, double* g, const uint8_t* skip) {
    uint64_t* key = c.k64a.ensure(ns);
    uint32_t* slot = c.u32c.ensure(ns);
    hipLaunchKernelGGL(k_grad_keys, dim3(nblocks(ns)), dim3(kBlock), 0, c.s, gv, ns, c.nV, skip, key);
    CIPC_LAUNCH_CHECK();
    fill_iota(slot, ns, c.s);
    radix_sort_pairs(key, slot, ns, bits_for((uint64_t)c.nV + 1), c.pw, c.s);
    uint32_t* st = c.u32a.ensure(c.nV);
    uint32_t* en = c.u32b.ensure(c.nV);
    CIPC_HIP(hipMemsetAsync(st, 0, c.nV * sizeof(uint32_t), c.s));
    CIPC_HIP(hipMemsetAsync(en, 0, c.nV * sizeof(uint32_t), c.s));
    hipLaunchKernelGGL(k_vert_bounds, dim3(nblocks(ns)), dim3(kBlock), 0, c.s, key, ns, c.nV, st, en);
    CIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_grad_gather, dim3(nblocks(c.nV)), dim3(kBlock), 0, c.s, gl, ns, slot, st, en, c.nV, g);
    CIPC_LAUNCH_CHECK();
}

// Blocks in a Dirichlet row or column are zero after Project_DBC (Math/CSR_MATRIX.h:
// 130-141) and are not stored: a Dirichlet vertex of a static obstacle (a ground corner)
// touches every stencil near it, and its row would otherwise hold a block per contact.
__global__ void k_hess_keys(const int* __restrict__ C, int nC, int nV, const uint8_t* __restrict__ dbc, uint64_t* __restrict__ key) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 16l * nC) return;
    int i = (int)(q / 16), a = (int)((q / 4) % 4), b = (int)(q % 4);
    int cc[4], vid[4], kind, nv;
    double mult;
    decode(C, nC, 1, i, cc, kind, vid, nv, mult);
    const bool keep = a < nv && b < nv && !dbc[vid[a]] && !dbc[vid[b]];
    key[q] = keep ? (uint64_t)vid[a] * (uint64_t)nV + (uint64_t)vid[b] : (uint64_t)nV * (uint64_t)nV;
}
__global__ void k_bsr_from_sorted(const uint64_t* __restrict__ key, long n, int nV, uint32_t* __restrict__ head) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (key[i] < (uint64_t)nV * nV && (i == 0 || key[i] != key[i - 1])) ? 1u : 0u;
}
// sum the local blocks of each unique (row, col) into the contact BSR; Dirichlet
// rows/columns are zeroed (Project_DBC)
__global__ void k_bsr_sum(const uint64_t* __restrict__ key, long n, int nV, const uint32_t* __restrict__ pos,
                          const uint32_t* __restrict__ slot, const double* __restrict__ hl, long ns, const uint8_t* __restrict__ dbc,
                          int nnzb, int* __restrict__ rowOf, int* __restrict__ col, double* __restrict__ val) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = key[i];
    if (k >= (uint64_t)nV * nV) return;
    if (!(i == 0 || key[i - 1] != k)) return;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (long j = i; j < n && key[j] == k; ++j) {
        long q = slot[j];
#pragma unroll
        for (int m = 0; m < 9; ++m) acc[m] += hl[m * ns + q];
    }
    int r = (int)(k / (uint64_t)nV), cl = (int)(k % (uint64_t)nV);
    bool z = dbc[r] || dbc[cl];
    int o = (int)pos[i];
    rowOf[o] = r;
    col[o] = cl;
#pragma unroll
    for (int m = 0; m < 9; ++m) val[(long)m * nnzb + o] = z ? 0.0 : acc[m];
}
// rowptr[r] = first block of a row >= r (rowOf sorted): one binary search per row, so
// long runs of empty rows (Dirichlet obstacles) cost no serial loop
__global__ void k_rowptr(const int* __restrict__ rowOf, int nnzb, int nrow, int* __restrict__ rowptr) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nrow) return;
    int lo = 0, hi = nnzb;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rowOf[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    rowptr[r] = lo;
}

// contact BSR: barrier blocks of C and, when fricXn is given, the friction blocks of the
// lagged friction stencils (slots 16 nC onwards), summed per (row, col)
void barrier_hessian(Ctx& c, const double* x, const int* C, int nC, double kappa, const double* fricXn) {
    BSR& A = c.Acon;
    A.nrow = c.nV;
    const int nF = fricXn ? c.nF : 0;
    if (nC + nF == 0) {
        A.nnzb = 0;
        A.rowptr.ensure(c.nV + 1);
        CIPC_HIP(hipMemsetAsync(A.rowptr.p, 0, (c.nV + 1) * sizeof(int), c.s));
        return;
    }
    const double xi2 = c.P.xi * c.P.xi, shat = c.dHat2 + 2 * std::sqrt(c.dHat2) * c.P.xi;
    long ns = 16l * (nC + nF);
    double* hl = c.conLocal.ensure(9 * ns);
    uint64_t* key = c.k64b.ensure(ns);
    uint32_t* slot = c.u32d.ensure(ns);
    if (nC) {
        double* Hc = c.conH.ensure(144 * (size_t)nC);
        int* nvv = c.conNv.ensure((size_t)nC);
        hipLaunchKernelGGL(k_barrier_hess_build, dim3(nblocks(nC, 64)), dim3(64), 0, c.s, x, c.X0.p, C, nC, kappa, xi2, shat, Hc, nvv);
        hipLaunchKernelGGL(k_barrier_hess_proj, dim3(nblocks(nC, 64)), dim3(64), 0, c.s, (const double*)Hc, (const int*)nvv, nC, hl, ns);
        CIPC_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_hess_keys, dim3(nblocks(16l * nC)), dim3(kBlock), 0, c.s, C, nC, c.nV, proj_mask(c), key);
        CIPC_LAUNCH_CHECK();
    }
    if (nF) {
        friction_hess_local(c, x, fricXn, hl, ns, 16l * nC);
        friction_hess_keys(c, key + 16l * nC);
    }
    fill_iota(slot, ns, c.s);
    radix_sort_pairs(key, slot, ns, bits_for((uint64_t)c.nV * c.nV + 1), c.pw, c.s);
    uint32_t* head = c.u32a.ensure(ns);
    uint32_t* pos = c.u32b.ensure(ns);
    hipLaunchKernelGGL(k_bsr_from_sorted, dim3(nblocks(ns)), dim3(kBlock), 0, c.s, key, ns, c.nV, head);
    CIPC_LAUNCH_CHECK();
    scan_exclusive(head, pos, ns, c.counters.p + 5, c.pw, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hCnt + 5, c.counters.p + 5, sizeof(uint32_t), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    int nnzb = (int)c.hCnt[5];
    A.nnzb = nnzb;
    A.rowOf.ensure(std::max(nnzb, 1));
    A.col.ensure(std::max(nnzb, 1));
    A.val.ensure(9l * std::max(nnzb, 1));
    A.rowptr.ensure(c.nV + 1);
    hipLaunchKernelGGL(k_bsr_sum, dim3(nblocks(ns)), dim3(kBlock), 0, c.s, key, ns, c.nV, pos, slot, hl, ns, proj_mask(c), nnzb,
                       A.rowOf.p, A.col.p, A.val.p);
    CIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rowptr, dim3(nblocks(c.nV + 1)), dim3(kBlock), 0, c.s, A.rowOf.p, nnzb, c.nV, A.rowptr.p);
    CIPC_LAUNCH_CHECK();
}

double min_dist2(Ctx& c, const double* x, const int* C, int nC, double* d2out) {
    if (nC == 0) return 1e300;
    int nb = nblocks(nC);
    double* part = c.red.ensure(nb);
    hipLaunchKernelGGL(k_con_dist2, dim3(nb), dim3(kBlock), 0, c.s, x, C, nC, d2out, part);
    CIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_min_partials, dim3(1), dim3(64), 0, c.s, part, nb, c.scal.p + 2);
    CIPC_LAUNCH_CHECK();
    CIPC_HIP(hipMemcpyAsync(c.hScal + 2, c.scal.p + 2, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return c.hScal[2] - c.P.xi * c.P.xi;
}

}  // namespace cipc
