// Two-level correction of the learned preconditioner.  The learned factor F (M_s^-1 =
// F^T F) is local: it reaches one ring of blocks, and the PCG spends most of its
// iterations on the smooth, low-energy deformations of the sheets (bending and rigid
// motion of whole patches).  A coarse space of rigid-body modes on aggregates of the
// elastic vertex graph captures those: per aggregate three translations and three
// rotations about its centroid, T_i = [I | R(d_i)] with d_i = (x_i - c_a) / rho_a and
// R(d) e_k = e_k x d.  The preconditioner becomes
//     M^-1 = F^T F + P A_c^-1 P^T,   A_c = P^T H P   (Galerkin, H = elastic + contact),
// which stays symmetric positive definite.  The coarse problem (6 unknowns per aggregate,
// at most kCoarseMaxDof) is dense: its Cholesky factor L and the explicit inverse L^-1
// are computed at a step's first solve and then every `every` solves (tiled kernels of
// direct.hip), and every PCG iteration applies A_c^-1 ~ W^T W, W = L^-1 rounded to FP32,
// as two triangular products (FP64 accumulation).  W^T W is positive semidefinite
// whatever the rounding, so M^-1 stays SPD; the rounding error of W acts like
// sqrt(cond(A_c)) x the unit roundoff in the energy norm, and contact makes A_c
// ill-conditioned (with BF16 storage the PCG stalled in deep contact), hence FP32.
// No counterpart in the reference, which factorises H directly (DIRECT_SOLVER.h:14-66).
//
// Aggregates are built once per elastic pattern on the host: greedy root-and-neighbours
// aggregation of the vertex graph (a vertex all of whose neighbours are free becomes a
// root and takes them; the rest join the neighbouring aggregate they share most edges
// with), then the same once on the aggregate graph (about 10^2 vertices per aggregate),
// then pairwise matching of neighbouring aggregates while the coarse problem is above the
// size asked for (the set-up grows with its cube, each iteration with its square); if
// separate components keep it above (thousands of rods), only the largest aggregates stay.  Dirichlet vertices and vertices without elastic neighbours (free
// particles) stay outside the coarse space; the smoother alone treats them.
#include "ctx.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <set>

namespace cipc {

constexpr int kCgBlock = 256;
constexpr int kCgTile = 64;

// greedy aggregation of graph (ptr, adj) over active vertices; isolated active vertices
// become aggregates of their own when keepIsolated, else stay unassigned (-1)
static int greedy_aggregate(int n, const std::vector<int>& ptr, const std::vector<int>& adj, const std::vector<char>& act, bool keepIsolated,
                     std::vector<int>& agg) {
    agg.assign(n, -1);
    int na = 0;
    for (int v = 0; v < n; ++v) {
        if (!act[v] || agg[v] >= 0) continue;
        bool ok = true;
        int nAct = 0;
        for (int e = ptr[v]; e < ptr[v + 1]; ++e) {
            const int u = adj[e];
            if (!act[u]) continue;
            ++nAct;
            if (agg[u] >= 0) { ok = false; break; }
        }
        if (!ok) continue;
        if (nAct == 0) {
            if (keepIsolated) agg[v] = na++;
            continue;
        }
        agg[v] = na;
        for (int e = ptr[v]; e < ptr[v + 1]; ++e)
            if (act[adj[e]]) agg[adj[e]] = na;
        ++na;
    }
    std::vector<int> cnt;
    for (int v = 0; v < n; ++v) {
        if (!act[v] || agg[v] >= 0) continue;
        int best = -1, bestN = 0;
        cnt.clear();
        for (int e = ptr[v]; e < ptr[v + 1]; ++e) {
            const int u = adj[e];
            if (act[u] && agg[u] >= 0) cnt.push_back(agg[u]);
        }
        std::sort(cnt.begin(), cnt.end());
        for (size_t i = 0; i < cnt.size();) {
            size_t j = i;
            while (j < cnt.size() && cnt[j] == cnt[i]) ++j;
            if ((int)(j - i) > bestN) { bestN = (int)(j - i); best = cnt[i]; }
            i = j;
        }
        if (best >= 0) agg[v] = best;
        else if (keepIsolated) agg[v] = na++;
    }
    return na;
}

// aggregates of the elastic vertex graph (host), uploaded with their member lists
static void coarse_build(Ctx& c) {
    CoarseSpace& g = c.cg;
    const int nV = c.nV;
    BSR& E = c.Ael;
    std::vector<int> rp(nV + 1), col(E.nnzb);
    CIPC_HIP(hipMemcpyAsync(rp.data(), E.rowptr.p, (nV + 1) * sizeof(int), hipMemcpyDeviceToHost, c.s));
    if (E.nnzb) CIPC_HIP(hipMemcpyAsync(col.data(), E.col.p, E.nnzb * sizeof(int), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    std::vector<int> ptr(nV + 1, 0), adj;
    adj.reserve(E.nnzb);
    for (int v = 0; v < nV; ++v) {
        for (int b = rp[v]; b < rp[v + 1]; ++b)
            if (col[b] != v) adj.push_back(col[b]);
        ptr[v + 1] = (int)adj.size();
    }
    std::vector<char> act(nV, 1);
    for (int v : c.hDbcIdx) act[v] = 0;
    std::vector<int> agg;
    int na = greedy_aggregate(nV, ptr, adj, act, false, agg);
    // coarsen the aggregate graph: once always (aggregates of ~10^2 vertices), then while
    // the coarse problem is too large
    const int maxDof = g.maxDof > 0 ? std::min(g.maxDof, kCoarseMaxDof) : kCoarseMaxDof;
    for (int pass = 0; pass < 12 && na > 0 && (pass == 0 || 6l * na > maxDof); ++pass) {
        std::vector<std::set<int>> nb(na);
        std::vector<std::map<int, int>> nbw(na);   // shared vertex-graph edges
        for (int v = 0; v < nV; ++v) {
            if (agg[v] < 0) continue;
            for (int e = ptr[v]; e < ptr[v + 1]; ++e) {
                const int b = agg[adj[e]];
                if (b >= 0 && b != agg[v]) {
                    nb[agg[v]].insert(b);
                    nbw[agg[v]][b]++;
                }
            }
        }
        std::vector<int> aptr(na + 1, 0), aadj;
        for (int a = 0; a < na; ++a) {
            aadj.insert(aadj.end(), nb[a].begin(), nb[a].end());
            aptr[a + 1] = (int)aadj.size();
        }
        std::vector<int> up;
        int na2;
        if (pass == 0) {
            na2 = greedy_aggregate(na, aptr, aadj, std::vector<char>(na, 1), true, up);
        } else {
            // pairwise matching: each free aggregate takes its free neighbour with the most
            // shared edges (first in order on ties), or stays alone
            up.assign(na, -1);
            na2 = 0;
            for (int a = 0; a < na; ++a) {
                if (up[a] >= 0) continue;
                int best = -1, bestW = -1;
                for (int e = aptr[a]; e < aptr[a + 1]; ++e) {
                    const int b = aadj[e];
                    if (up[b] >= 0) continue;
                    const int w = (int)nbw[a].count(b) ? nbw[a][b] : 0;
                    if (w > bestW) { bestW = w; best = b; }
                }
                up[a] = na2;
                if (best >= 0) up[best] = na2;
                ++na2;
            }
            if (na2 == na) break;   // nothing left to merge
        }
        for (int v = 0; v < nV; ++v)
            if (agg[v] >= 0) agg[v] = up[agg[v]];
        na = na2;
    }
    // still too many (components that cannot merge: thousands of separate rods): keep the
    // largest aggregates, the rest stay with the smoother
    if (na > 0 && 6l * na > maxDof) {
        std::vector<int> size(na, 0);
        for (int v = 0; v < nV; ++v)
            if (agg[v] >= 0) size[agg[v]]++;
        std::vector<int> order(na);
        for (int a = 0; a < na; ++a) order[a] = a;
        std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return size[p] > size[q]; });
        const int keep = maxDof / 6;
        std::vector<int> remap(na, -1);
        for (int r = 0; r < keep; ++r) remap[order[r]] = r;
        // renumber the kept ones in their original order (deterministic, locality kept)
        std::vector<int> kept;
        for (int a = 0; a < na; ++a)
            if (remap[a] >= 0) kept.push_back(a);
        for (int r = 0; r < (int)kept.size(); ++r) remap[kept[r]] = r;
        for (int v = 0; v < nV; ++v)
            if (agg[v] >= 0) agg[v] = remap[agg[v]];
        na = (int)kept.size();
    }
    g.nAgg = na;
    g.nVBuilt = nV;
    g.nnzbBuilt = E.nnzb;
    g.builtMaxDof = g.maxDof;
    g.built = true;
    g.ready = false;
    if (g.nAgg == 0) return;
    // member lists (counting sort by aggregate, vertex order inside)
    std::vector<int> aptr(na + 1, 0), members;
    for (int v = 0; v < nV; ++v)
        if (agg[v] >= 0) aptr[agg[v] + 1]++;
    for (int a = 0; a < na; ++a) aptr[a + 1] += aptr[a];
    members.resize(aptr[na]);
    std::vector<int> fill(aptr.begin(), aptr.end() - 1);
    for (int v = 0; v < nV; ++v)
        if (agg[v] >= 0) members[fill[agg[v]]++] = v;
    int maxM = 0;
    for (int a = 0; a < na; ++a) maxM = std::max(maxM, aptr[a + 1] - aptr[a]);
    if (maxM > kCoarseMaxVerts) {
        g.nAgg = 0;
        return;
    }
    g.nc = 6 * na;
    g.npad = (g.nc + kCgTile - 1) / kCgTile * kCgTile;
    upload(g.agg, agg, c.s);
    upload(g.aggPtr, aptr, c.s);
    upload(g.aggVerts, members, c.s);
    g.nMembers = (int)members.size();
    const size_t np2 = (size_t)g.npad * g.npad;
    g.A.ensure_exact(np2);
    g.W.ensure_exact(np2);
    if (g.bf16) {
        g.W16.ensure_exact(np2);
        g.WT16.ensure_exact(np2);
    } else {
        g.W32.ensure_exact(np2);
        g.WT32.ensure_exact(np2);
    }
    g.tileInv.ensure_exact((size_t)(g.npad / kCgTile) * kCgTile * kCgTile);
    g.diag0.ensure_exact(g.npad);
    g.dropped.ensure_exact(g.npad);
    g.rc.ensure_exact(g.npad);
    g.wc.ensure_exact(g.npad);
    g.zc.ensure_exact(g.npad);
    g.cent.ensure_exact(4l * na);
    g.d.ensure_exact(3l * nV);
    g.dm.ensure_exact(3l * (size_t)members.size());
    g.flag.ensure_exact(1);
    CIPC_HIP(hipMemsetAsync(g.rc.p, 0, g.npad * sizeof(double), c.s));
    CIPC_HIP(hipMemsetAsync(g.d.p, 0, 3l * nV * sizeof(double), c.s));
}

// per aggregate (one wave): centroid and 1 / rms distance of the members from it
__global__ __launch_bounds__(kCgBlock) void k_cg_centroids(int nAgg, const int* __restrict__ aggPtr, const int* __restrict__ aggVerts,
                                                           const double* __restrict__ x, double* __restrict__ cent) {
    const int a = blockIdx.x * (kCgBlock / 64) + threadIdx.x / 64, lane = threadIdx.x & 63;
    if (a >= nAgg) return;
    const int q0 = aggPtr[a], q1 = aggPtr[a + 1];
    double s0 = 0, s1 = 0, s2 = 0;
    for (int q = q0 + lane; q < q1; q += 64) {
        const int v = aggVerts[q];
        s0 += x[3 * v];
        s1 += x[3 * v + 1];
        s2 += x[3 * v + 2];
    }
    const double m = (double)(q1 - q0);
    const double c0 = wave_sum(s0) / m, c1 = wave_sum(s1) / m, c2 = wave_sum(s2) / m;
    double r2 = 0;
    for (int q = q0 + lane; q < q1; q += 64) {
        const int v = aggVerts[q];
        const double e0 = x[3 * v] - c0, e1 = x[3 * v + 1] - c1, e2 = x[3 * v + 2] - c2;
        r2 += e0 * e0 + e1 * e1 + e2 * e2;
    }
    r2 = wave_sum(r2) / m;
    if (lane == 0) {
        cent[4 * a] = c0;
        cent[4 * a + 1] = c1;
        cent[4 * a + 2] = c2;
        cent[4 * a + 3] = r2 > 0 ? 1.0 / std::sqrt(r2) : 1.0;
    }
}

__global__ void k_cg_offsets(int nV, const int* __restrict__ agg, const double* __restrict__ x, const double* __restrict__ cent,
                             double* __restrict__ d) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    const int a = agg[v];
    if (a < 0) return;
    const double s = cent[4 * a + 3];
#pragma unroll
    for (int k = 0; k < 3; ++k) d[3 * v + k] = (x[3 * v + k] - cent[4 * a + k]) * s;
}

__global__ void k_cg_member_offsets(int nM, const int* __restrict__ aggVerts, const double* __restrict__ d, double* __restrict__ dm) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nM) return;
    const int v = aggVerts[q];
#pragma unroll
    for (int k = 0; k < 3; ++k) dm[3 * q + k] = d[3 * v + k];
}

__global__ void k_cg_zero_identity(double* __restrict__ a, long n2, int nc, int npad) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x) {
        const long r = i / npad, q = i % npad;
        a[i] = (r == q && r >= nc) ? 1.0 : 0.0;
    }
}

// item -> (member index) by binary search over the members' block prefix
__device__ __forceinline__ int cg_member_of(const int* pre, int m, int item) {
    int lo = 0, hi = m - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Galerkin row block of aggregate a: A_c[a, b] = sum over blocks H_ij (i in a, j in b) of
// T_i^T H_ij T_j, for every neighbour aggregate b.  One workgroup per aggregate: the
// distinct neighbours are collected in an LDS hash set and sorted; wave w then sums the
// contributions of neighbours w, w + 4, ... over the aggregate's blocks in a fixed lane
// order with a fixed reduction, so the matrix is the same bits on every run.  Row block a
// is written by workgroup a only.  flag bit 1: more than kCoarseMaxNb neighbours.
__global__ __launch_bounds__(kCgBlock) void k_cg_galerkin(const int* __restrict__ aggPtr, const int* __restrict__ aggVerts,
                                                          const int* __restrict__ agg, const double* __restrict__ d,
                                                          const int* __restrict__ Erp, const int* __restrict__ Ecol,
                                                          const double* __restrict__ Eval, int Ennzb, const int* __restrict__ Krp,
                                                          const int* __restrict__ Kcol, const double* __restrict__ Kval, int Knnzb,
                                                          double* __restrict__ A, long ld, int* __restrict__ flag) {
    __shared__ int setKey[2 * kCoarseMaxNb];
    __shared__ int nbSorted[kCoarseMaxNb];
    __shared__ int pre[kCoarseMaxVerts + 1];
    __shared__ int scanTmp[kCgBlock];
    __shared__ int nNb;
    const int a = blockIdx.x;
    const int m0 = aggPtr[a], m = aggPtr[a + 1] - m0;
    for (int t = threadIdx.x; t < 2 * kCoarseMaxNb; t += kCgBlock) setKey[t] = -1;
    if (threadIdx.x == 0) nNb = 0;
    // block prefix over the members' rows (elastic + contact)
    const int per = (m + kCgBlock - 1) / kCgBlock;
    const int lo = min((int)threadIdx.x * per, m), hi = min(lo + per, m);
    int sum = 0;
    for (int q = lo; q < hi; ++q) {
        const int v = aggVerts[m0 + q];
        sum += (Erp[v + 1] - Erp[v]) + (Knnzb ? Krp[v + 1] - Krp[v] : 0);
    }
    scanTmp[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int t = 0; t < kCgBlock; ++t) {
            const int v = scanTmp[t];
            scanTmp[t] = acc;
            acc += v;
        }
        pre[m] = acc;
    }
    __syncthreads();
    int run = scanTmp[threadIdx.x];
    for (int q = lo; q < hi; ++q) {
        pre[q] = run;
        const int v = aggVerts[m0 + q];
        run += (Erp[v + 1] - Erp[v]) + (Knnzb ? Krp[v + 1] - Krp[v] : 0);
    }
    __syncthreads();
    const int nItems = pre[m];
    // distinct neighbour aggregates
    for (int it = threadIdx.x; it < nItems; it += kCgBlock) {
        const int q = cg_member_of(pre, m, it);
        const int v = aggVerts[m0 + q], off = it - pre[q], lenE = Erp[v + 1] - Erp[v];
        const int cj = off < lenE ? Ecol[Erp[v] + off] : Kcol[Krp[v] + off - lenE];
        const int b = agg[cj];
        if (b < 0) continue;
        const unsigned h = ((unsigned)b * 2654435761u) & (2 * kCoarseMaxNb - 1);
        for (int probe = 0; probe < 2 * kCoarseMaxNb; ++probe) {
            const int slot = (int)((h + probe) & (2 * kCoarseMaxNb - 1));
            const int prev = atomicCAS(&setKey[slot], -1, b);
            if (prev == -1) {
                atomicAdd(&nNb, 1);
                break;
            }
            if (prev == b) break;
        }
    }
    __syncthreads();
    const int nb = nNb;
    if (nb > kCoarseMaxNb) {
        if (threadIdx.x == 0) atomicOr(flag, 1);
        return;
    }
    for (int t = threadIdx.x; t < 2 * kCoarseMaxNb; t += kCgBlock) {
        const int key = setKey[t];
        if (key < 0) continue;
        int rank = 0;
        for (int u = 0; u < 2 * kCoarseMaxNb; ++u) {
            const int o = setKey[u];
            rank += (o >= 0 && o < key) ? 1 : 0;
        }
        nbSorted[rank] = key;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int s = wv; s < nb; s += kCgBlock / 64) {
        const int b = nbSorted[s];
        double acc[36];
#pragma unroll
        for (int k = 0; k < 36; ++k) acc[k] = 0.0;
        for (int it = lane; it < nItems; it += 64) {
            const int q = cg_member_of(pre, m, it);
            const int v = aggVerts[m0 + q], off = it - pre[q], lenE = Erp[v + 1] - Erp[v];
            const bool el = off < lenE;
            const int blk = el ? Erp[v] + off : Krp[v] + off - lenE;
            const int cj = el ? Ecol[blk] : Kcol[blk];
            if (agg[cj] != b) continue;
            const double* val = el ? Eval : Kval;
            const long nnzb = el ? Ennzb : Knnzb;
            double h[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) h[k] = val[k * nnzb + blk];
            const double di0 = d[3 * v], di1 = d[3 * v + 1], di2 = d[3 * v + 2];
            const double dj0 = d[3 * cj], dj1 = d[3 * cj + 1], dj2 = d[3 * cj + 2];
            // R(d): columns e_k x d
            const double Ri[9] = {0, di2, -di1, -di2, 0, di0, di1, -di0, 0};
            const double Rj[9] = {0, dj2, -dj1, -dj2, 0, dj0, dj1, -dj0, 0};
            double M1[9];   // H R_j
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int k = 0; k < 3; ++k) M1[3 * p + k] = h[3 * p] * Rj[k] + h[3 * p + 1] * Rj[3 + k] + h[3 * p + 2] * Rj[6 + k];
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int q2 = 0; q2 < 3; ++q2) {
                    acc[6 * p + q2] += h[3 * p + q2];                // translations x translations
                    acc[6 * p + 3 + q2] += M1[3 * p + q2];           // translations x rotations
                    // rotations x translations: (R_i^T H)[p][q2]
                    acc[6 * (3 + p) + q2] += Ri[p] * h[q2] + Ri[3 + p] * h[3 + q2] + Ri[6 + p] * h[6 + q2];
                    // rotations x rotations: (R_i^T H R_j)[p][q2]
                    acc[6 * (3 + p) + 3 + q2] += Ri[p] * M1[q2] + Ri[3 + p] * M1[3 + q2] + Ri[6 + p] * M1[6 + q2];
                }
        }
#pragma unroll
        for (int k = 0; k < 36; ++k) acc[k] = wave_sum(acc[k]);
#pragma unroll
        for (int k = 0; k < 36; ++k)
            if (lane == k) A[(6l * a + k / 6) * ld + 6l * b + k % 6] = acc[k];
    }
}

__global__ void k_cg_diag(const double* __restrict__ A, long ld, int npad, double* __restrict__ diag0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npad) diag0[i] = A[(long)i * ld + i];
}

__device__ __forceinline__ float cg_store(double v, float*) { return (float)v; }
__device__ __forceinline__ uint16_t cg_store(double v, uint16_t*) {
    const uint32_t u = __float_as_uint((float)v);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);   // BF16, round to nearest even
}

// W = L^-1 (lower tiles) and WT = its transpose (upper tiles) in the storage type S (FP32
// or BF16); one workgroup per lower tile, staged through LDS so that both writes are
// coalesced
template <class S>
__global__ __launch_bounds__(kCgBlock) void k_cg_store(const double* __restrict__ W, long ld, S* __restrict__ Ws, S* __restrict__ WTs) {
    __shared__ S T[kCgTile][kCgTile + 2];
    int b = blockIdx.x, ti = 0;
    while (b >= ti + 1) { b -= ti + 1; ++ti; }
    const int tj = b;
    const long oi = (long)ti * kCgTile, oj = (long)tj * kCgTile;
    for (int e = threadIdx.x; e < kCgTile * kCgTile; e += kCgBlock) {
        const int r = e / kCgTile, q = e % kCgTile;
        const S v = cg_store(W[(oi + r) * ld + oj + q], (S*)nullptr);
        T[r][q] = v;
        Ws[(oi + r) * ld + oj + q] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kCgTile * kCgTile; e += kCgBlock) {
        const int r = e / kCgTile, q = e % kCgTile;
        WTs[(oj + r) * ld + oi + q] = T[q][r];
    }
}

// rc = P^T r: per aggregate (one workgroup) the sums of r_i and d_i x r_i, reduced in a
// fixed order
template <class V>
__global__ __launch_bounds__(kCgBlock) void k_cg_restrict(int nAgg, const int* __restrict__ aggPtr, const int* __restrict__ aggVerts,
                                                          const double* __restrict__ dm, const V* __restrict__ r, double* __restrict__ rc,
                                                          const int* __restrict__ done) {
    if (done && *done) return;
    __shared__ double part[kCgBlock / 64][6];
    const int a = blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int q1 = aggPtr[a + 1];
    double s[6] = {0, 0, 0, 0, 0, 0};
    // members q, q + kCgBlock, ... of this lane, two per pass with all loads issued first
    // (the offsets are stored in member order, so only r is gathered); summed in member
    // order
    for (int q = aggPtr[a] + threadIdx.x; q < q1; q += 2 * kCgBlock) {
        const int qb = q + kCgBlock;
        const bool hb = qb < q1;
        const int va = aggVerts[q], vb = aggVerts[hb ? qb : q];
        double dv[2][3], rv[2][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            dv[0][k] = dm[3 * q + k];
            dv[1][k] = dm[3 * (hb ? qb : q) + k];
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            rv[0][k] = (double)r[3 * va + k];
            rv[1][k] = (double)r[3 * vb + k];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (u == 1 && !hb) break;
            const double* dd = dv[u];
            const double* rr = rv[u];
            s[0] += rr[0];
            s[1] += rr[1];
            s[2] += rr[2];
            s[3] += dd[1] * rr[2] - dd[2] * rr[1];
            s[4] += dd[2] * rr[0] - dd[0] * rr[2];
            s[5] += dd[0] * rr[1] - dd[1] * rr[0];
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = wave_sum(s[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) part[wv][k] = s[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        rc[6 * a + k] = (part[0][k] + part[1][k]) + (part[2][k] + part[3][k]);
    }
}

// 16 bytes of a row as doubles: four FP32 or eight BF16 entries
__device__ __forceinline__ void cg_unpack(const float4& w, const float*, double* e) {
    e[0] = w.x; e[1] = w.y; e[2] = w.z; e[3] = w.w;
}
__device__ __forceinline__ void cg_unpack(const float4& w, const uint16_t*, double* e) {
    const uint32_t words[4] = {__float_as_uint(w.x), __float_as_uint(w.y), __float_as_uint(w.z), __float_as_uint(w.w)};
#pragma unroll
    for (int k = 0; k < 8; ++k) e[k] = (double)__uint_as_float((k & 1) ? (words[k >> 1] & 0xFFFF0000u) : (words[k >> 1] << 16));
}

// y = W x over the lower (W = L^-1) or upper (W = L^-T) triangle of a row-major matrix of
// FP32 or BF16 entries, FP64 accumulation.  One workgroup per pair of bands of kCgBand
// rows, (kCgBand b ..) and (npad - kCgBand (b+1) ..), whose triangle parts together hold
// about the same number of entries for every b; 16 bytes per lane load, kCgChunk of them
// per row issued before any is used.  Each lane's slice of x is loaded once and used for
// the band's kCgBand rows: with one row per workgroup the x loads (twice the matrix's
// bytes for FP32, four times for BF16, served by L2) cost more than the matrix stream.
// Fixed-order reduction.
constexpr int kCgBand = 4;
static_assert(kCgTile % (2 * kCgBand) == 0, "the padded size must split into band pairs");
template <class S, bool UPPER>
__global__ __launch_bounds__(kCgBlock) void k_cg_gemv(const S* __restrict__ W, int npad, const double* __restrict__ x,
                                                      double* __restrict__ y, const int* __restrict__ done) {
    if (done && *done) return;
    constexpr int E = 16 / sizeof(S);   // entries per 16-byte group
    constexpr int R = kCgBand;
    constexpr int C = sizeof(S) == 4 ? 1 : 2;
    typedef float f4 __attribute__((ext_vector_type(4)));
    __shared__ double part[2 * R][kCgBlock / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rA = R * (int)blockIdx.x, rB = npad - R - R * (int)blockIdx.x;
    // each band's groups: the union of its rows' triangle parts (entries outside a row's
    // part are masked below)
    const int a0 = (UPPER ? rA : 0) / E, a1 = (UPPER ? npad - 1 : rA + R - 1) / E + 1;
    const int b0 = (UPPER ? rB : 0) / E, b1 = (UPPER ? npad - 1 : rB + R - 1) / E + 1;
    const int nA = a1 - a0, T = nA + (b1 - b0);
    double accA[R], accB[R];
#pragma unroll
    for (int q = 0; q < R; ++q) accA[q] = accB[q] = 0;
    for (int base = 0; base < T; base += C * kCgBlock) {
        f4 raw[C][R];
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int t = base + u * kCgBlock + (int)threadIdx.x;
            const int r = t < nA ? rA : rB;
            const int g = t < nA ? a0 + t : b0 + t - nA;
            // streamed once per iteration: non-temporal, so that it does not push the
            // SpMV's matrix out of the last-level cache
#pragma unroll
            for (int q = 0; q < R; ++q)
                raw[u][q] = t < T ? __builtin_nontemporal_load(reinterpret_cast<const f4*>(W + (long)(r + q) * npad) + g) : f4{0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int t = base + u * kCgBlock + (int)threadIdx.x;
            if (t < T) {
                const bool isA = t < nA;
                const int r = isA ? rA : rB;
                const int j = E * (isA ? a0 + t : b0 + t - nA);
                double xv[E];
#pragma unroll
                for (int k = 0; k < E; ++k) xv[k] = x[j + k];
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    double e[E];
                    cg_unpack(make_float4(raw[u][q].x, raw[u][q].y, raw[u][q].z, raw[u][q].w), (const S*)nullptr, e);
                    double s = 0;
#pragma unroll
                    for (int k = 0; k < E; ++k) {
                        const bool in = UPPER ? (j + k >= r + q) : (j + k <= r + q);
                        if (in) s += e[k] * xv[k];
                    }
                    if (isA) accA[q] += s;
                    else accB[q] += s;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const double va = wave_sum(accA[q]);
        const double vb = wave_sum(accB[q]);
        if (lane == 0) {
            part[q][wv] = va;
            part[R + q][wv] = vb;
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * R) {
        const int q = threadIdx.x;
        y[q < R ? rA + q : rB + q - R] = (part[q][0] + part[q][1]) + (part[q][2] + part[q][3]);
    }
}

static double wall_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool coarse_prepare(Ctx& c, const double* x) {
    const double t0 = wall_ms();
    CoarseSpace& g = c.cg;
    if (!g.built || g.nVBuilt != c.nV || g.nnzbBuilt != c.Ael.nnzb || g.builtMaxDof != g.maxDof) coarse_build(c);
    g.ready = false;
    if (g.nAgg == 0) return false;
    BSR& E = c.Ael;
    BSR& K = c.Acon;
    if (K.nnzb == 0) K.rowptr.ensure(c.nV + 1);
    const long ld = g.npad;
    hipLaunchKernelGGL(k_cg_centroids, dim3(nblocks(g.nAgg, kCgBlock / 64)), dim3(kCgBlock), 0, c.s, g.nAgg, (const int*)g.aggPtr.p,
                       (const int*)g.aggVerts.p, x, g.cent.p);
    hipLaunchKernelGGL(k_cg_offsets, dim3(nblocks(c.nV)), dim3(kBlock), 0, c.s, c.nV, (const int*)g.agg.p, x, (const double*)g.cent.p, g.d.p);
    hipLaunchKernelGGL(k_cg_member_offsets, dim3(nblocks(g.nMembers)), dim3(kBlock), 0, c.s, g.nMembers, (const int*)g.aggVerts.p,
                       (const double*)g.d.p, g.dm.p);
    hipLaunchKernelGGL(k_cg_zero_identity, dim3(2048), dim3(kBlock), 0, c.s, g.A.p, (long)g.npad * g.npad, g.nc, g.npad);
    CIPC_HIP(hipMemsetAsync(g.flag.p, 0, sizeof(int), c.s));
    CIPC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cg_galerkin, dim3(g.nAgg), dim3(kCgBlock), 0, c.s, (const int*)g.aggPtr.p, (const int*)g.aggVerts.p,
                       (const int*)g.agg.p, (const double*)g.d.p, (const int*)E.rowptr.p, (const int*)E.col.p, (const double*)E.val.p,
                       E.nnzb, (const int*)K.rowptr.p, (const int*)K.col.p, (const double*)K.val.p, K.nnzb, g.A.p, ld, g.flag.p);
    hipLaunchKernelGGL(k_cg_diag, dim3(nblocks(g.npad)), dim3(kBlock), 0, c.s, (const double*)g.A.p, ld, g.npad, g.diag0.p);
    CIPC_LAUNCH_CHECK();
    // a pivot below 1e-9 of its diagonal is a mode the matrix does not see (one vertex,
    // or collinear members: rotation about their line)
    dense_cholesky(c, g.A.p, g.npad, g.flag.p, g.tileInv.p, g.diag0.p, g.dropped.p, 1e-9);
    dense_lower_inverse(c, g.A.p, g.W.p, g.npad, g.tileInv.p);
    const int nt = g.npad / kCgTile;
    if (g.bf16) {
        g.W16.ensure_exact((size_t)g.npad * g.npad);
        g.WT16.ensure_exact((size_t)g.npad * g.npad);
        hipLaunchKernelGGL(k_cg_store<uint16_t>, dim3(nt * (nt + 1) / 2), dim3(kCgBlock), 0, c.s, (const double*)g.W.p, ld, g.W16.p,
                           g.WT16.p);
    } else {
        g.W32.ensure_exact((size_t)g.npad * g.npad);
        g.WT32.ensure_exact((size_t)g.npad * g.npad);
        hipLaunchKernelGGL(k_cg_store<float>, dim3(nt * (nt + 1) / 2), dim3(kCgBlock), 0, c.s, (const double*)g.W.p, ld, g.W32.p,
                           g.WT32.p);
    }
    CIPC_LAUNCH_CHECK();
    CIPC_HIP(hipMemcpyAsync(c.hCnt + 49, g.flag.p, sizeof(int), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    g.setups++;
    g.lastSetupMs = wall_ms() - t0;
    if (c.hCnt[49] & 1) {   // an aggregate with too many neighbours: no coarse level this time
        g.failures++;
        return false;
    }
    g.ready = true;
    return true;
}

template <class V>
void coarse_restrict_solve(Ctx& c, const V* r, const int* done) {
    CoarseSpace& g = c.cg;
    hipLaunchKernelGGL((k_cg_restrict<V>), dim3(g.nAgg), dim3(kCgBlock), 0, c.s, g.nAgg, (const int*)g.aggPtr.p,
                       (const int*)g.aggVerts.p, (const double*)g.dm.p, r, g.rc.p, done);
    if (g.bf16) {
        hipLaunchKernelGGL((k_cg_gemv<uint16_t, false>), dim3(g.npad / (2 * kCgBand)), dim3(kCgBlock), 0, c.s, (const uint16_t*)g.W16.p, g.npad,
                           (const double*)g.rc.p, g.wc.p, done);
        hipLaunchKernelGGL((k_cg_gemv<uint16_t, true>), dim3(g.npad / (2 * kCgBand)), dim3(kCgBlock), 0, c.s, (const uint16_t*)g.WT16.p, g.npad,
                           (const double*)g.wc.p, g.zc.p, done);
    } else {
        hipLaunchKernelGGL((k_cg_gemv<float, false>), dim3(g.npad / (2 * kCgBand)), dim3(kCgBlock), 0, c.s, (const float*)g.W32.p, g.npad,
                           (const double*)g.rc.p, g.wc.p, done);
        hipLaunchKernelGGL((k_cg_gemv<float, true>), dim3(g.npad / (2 * kCgBand)), dim3(kCgBlock), 0, c.s, (const float*)g.WT32.p, g.npad,
                           (const double*)g.wc.p, g.zc.p, done);
    }
    CIPC_LAUNCH_CHECK();
}
template void coarse_restrict_solve<double>(Ctx&, const double*, const int*);
template void coarse_restrict_solve<float>(Ctx&, const float*, const int*);

}  // namespace cipc
