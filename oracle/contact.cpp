// Oracle: IPC contact -- constraint set (broad + narrow phase, dedup with
// multiplicity), barrier energy / gradient / Hessian with per-stencil PSD projection,
// minimum distance, and the CCD step size with the spatial-hash step clamp.
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Restates FEM/IPC.h:19-740 (Compute_Constraint_Set, shell path with the rod and
// particle branches :271-353), :742-941 (Compute_Barrier), :943-1256 (Compute_Barrier_Gradient),
// :1258-1731 (Compute_Barrier_Hessian), :1879-2244 (Compute_Intersection_Free_StepSize),
// :2246-2388 (Compute_Min_Dist2), Grid/SPATIAL_HASH.h:432-622 (CCD hash incl. the
// step clamp) and Math/Distance/CCD.h:150-483.
//
// Ordering: the reference orders PT constraints of one point by unordered_set
// iteration (unpinned); here they are ordered by triangle index.  EE constraints keep
// the reference's (eI, eJ ascending) order (Query_Edge_For_Edges sorts), and the
// deduplicated PP/PE entries keep std::map<VECTOR<int,4>> (lexicographic) order.
//
// Parity pins (tests/test_oracle_contact.py): the hashed constraint set equals an
// O(n^2) brute-force evaluation of the same predicates; every active pair has
// distance < (dhat+xi)^2 and every rejected pair within the AABB filter does not;
// CCD step sizes never exceed the analytic time of impact of closed-form cases and
// every accepted state is intersection free (exact segment-triangle test).
#include "oracle.h"
#include <algorithm>
#include <map>
#include <unordered_map>
#include <cstdio>

namespace orc {

static inline V3 X3(const double* x, int i) { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }

// CCD.h:150-165
static bool pt_cd_broadphase(V3 p, V3 t0, V3 t1, V3 t2, double dist) {
    double mx[3] = {std::max(std::max(t0.x, t1.x), t2.x), std::max(std::max(t0.y, t1.y), t2.y), std::max(std::max(t0.z, t1.z), t2.z)};
    double mn[3] = {std::min(std::min(t0.x, t1.x), t2.x), std::min(std::min(t0.y, t1.y), t2.y), std::min(std::min(t0.z, t1.z), t2.z)};
    double pp[3] = {p.x, p.y, p.z};
    for (int k = 0; k < 3; ++k) if (pp[k] - mx[k] > dist || mn[k] - pp[k] > dist) return false;
    return true;
}
// CCD.h:15-30
static bool pe_cd_broadphase(V3 p, V3 e0, V3 e1, double dist) {
    if (p.x - std::max(e0.x, e1.x) > dist || std::min(e0.x, e1.x) - p.x > dist) return false;
    if (p.y - std::max(e0.y, e1.y) > dist || std::min(e0.y, e1.y) - p.y > dist) return false;
    if (p.z - std::max(e0.z, e1.z) > dist || std::min(e0.z, e1.z) - p.z > dist) return false;
    return true;
}
// CCD.h:168-185
static bool ee_cd_broadphase(V3 a0, V3 a1, V3 b0, V3 b1, double dist) {
    double A0[3] = {a0.x, a0.y, a0.z}, A1[3] = {a1.x, a1.y, a1.z}, B0[3] = {b0.x, b0.y, b0.z}, B1[3] = {b1.x, b1.y, b1.z};
    for (int k = 0; k < 3; ++k) {
        double maxa = std::max(A0[k], A1[k]), mina = std::min(A0[k], A1[k]);
        double maxb = std::max(B0[k], B1[k]), minb = std::min(B0[k], B1[k]);
        if (mina - maxb > dist || minb - maxa > dist) return false;
    }
    return true;
}

// ----------------------------------------------------------------- uniform grid
// A complete broad phase over AABBs; the candidate set it yields is filtered by the
// exact predicates, so its cell size only affects speed.
struct Grid {
    double org[3], inv;
    std::unordered_map<long long, std::vector<int>> cells;
    long long key(long long i, long long j, long long k) const { return (i * 73856093LL) ^ (j * 19349663LL) ^ (k * 83492791LL); }
    void cell(const double* p, long long* c) const { for (int d = 0; d < 3; ++d) c[d] = (long long)std::floor((p[d] - org[d]) * inv); }
    void insert(const double* lo, const double* hi, int id) {
        long long a[3], b[3]; cell(lo, a); cell(hi, b);
        for (long long i = a[0]; i <= b[0]; ++i)
            for (long long j = a[1]; j <= b[1]; ++j)
                for (long long k = a[2]; k <= b[2]; ++k) cells[key(i, j, k)].push_back(id);
    }
    void query(const double* lo, const double* hi, std::vector<int>& out) const {
        long long a[3], b[3]; cell(lo, a); cell(hi, b);
        out.clear();
        for (long long i = a[0]; i <= b[0]; ++i)
            for (long long j = a[1]; j <= b[1]; ++j)
                for (long long k = a[2]; k <= b[2]; ++k) {
                    auto it = cells.find(key(i, j, k));
                    if (it != cells.end()) out.insert(out.end(), it->second.begin(), it->second.end());
                }
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    }
};

static double mean_bedge_len(const Scene& S, const double* x);

// ----------------------------------------------------------------- constraint set
void constraint_set(const Scene& S, const double* x, std::vector<Constraint>& out, bool bruteForce) {
    out.clear();
    const double dHat = std::sqrt(S.dHat2) + S.P.xi;   // IPC.h:53-54
    const double dHat2 = dHat * dHat;
    const int nBN = (int)S.bNode.size(), nBE = (int)S.bEdge.size() / 2, nBT = (int)S.bTri.size() / 3;
    const double* X0 = S.X0.data();
    const char* dbc = S.isDBC.data();

    const int nRod = (int)S.rod.size() / 2;
    Grid gT, gE, gP;
    if (!bruteForce) {
        double cs = mean_bedge_len(S, x);
        if (!(cs > 0)) cs = 1;
        double lo[3] = {1e300, 1e300, 1e300};
        for (int i = 0; i < S.nV; ++i) for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], x[3 * i + d]);
        for (Grid* g : {&gT, &gE, &gP}) { for (int d = 0; d < 3; ++d) g->org[d] = lo[d]; g->inv = 1.0 / cs; }
        for (int s = S.codimBN1; s < nBN; ++s) {
            double q[3] = {x[3 * S.bNode[s]], x[3 * S.bNode[s] + 1], x[3 * S.bNode[s] + 2]};
            gP.insert(q, q, s);
        }
        for (int t = 0; t < nBT; ++t) {
            double l[3], h[3];
            for (int d = 0; d < 3; ++d) {
                double a = x[3 * S.bTri[3 * t] + d], b = x[3 * S.bTri[3 * t + 1] + d], c = x[3 * S.bTri[3 * t + 2] + d];
                l[d] = std::min(std::min(a, b), c); h[d] = std::max(std::max(a, b), c);
            }
            gT.insert(l, h, t);
        }
        for (int e = 0; e < nBE; ++e) {
            double l[3], h[3];
            for (int d = 0; d < 3; ++d) {
                double a = x[3 * S.bEdge[2 * e] + d], b = x[3 * S.bEdge[2 * e + 1] + d];
                l[d] = std::min(a, b); h[d] = std::max(a, b);
            }
            gE.insert(l, h, e);
        }
    }

    // ---- point-triangle, IPC.h:144-355
    std::vector<std::vector<Constraint>> keptPT(nBN), dupPT(nBN);
#pragma omp parallel for schedule(dynamic, 64) num_threads(S.P.threads)
    for (int svI = 0; svI < nBN; ++svI) {
        int vI = S.bNode[svI];
        V3 p = X3(x, vI);
        std::vector<int> cand;
        if (bruteForce) { cand.resize(nBT); for (int t = 0; t < nBT; ++t) cand[t] = t; }
        else {
            double l[3] = {p.x - dHat, p.y - dHat, p.z - dHat}, h[3] = {p.x + dHat, p.y + dHat, p.z + dHat};
            gT.query(l, h, cand);
        }
        for (int sfI : cand) {
            const int* t = &S.bTri[3 * sfI];
            if (vI == t[0] || vI == t[1] || vI == t[2]) continue;
            if (dbc[vI] && dbc[t[0]] && dbc[t[1]] && dbc[t[2]]) continue;
            V3 t0 = X3(x, t[0]), t1 = X3(x, t[1]), t2 = X3(x, t[2]);
            if (!pt_cd_broadphase(p, t0, t1, t2, dHat)) continue;
            double d;
            Constraint c{{-vI - 1, -1, -1, -1}};
            bool isPT = false;
            switch (pt_type(p, t0, t1, t2)) {
                case 0: d = dist2_pp(p, t0); c.c[1] = t[0]; break;
                case 1: d = dist2_pp(p, t1); c.c[1] = t[1]; break;
                case 2: d = dist2_pp(p, t2); c.c[1] = t[2]; break;
                case 3: d = dist2_pe(p, t0, t1); c.c[1] = t[0]; c.c[2] = t[1]; break;
                case 4: d = dist2_pe(p, t1, t2); c.c[1] = t[1]; c.c[2] = t[2]; break;
                case 5: d = dist2_pe(p, t2, t0); c.c[1] = t[2]; c.c[2] = t[0]; break;
                default: d = dist2_pt(p, t0, t1, t2); c.c[1] = t[0]; c.c[2] = t[1]; c.c[3] = t[2]; isPT = true; break;
            }
            if (d < dHat2) (isPT ? keptPT : dupPT)[svI].push_back(c);
        }
        if (svI < S.codimBN0) continue;
        // rod nodes and particles against rod segments, IPC.h:271-317
        std::vector<int> ecand;
        if (bruteForce) { for (int e = nBE - nRod; e < nBE; ++e) ecand.push_back(e); }
        else {
            double l[3] = {p.x - dHat, p.y - dHat, p.z - dHat}, h[3] = {p.x + dHat, p.y + dHat, p.z + dHat};
            gE.query(l, h, ecand);
        }
        for (int eI : ecand) {
            if (eI < nBE - nRod) continue;
            const int* e = &S.bEdge[2 * eI];
            if (vI == e[0] || vI == e[1]) continue;
            if (dbc[vI] && dbc[e[0]] && dbc[e[1]]) continue;
            V3 e0 = X3(x, e[0]), e1 = X3(x, e[1]);
            if (!pe_cd_broadphase(p, e0, e1, dHat)) continue;
            double d;
            Constraint c{{-vI - 1, -1, -1, -1}};
            switch (pe_type(p, e0, e1)) {
                case 0: d = dist2_pp(p, e0); c.c[1] = e[0]; break;
                case 1: d = dist2_pp(p, e1); c.c[1] = e[1]; break;
                default: d = dist2_pe(p, e0, e1); c.c[1] = e[0]; c.c[2] = e[1]; break;
            }
            if (d < dHat2) dupPT[svI].push_back(c);
        }
        if (svI < S.codimBN1) continue;
        // particle against the boundary nodes after it (all particles), IPC.h:319-349
        std::vector<int> pcand;
        if (bruteForce) { for (int j = svI + 1; j < nBN; ++j) pcand.push_back(j); }
        else {
            double l[3] = {p.x - dHat, p.y - dHat, p.z - dHat}, h[3] = {p.x + dHat, p.y + dHat, p.z + dHat};
            gP.query(l, h, pcand);
        }
        for (int svJ : pcand) {
            if (svJ <= svI) continue;
            int vJ = S.bNode[svJ];
            if (dbc[vI] && dbc[vJ]) continue;
            double d = dist2_pp(p, X3(x, vJ));
            if (d < dHat2) dupPT[svI].push_back(Constraint{{-vI - 1, vJ, -1, -1}});
        }
    }

    // ---- edge-edge, IPC.h:357-568
    std::vector<std::vector<Constraint>> keptEE(nBE), dupEE(nBE);
#pragma omp parallel for schedule(dynamic, 64) num_threads(S.P.threads)
    for (int eI = 0; eI < nBE; ++eI) {
        const int* ei = &S.bEdge[2 * eI];
        V3 ea0 = X3(x, ei[0]), ea1 = X3(x, ei[1]);
        std::vector<int> cand;
        if (bruteForce) { for (int e = eI + 1; e < nBE; ++e) cand.push_back(e); }
        else {
            double l[3] = {std::min(ea0.x, ea1.x) - dHat, std::min(ea0.y, ea1.y) - dHat, std::min(ea0.z, ea1.z) - dHat};
            double h[3] = {std::max(ea0.x, ea1.x) + dHat, std::max(ea0.y, ea1.y) + dHat, std::max(ea0.z, ea1.z) + dHat};
            gE.query(l, h, cand);
        }
        for (int eJ : cand) {
            if (eJ <= eI) continue;
            const int* ej = &S.bEdge[2 * eJ];
            if (ei[0] == ej[0] || ei[0] == ej[1] || ei[1] == ej[0] || ei[1] == ej[1]) continue;
            if (dbc[ei[0]] && dbc[ei[1]] && dbc[ej[0]] && dbc[ej[1]]) continue;
            V3 eb0 = X3(x, ej[0]), eb1 = X3(x, ej[1]);
            if (!ee_cd_broadphase(ea0, ea1, eb0, eb1, dHat)) continue;
            double crossN = ee_cross_norm2(ea0, ea1, eb0, eb1);
            double eps = ee_mollifier_threshold(X3(X0, ei[0]), X3(X0, ei[1]), X3(X0, ej[0]), X3(X0, ej[1]));
            bool moll = crossN < eps;
            double d;
            Constraint c;
            int a0 = ei[0], a1 = ei[1], b0 = ej[0], b1 = ej[1];
            switch (ee_type(ea0, ea1, eb0, eb1)) {
                case 0: d = dist2_pp(ea0, eb0);
                    c = moll ? Constraint{{a0, b0, -a1 - 1, -b1 - 1}} : Constraint{{-a0 - 1, b0, -1, -1}}; break;
                case 1: d = dist2_pp(ea0, eb1);
                    c = moll ? Constraint{{a0, b1, -a1 - 1, -b0 - 1}} : Constraint{{-a0 - 1, b1, -1, -1}}; break;
                case 2: d = dist2_pe(ea0, eb0, eb1);
                    c = moll ? Constraint{{a0, b0, b1, -a1 - 1}} : Constraint{{-a0 - 1, b0, b1, -1}}; break;
                case 3: d = dist2_pp(ea1, eb0);
                    c = moll ? Constraint{{a1, b0, -a0 - 1, -b1 - 1}} : Constraint{{-a1 - 1, b0, -1, -1}}; break;
                case 4: d = dist2_pp(ea1, eb1);
                    c = moll ? Constraint{{a1, b1, -a0 - 1, -b0 - 1}} : Constraint{{-a1 - 1, b1, -1, -1}}; break;
                case 5: d = dist2_pe(ea1, eb0, eb1);
                    c = moll ? Constraint{{a1, b0, b1, -a0 - 1}} : Constraint{{-a1 - 1, b0, b1, -1}}; break;
                case 6: d = dist2_pe(eb0, ea0, ea1);
                    c = moll ? Constraint{{b0, a0, a1, -b1 - 1}} : Constraint{{-b0 - 1, a0, a1, -1}}; break;
                case 7: d = dist2_pe(eb1, ea0, ea1);
                    c = moll ? Constraint{{b1, a0, a1, -b0 - 1}} : Constraint{{-b1 - 1, a0, a1, -1}}; break;
                default: d = dist2_ee(ea0, ea1, eb0, eb1);
                    c = moll ? Constraint{{a0, a1, -b0 - 1, b1}} : Constraint{{a0, a1, b0, b1}}; break;
            }
            if (d < dHat2) (c.c[0] < 0 ? dupEE : keptEE)[eI].push_back(c);
        }
    }

    // ---- merge with PP/PE multiplicity, IPC.h:571-661
    std::map<std::array<int, 4>, int> counter;
    for (int svI = 0; svI < nBN; ++svI) {
        for (auto& c : keptPT[svI]) out.push_back(c);
        for (auto& c : dupPT[svI]) ++counter[{c.c[0], c.c[1], c.c[2], c.c[3]}];
    }
    for (int eI = 0; eI < nBE; ++eI) {
        for (auto& c : keptEE[eI]) out.push_back(c);
        for (auto& c : dupEE[eI]) ++counter[{c.c[0], c.c[1], c.c[2], c.c[3]}];
    }
    for (auto& kv : counter) out.push_back(Constraint{{kv.first[0], kv.first[1], kv.first[2], -kv.second}});
    // stencil weights are all 1 on this path (IPC.h:656-660)
}

// ----------------------------------------------------------------- per-constraint evaluation
struct Local { int nv; int vid[4]; double E; double g[12]; double H[144]; double d2; };

// embed a distance derivative over a vertex subset into the 4-vertex stencil
static void embed(int nsub, const int* slots, const double* gs, const double* Hs, double* g12, double* H12) {
    for (int i = 0; i < 12; ++i) g12[i] = 0;
    for (int i = 0; i < 144; ++i) H12[i] = 0;
    int n = nsub * 3;
    for (int a = 0; a < nsub; ++a)
        for (int k = 0; k < 3; ++k) {
            g12[slots[a] * 3 + k] = gs[a * 3 + k];
            for (int b = 0; b < nsub; ++b)
                for (int l = 0; l < 3; ++l) H12[(slots[a] * 3 + k) * 12 + slots[b] * 3 + l] = Hs[(a * 3 + k) * n + b * 3 + l];
        }
}

// returns false when d^2 - xi^2 <= 0
static bool eval_constraint(const Scene& S, const double* x, const Constraint& cc, double kappa, int want, Local& L) {
    const int* c = cc.c;
    const double xi2 = S.P.xi * S.P.xi;
    const double shat = S.dHat2 + 2 * std::sqrt(S.dHat2) * S.P.xi;   // IPC.h:756-757
    double gd[12], Hd[144];
    double mult = 1;
    bool moll = false;
    if (c[0] >= 0) {
        if (c[2] >= 0 && c[3] >= 0) {
            L.nv = 4; for (int i = 0; i < 4; ++i) L.vid[i] = c[i];
            V3 a0 = X3(x, c[0]), a1 = X3(x, c[1]), b0 = X3(x, c[2]), b1 = X3(x, c[3]);
            L.d2 = dist2_ee(a0, a1, b0, b1);
            if (want) dist2_ee_gh(a0, a1, b0, b1, gd, want > 1 ? Hd : nullptr);
        } else {
            moll = true;
            L.nv = 4;
            double gs[9], Hs[81];
            if (c[3] >= 0) {
                int v[4] = {c[0], c[1], -c[2] - 1, c[3]};
                for (int i = 0; i < 4; ++i) L.vid[i] = v[i];
                L.d2 = dist2_ee(X3(x, v[0]), X3(x, v[1]), X3(x, v[2]), X3(x, v[3]));
                if (want) dist2_ee_gh(X3(x, v[0]), X3(x, v[1]), X3(x, v[2]), X3(x, v[3]), gd, want > 1 ? Hd : nullptr);
            } else if (c[2] >= 0) {
                int v[4] = {c[0], -c[3] - 1, c[1], c[2]};
                for (int i = 0; i < 4; ++i) L.vid[i] = v[i];
                L.d2 = dist2_pe(X3(x, v[0]), X3(x, v[2]), X3(x, v[3]));
                if (want) {
                    dist2_pe_gh(X3(x, v[0]), X3(x, v[2]), X3(x, v[3]), gs, Hs);
                    int sl[3] = {0, 2, 3};
                    embed(3, sl, gs, Hs, gd, Hd);
                }
            } else {
                int v[4] = {c[0], -c[2] - 1, c[1], -c[3] - 1};
                for (int i = 0; i < 4; ++i) L.vid[i] = v[i];
                L.d2 = dist2_pp(X3(x, v[0]), X3(x, v[2]));
                if (want) {
                    dist2_pp_gh(X3(x, v[0]), X3(x, v[2]), gs, Hs);
                    int sl[2] = {0, 2};
                    embed(2, sl, gs, Hs, gd, Hd);
                }
            }
        }
    } else {
        int p = -c[0] - 1;
        V3 P = X3(x, p);
        if (c[3] >= 0) {
            L.nv = 4; L.vid[0] = p; L.vid[1] = c[1]; L.vid[2] = c[2]; L.vid[3] = c[3];
            L.d2 = dist2_pt(P, X3(x, c[1]), X3(x, c[2]), X3(x, c[3]));
            if (want) dist2_pt_gh(P, X3(x, c[1]), X3(x, c[2]), X3(x, c[3]), gd, want > 1 ? Hd : nullptr);
        } else if (c[2] >= 0) {
            L.nv = 3; L.vid[0] = p; L.vid[1] = c[1]; L.vid[2] = c[2];
            L.d2 = dist2_pe(P, X3(x, c[1]), X3(x, c[2]));
            if (want) dist2_pe_gh(P, X3(x, c[1]), X3(x, c[2]), gd, want > 1 ? Hd : nullptr);
            mult = -c[3];
        } else {
            L.nv = 2; L.vid[0] = p; L.vid[1] = c[1];
            L.d2 = dist2_pp(P, X3(x, c[1]));
            if (want) dist2_pp_gh(P, X3(x, c[1]), gd, want > 1 ? Hd : nullptr);
            mult = -c[3];
        }
    }
    double s = L.d2 - xi2;
    if (s <= 0) return false;
    double b = barrier(s, shat, kappa);
    int n = L.nv * 3;
    if (!moll) {
        L.E = mult * b;
        if (want >= 1) {
            double bg = barrier_g(s, shat, kappa);
            for (int i = 0; i < n; ++i) L.g[i] = mult * bg * gd[i];
            if (want >= 2) {
                double bh = barrier_h(s, shat, kappa);
                for (int i = 0; i < n; ++i)
                    for (int j = 0; j < n; ++j) L.H[i * n + j] = mult * (bh * gd[i] * gd[j] + bg * Hd[i * n + j]);
                if (S.P.projectPD) make_pd(n, L.H);
            }
        }
    } else {
        const double* X0 = S.X0.data();
        double eps = ee_mollifier_threshold(X3(X0, L.vid[0]), X3(X0, L.vid[1]), X3(X0, L.vid[2]), X3(X0, L.vid[3]));
        double e, ge[12], He[144];
        ee_mollifier(X3(x, L.vid[0]), X3(x, L.vid[1]), X3(x, L.vid[2]), X3(x, L.vid[3]), eps, &e,
                     want ? ge : nullptr, want > 1 ? He : nullptr);
        L.E = e * b;
        if (want >= 1) {
            double bg = barrier_g(s, shat, kappa);
            for (int i = 0; i < 12; ++i) L.g[i] = e * bg * gd[i] + b * ge[i];
            if (want >= 2) {
                double bh = barrier_h(s, shat, kappa);
                for (int i = 0; i < 12; ++i)
                    for (int j = 0; j < 12; ++j)
                        L.H[i * 12 + j] = bg * gd[i] * ge[j] + bg * ge[i] * gd[j] + b * He[i * 12 + j]
                                        + e * bh * gd[i] * gd[j] + e * bg * Hd[i * 12 + j];
                if (S.P.projectPD) make_pd(12, L.H);
            }
        }
    }
    return true;
}

double barrier_energy(const Scene& S, const double* x, const std::vector<Constraint>& C, double kappa, int* bad) {
    double E = 0;
    if (bad) *bad = 0;
    for (size_t i = 0; i < C.size(); ++i) {
        Local L;
        if (!eval_constraint(S, x, C[i], kappa, 0, L)) { if (bad) *bad = 1; continue; }
        E += L.E;
    }
    return E;
}

void barrier_grad(const Scene& S, const double* x, const std::vector<Constraint>& C, double kappa, double* g) {
    for (size_t i = 0; i < C.size(); ++i) {
        Local L;
        if (!eval_constraint(S, x, C[i], kappa, 1, L)) continue;
        for (int a = 0; a < L.nv; ++a)
            for (int k = 0; k < 3; ++k) g[3 * L.vid[a] + k] += L.g[a * 3 + k];
    }
}

void barrier_hess(const Scene& S, const double* x, const std::vector<Constraint>& C, double kappa, std::vector<Block>& out) {
    std::vector<std::vector<Block>> per(C.size());
#pragma omp parallel for schedule(dynamic, 32) num_threads(S.P.threads)
    for (int i = 0; i < (int)C.size(); ++i) {
        Local L;
        if (!eval_constraint(S, x, C[i], kappa, 2, L)) continue;
        int n = L.nv * 3;
        for (int a = 0; a < L.nv; ++a)
            for (int b = 0; b < L.nv; ++b) {
                Block B; B.r = L.vid[a]; B.c = L.vid[b];
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) B.b[k * 3 + l] = L.H[(a * 3 + k) * n + b * 3 + l];
                per[i].push_back(B);
            }
    }
    for (auto& v : per) out.insert(out.end(), v.begin(), v.end());
}

// IPC.h:2246-2388 (per-constraint d^2, no offset)
void constraint_dist2(const Scene& S, const double* x, const std::vector<Constraint>& C, std::vector<double>& d2) {
    d2.resize(C.size());
    for (size_t i = 0; i < C.size(); ++i) {
        Local L;
        eval_constraint(S, x, C[i], 1.0, 0, L);
        d2[i] = L.d2;
    }
}

// ----------------------------------------------------------------- CCD
// CCD.h:188-235
static bool pt_ccd_broadphase(V3 p, V3 t0, V3 t1, V3 t2, V3 dp, V3 dt0, V3 dt1, V3 dt2, double dist) {
    V3 pe = p + dp, q0 = t0 + dt0, q1 = t1 + dt1, q2 = t2 + dt2;
    double P[3] = {p.x, p.y, p.z}, PE[3] = {pe.x, pe.y, pe.z};
    V3 T[6] = {t0, t1, t2, q0, q1, q2};
    for (int k = 0; k < 3; ++k) {
        double maxp = std::max(P[k], PE[k]), minp = std::min(P[k], PE[k]);
        double mx = -1e300, mn = 1e300;
        for (auto& t : T) { double v = k == 0 ? t.x : (k == 1 ? t.y : t.z); mx = std::max(mx, v); mn = std::min(mn, v); }
        if (minp - mx > dist || mn - maxp > dist) return false;
    }
    return true;
}
static bool ee_ccd_broadphase(V3 a0, V3 a1, V3 b0, V3 b1, V3 da0, V3 da1, V3 db0, V3 db1, double dist) {
    V3 A[4] = {a0, a1, a0 + da0, a1 + da1}, B[4] = {b0, b1, b0 + db0, b1 + db1};
    for (int k = 0; k < 3; ++k) {
        double mxa = -1e300, mna = 1e300, mxb = -1e300, mnb = 1e300;
        for (int i = 0; i < 4; ++i) {
            double va = k == 0 ? A[i].x : (k == 1 ? A[i].y : A[i].z);
            double vb = k == 0 ? B[i].x : (k == 1 ? B[i].y : B[i].z);
            mxa = std::max(mxa, va); mna = std::min(mna, va); mxb = std::max(mxb, vb); mnb = std::min(mnb, vb);
        }
        if (mna - mxb > dist || mnb - mxa > dist) return false;
    }
    return true;
}
// CCD.h:237-277
static bool pe_ccd_broadphase(V3 p, V3 e0, V3 e1, V3 dp, V3 de0, V3 de1, double dist) {
    V3 pe = p + dp, q0 = e0 + de0, q1 = e1 + de1;
    double P[2][3] = {{p.x, p.y, p.z}, {pe.x, pe.y, pe.z}};
    V3 E[4] = {e0, e1, q0, q1};
    for (int k = 0; k < 3; ++k) {
        double maxp = std::max(P[0][k], P[1][k]), minp = std::min(P[0][k], P[1][k]);
        double mx = -1e300, mn = 1e300;
        for (auto& e : E) { double v = k == 0 ? e.x : (k == 1 ? e.y : e.z); mx = std::max(mx, v); mn = std::min(mn, v); }
        if (minp - mx > dist || mn - maxp > dist) return false;
    }
    return true;
}
static bool pp_ccd_broadphase(V3 p0, V3 p1, V3 dp0, V3 dp1, double dist) {
    V3 q0 = p0 + dp0, q1 = p1 + dp1;
    double A[2][3] = {{p0.x, p0.y, p0.z}, {q0.x, q0.y, q0.z}}, B[2][3] = {{p1.x, p1.y, p1.z}, {q1.x, q1.y, q1.z}};
    for (int k = 0; k < 3; ++k) {
        double mxa = std::max(A[0][k], A[1][k]), mna = std::min(A[0][k], A[1][k]);
        double mxb = std::max(B[0][k], B[1][k]), mnb = std::min(B[0][k], B[1][k]);
        if (mna - mxb > dist || mnb - mxa > dist) return false;
    }
    return true;
}
static const long ACCD_MAX_ITERS = 1000000;

// CCD.h:397-441 (additive CCD, point-edge)
static bool pe_accd(V3 p, V3 e0, V3 e1, V3 dp, V3 de0, V3 de1, double eta, double thickness, double* toc) {
    V3 mov = V3{(dp.x + de0.x + de1.x) / 3, (dp.y + de0.y + de1.y) / 3, (dp.z + de0.z + de1.z) / 3};
    de0 = de0 - mov; de1 = de1 - mov; dp = dp - mov;
    double maxDispMag = std::sqrt(sq(dp)) + std::sqrt(std::max(sq(de0), sq(de1)));
    if (maxDispMag == 0) return false;
    double dist2_cur = dist2_pe_unclassified(p, e0, e1);
    double dist_cur = std::sqrt(dist2_cur);
    double gap = eta * (dist2_cur - thickness * thickness) / (dist_cur + thickness);
    double toc_prev = *toc;
    double t = 0;
    for (long it = 0;; ++it) {
        if (it >= ACCD_MAX_ITERS) { *toc = t; return true; }
        double tocLowerBound = (1 - eta) * (dist2_cur - thickness * thickness) / ((dist_cur + thickness) * maxDispMag);
        p = p + tocLowerBound * dp; e0 = e0 + tocLowerBound * de0; e1 = e1 + tocLowerBound * de1;
        dist2_cur = dist2_pe_unclassified(p, e0, e1);
        dist_cur = std::sqrt(dist2_cur);
        if (t != 0 && (dist2_cur - thickness * thickness) / (dist_cur + thickness) < gap) break;
        t += tocLowerBound;
        if (t > toc_prev) { *toc = t; return false; }
    }
    *toc = t;
    return true;
}
// CCD.h:443-483 (additive CCD, point-point)
static bool pp_accd(V3 p0, V3 p1, V3 dp0, V3 dp1, double eta, double thickness, double* toc) {
    V3 mov = V3{(dp0.x + dp1.x) / 2, (dp0.y + dp1.y) / 2, (dp0.z + dp1.z) / 2};
    dp1 = dp1 - mov; dp0 = dp0 - mov;
    double maxDispMag = std::sqrt(sq(dp0)) + std::sqrt(sq(dp1));
    if (maxDispMag == 0) return false;
    double dist2_cur = dist2_pp(p0, p1);
    double dist_cur = std::sqrt(dist2_cur);
    double gap = eta * (dist2_cur - thickness * thickness) / (dist_cur + thickness);
    double toc_prev = *toc;
    double t = 0;
    for (long it = 0;; ++it) {
        if (it >= ACCD_MAX_ITERS) { *toc = t; return true; }
        double tocLowerBound = (1 - eta) * (dist2_cur - thickness * thickness) / ((dist_cur + thickness) * maxDispMag);
        p0 = p0 + tocLowerBound * dp0; p1 = p1 + tocLowerBound * dp1;
        dist2_cur = dist2_pp(p0, p1);
        dist_cur = std::sqrt(dist2_cur);
        if (t != 0 && (dist2_cur - thickness * thickness) / (dist_cur + thickness) < gap) break;
        t += tocLowerBound;
        if (t > toc_prev) { *toc = t; return false; }
    }
    *toc = t;
    return true;
}

// CCD.h:280-328 (additive CCD, point-triangle); toc in/out as in the reference
static bool pt_accd(V3 p, V3 t0, V3 t1, V3 t2, V3 dp, V3 dt0, V3 dt1, V3 dt2, double eta, double thickness, double* toc) {
    V3 mov = V3{(dt0.x + dt1.x + dt2.x + dp.x) / 4, (dt0.y + dt1.y + dt2.y + dp.y) / 4, (dt0.z + dt1.z + dt2.z + dp.z) / 4};
    dt0 = dt0 - mov; dt1 = dt1 - mov; dt2 = dt2 - mov; dp = dp - mov;
    double m2 = std::max(std::max(sq(dt0), sq(dt1)), sq(dt2));
    double maxDispMag = std::sqrt(sq(dp)) + std::sqrt(m2);
    if (maxDispMag == 0) return false;
    double dist2_cur = dist2_pt_unclassified(p, t0, t1, t2);
    double dist_cur = std::sqrt(dist2_cur);
    double gap = eta * (dist2_cur - thickness * thickness) / (dist_cur + thickness);
    double toc_prev = *toc;
    double t = 0;
    for (long it = 0;; ++it) {
        // guard for touching stencils where the advance is zero (the reference would
        // spin forever); the accumulated t is a valid conservative bound
        if (it >= ACCD_MAX_ITERS) { *toc = t; return true; }
        double tocLowerBound = (1 - eta) * (dist2_cur - thickness * thickness) / ((dist_cur + thickness) * maxDispMag);
        p = p + tocLowerBound * dp; t0 = t0 + tocLowerBound * dt0; t1 = t1 + tocLowerBound * dt1; t2 = t2 + tocLowerBound * dt2;
        dist2_cur = dist2_pt_unclassified(p, t0, t1, t2);
        dist_cur = std::sqrt(dist2_cur);
        if (t != 0 && ((dist2_cur - thickness * thickness) / (dist_cur + thickness) < gap)) break;
        t += tocLowerBound;
        if (t > toc_prev) { *toc = t; return false; }
    }
    *toc = t;
    return true;
}
// CCD.h:331-395
static bool ee_accd(V3 ea0, V3 ea1, V3 eb0, V3 eb1, V3 dea0, V3 dea1, V3 deb0, V3 deb1, double eta, double thickness, double* toc) {
    V3 mov = V3{(dea0.x + dea1.x + deb0.x + deb1.x) / 4, (dea0.y + dea1.y + deb0.y + deb1.y) / 4, (dea0.z + dea1.z + deb0.z + deb1.z) / 4};
    dea0 = dea0 - mov; dea1 = dea1 - mov; deb0 = deb0 - mov; deb1 = deb1 - mov;
    double maxDispMag = std::sqrt(std::max(sq(dea0), sq(dea1))) + std::sqrt(std::max(sq(deb0), sq(deb1)));
    if (maxDispMag == 0) return false;
    double th2 = thickness * thickness;
    auto dist_fallback = [&](double& d2) {
        double dFunc = d2 - th2;
        if (dFunc <= 0) {
            // nearly parallel far-away edges: fall back to the closest endpoint pair
            double l[4] = {sq(ea0 - eb0), sq(ea0 - eb1), sq(ea1 - eb0), sq(ea1 - eb1)};
            double m = l[0]; for (int i = 1; i < 4; ++i) if (l[i] < m) m = l[i];
            d2 = m; dFunc = d2 - th2;
        }
        return dFunc;
    };
    double dist2_cur = dist2_ee_unclassified(ea0, ea1, eb0, eb1);
    double dFunc = dist_fallback(dist2_cur);
    double dist_cur = std::sqrt(dist2_cur);
    double gap = eta * dFunc / (dist_cur + thickness);
    double toc_prev = *toc;
    double t = 0;
    for (long it = 0;; ++it) {
        if (it >= ACCD_MAX_ITERS) { *toc = t; return true; }
        double tocLowerBound = (1 - eta) * dFunc / ((dist_cur + thickness) * maxDispMag);
        ea0 = ea0 + tocLowerBound * dea0; ea1 = ea1 + tocLowerBound * dea1;
        eb0 = eb0 + tocLowerBound * deb0; eb1 = eb1 + tocLowerBound * deb1;
        dist2_cur = dist2_ee_unclassified(ea0, ea1, eb0, eb1);
        dFunc = dist_fallback(dist2_cur);
        dist_cur = std::sqrt(dist2_cur);
        if (t != 0 && (dFunc / (dist_cur + thickness) < gap)) break;
        t += tocLowerBound;
        if (t > toc_prev) { *toc = t; return false; }
    }
    *toc = t;
    return true;
}

// Sum in the fixed order the HIP reduction uses: consecutive chunks of 256 values,
// each chunk folded by halving strides (a[i] += a[i + s], s = 128 ... 1), chunk
// partials added in chunk order.  Used for the two scalars that set the CCD grid
// and the step clamp, so the oracle and the kernels round them identically.
static double chunk_sum(const std::vector<double>& v) {
    double total = 0;
    size_t n = v.size();
    for (size_t b = 0; b < n; b += 256) {
        double a[256];
        for (int t = 0; t < 256; ++t) a[t] = (b + t < n) ? v[b + t] : 0.0;
        for (int s = 128; s > 0; s >>= 1)
            for (int t = 0; t < s; ++t) a[t] += a[t + s];
        total += a[0];
    }
    return total;
}
static double mean_bedge_len(const Scene& S, const double* x) {
    int nBE = (int)S.bEdge.size() / 2;
    if (!nBE) return 0;
    std::vector<double> len(nBE);
    for (int e = 0; e < nBE; ++e) len[e] = std::sqrt(sq(X3(x, S.bEdge[2 * e]) - X3(x, S.bEdge[2 * e + 1])));
    return chunk_sum(len) / nBE;
}

// IPC.h:1879-2244 with the CCD hash of SPATIAL_HASH.h:432-622.  The candidate pairs
// are those whose swept voxel ranges overlap on the reference's voxel grid; the
// voxel grid and the step clamp are reproduced exactly.
double ccd_stepsize(const Scene& S, const double* x, const double* p, double alpha, bool bruteForce, double* clampedAlpha) {
    const int nBN = (int)S.bNode.size(), nBE = (int)S.bEdge.size() / 2, nBT = (int)S.bTri.size() / 3;
    const double xi = S.P.xi;
    const char* dbc = S.isDBC.data();
    // ---- step clamp, SPATIAL_HASH.h:454-482
    double voxelSize = 1.0;
    if (nBE) voxelSize *= mean_bedge_len(S, x);
    std::vector<double> pabs(nBN);
    for (int s = 0; s < nBN; ++s) {
        int v = S.bNode[s];
        pabs[s] = (std::fabs(p[3 * v]) + std::fabs(p[3 * v + 1])) + std::fabs(p[3 * v + 2]);
    }
    double pSize = chunk_sum(pabs) / (nBN * 3);
    double spanSize = alpha * pSize / voxelSize;
    if (spanSize > 1) alpha /= spanSize;
    if (clampedAlpha) *clampedAlpha = alpha;
    // ---- voxel grid, SPATIAL_HASH.h:484-544
    std::vector<int> vmin(3 * nBN), vmax(3 * nBN);
    std::vector<int> v2s(S.nV, -1);
    double lb[3] = {1e300, 1e300, 1e300}, rt[3] = {-1e300, -1e300, -1e300};
    for (int s = 0; s < nBN; ++s) {
        int v = S.bNode[s]; v2s[v] = s;
        for (int d = 0; d < 3; ++d) {
            double a = x[3 * v + d], b = x[3 * v + d] + alpha * p[3 * v + d];
            lb[d] = std::min(lb[d], std::min(a, b)); rt[d] = std::max(rt[d], std::max(a, b));
        }
    }
    for (int d = 0; d < 3; ++d) { lb[d] -= xi / 2; rt[d] += xi / 2; }
    double range[3] = {rt[0] - lb[0], rt[1] - lb[1], rt[2] - lb[2]};
    double inv = 1.0 / voxelSize;
    double vcl = 1;
    for (int d = 0; d < 3; ++d) vcl *= std::max(1.0, std::ceil(range[d] * inv));
    if (vcl > 1e9) { voxelSize *= std::pow(vcl / 1.0e9, 1.0 / 3); inv = 1.0 / voxelSize; }
    for (int s = 0; s < nBN; ++s) {
        int v = S.bNode[s];
        for (int d = 0; d < 3; ++d) {
            double a = x[3 * v + d], b = x[3 * v + d] + alpha * p[3 * v + d];
            double mn = std::min(a, b) - xi / 2, mx = std::max(a, b) + xi / 2;
            vmin[3 * s + d] = (int)std::floor((mn - lb[d]) * inv);
            vmax[3 * s + d] = (int)std::floor((mx - lb[d]) * inv);
        }
    }
    auto prim_range = [&](const int* vs, int n, int* lo, int* hi) {
        for (int d = 0; d < 3; ++d) { lo[d] = 1 << 30; hi[d] = -(1 << 30); }
        for (int i = 0; i < n; ++i) {
            int s = v2s[vs[i]];
            for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], vmin[3 * s + d]); hi[d] = std::max(hi[d], vmax[3 * s + d]); }
        }
    };
    auto overlap = [](const int* alo, const int* ahi, const int* blo, const int* bhi) {
        for (int d = 0; d < 3; ++d) if (alo[d] > bhi[d] || blo[d] > ahi[d]) return false;
        return true;
    };
    std::vector<int> tlo(3 * nBT), thi(3 * nBT), elo(3 * nBE), ehi(3 * nBE);
    for (int t = 0; t < nBT; ++t) prim_range(&S.bTri[3 * t], 3, &tlo[3 * t], &thi[3 * t]);
    for (int e = 0; e < nBE; ++e) prim_range(&S.bEdge[2 * e], 2, &elo[3 * e], &ehi[3 * e]);
    (void)bruteForce;

    // index grid over voxel ranges to enumerate overlapping pairs quickly
    const int nRod = (int)S.rod.size() / 2;
    Grid gT, gE, gP;
    for (Grid* g : {&gT, &gE, &gP}) { g->org[0] = g->org[1] = g->org[2] = 0; g->inv = 1.0 / 4.0; }
    for (int s = S.codimBN1; s < nBN; ++s) {
        double l[3] = {(double)vmin[3 * s], (double)vmin[3 * s + 1], (double)vmin[3 * s + 2]};
        double h[3] = {(double)vmax[3 * s], (double)vmax[3 * s + 1], (double)vmax[3 * s + 2]};
        gP.insert(l, h, s);
    }
    for (int t = 0; t < nBT; ++t) {
        double l[3] = {(double)tlo[3 * t], (double)tlo[3 * t + 1], (double)tlo[3 * t + 2]};
        double h[3] = {(double)thi[3 * t], (double)thi[3 * t + 1], (double)thi[3 * t + 2]};
        gT.insert(l, h, t);
    }
    for (int e = 0; e < nBE; ++e) {
        double l[3] = {(double)elo[3 * e], (double)elo[3 * e + 1], (double)elo[3 * e + 2]};
        double h[3] = {(double)ehi[3 * e], (double)ehi[3 * e + 1], (double)ehi[3 * e + 2]};
        gE.insert(l, h, e);
    }

    // ---- point-triangle, IPC.h:1959-2164
    std::vector<double> lpt(nBN, alpha);
    int zeroToi = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(S.P.threads)
    for (int s = 0; s < nBN; ++s) {
        int vI = S.bNode[s];
        V3 P = X3(x, vI), dP = X3(p, vI);
        std::vector<int> cand;
        double l[3] = {(double)vmin[3 * s], (double)vmin[3 * s + 1], (double)vmin[3 * s + 2]};
        double h[3] = {(double)vmax[3 * s], (double)vmax[3 * s + 1], (double)vmax[3 * s + 2]};
        gT.query(l, h, cand);
        for (int t : cand) {
            if (!overlap(&vmin[3 * s], &vmax[3 * s], &tlo[3 * t], &thi[3 * t])) continue;
            const int* tv = &S.bTri[3 * t];
            if (vI == tv[0] || vI == tv[1] || vI == tv[2]) continue;
            if (dbc[vI] && dbc[tv[0]] && dbc[tv[1]] && dbc[tv[2]]) continue;
            V3 t0 = X3(x, tv[0]), t1 = X3(x, tv[1]), t2 = X3(x, tv[2]);
            V3 d0 = X3(p, tv[0]), d1 = X3(p, tv[1]), d2 = X3(p, tv[2]);
            if (!pt_ccd_broadphase(P, t0, t1, t2, dP, d0, d1, d2, xi)) continue;
            double la = lpt[s];
            if (pt_accd(P, t0, t1, t2, dP, d0, d1, d2, 0.1, xi, &la)) {
                if (lpt[s] > la) lpt[s] = la;
                if (la == 0) {
#pragma omp atomic write
                    zeroToi = 1;
                }
            }
        }
        if (s < S.codimBN1) continue;
        // particle against rod segments, IPC.h:2095-2133
        gE.query(l, h, cand);
        for (int e : cand) {
            if (e < nBE - nRod) continue;
            if (!overlap(&vmin[3 * s], &vmax[3 * s], &elo[3 * e], &ehi[3 * e])) continue;
            const int* ev = &S.bEdge[2 * e];
            if (vI == ev[0] || vI == ev[1]) continue;
            if (dbc[vI] && dbc[ev[0]] && dbc[ev[1]]) continue;
            V3 e0 = X3(x, ev[0]), e1 = X3(x, ev[1]), de0 = X3(p, ev[0]), de1 = X3(p, ev[1]);
            if (!pe_ccd_broadphase(P, e0, e1, dP, de0, de1, xi)) continue;
            double la = lpt[s];
            if (pe_accd(P, e0, e1, dP, de0, de1, 0.1, xi, &la)) {
                if (lpt[s] > la) lpt[s] = la;
            }
        }
        // particle against the particles after it, IPC.h:2135-2162
        gP.query(l, h, cand);
        for (int sJ : cand) {
            if (sJ <= s) continue;
            if (!overlap(&vmin[3 * s], &vmax[3 * s], &vmin[3 * sJ], &vmax[3 * sJ])) continue;
            int vJ = S.bNode[sJ];
            if (dbc[vI] && dbc[vJ]) continue;
            V3 Q = X3(x, vJ), dQ = X3(p, vJ);
            if (!pp_ccd_broadphase(P, Q, dP, dQ, xi)) continue;
            double la = lpt[s];
            if (pp_accd(P, Q, dP, dQ, 0.1, xi, &la)) {
                if (lpt[s] > la) lpt[s] = la;
            }
        }
    }
    for (double a : lpt) alpha = std::min(alpha, a);

    // ---- edge-edge, IPC.h:2166-2242
    std::vector<double> lee(nBE, alpha);
#pragma omp parallel for schedule(dynamic, 64) num_threads(S.P.threads)
    for (int eI = 0; eI < nBE; ++eI) {
        const int* ei = &S.bEdge[2 * eI];
        V3 a0 = X3(x, ei[0]), a1 = X3(x, ei[1]), da0 = X3(p, ei[0]), da1 = X3(p, ei[1]);
        std::vector<int> cand;
        double l[3] = {(double)elo[3 * eI], (double)elo[3 * eI + 1], (double)elo[3 * eI + 2]};
        double h[3] = {(double)ehi[3 * eI], (double)ehi[3 * eI + 1], (double)ehi[3 * eI + 2]};
        gE.query(l, h, cand);
        for (int eJ : cand) {
            if (eJ <= eI) continue;
            if (!overlap(&elo[3 * eI], &ehi[3 * eI], &elo[3 * eJ], &ehi[3 * eJ])) continue;
            const int* ej = &S.bEdge[2 * eJ];
            if (ei[0] == ej[0] || ei[0] == ej[1] || ei[1] == ej[0] || ei[1] == ej[1]) continue;
            if (dbc[ei[0]] && dbc[ei[1]] && dbc[ej[0]] && dbc[ej[1]]) continue;
            V3 b0 = X3(x, ej[0]), b1 = X3(x, ej[1]), db0 = X3(p, ej[0]), db1 = X3(p, ej[1]);
            if (!ee_ccd_broadphase(a0, a1, b0, b1, da0, da1, db0, db1, xi)) continue;
            double la = lee[eI];
            if (ee_accd(a0, a1, b0, b1, da0, da1, db0, db1, 0.1, xi, &la)) {
                if (lee[eI] > la) lee[eI] = la;
            }
        }
    }
    for (double a : lee) alpha = std::min(alpha, a);
    if (zeroToi) return 0.0;
    return alpha;
}

}  // namespace orc
