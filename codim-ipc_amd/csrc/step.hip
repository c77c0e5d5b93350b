// Scene set-up and the implicit-Euler step driven from the host.
//
// ctx_setup restates Initialize_Discrete_Shell<double,3,KL=false>
// (FEM/Shell/DISCRETE_SHELL.h:218-357), surface primitive extraction
// (Utils/MESHIO.h:728-835) and Initialize_EIPC<double,false> (DISCRETE_SHELL.h:555-577).
// advance_one_step restates Advance_One_Step_IE_Discrete_Shell<double,3,false,false>
// (FEM/Shell/IMPLICIT_EULER.h:151-902) and Line_Search (:9-149) with moving Dirichlet
// nodes (the penalty continuation when the CCD blocks their motion), isotropic strain
// limiting (INEXT.h) and lagged
// friction (basis at the step start, re-lagged after convergence up to fricIterAmt
// times, IMPLICIT_EULER.h:412-437, 770-859);
// every array operation runs on the device, the host only sequences the stages and
// reads back the scalars the control flow branches on (energies, step sizes,
// residuals, counts).
#include "ctx.hpp"
#include "dev_math.hpp"
#include <algorithm>
#include <cstdio>
#include <array>
#include <stdexcept>
#include <string>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <set>

namespace cipc {

void build_elastic_plans(Ctx& c);
void elastic_gradient_only(Ctx& c, const double* x, double* g);

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------------ set-up (host)
static inline void hv(const std::vector<double>& X, int i, double* o) { o[0] = X[3 * i]; o[1] = X[3 * i + 1]; o[2] = X[3 * i + 2]; }
static inline double hdot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void hsub(const double* a, const double* b, double* o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static inline void hcross(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0];
}
static double h_dihedral(const double* x0, const double* x1, const double* x2, const double* x3) {
    double a[3], b[3], n1[3], n2[3], t[3];
    hsub(x1, x0, a); hsub(x2, x0, b); hcross(a, b, n1);
    hsub(x2, x3, a); hsub(x1, x3, b); hcross(a, b, n2);
    double cth = hdot(n1, n2) / std::sqrt(hdot(n1, n1) * hdot(n2, n2));
    cth = std::max(-1.0, std::min(1.0, cth));
    double th = std::acos(cth);
    hcross(n2, n1, t);
    hsub(x1, x2, a);
    if (hdot(t, a) < 0) th = -th;
    return th;
}

void ctx_setup(Ctx& c, const std::vector<double>& X, const std::vector<int>& triIn) {
    const Params& P = c.P;
    // degenerate filter, DISCRETE_SHELL.h:228-244
    std::vector<int> tri;
    for (size_t t = 0; t < triIn.size() / 3; ++t) {
        double a[3], b[3], cc[3], e1[3], e2[3];
        hv(X, triIn[3 * t], a); hv(X, triIn[3 * t + 1], b); hv(X, triIn[3 * t + 2], cc);
        hsub(b, a, e1); hsub(cc, a, e2);
        double I00 = hdot(e1, e1), I01 = hdot(e1, e2), I11 = hdot(e2, e2);
        if (I00 * I11 - I01 * I01 != 0) for (int k = 0; k < 3; ++k) tri.push_back(triIn[3 * t + k]);
    }
    c.hTri = tri;
    c.nT = (int)tri.size() / 3;
    // contact triangles: the surface of the tets (Find_Surface_TriMesh, MESHIO.h:486-544:
    // faces keyed as written in a std::map, kept when no rotation of the reverse exists)
    // followed by the shell (IMPLICIT_EULER.h:235-241)
    c.nTet = (int)c.hTet.size() / 4;
    c.hBTri.clear();
    {
        std::map<std::array<int, 3>, int> faces;
        for (int t = 0; t < c.nTet; ++t) {
            const int* v = &c.hTet[4 * t];
            faces[{v[0], v[2], v[1]}] = t; faces[{v[0], v[3], v[2]}] = t;
            faces[{v[0], v[1], v[3]}] = t; faces[{v[1], v[2], v[3]}] = t;
        }
        for (const auto& f : faces) {
            const auto& k = f.first;
            if (faces.count({k[2], k[1], k[0]}) || faces.count({k[1], k[0], k[2]}) || faces.count({k[0], k[2], k[1]})) continue;
            c.hBTri.insert(c.hBTri.end(), k.begin(), k.end());
        }
    }
    c.hBTri.insert(c.hBTri.end(), tri.begin(), tri.end());
    c.nBT = (int)c.hBTri.size() / 3;
    // edge2tri + primitives, MESHIO.h:728-766 (hinges from the shell's edges, primitives
    // from every contact triangle)
    std::map<std::pair<int, int>, int> edge2tri;
    std::set<std::pair<int, int>> edgeSet, primEdgeSet;
    std::vector<char> isBN(c.nV, 0);
    for (int t = 0; t < c.nT; ++t) {
        const int* v = &tri[3 * t];
        edge2tri[{v[0], v[1]}] = t; edge2tri[{v[1], v[2]}] = t; edge2tri[{v[2], v[0]}] = t;
    }
    auto add_edges = [](std::set<std::pair<int, int>>& es, const int* v) {
        if (!es.count({v[1], v[0]})) es.insert({v[0], v[1]});
        if (!es.count({v[2], v[1]})) es.insert({v[1], v[2]});
        if (!es.count({v[0], v[2]})) es.insert({v[2], v[0]});
    };
    for (int t = 0; t < c.nT; ++t) add_edges(edgeSet, &tri[3 * t]);
    for (int t = 0; t < c.nBT; ++t) {
        const int* v = &c.hBTri[3 * t];
        add_edges(primEdgeSet, v);
        isBN[v[0]] = isBN[v[1]] = isBN[v[2]] = 1;
    }
    c.hBNode.clear(); c.hBEdge.clear();
    for (int i = 0; i < c.nV; ++i) if (isBN[i]) c.hBNode.push_back(i);
    for (auto& e : primEdgeSet) { c.hBEdge.push_back(e.first); c.hBEdge.push_back(e.second); }
    c.nBN = (int)c.hBNode.size();
    c.nBE = (int)c.hBEdge.size() / 2;
    // inverse rest metric, hinges, DISCRETE_SHELL.h:132-213
    c.hBinv.resize(3 * c.nT);
    for (int t = 0; t < c.nT; ++t) {
        double a[3], b[3], cc[3], e1[3], e2[3];
        hv(X, tri[3 * t], a); hv(X, tri[3 * t + 1], b); hv(X, tri[3 * t + 2], cc);
        hsub(b, a, e1); hsub(cc, a, e2);
        double I00 = hdot(e1, e1), I01 = hdot(e1, e2), I11 = hdot(e2, e2);
        double det = I00 * I11 - I01 * I01;
        c.hBinv[3 * t] = I11 / det; c.hBinv[3 * t + 1] = -I01 / det; c.hBinv[3 * t + 2] = I00 / det;
    }
    // strain limiting: inverse of the upper-triangular factor B of the rest metric
    // (B^T B = I_rest, INEXT.h:145-151): (B^-1)_00, (B^-1)_01, (B^-1)_11
    c.hSlIB.resize(3 * c.nT);
    for (int t = 0; t < c.nT; ++t) {
        double a[3], b[3], cc[3], e1[3], e2[3];
        hv(X, tri[3 * t], a); hv(X, tri[3 * t + 1], b); hv(X, tri[3 * t + 2], cc);
        hsub(b, a, e1); hsub(cc, a, e2);
        double I00 = hdot(e1, e1), I01 = hdot(e1, e2), I11 = hdot(e2, e2);
        double B00 = std::sqrt(I00), B01 = I01 / B00, B11 = std::sqrt(I11 - I01 * I01 / I00);
        double inv = 1.0 / (B00 * B11);
        c.hSlIB[3 * t] = B11 * inv; c.hSlIB[3 * t + 1] = -B01 * inv; c.hSlIB[3 * t + 2] = B00 * inv;
    }
    // fibre model: inverse of the rest x-z coordinate matrix (ANISO_INEXT.h:64-68)
    c.hFbIB.resize(4 * c.nT);
    for (int t = 0; t < c.nT; ++t) {
        const int* v = &tri[3 * t];
        const double b00 = X[3 * v[1]] - X[3 * v[0]], b01 = X[3 * v[2]] - X[3 * v[0]];
        const double b10 = X[3 * v[1] + 2] - X[3 * v[0] + 2], b11 = X[3 * v[2] + 2] - X[3 * v[0] + 2];
        const double inv = 1.0 / (b00 * b11 - b01 * b10);
        c.hFbIB[4 * t] = b11 * inv; c.hFbIB[4 * t + 1] = -b01 * inv; c.hFbIB[4 * t + 2] = -b10 * inv; c.hFbIB[4 * t + 3] = b00 * inv;
    }
    c.hHinge.clear(); c.hHingeInfo.clear();
    for (auto& e : edgeSet) {
        auto fa = edge2tri.find({e.first, e.second});
        if (fa == edge2tri.end()) continue;
        const int* A = &tri[3 * fa->second];
        int v0 = -1;
        for (int j = 0; j < 3; ++j) if (A[j] == e.second) { v0 = A[(j + 1) % 3]; break; }
        auto fb = edge2tri.find({e.second, e.first});
        if (fb == edge2tri.end()) continue;
        const int* B = &tri[3 * fb->second];
        int v3 = B[0];
        for (int j = 0; j < 3; ++j) if (B[j] == e.first) { v3 = B[(j + 1) % 3]; break; }
        int st[4] = {v0, e.first, e.second, v3};
        for (int k = 0; k < 4; ++k) c.hHinge.push_back(st[k]);
        double p0[3], p1[3], p2[3], p3[3], a[3], b[3], n1[3], n2[3];
        hv(X, st[0], p0); hv(X, st[1], p1); hv(X, st[2], p2); hv(X, st[3], p3);
        double th = h_dihedral(p0, p1, p2, p3);
        hsub(p1, p2, a);
        double el = std::sqrt(hdot(a, a));
        hsub(p1, p0, a); hsub(p2, p0, b); hcross(a, b, n1);
        hsub(p2, p3, a); hsub(p1, p3, b); hcross(a, b, n2);
        c.hHingeInfo.push_back(th);
        c.hHingeInfo.push_back(el);
        c.hHingeInfo.push_back((std::sqrt(hdot(n1, n1)) + std::sqrt(hdot(n2, n2))) / (el * 6));
    }
    c.nH = (int)c.hHinge.size() / 4;
    // lumped mass, DISCRETE_SHELL.h:279-299
    c.hMass.assign(c.nV, 0.0);
    c.hVol.resize(c.nT);
    for (int t = 0; t < c.nT; ++t) {
        const int* v = &tri[3 * t];
        double a[3], b[3], cc[3], e1[3], e2[3], n[3];
        hv(X, v[0], a); hv(X, v[1], b); hv(X, v[2], cc);
        hsub(b, a, e1); hsub(cc, a, e2); hcross(e1, e2, n);
        double area = std::sqrt(hdot(n, n)) / 2;
        double mp = area * P.thick * P.rho / 3;
        c.hMass[v[0]] += mp; c.hMass[v[1]] += mp; c.hMass[v[2]] += mp;
        c.hVol[t] = area * P.thick;
    }
    // tets: rest basis Dm^-1, volume det(Dm) / 6, lumped mass vol rho / 4 per node
    // (FEM/DEFORMATION_GRADIENT.h:39-77, FEM/ELEM_TO_NODE.h:11-46), Lame parameters
    // (Physics/NEOHOOKEAN.h:46-47)
    std::vector<double> tetIB(9 * (size_t)c.nTet), tetVol(c.nTet);
    for (int t = 0; t < c.nTet; ++t) {
        const int* v = &c.hTet[4 * t];
        double p0[3], q[3][3];
        hv(X, v[0], p0);
        for (int k = 0; k < 3; ++k) { double pk[3]; hv(X, v[k + 1], pk); hsub(pk, p0, q[k]); }
        // Dm row-major: column k = q[k]
        const double M[9] = {q[0][0], q[1][0], q[2][0], q[0][1], q[1][1], q[2][1], q[0][2], q[1][2], q[2][2]};
        const double det = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) + M[2] * (M[3] * M[7] - M[4] * M[6]);
        if (!(det > 0)) throw std::invalid_argument("tet " + std::to_string(t) + " is inverted or degenerate in the rest shape");
        double* I = &tetIB[9 * (size_t)t];
        I[0] = (M[4] * M[8] - M[5] * M[7]) / det; I[1] = (M[2] * M[7] - M[1] * M[8]) / det; I[2] = (M[1] * M[5] - M[2] * M[4]) / det;
        I[3] = (M[5] * M[6] - M[3] * M[8]) / det; I[4] = (M[0] * M[8] - M[2] * M[6]) / det; I[5] = (M[2] * M[3] - M[0] * M[5]) / det;
        I[6] = (M[3] * M[7] - M[4] * M[6]) / det; I[7] = (M[1] * M[6] - M[0] * M[7]) / det; I[8] = (M[0] * M[4] - M[1] * M[3]) / det;
        tetVol[t] = det / 6.0;
        const double mt = tetVol[t] * P.tetRho / 4;
        for (int k = 0; k < 4; ++k) c.hMass[v[k]] += mt;
    }
    c.tetLam = P.tetE * P.tetNu / ((1.0 + P.tetNu) * (1.0 - 2.0 * P.tetNu));
    c.tetMu = P.tetE / (2.0 * (1.0 + P.tetNu));
    c.lam = P.E * P.nu / (1.0 - P.nu * P.nu);
    c.mu = P.E / (2.0 * (1.0 + P.nu));
    c.kBend = P.bendMult * (P.E * std::pow(P.thick, 3) / (24 * (1.0 - P.nu * P.nu)));
    // rods, DISCRETE_ROD.h:87-165: per-segment E, rest length, thickness and lumped mass;
    // a bending hinge per pair of segments sharing a node (std::map node order)
    c.nRod = (int)c.hRod.size() / 2;
    c.hRodInfo.assign(3 * (size_t)c.nRod, 0.0);
    std::map<int, std::vector<int>> nbr;
    for (int r = 0; r < c.nRod; ++r) {
        int a = c.hRod[2 * r], b = c.hRod[2 * r + 1];
        nbr[a].push_back(b);
        nbr[b].push_back(a);
        double pa[3], pb[3], d[3];
        hv(X, a, pa); hv(X, b, pb); hsub(pa, pb, d);
        double l = std::sqrt(hdot(d, d));
        c.hRodInfo[3 * r] = P.rodE;
        c.hRodInfo[3 * r + 1] = l;
        c.hRodInfo[3 * r + 2] = P.rodThick;
        double mp = l * M_PI * P.rodThick * P.rodThick / 4 * P.rodRho / 2;
        c.hMass[a] += mp;
        c.hMass[b] += mp;
    }
    c.hRodHinge.clear(); c.hRodHingeInfo.clear();
    for (auto& kv : nbr) {
        const std::vector<int>& n = kv.second;
        for (size_t i = 0; i < n.size(); ++i)
            for (size_t j = i + 1; j < n.size(); ++j) {
                c.hRodHinge.push_back(n[i]); c.hRodHinge.push_back(kv.first); c.hRodHinge.push_back(n[j]);
                double p0[3], p1[3], p2[3], d0[3], d1[3];
                hv(X, n[i], p0); hv(X, kv.first, p1); hv(X, n[j], p2);
                hsub(p1, p0, d0); hsub(p2, p1, d1);
                c.hRodHingeInfo.push_back(P.rodE * P.rodBendMult);
                c.hRodHingeInfo.push_back(std::sqrt(hdot(d0, d0)) + std::sqrt(hdot(d1, d1)));
                c.hRodHingeInfo.push_back(P.rodThick);
            }
    }
    c.nRH = (int)c.hRodHinge.size() / 3;
    // particles, DISCRETE_PARTICLE.h:52-75
    c.nPart = (int)c.hPart.size();
    {
        const double pm = 4.0 / 3.0 * M_PI * P.partThick * P.partThick * P.partThick / 8 * P.partRho;
        for (int v : c.hPart) c.hMass[v] += pm;
    }
    // contact primitives, IMPLICIT_EULER.h:243-277: rod segments after the surface edges,
    // rod nodes (ascending) then particles after the surface nodes
    for (int r = 0; r < c.nRod; ++r) { c.hBEdge.push_back(c.hRod[2 * r]); c.hBEdge.push_back(c.hRod[2 * r + 1]); }
    c.codimBN0 = (int)c.hBNode.size();
    for (auto& kv : nbr) c.hBNode.push_back(kv.first);
    c.codimBN1 = (int)c.hBNode.size();
    for (int v : c.hPart) c.hBNode.push_back(v);
    c.nBN = (int)c.hBNode.size();
    c.nBE = (int)c.hBEdge.size() / 2;
    // Initialize_EIPC<double,false>, DISCRETE_SHELL.h:563-574
    c.dHat2 = P.dHat * P.dHat;
    double t2 = 1.0e-16 - c.dHat2;
    double Hb = 1.0 * ((std::log(1.0e-16 / c.dHat2) * -2.0 - t2 * 4.0 / 1.0e-16) + 1.0 / (1.0e-16 * 1.0e-16) * (t2 * t2));
    double mmean = 0;
    for (int i = 0; i < c.nV; ++i) mmean += 3 * c.hMass[i];
    mmean /= 3 * c.nV;
    c.kappa = P.kappaStiffMult * 1.0e11 * mmean * 3 / (4.0e-16 * Hb);
    c.kappaMax = 100 * c.kappa;
    // broad-phase cell size: mean rest boundary-edge length
    double sum = 0;
    for (int e = 0; e < c.nBE; ++e) {
        double a[3], b[3], d[3];
        hv(X, c.hBEdge[2 * e], a); hv(X, c.hBEdge[2 * e + 1], b); hsub(a, b, d);
        sum += std::sqrt(hdot(d, d));
    }
    c.cellSize = c.nBE ? sum / c.nBE : 1.0;
    if (!(c.cellSize > 0)) c.cellSize = 1.0;

    // ---- uploads (SoA element attributes)
    auto soa_i = [](const std::vector<int>& a, int n, int k) {
        std::vector<int> o((size_t)n * k);
        for (int i = 0; i < n; ++i) for (int j = 0; j < k; ++j) o[(size_t)j * n + i] = a[(size_t)k * i + j];
        return o;
    };
    auto soa_d = [](const std::vector<double>& a, int n, int k) {
        std::vector<double> o((size_t)n * k);
        for (int i = 0; i < n; ++i) for (int j = 0; j < k; ++j) o[(size_t)j * n + i] = a[(size_t)k * i + j];
        return o;
    };
    hipStream_t s = c.s;
    upload(c.X0, X, s);
    upload(c.mass, c.hMass, s);
    std::vector<uint8_t> dbc(c.nV, 0);
    for (int d : c.hDbcIdx) dbc[d] = 1;
    upload(c.dbc, dbc, s);
    upload(c.tri, soa_i(tri, c.nT, 3), s);
    upload(c.Binv, soa_d(c.hBinv, c.nT, 3), s);
    upload(c.slIB, soa_d(c.hSlIB, c.nT, 3), s);
    upload(c.fbIB, soa_d(c.hFbIB, c.nT, 4), s);
    upload(c.vol, c.hVol, s);
    upload(c.hinge, soa_i(c.hHinge, c.nH, 4), s);
    upload(c.hInfo, soa_d(c.hHingeInfo, c.nH, 3), s);
    upload(c.bNode, c.hBNode, s);
    upload(c.btri, soa_i(c.hBTri, c.nBT, 3), s);
    if (c.nTet) {
        upload(c.tetIdx, soa_i(c.hTet, c.nTet, 4), s);
        upload(c.tetIB, soa_d(tetIB, c.nTet, 9), s);
        upload(c.tetVol, tetVol, s);
    }
    upload(c.bEdge, soa_i(c.hBEdge, c.nBE, 2), s);
    if (c.nRod) {
        // spring weight h^2 pi t^2 / 4 (ROD_STRETCHING.h:22-24), bending weight
        // h^2 / lsum * E t^4 pi / 64 (ROD_BENDING.h:55-56)
        std::vector<double> par(3 * (size_t)c.nRod);
        for (int r = 0; r < c.nRod; ++r) {
            double t = c.hRodInfo[3 * r + 2];
            par[r] = c.hRodInfo[3 * r];
            par[(size_t)c.nRod + r] = c.hRodInfo[3 * r + 1];
            par[2 * (size_t)c.nRod + r] = P.dt * P.dt * M_PI * t * t / 4;
        }
        upload(c.rod, soa_i(c.hRod, c.nRod, 2), s);
        upload(c.rodPar, par, s);
    }
    if (c.nRH) {
        std::vector<double> w(c.nRH);
        for (int h = 0; h < c.nRH; ++h) {
            const double* I = &c.hRodHingeInfo[3 * h];
            w[h] = P.dt * P.dt / I[1] * I[0] * std::pow(I[2], 4) * M_PI / 64;
        }
        upload(c.rodHinge, soa_i(c.hRodHinge, c.nRH, 3), s);
        upload(c.rodHW, w, s);
    }
    if (c.nPart) upload(c.partSlot, c.hPart, s);
    const size_t n3 = 3 * (size_t)c.nV;
    for (DevBuf<double>* b : {&c.x, &c.xn, &c.xt, &c.v, &c.xprev, &c.p, &c.g, &c.rhs}) b->ensure_exact(n3);
    c.scal.ensure_exact(64);
    c.counters.ensure_exact(64);
    CIPC_HIP(hipMemsetAsync(c.counters.p, 0, 64 * sizeof(uint32_t), s));
    build_elastic_plans(c);
}

// ------------------------------------------------------------------ vector kernels
__global__ void k_predict(int nV, const double* __restrict__ x, const double* __restrict__ v, const double* __restrict__ m, double h,
                          double gx, double gy, double gz, double* __restrict__ xt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nV) return;
    double g[3] = {gx, gy, gz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double a = m[i] > 0 ? g[k] : 0.0;
        xt[3 * i + k] = x[3 * i + k] + h * v[3 * i + k] + h * h * a;
    }
}
__global__ void k_dbc_disp(int n, const int* __restrict__ idx, const double* __restrict__ tgt, const double* __restrict__ x,
                           double* __restrict__ disp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = idx[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) disp[3 * v + k] = tgt[3 * i + k] - x[3 * v + k];
}
__global__ void k_dbc_set(int n, const int* __restrict__ idx, const double* __restrict__ tgt, double* __restrict__ x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = idx[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) x[3 * v + k] = tgt[3 * i + k];
}
// Dirichlet nodes moved the fraction alpha of the way from x_n to their targets
__global__ void k_dbc_blend(int n, const int* __restrict__ idx, const double* __restrict__ tgt, const double* __restrict__ xn,
                            double alpha, double* __restrict__ x) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = idx[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) x[3 * v + k] = xn[3 * v + k] + alpha * (tgt[3 * i + k] - xn[3 * v + k]);
}
// rhs = -g with Dirichlet rows zeroed (IMPLICIT_EULER.h:476-491)
__global__ void k_rhs(int nV, double* __restrict__ g, const uint8_t* __restrict__ dbc, double* __restrict__ rhs) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nV) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double gv = dbc[i] ? 0.0 : g[3 * i + k];
        g[3 * i + k] = gv;
        rhs[3 * i + k] = -gv;
    }
}
__global__ void k_scale_to(long n, double s, const double* __restrict__ a, double* __restrict__ out) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = s * a[i];
}
__global__ void k_axpy_to(long n, const double* __restrict__ a, double s, const double* __restrict__ b, double* __restrict__ out) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] + s * b[i];
}
__global__ void k_sumsq(long n, const double* __restrict__ a, double* __restrict__ partial) {
    __shared__ double lds[4];
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double v = i < n ? a[i] * a[i] : 0.0;
    double s = block_sum_256(v, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}
__global__ void k_velocity(long n, const double* __restrict__ x, const double* __restrict__ xn, double h, double* __restrict__ v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (x[i] - xn[i]) / h;
}
__global__ void k_kappa_flag(int n, const double* __restrict__ d2prev, const double* __restrict__ d2cur, uint32_t* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (d2prev[i] < 1e-18 && d2cur[i] < d2prev[i]) atomicOr(flag, 1u);
}

// residual.txt statistics of a Newton direction: per workgroup the sum of the nodes'
// |p_i| and the largest |p component| (IMPLICIT_EULER.h:648-663)
__global__ void k_res_stats(int nV, const double* __restrict__ p, double* __restrict__ part) {
    __shared__ double lds[4];
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    double mag = 0, mx = 0;
    if (v < nV) {
        const double a = p[3 * v], b = p[3 * v + 1], c = p[3 * v + 2];
        mag = sqrt(a * a + b * b + c * c);
        mx = fmax(fabs(a), fmax(fabs(b), fabs(c)));
    }
    const double s = block_sum_256(mag, lds);
    double m = mx;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    __shared__ double lm[4];
    if ((threadIdx.x & 63) == 0) lm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s;
        part[2 * blockIdx.x + 1] = fmax(fmax(lm[0], lm[1]), fmax(lm[2], lm[3]));
    }
}
// stretch.txt statistics (Compute_Max_And_Avg_Stretch, DISCRETE_SHELL.h:602-674): per
// triangle not all Dirichlet, F = A B^-1 with A the deformed edge frame
// [[|e1|, e1.e2/|e1|], [0, |e1 x e2|/|e1|]] and B the factor of the rest metric; per
// workgroup: max sigma_max, sum and count of sigma_max > 1, min sigma_min, sum and count
// of sigma_min < 1
__global__ void k_stretch_stats(const double* __restrict__ x, const int* __restrict__ tri, int nT, const double* __restrict__ IB,
                                const uint8_t* __restrict__ dbc, double* __restrict__ part) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double smax = 1.0, ssum = 0, scnt = 0, cmin = 1.0, csum = 0, ccnt = 0;
    if (t < nT) {
        const int a = tri[t], b = tri[nT + t], c = tri[2 * nT + t];
        if (!(dbc[a] && dbc[b] && dbc[c])) {
            const d3 e1 = ld3(x, b) - ld3(x, a), e2 = ld3(x, c) - ld3(x, a);
            const double l = sqrt(sq(e1));
            const double A00 = l, A01 = dot(e1, e2) / l, A11 = sqrt(sq(cross(e1, e2))) / l;
            const double i00 = IB[t], i01 = IB[nT + t], i11 = IB[2 * nT + t];   // B^-1 (upper triangular)
            const double F00 = A00 * i00, F01 = A00 * i01 + A01 * i11, F11 = A11 * i11;
            const double p = F00 * F00, q = F00 * F01, r = F01 * F01 + F11 * F11;   // F^T F
            const double hm = 0.5 * (p - r), rt = sqrt(hm * hm + q * q);
            const double s0 = sqrt(0.5 * (p + r) + rt), s1 = sqrt(fmax(0.5 * (p + r) - rt, 0.0));
            smax = s0;
            if (s0 > 1) { ssum = s0; scnt = 1; }
            cmin = s1;
            if (s1 < 1) { csum = s1; ccnt = 1; }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        smax = fmax(smax, __shfl_xor(smax, o, 64));
        cmin = fmin(cmin, __shfl_xor(cmin, o, 64));
        ssum += __shfl_xor(ssum, o, 64); scnt += __shfl_xor(scnt, o, 64);
        csum += __shfl_xor(csum, o, 64); ccnt += __shfl_xor(ccnt, o, 64);
    }
    __shared__ double L[4][6];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { L[w][0] = smax; L[w][1] = ssum; L[w][2] = scnt; L[w][3] = cmin; L[w][4] = csum; L[w][5] = ccnt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double o6[6] = {L[0][0], L[0][1], L[0][2], L[0][3], L[0][4], L[0][5]};
        for (int k = 1; k < 4; ++k) {
            o6[0] = fmax(o6[0], L[k][0]); o6[1] += L[k][1]; o6[2] += L[k][2];
            o6[3] = fmin(o6[3], L[k][3]); o6[4] += L[k][4]; o6[5] += L[k][5];
        }
        for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = o6[k];
    }
}

static void append_line(const Ctx& c, const char* file, const char* line) {
    const std::string path = c.outputFolder + "/" + file;
    FILE* f = std::fopen(path.c_str(), "a+");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::fputs(line, f);
    std::fclose(f);
}
// residual.txt line of a Newton iteration: PNIter avgResMag maxRes Eprev L2Norm
static void log_residual(Ctx& c, int pnIter, double E, double L2, double hres) {
    const int nb = nblocks(c.nV);
    double* part = c.red.ensure(2 * (size_t)nb);
    hipLaunchKernelGGL(k_res_stats, dim3(nb), dim3(kBlock), 0, c.s, c.nV, c.p.p, part);
    CIPC_LAUNCH_CHECK();
    std::vector<double> h(2 * (size_t)nb);
    CIPC_HIP(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    double sum = 0, mx = 0;
    for (int b = 0; b < nb; ++b) { sum += h[2 * b]; mx = std::max(mx, h[2 * b + 1]); }
    char line[256];
    std::snprintf(line, sizeof(line), "%d %le %le %le %le\n", pnIter, sum / (c.nV - c.nDBC) / hres, mx / hres, E, L2);
    append_line(c, "residual.txt", line);
}
// counter.txt (PNIter, contacts) and stretch.txt (maxs avgs minc avgc) at the step's end
static void log_step_end(Ctx& c, int pnIter) {
    char line[256];
    if (c.P.collision) std::snprintf(line, sizeof(line), "%d %d\n", pnIter, c.nC);
    else std::snprintf(line, sizeof(line), "%d\n", pnIter);
    append_line(c, "counter.txt", line);
    if (!c.nT) return;
    const int nb = nblocks(c.nT);
    double* part = c.red.ensure(6 * (size_t)nb);
    hipLaunchKernelGGL(k_stretch_stats, dim3(nb), dim3(kBlock), 0, c.s, c.x.p, c.tri.p, c.nT, c.slIB.p, c.dbc.p, part);
    CIPC_LAUNCH_CHECK();
    std::vector<double> h(6 * (size_t)nb);
    CIPC_HIP(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    double maxs = 1, ssum = 0, scnt = 0, minc = 1, csum = 0, ccnt = 0;
    for (int b = 0; b < nb; ++b) {
        const double* o = &h[6 * (size_t)b];
        maxs = std::max(maxs, o[0]); ssum += o[1]; scnt += o[2];
        minc = std::min(minc, o[3]); csum += o[4]; ccnt += o[5];
    }
    std::snprintf(line, sizeof(line), "%le %le %le %le\n", maxs, scnt ? ssum / scnt : 0.0, minc, ccnt ? csum / ccnt : 0.0);
    append_line(c, "stretch.txt", line);
}

static double sumsq(Ctx& c, const double* a, long n) {
    int nb = nblocks(n);
    double* part = c.red.ensure(nb);
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(kBlock), 0, c.s, n, a, part);
    CIPC_LAUNCH_CHECK();
    sum_partials(part, nb, c.scal.p + 6, c.s);
    CIPC_HIP(hipMemcpyAsync(c.hScal + 6, c.scal.p + 6, sizeof(double), hipMemcpyDeviceToHost, c.s));
    CIPC_HIP(hipStreamSynchronize(c.s));
    return c.hScal[6];
}

// incremental potential at x with constraint set C (INC_POTENTIAL.h:15-152)
// +inf when a strain limit is reached (INC_POTENTIAL.h:75-80: the potential is invalid)
static double inc_potential(Ctx& c, const double* x, const int* C, int nC, int* bad) {
    if (c.nTet && !tets_valid(c, x)) { if (bad) *bad = 0; return INFINITY; }   // an inverted tet, INC_POTENTIAL.h:98-107
    double Esl = strain_limit_energy(c, x);
    if (fiber_on(c)) Esl += fiber_energy(c, x);   // Check_Fiber_Feasibility + Compute_Fiber_Energy
    if (std::isinf(Esl)) { if (bad) *bad = 0; return Esl; }
    double E = Esl + elastic_energy(c, x, nullptr, nullptr) + inertia_energy(c, x);
    if (c.P.collision) E += barrier_energy(c, x, C, nC, c.kappa, bad);
    else if (bad) *bad = 0;
    if (c.fricOn) E += friction_energy(c, x, c.xn.p);   // IMPLICIT_EULER.h:436
    if (c.dbcStiff) E += dbc_penalty_energy(c, x);      // IMPLICIT_EULER.h:438-440
    return E;
}

// ------------------------------------------------------------------ the step
int advance_one_step(Ctx& c) {
    struct PenaltyOff {   // the Dirichlet penalty never outlives a step (early returns included)
        Ctx& c;
        ~PenaltyOff() { c.dbcStiff = 0; }
    } penaltyOff{c};
    const Params& P = c.P;
    const long n = 3l * c.nV;
    const double h = P.dt;
    cipc_step_stats& st = c.stats;
    st = cipc_step_stats{};
    c.lpStepStart = true;
    c.cg.stepStart = true;
    c.lpOffThisStep = false;
    c.fp64ThisStep = false;
    c.kappaSDoublings = 0;
    double t0 = now_ms(), ta;
    hipStream_t s = c.s;
    static_assert(sizeof(int) == 4, "");
    // x_n, x_tilde (IMPLICIT_EULER.h:195-217)
    const bool stat = P.staticSolve != 0;
    if (stat && P.collision && (P.fricMu > 0 || c.nFricComp > 0)) {
        c.lastError = "a static solve has no previous state for lagged friction (the reference leaves Xn unset)";
        return CIPC_E_UNSUPPORTED;
    }
    const double hres = stat ? 1.0 : h;   // residual scale, IMPLICIT_EULER.h:662-665
    CIPC_HIP(hipMemcpyAsync(c.xn.p, c.x.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (!stat) {
        hipLaunchKernelGGL(k_predict, dim3(nblocks(c.nV)), dim3(kBlock), 0, s, c.nV, c.x.p, c.v.p, c.mass.p, h, P.grav[0], P.grav[1],
                           P.grav[2], c.xt.p);
        CIPC_LAUNCH_CHECK();
    }
    // Dirichlet motion and its CCD (IMPLICIT_EULER.h:304-401): move the nodes when the CCD
    // allows the whole displacement, else keep them and pull them with the penalty
    c.dbcStiff = 0;
    double dbcPenXn = 0;
    if (c.nDBC) {
        dirichlet_device_init(c);
        CIPC_HIP(hipMemsetAsync(c.p.p, 0, n * sizeof(double), s));
        hipLaunchKernelGGL(k_dbc_disp, dim3(nblocks(c.nDBC)), dim3(kBlock), 0, s, c.nDBC, c.dbcIdxD.p, c.dbcTgt.p, c.x.p, c.p.p);
        CIPC_LAUNCH_CHECK();
        dbc_fixed_mask(c, c.x.p);
        double dbcAlpha = 1;
        if (P.collision) {
            int z = 0;
            dbcAlpha = ccd_stepsize(c, c.x.p, c.p.p, 1.0, nullptr, &z);
        }
        // a motion that would break a strain limit or the fibre limits is cut back by
        // halving until the moved nodes leave every triangle feasible (IMPLICIT_EULER.h:
        // 336-376); only dbcAlpha survives, the nodes stay at x_n here
        if ((strain_limit_on(c) || (c.P.fiberK[0] > 0 && c.nT) || c.nTet)) {
            for (int it = 0; it < 64; ++it) {
                hipLaunchKernelGGL(k_dbc_blend, dim3(nblocks(c.nDBC)), dim3(kBlock), 0, s, c.nDBC, c.dbcIdxD.p, c.dbcTgt.p, c.xn.p,
                                   dbcAlpha, c.x.p);
                CIPC_LAUNCH_CHECK();
                double Ef = strain_limit_energy(c, c.x.p);
                if (fiber_on(c)) Ef += fiber_energy(c, c.x.p);
                if (c.nTet && !tets_valid(c, c.x.p)) Ef = INFINITY;   // IMPLICIT_EULER.h:364-374
                if (!std::isinf(Ef)) break;
                dbcAlpha /= 2;
            }
            CIPC_HIP(hipMemcpyAsync(c.x.p, c.xn.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        if (dbcAlpha == 1) {
            hipLaunchKernelGGL(k_dbc_set, dim3(nblocks(c.nDBC)), dim3(kBlock), 0, s, c.nDBC, c.dbcIdxD.p, c.dbcTgt.p, c.x.p);
            CIPC_LAUNCH_CHECK();
        } else {
            c.dbcStiff = 1e6;
            dbcPenXn = dbc_dist2(c, c.xn.p);
        }
    }
    c.nC = 0;
    c.nCprev = 0;
    if (P.collision) { ta = now_ms(); constraint_set(c, c.x.p, c.C, c.nC); st.t_constraint_set_ms += now_ms() - ta; }
    // lagged friction stencils at the step start (IMPLICIT_EULER.h:419-427)
    c.fricOn = P.collision && (P.fricMu > 0 || c.nFricComp > 0);
    c.nF = 0;
    if (c.fricOn) friction_basis(c, c.x.p, c.C.p, c.nC, c.kappa);
    int bad = 0;
    double Eprev = inc_potential(c, c.x.p, c.C.p, c.nC, &bad);
    if (bad) { c.lastError = "non-positive distance in the initial state"; return CIPC_E_PENETRATION; }
    if (std::isinf(Eprev)) { c.lastError = "strain limit violated by the initial state"; return CIPC_E_UNSUPPORTED; }
    std::deque<double> resRecord;
    bool useGD = false;
    double L2Norm = 0;
    int PNIter = 0;
    bool havePrev = false;     // c.p holds the last Newton direction, accepted at alphaPrev
    double alphaPrev = 1.0;
    int fricIterI = 0;
    bool reuseDir = false;     // c.p already holds this iterate's direction (friction re-lag)
    // gradient, Hessian and Newton direction into c.p; returns false when the solve
    // found the matrix not positive definite
    auto newton_direction = [&]() -> bool {
        // gradient (INC_POTENTIAL.h:154-275)
        ta = now_ms();
        elastic_gradient(c, c.x.p, c.g.p);
        // rows that k_rhs projects out need no barrier / friction terms
        if (P.collision) barrier_gradient(c, c.x.p, c.C.p, c.nC, c.kappa, c.g.p, proj_mask(c));
        if (c.fricOn) friction_gradient(c, c.x.p, c.xn.p, c.g.p, proj_mask(c));
        dbc_penalty_grad(c, c.x.p, c.g.p);   // IMPLICIT_EULER.h:467-475
        hipLaunchKernelGGL(k_rhs, dim3(nblocks(c.nV)), dim3(kBlock), 0, s, c.nV, c.g.p, proj_mask(c), c.rhs.p);
        CIPC_LAUNCH_CHECK();
        CIPC_HIP(hipStreamSynchronize(s));
        st.t_gradient_ms += now_ms() - ta;
        // Hessian and solve (IMPLICIT_EULER.h:494-554)
        if (!useGD) {
            ta = now_ms();
            elastic_hessian(c, c.x.p);
            dbc_penalty_hess(c);   // INC_POTENTIAL.h:379-381
            if (P.collision) barrier_hessian(c, c.x.p, c.C.p, c.nC, c.kappa, c.fricOn ? c.xn.p : nullptr);
            else c.Acon.nnzb = 0;
            CIPC_HIP(hipStreamSynchronize(s));
            st.t_hessian_ms += now_ms() - ta;
            ta = now_ms();
            int it = 0;
            double rr = 0;
            // after a line search that stopped at alpha < 1, the quadratic model predicts
            // the next direction (1 - alpha) p (the gradient shrank to (1 - alpha) g); it
            // seeds the solve, whose stopping rule stays ||b - H p|| <= tol ||b||
            const double* x0 = nullptr;
            if (c.warmStart && havePrev && alphaPrev < 1.0) {
                double* wx = c.warmX.ensure(n);
                hipLaunchKernelGGL(k_scale_to, dim3(nblocks(n)), dim3(kBlock), 0, s, n, 1.0 - alphaPrev, c.p.p, wx);
                CIPC_LAUNCH_CHECK();
                x0 = wx;
            }
            c.solveX = c.x.p;
            int rc = pcg_solve(c, c.rhs.p, c.p.p, P.pcgTol, P.pcgMaxIter, &it, &rr, x0);
            st.pcg_iters += it;
            st.t_solve_ms += now_ms() - ta;
            if (rc == 1) { useGD = true; st.gd_fallbacks++; return false; }
        }
        return true;
    };
    do {
        if (!reuseDir) {
            newton_direction();
            if (useGD) CIPC_HIP(hipMemcpyAsync(c.p.p, c.rhs.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        reuseDir = false;
        // line search (IMPLICIT_EULER.h:9-149)
        CIPC_HIP(hipMemcpyAsync(c.xprev.p, c.x.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        double alpha = 1.0;
        // feasibility first -- strain limits, fibre limits, uninverted tets
        // (IMPLICIT_EULER.h:74-92) -- then the CCD from there
        if (strain_limit_on(c) || (c.P.fiberK[0] > 0 && c.nT) || c.nTet) {
            for (;;) {
                hipLaunchKernelGGL(k_axpy_to, dim3(nblocks(n)), dim3(kBlock), 0, s, n, c.xprev.p, alpha, c.p.p, c.x.p);
                CIPC_LAUNCH_CHECK();
                double Ef = strain_limit_energy(c, c.x.p);
                if (fiber_on(c)) Ef += fiber_energy(c, c.x.p);
                if (c.nTet && !tets_valid(c, c.x.p)) Ef = INFINITY;
                if (!std::isinf(Ef)) break;
                alpha /= 2;
            }
            CIPC_HIP(hipMemcpyAsync(c.x.p, c.xprev.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        if (P.collision) {
            ta = now_ms();
            int z = 0;
            alpha = ccd_stepsize(c, c.x.p, c.p.p, alpha, nullptr, &z);
            st.t_ccd_ms += now_ms() - ta;
            if (alpha == 0 || z) { c.lastError = "CCD returned a zero step"; return CIPC_E_ZERO_TOI; }
        }
        double feasibleAlpha = alpha;
        double E = Eprev;
        do {
            hipLaunchKernelGGL(k_axpy_to, dim3(nblocks(n)), dim3(kBlock), 0, s, n, c.xprev.p, alpha, c.p.p, c.x.p);
            CIPC_LAUNCH_CHECK();
            st.ls_trials++;
            if (P.collision) {
                ta = now_ms();
                constraint_set(c, c.x.p, c.C, c.nC);
                st.t_constraint_set_ms += now_ms() - ta;
                if (c.nC) {
                    double mn = min_dist2(c, c.x.p, c.C.p, c.nC, nullptr);
                    if (mn <= 0) { alpha /= 2; E = Eprev + 1; continue; }
                }
            }
            ta = now_ms();
            E = inc_potential(c, c.x.p, c.C.p, c.nC, &bad);
            st.t_energy_ms += now_ms() - ta;
            if (bad) { c.lastError = "non-positive distance in line search"; return CIPC_E_PENETRATION; }
            alpha /= 2.0;
        } while (E > Eprev);
        Eprev = E;
        alphaPrev = 2.0 * alpha;   // the accepted step
        havePrev = !useGD;
        // kappa adaptation (IMPLICIT_EULER.h:566-595)
        if (c.nCprev > 0) {
            double* cur = c.dscr2.ensure(c.nCprev);
            min_dist2(c, c.x.p, c.Cprev.p, c.nCprev, cur);
            CIPC_HIP(hipMemsetAsync(c.counters.p + 14, 0, sizeof(uint32_t), s));
            hipLaunchKernelGGL(k_kappa_flag, dim3(nblocks(c.nCprev)), dim3(kBlock), 0, s, c.nCprev, c.d2prev.p, cur, c.counters.p + 14);
            CIPC_LAUNCH_CHECK();
            CIPC_HIP(hipMemcpyAsync(c.hCnt + 14, c.counters.p + 14, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            CIPC_HIP(hipStreamSynchronize(s));
            if (c.hCnt[14] && c.kappa < c.kappaMax) {
                c.kappa *= 2;
                Eprev = inc_potential(c, c.x.p, c.C.p, c.nC, &bad);
            }
        }
        // remember the constraint set and its distances
        c.nCprev = c.nC;
        if (c.nC) {
            c.Cprev.ensure(4l * c.nC);
            CIPC_HIP(hipMemcpyAsync(c.Cprev.p, c.C.p, 4l * c.nC * sizeof(int), hipMemcpyDeviceToDevice, s));
            c.d2prev.ensure(c.nC);
            min_dist2(c, c.x.p, c.C.p, c.nC, c.d2prev.p);
        }
        // kappa_s adaptation (IMPLICIT_EULER.h:606-645): double the upper strain-limit
        // stiffness, up to 1e5, when a triangle's largest singular value stayed within
        // 1e-4 (s - sHat) of the limit at this and the previous Newton iterate; the
        // caller's kappa_s is updated in place (cipc_get_strain_limit reads it back)
        if (strain_limit_on(c)) {
            const bool upd = strain_limit_near_twice(c, c.x.p, PNIter == 0);
            if (upd && c.P.slKappa[0] < 1e5) {
                c.P.slKappa[0] *= 2;
                c.kappaSDoublings++;
                Eprev = inc_potential(c, c.x.p, c.C.p, c.nC, &bad);
            }
        }
        // stopping rule (IMPLICIT_EULER.h:647-710, 866)
        double ss = sumsq(c, c.p.p, n);
        L2Norm = std::sqrt(ss / (c.nV - c.nDBC)) / hres;
        PNIter++;
        if (!c.outputFolder.empty()) log_residual(c, PNIter, Eprev, L2Norm, hres);
        resRecord.push_back(L2Norm);
        if (resRecord.size() > 3) resRecord.pop_front();
        double cur = L2Norm;
        L2Norm = *std::max_element(resRecord.begin(), resRecord.end());
        if (useGD) L2Norm = P.PNTol * 10;
        if (alpha * 2 < 1e-8 && feasibleAlpha > 1e-8) {
            if (!useGD) {
                if (cur < P.PNTol) resRecord.assign(3, cur);
                else useGD = true;
            }
        } else {
            useGD = false;
        }
        st.last_residual = cur;
        if (PNIter > 2000) { c.lastError = "Newton iteration limit"; return CIPC_E_NO_CONVERGENCE; }
        // Dirichlet penalty continuation (IMPLICIT_EULER.h:712-768)
        if (c.dbcStiff) {
            const double progress = 1 - std::sqrt(dbc_dist2(c, c.x.p) / dbcPenXn);
            if (progress < 0.99) {
                if (L2Norm < P.PNTol * 10 && c.dbcStiff < 1e8) {
                    c.dbcStiff *= 2;
                    Eprev = inc_potential(c, c.x.p, c.C.p, c.nC, &bad);
                }
                L2Norm = P.PNTol * 10;
            } else {
                c.dbcStiff = 0;
                Eprev = inc_potential(c, c.x.p, c.C.p, c.nC, &bad);
            }
        }
        // friction re-lagged at a converged iterate (IMPLICIT_EULER.h:770-859): new basis
        // and normal forces; the Newton direction of the updated potential decides whether
        // the loop goes on.  The reference recomputes that same direction at the top of
        // the next iteration; here it is kept in c.p and used directly.
        if (c.fricOn && resRecord.size() == 3 && L2Norm <= P.PNTol) {
            ++fricIterI;
            if (fricIterI < P.fricIterAmt || P.fricIterAmt <= 0) {
                friction_basis(c, c.x.p, c.C.p, c.nC, c.kappa);
                havePrev = false;
                if (!newton_direction()) { c.lastError = "friction update: Newton matrix not positive definite"; return CIPC_E_NO_CONVERGENCE; }
                L2Norm = std::sqrt(sumsq(c, c.p.p, n) / (c.nV - c.nDBC)) / hres;
                if (L2Norm > P.PNTol) {
                    Eprev = inc_potential(c, c.x.p, c.C.p, c.nC, &bad);
                    reuseDir = true;
                }
            }
        }
    } while (resRecord.size() < 3 || L2Norm > P.PNTol);
    c.fricOn = false;
    if (!stat) {   // velocities stay as they are after a static solve (IMPLICIT_EULER.h:893-899)
        hipLaunchKernelGGL(k_velocity, dim3(nblocks(n)), dim3(kBlock), 0, s, n, c.x.p, c.xn.p, h, c.v.p);
        CIPC_LAUNCH_CHECK();
    }
    CIPC_HIP(hipStreamSynchronize(s));
    if (!c.outputFolder.empty()) log_step_end(c, PNIter);
    st.newton_iters = PNIter;
    st.contacts = c.nC;
    st.kappa = c.kappa;
    st.kappa_s = c.P.slKappa[0];
    st.kappa_s_doublings = c.kappaSDoublings;
    st.energy = Eprev;
    st.t_total_ms = now_ms() - t0;
    st.status = CIPC_OK;
    return CIPC_OK;
}

}  // namespace cipc
