// Oracle: scene set-up and one implicit-Euler step by projected Newton.
// TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// scene_init restates Initialize_Discrete_Shell<double,3,KL=false>
// (FEM/Shell/DISCRETE_SHELL.h:218-357: degenerate-triangle filter, edge2tri, hinge
// stencils and rest data :169-213, lumped mass, lambda/mu, hinge k), the surface
// primitive extraction (Utils/MESHIO.h:728-835) and Initialize_EIPC<double,false>
// (DISCRETE_SHELL.h:555-577: kappa from the mean lumped mass).
//
// advance_one_step restates Advance_One_Step_IE_Discrete_Shell<double,3,false,false,
// false> (FEM/Shell/IMPLICIT_EULER.h:151-902) for the configuration this build
// covers: isotropic strain limiting (INEXT.h), no fibres, no tets, no stitches, static
// or CCD-unblocked Dirichlet nodes, lagged friction (basis at the step start, re-lagged
// after convergence up to fricIterAmt times, IMPLICIT_EULER.h:412-437, 770-859), and
// Line_Search (IMPLICIT_EULER.h:9-149).
// Parity pins: energy non-increase per accepted step, intersection-free states,
// stopping rule; trajectories are the reference for the HIP path.
#include "oracle.h"
#include <algorithm>
#include <map>
#include <set>
#include <chrono>
#include <deque>
#include <cmath>
#include <cstdio>

namespace orc {

static inline V3 X3(const double* x, int i) { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }
static double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

void scene_init(Scene& S) {
    const Params& P = S.P;
    // degenerate filter, DISCRETE_SHELL.h:228-244
    std::vector<int> kept;
    for (int t = 0; t < (int)S.tri.size() / 3; ++t) {
        V3 a = X3(S.X0.data(), S.tri[3 * t]), b = X3(S.X0.data(), S.tri[3 * t + 1]), c = X3(S.X0.data(), S.tri[3 * t + 2]);
        V3 e1 = b - a, e2 = c - a;
        double I00 = sq(e1), I01 = dot(e1, e2), I11 = sq(e2);
        if (I00 * I11 - I01 * I01 != 0) for (int k = 0; k < 3; ++k) kept.push_back(S.tri[3 * t + k]);
    }
    S.tri = kept;
    S.nT = (int)S.tri.size() / 3;
    // contact triangles: the tets' surface first, then the shell (IMPLICIT_EULER.h:235-241)
    std::vector<int> primTri;
    tet_surface(S, primTri);
    primTri.insert(primTri.end(), S.tri.begin(), S.tri.end());
    if (S.x.empty()) S.x = S.X0;
    if (S.v.empty()) S.v.assign(3 * S.nV, 0.0);
    S.isDBC.assign(S.nV, 0);
    for (int d : S.dbcIdx) S.isDBC[d] = 1;
    // edge2tri and surface primitives, MESHIO.h:728-766
    std::map<std::pair<int, int>, int> edge2tri;
    std::set<std::pair<int, int>> edgeSet, primEdgeSet;
    std::vector<char> isBN(S.nV, 0);
    for (int t = 0; t < S.nT; ++t) {
        const int* v = &S.tri[3 * t];
        edge2tri[{v[0], v[1]}] = t; edge2tri[{v[1], v[2]}] = t; edge2tri[{v[2], v[0]}] = t;
    }
    auto add_edges = [](std::set<std::pair<int, int>>& es, const int* v) {
        if (!es.count({v[1], v[0]})) es.insert({v[0], v[1]});
        if (!es.count({v[2], v[1]})) es.insert({v[1], v[2]});
        if (!es.count({v[0], v[2]})) es.insert({v[2], v[0]});
    };
    for (int t = 0; t < S.nT; ++t) add_edges(edgeSet, &S.tri[3 * t]);        // hinges: the shell's edges
    for (size_t t = 0; t < primTri.size() / 3; ++t) {                        // primitives: shell + tet surface
        const int* v = &primTri[3 * t];
        add_edges(primEdgeSet, v);
        isBN[v[0]] = isBN[v[1]] = isBN[v[2]] = 1;
    }
    S.bNode.clear(); S.bEdge.clear();
    for (int i = 0; i < S.nV; ++i) if (isBN[i]) S.bNode.push_back(i);
    for (auto& e : primEdgeSet) { S.bEdge.push_back(e.first); S.bEdge.push_back(e.second); }
    S.bTri = primTri;
    // rest first fundamental form (inverted) and hinges, DISCRETE_SHELL.h:132-213
    S.Binv.resize(3 * S.nT);
    for (int t = 0; t < S.nT; ++t) {
        V3 a = X3(S.X0.data(), S.tri[3 * t]), b = X3(S.X0.data(), S.tri[3 * t + 1]), c = X3(S.X0.data(), S.tri[3 * t + 2]);
        V3 e1 = b - a, e2 = c - a;
        double I00 = sq(e1), I01 = dot(e1, e2), I11 = sq(e2);
        double det = I00 * I11 - I01 * I01;
        S.Binv[3 * t] = I11 / det; S.Binv[3 * t + 1] = -I01 / det; S.Binv[3 * t + 2] = I00 / det;
    }
    S.hinge.clear(); S.hingeInfo.clear();
    for (auto& e : edgeSet) {
        auto fa = edge2tri.find({e.first, e.second});
        if (fa == edge2tri.end()) continue;
        const int* A = &S.tri[3 * fa->second];
        int v0 = -1;
        for (int j = 0; j < 3; ++j) if (A[j] == e.second) { v0 = A[(j + 1) % 3]; break; }
        auto fb = edge2tri.find({e.second, e.first});
        if (fb == edge2tri.end()) continue;
        const int* B = &S.tri[3 * fb->second];
        int v3 = B[0];
        for (int j = 0; j < 3; ++j) if (B[j] == e.first) { v3 = B[(j + 1) % 3]; break; }
        int st[4] = {v0, e.first, e.second, v3};
        for (int k = 0; k < 4; ++k) S.hinge.push_back(st[k]);
        V3 X0 = X3(S.X0.data(), st[0]), X1 = X3(S.X0.data(), st[1]), X2 = X3(S.X0.data(), st[2]), Xq = X3(S.X0.data(), st[3]);
        double th = dihedral_angle(X0, X1, X2, Xq);
        double el = std::sqrt(sq(X1 - X2));
        V3 n1 = cross(X1 - X0, X2 - X0), n2 = cross(X2 - Xq, X1 - Xq);
        S.hingeInfo.push_back(th);
        S.hingeInfo.push_back(el);
        S.hingeInfo.push_back((std::sqrt(sq(n1)) + std::sqrt(sq(n2))) / (el * 6));
    }
    // lumped mass, DISCRETE_SHELL.h:279-299
    S.mass.assign(S.nV, 0.0);
    S.vol.resize(S.nT);
    for (int t = 0; t < S.nT; ++t) {
        const int* v = &S.tri[3 * t];
        V3 a = X3(S.X0.data(), v[0]), b = X3(S.X0.data(), v[1]), c = X3(S.X0.data(), v[2]);
        double area = std::sqrt(sq(cross(b - a, c - a))) / 2;
        double mp = area * P.thick * P.rho / 3;
        S.mass[v[0]] += mp; S.mass[v[1]] += mp; S.mass[v[2]] += mp;
        S.vol[t] = area * P.thick;
    }
    S.lam = P.E * P.nu / (1.0 - P.nu * P.nu);
    S.mu = P.E / (2.0 * (1.0 + P.nu));
    S.kBend = P.bendMult * (P.E * std::pow(P.thick, 3) / (24 * (1.0 - P.nu * P.nu)));
    // tets (initialize_added_objects), rods and particles (initialize_rod /
    // initialize_particle) run before initialize_OIPC
    if (!tet_setup(S)) { S.setupError = "a rest tet is inverted or degenerate"; }
    codim_init(S);
    strain_limit_setup(S);
    fiber_setup(S);
    // Initialize_EIPC<double,false>, DISCRETE_SHELL.h:563-574
    S.dHat2 = P.dHat * P.dHat;
    double Hb = barrier_h(1.0e-16, S.dHat2, 1.0);
    double mmean = 0;
    for (int i = 0; i < S.nV; ++i) mmean += 3 * S.mass[i];
    mmean /= 3 * S.nV;
    S.kappa = P.kappaStiffMult * 1.0e11 * mmean * 3 / (4.0e-16 * Hb);
    S.kappaMax = 100 * S.kappa;
}

void step_dirichlet(Scene& S) {
    const double h = S.P.dt;
    for (const Scene::DBCMotion& M : S.dbcMotion) {
        const double th = M.angVelDeg / 180 * M_PI * h;
        const double c = std::cos(th), sn = std::sin(th);
        const double* a = M.axis;
        // rotation by th about axis a (Rodrigues, axis taken as given)
        const double R[9] = {c + (1 - c) * a[0] * a[0], (1 - c) * a[0] * a[1] - sn * a[2], (1 - c) * a[0] * a[2] + sn * a[1],
                             (1 - c) * a[1] * a[0] + sn * a[2], c + (1 - c) * a[1] * a[1], (1 - c) * a[1] * a[2] - sn * a[0],
                             (1 - c) * a[2] * a[0] - sn * a[1], (1 - c) * a[2] * a[1] + sn * a[0], c + (1 - c) * a[2] * a[2]};
        for (int i = M.begin; i < M.end; ++i) {
            double* t = &S.dbcTarget[3 * i];
            if (M.angVelDeg != 0) {
                const double d[3] = {t[0] - M.center[0], t[1] - M.center[1], t[2] - M.center[2]};
                for (int k = 0; k < 3; ++k) t[k] = R[3 * k] * d[0] + R[3 * k + 1] * d[1] + R[3 * k + 2] * d[2] + M.center[k];
            }
            for (int k = 0; k < 3; ++k) t[k] += M.v[k] * h;
        }
    }
}

// 1/2 k sum_DBC m |x - target|^2 (Shell/DIRICHLET.h:26-48)
static double dbc_energy(const Scene& S, const double* x, double k) {
    double E = 0;
    for (size_t i = 0; i < S.dbcIdx.size(); ++i) {
        const int v = S.dbcIdx[i];
        double d2 = 0;
        for (int c = 0; c < 3; ++c) d2 += (S.dbcTarget[3 * i + c] - x[3 * v + c]) * (S.dbcTarget[3 * i + c] - x[3 * v + c]);
        E += d2 * S.mass[v];
    }
    return 0.5 * k * E;
}
static double dbc_dist2(const Scene& S, const double* x) {
    double E = 0;
    for (size_t i = 0; i < S.dbcIdx.size(); ++i)
        for (int c = 0; c < 3; ++c) E += (S.dbcTarget[3 * i + c] - x[3 * S.dbcIdx[i] + c]) * (S.dbcTarget[3 * i + c] - x[3 * S.dbcIdx[i] + c]);
    return E;
}

// full incremental potential at x for constraint set C (+ friction with the lagged set F)
// (*invalid set when a strain limit is reached, INC_POTENTIAL.h:75-80)
static double inc_potential(const Scene& S, const double* x, const double* xt, const std::vector<Constraint>& C,
                            double kappa, bool withContact, int* bad, const double* xn = nullptr,
                            const FricSet* F = nullptr, bool* invalid = nullptr, double dbcStiff = 0) {
    double E = 0;
    if (!S.tet.empty() && !tets_valid(S, x)) {   // an inverted tet, INC_POTENTIAL.h:98-107
        if (invalid) *invalid = true;
        if (bad) *bad = 0;
        return INFINITY;
    }
    bool ok = strain_limit_energy(S, x, &E);
    if (fiber_active(S)) ok = fiber_energy(S, x, &E) && ok;   // Check_Fiber_Feasibility, INC_POTENTIAL.h:81-88
    if (invalid) *invalid = !ok;
    E += elastic_energy(S, x) + inertia_energy(S, x, xt);
    if (withContact) E += barrier_energy(S, x, C, kappa, bad);
    if (F) E += friction_energy(S, x, xn, *F);
    if (dbcStiff) E += dbc_energy(S, x, dbcStiff);
    return E;
}

int advance_one_step(Scene& S, StepStats& st) {
    const Params& P = S.P;
    const int n = 3 * S.nV;
    const double h = P.dt;
    double t0 = now();
    st = StepStats();
    std::vector<double>& x = S.x;
    std::vector<double> xn = x, xt(n, 0.0);
    const bool stat = P.staticSolve != 0;
    if (stat && P.collision && (P.fricMu > 0 || friction_by_component(S))) return -7;   // no previous state for friction
    const double hres = stat ? 1.0 : h;   // residual scale, IMPLICIT_EULER.h:662-665
    // x_tilde = x + h v + h^2 M^{-1} b, IMPLICIT_EULER.h:195-217 (b = m g); unused in a static solve
    for (int i = 0; i < S.nV && !stat; ++i)
        for (int k = 0; k < 3; ++k) {
            double a = S.mass[i] > 0 ? P.grav[k] : 0.0;
            xt[3 * i + k] = x[3 * i + k] + h * S.v[3 * i + k] + h * h * a;
        }
    // Dirichlet targets, IMPLICIT_EULER.h:304-401: move the nodes if the CCD allows the
    // whole motion, else pull them with a mass-weighted penalty (DBCStiff) that is raised
    // until 99% of the distance is covered; fixed nodes (zero motion) stay projected
    std::vector<double> disp(n, 0.0);
    std::vector<char> fixedDBC(S.nV, 0);
    for (size_t k = 0; k < S.dbcIdx.size(); ++k) {
        int v = S.dbcIdx[k];
        for (int d = 0; d < 3; ++d) disp[3 * v + d] = S.dbcTarget[3 * k + d] - x[3 * v + d];
        if (!(disp[3 * v] || disp[3 * v + 1] || disp[3 * v + 2])) fixedDBC[v] = 1;
    }
    double dbcAlpha = 1;
    if (P.collision) dbcAlpha = ccd_stepsize(S, x.data(), disp.data(), 1.0);
    // a Dirichlet motion that would break a strain limit or the fibre limits is cut back
    // by halving until the moved nodes leave every triangle feasible, IMPLICIT_EULER.h:
    // 336-376 (only dbcAlpha survives: the nodes return to x_n below)
    if ((P.slKappa[0] > 0 || P.fiberK[0] > 0 || !S.tet.empty()) && !S.dbcIdx.empty()) {
        for (int it = 0; it < 64; ++it) {
            for (size_t k = 0; k < S.dbcIdx.size(); ++k)
                for (int d = 0; d < 3; ++d) {
                    const int i = 3 * S.dbcIdx[k] + d;
                    x[i] = xn[i] + dbcAlpha * (S.dbcTarget[3 * k + d] - xn[i]);
                }
            double Et = 0;
            bool ok = strain_limit_energy(S, x.data(), &Et);
            if (fiber_active(S)) ok = fiber_energy(S, x.data(), &Et) && ok;
            if (!S.tet.empty()) ok = tets_valid(S, x.data()) && ok;   // IMPLICIT_EULER.h:364-374
            if (ok) break;
            dbcAlpha /= 2;
        }
        x = xn;
    }
    double dbcStiff = 0, dbcPenXn = 0;
    if (dbcAlpha == 1) {
        for (size_t k = 0; k < S.dbcIdx.size(); ++k)
            for (int d = 0; d < 3; ++d) x[3 * S.dbcIdx[k] + d] = S.dbcTarget[3 * k + d];
    } else {
        dbcStiff = 1e6;
        dbcPenXn = dbc_dist2(S, xn.data());
    }

    std::vector<Constraint> C, Cprev;
    std::vector<double> d2prev, strainPrev;
    if (P.collision) constraint_set(S, x.data(), C);
    // lagged friction set at the step start, IMPLICIT_EULER.h:419-437
    const bool fric = P.collision && (P.fricMu > 0 || friction_by_component(S));
    FricSet F;
    const FricSet* Fp = fric ? &F : nullptr;
    if (fric) friction_basis(S, x.data(), C, S.kappa, F);
    int bad = 0;
    bool invalid = false;
    double Eprev = inc_potential(S, x.data(), xt.data(), C, S.kappa, P.collision, &bad, xn.data(), Fp, &invalid, dbcStiff);
    if (bad) return -1;
    if (invalid) return -6;   // strain limit violated by the initial state
    int fricIterI = 0;
    S.eBefore.clear(); S.eAfter.clear();
    std::deque<double> resRecord;
    bool useGD = false;
    double L2Norm = 0;
    int PNIter = 0;
    std::vector<double> g(n), rhs(n), sol(n), xprev(n);
    CSR A;
    do {
        // gradient, INC_POTENTIAL.h:154-275, DBC rows zeroed IMPLICIT_EULER.h:476-491
        double ta = now();
        std::fill(g.begin(), g.end(), 0.0);
        elastic_grad(S, x.data(), g.data());
        strain_limit_grad(S, x.data(), g.data());
        if (fiber_active(S)) fiber_grad(S, x.data(), g.data());
        for (int i = 0; i < S.nV; ++i)
            for (int k = 0; k < 3; ++k)
                g[3 * i + k] += stat ? (S.mass[i] > 0 ? -S.mass[i] * P.grav[k] : 0.0) : S.mass[i] * (x[3 * i + k] - xt[3 * i + k]);
        if (P.collision) barrier_grad(S, x.data(), C, S.kappa, g.data());
        if (fric) friction_grad(S, x.data(), xn.data(), F, g.data());
        if (dbcStiff) {   // Shell/DIRICHLET.h:50-70; only fixed nodes projected
            for (size_t i = 0; i < S.dbcIdx.size(); ++i) {
                const int v = S.dbcIdx[i];
                for (int k = 0; k < 3; ++k) g[3 * v + k] += dbcStiff * S.mass[v] * (x[3 * v + k] - S.dbcTarget[3 * i + k]);
            }
            for (int v : S.dbcIdx) if (fixedDBC[v]) for (int k = 0; k < 3; ++k) g[3 * v + k] = 0;
        } else {
            for (int v : S.dbcIdx) for (int k = 0; k < 3; ++k) g[3 * v + k] = 0;
        }
        for (int i = 0; i < n; ++i) rhs[i] = -g[i];
        st.tGrad += now() - ta;
        // Hessian + solve, IMPLICIT_EULER.h:494-554
        bool solved = false;
        if (!useGD) {
            ta = now();
            std::vector<Block> blocks;
            elastic_hess(S, x.data(), blocks);
            strain_limit_hess(S, x.data(), blocks);
            if (fiber_active(S)) fiber_hess(S, x.data(), blocks);
            if (P.collision) barrier_hess(S, x.data(), C, S.kappa, blocks);
            if (fric) friction_hess(S, x.data(), xn.data(), F, blocks);
            if (dbcStiff)
                for (int v : S.dbcIdx) {
                    Block B{v, v, {0, 0, 0, 0, 0, 0, 0, 0, 0}};
                    B.b[0] = B.b[4] = B.b[8] = dbcStiff * S.mass[v];
                    blocks.push_back(B);
                }
            assemble(S, blocks, A, dbcStiff ? &fixedDBC : nullptr);
            st.tHessian += now() - ta;
            ta = now();
            int rc;
            if (P.solver == 0) rc = solve_direct(A, rhs.data(), sol.data());
            else { int it = 0; rc = solve_pcg(A, rhs.data(), sol.data(), P.pcgTol, P.pcgMaxIter, &it, P.threads); st.pcgIters += it; }
            st.tSolve += now() - ta;
            solved = (rc == 0);
            if (!solved) { useGD = true; st.useGDCount++; }
        }
        if (useGD) sol = rhs;
        // line search, IMPLICIT_EULER.h:9-149
        xprev = x;
        double alpha = 1.0;
        // feasibility first (strain limits, fibre limits, uninverted tets), IMPLICIT_EULER.h:74-92
        if (P.slKappa[0] > 0 || P.fiberK[0] > 0 || !S.tet.empty()) {
            for (;;) {
                for (int i = 0; i < n; ++i) x[i] = xprev[i] + alpha * sol[i];
                double Et = 0;
                bool ok = strain_limit_energy(S, x.data(), &Et);
                if (fiber_active(S)) ok = fiber_energy(S, x.data(), &Et) && ok;
                if (!S.tet.empty()) ok = tets_valid(S, x.data()) && ok;
                if (ok) break;
                alpha /= 2;
            }
            x = xprev;
        }
        if (P.collision) {
            ta = now();
            alpha = ccd_stepsize(S, x.data(), sol.data(), alpha);   // from the feasible alpha
            st.tCCD += now() - ta;
            if (alpha == 0) return -3;
        }
        double feasibleAlpha = alpha;
        double E;
        bool valid = true;
        do {
            for (int i = 0; i < n; ++i) x[i] = xprev[i] + alpha * sol[i];
            st.lsTrials++;
            if (P.collision) {
                ta = now();
                constraint_set(S, x.data(), C);
                st.tCS += now() - ta;
                if (!C.empty()) {
                    std::vector<double> d2;
                    constraint_dist2(S, x.data(), C, d2);
                    double mn = *std::min_element(d2.begin(), d2.end()) - P.xi * P.xi;
                    if (mn <= 0) { alpha /= 2; E = Eprev + 1; continue; }
                }
            }
            ta = now();
            E = inc_potential(S, x.data(), xt.data(), C, S.kappa, P.collision, &bad, xn.data(), Fp, &invalid, dbcStiff);
            valid = !invalid;
            st.tEnergy += now() - ta;
            if (bad && valid) return -1;
            alpha /= 2.0;
        } while (E > Eprev || !valid);
        S.eBefore.push_back(Eprev);
        S.eAfter.push_back(E);
        Eprev = E;
        // kappa adaptation, IMPLICIT_EULER.h:566-595
        if (!Cprev.empty()) {
            std::vector<double> cur;
            constraint_dist2(S, x.data(), Cprev, cur);
            bool upd = false;
            for (size_t i = 0; i < cur.size(); ++i)
                if (d2prev[i] < 1e-18 && cur[i] < d2prev[i]) { upd = true; break; }
            if (upd && S.kappa < S.kappaMax) {
                S.kappa *= 2;
                Eprev = inc_potential(S, x.data(), xt.data(), C, S.kappa, P.collision, &bad, xn.data(), Fp, nullptr, dbcStiff);
            }
        }
        std::vector<double> d2;
        if (!C.empty()) constraint_dist2(S, x.data(), C, d2);
        Cprev = C; d2prev = d2;
        // kappa_s adaptation, IMPLICIT_EULER.h:606-645: double the upper strain-limit
        // stiffness (up to 1e5) when a triangle's largest singular value stayed within
        // 1e-4 (s - sHat) of the limit in this and the previous Newton iteration
        if (P.slKappa[0] > 0) {
            std::vector<double> smax;
            strain_max_sv(S, x.data(), smax);
            const double thr = 1.0e-4 * (P.slS[0] - P.slSHat[0]);
            bool upd = false;
            if (!strainPrev.empty())
                for (int t = 0; t < S.nT && !upd; ++t)
                    upd = (P.slS[0] - strainPrev[t] < thr) && (P.slS[0] - smax[t] < thr);
            if (upd && P.slKappa[0] < 1e5) {
                S.P.slKappa[0] *= 2;
                Eprev = inc_potential(S, x.data(), xt.data(), C, S.kappa, P.collision, &bad, xn.data(), Fp, nullptr, dbcStiff);
                st.kappaSDoublings++;
            }
            strainPrev.swap(smax);
        }
        // stopping rule, IMPLICIT_EULER.h:647-710, 866
        double s = 0;
        for (int i = 0; i < n; ++i) s += sol[i] * sol[i];
        L2Norm = std::sqrt(s / (S.nV - (int)S.dbcIdx.size())) / hres;
        PNIter++;
        resRecord.push_back(L2Norm);
        if (resRecord.size() > 3) resRecord.pop_front();
        double cur = L2Norm;
        L2Norm = *std::max_element(resRecord.begin(), resRecord.end());
        if (useGD) L2Norm = P.PNTol * 10;
        if (alpha * 2 < 1e-8 && feasibleAlpha > 1e-8) {
            if (!useGD) {
                if (cur < P.PNTol) { resRecord.assign(3, cur); }
                else useGD = true;
            }
        } else {
            useGD = false;
        }
        st.lastRes = cur;
        if (PNIter > 2000) return -4;
        if (P.maxNewton > 0 && PNIter >= P.maxNewton) break;   // timed prefix: stop at this iterate
        // penalty continuation, IMPLICIT_EULER.h:712-768
        if (dbcStiff) {
            const double progress = 1 - std::sqrt(dbc_dist2(S, x.data()) / dbcPenXn);
            if (progress < 0.99) {
                if (L2Norm < P.PNTol * 10 && dbcStiff < 1e8) {
                    dbcStiff *= 2;
                    Eprev = inc_potential(S, x.data(), xt.data(), C, S.kappa, P.collision, &bad, xn.data(), Fp, nullptr, dbcStiff);
                }
                L2Norm = P.PNTol * 10;
            } else {
                dbcStiff = 0;
                Eprev = inc_potential(S, x.data(), xt.data(), C, S.kappa, P.collision, &bad, xn.data(), Fp, nullptr, 0.0);
            }
        }
        // friction re-lagged at a converged iterate, IMPLICIT_EULER.h:770-859: new basis and
        // normal forces, then the Newton direction of the updated potential decides
        if (fric && resRecord.size() == 3 && L2Norm <= P.PNTol) {
            ++fricIterI;
            if (fricIterI < P.fricIterAmt || P.fricIterAmt <= 0) {
                friction_basis(S, x.data(), C, S.kappa, F);
                std::fill(g.begin(), g.end(), 0.0);
                elastic_grad(S, x.data(), g.data());
                strain_limit_grad(S, x.data(), g.data());
                if (fiber_active(S)) fiber_grad(S, x.data(), g.data());
                for (int i = 0; i < S.nV; ++i)
                    for (int k = 0; k < 3; ++k) g[3 * i + k] += S.mass[i] * (x[3 * i + k] - xt[3 * i + k]);
                barrier_grad(S, x.data(), C, S.kappa, g.data());
                friction_grad(S, x.data(), xn.data(), F, g.data());
                if (dbcStiff) {
                    for (size_t i = 0; i < S.dbcIdx.size(); ++i) {
                        const int v = S.dbcIdx[i];
                        for (int k = 0; k < 3; ++k) g[3 * v + k] += dbcStiff * S.mass[v] * (x[3 * v + k] - S.dbcTarget[3 * i + k]);
                    }
                    for (int v : S.dbcIdx) if (fixedDBC[v]) for (int k = 0; k < 3; ++k) g[3 * v + k] = 0;
                } else {
                    for (int v : S.dbcIdx) for (int k = 0; k < 3; ++k) g[3 * v + k] = 0;
                }
                for (int i = 0; i < n; ++i) rhs[i] = -g[i];
                std::vector<Block> blocks;
                elastic_hess(S, x.data(), blocks);
                strain_limit_hess(S, x.data(), blocks);
                if (fiber_active(S)) fiber_hess(S, x.data(), blocks);
                barrier_hess(S, x.data(), C, S.kappa, blocks);
                friction_hess(S, x.data(), xn.data(), F, blocks);
                if (dbcStiff)
                    for (int v : S.dbcIdx) {
                        Block B{v, v, {0, 0, 0, 0, 0, 0, 0, 0, 0}};
                        B.b[0] = B.b[4] = B.b[8] = dbcStiff * S.mass[v];
                        blocks.push_back(B);
                    }
                assemble(S, blocks, A, dbcStiff ? &fixedDBC : nullptr);
                int rc;
                if (P.solver == 0) rc = solve_direct(A, rhs.data(), sol.data());
                else { int it = 0; rc = solve_pcg(A, rhs.data(), sol.data(), P.pcgTol, P.pcgMaxIter, &it, P.threads); st.pcgIters += it; }
                if (rc != 0) return -5;   // IMPLICIT_EULER.h:826-831 exits on a non-SPD matrix
                double s2 = 0;
                for (int i = 0; i < n; ++i) s2 += sol[i] * sol[i];
                L2Norm = std::sqrt(s2 / (S.nV - (int)S.dbcIdx.size())) / hres;
                if (L2Norm > P.PNTol) Eprev = inc_potential(S, x.data(), xt.data(), C, S.kappa, P.collision, &bad, xn.data(), Fp, nullptr, dbcStiff);
            }
        }
    } while (resRecord.size() < 3 || L2Norm > P.PNTol);
    if (!stat)   // velocities stay as they are after a static solve, IMPLICIT_EULER.h:893-899
        for (int i = 0; i < n; ++i) S.v[i] = (x[i] - xn[i]) / h;
    st.newtonIters = PNIter;
    st.nContacts = (int)C.size();
    st.kappa = S.kappa;
    st.E = Eprev;
    st.tTotal = now() - t0;
    return PNIter;
}

}  // namespace orc
