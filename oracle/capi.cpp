// Oracle C entry points for ctypes (tests/, smoke, bench cpu_baseline only).
// TEST INFRASTRUCTURE ONLY (see oracle.h).
#include "oracle.h"
#include <algorithm>
#include <cstring>

using namespace orc;

static V3 P3(const double* p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

extern "C" {

// ---------------------------------------------------------------- primitives
// kind: 0 PP (2 pts), 1 PE (3), 2 PT (4), 3 EE (4)
double orc_dist2(int kind, const double* p) {
    switch (kind) {
        case 0: return dist2_pp(P3(p, 0), P3(p, 1));
        case 1: return dist2_pe(P3(p, 0), P3(p, 1), P3(p, 2));
        case 2: return dist2_pt(P3(p, 0), P3(p, 1), P3(p, 2), P3(p, 3));
        default: return dist2_ee(P3(p, 0), P3(p, 1), P3(p, 2), P3(p, 3));
    }
}
void orc_dist2_gh(int kind, const double* p, double* g, double* H) {
    switch (kind) {
        case 0: dist2_pp_gh(P3(p, 0), P3(p, 1), g, H); break;
        case 1: dist2_pe_gh(P3(p, 0), P3(p, 1), P3(p, 2), g, H); break;
        case 2: dist2_pt_gh(P3(p, 0), P3(p, 1), P3(p, 2), P3(p, 3), g, H); break;
        default: dist2_ee_gh(P3(p, 0), P3(p, 1), P3(p, 2), P3(p, 3), g, H); break;
    }
}
// kind: 1 PE, 2 PT, 3 EE; 4 PT unclassified distance, 5 EE unclassified, 6 PE unclassified
int orc_type(int kind, const double* p) {
    switch (kind) {
        case 1: return pe_type(P3(p, 0), P3(p, 1), P3(p, 2));
        case 2: return pt_type(P3(p, 0), P3(p, 1), P3(p, 2), P3(p, 3));
        default: return ee_type(P3(p, 0), P3(p, 1), P3(p, 2), P3(p, 3));
    }
}
double orc_dist2_unclassified(int kind, const double* p) {
    switch (kind) {
        case 1: return dist2_pe_unclassified(P3(p, 0), P3(p, 1), P3(p, 2));
        case 2: return dist2_pt_unclassified(P3(p, 0), P3(p, 1), P3(p, 2), P3(p, 3));
        default: return dist2_ee_unclassified(P3(p, 0), P3(p, 1), P3(p, 2), P3(p, 3));
    }
}
void orc_mollifier(const double* p, const double* prest, double* e, double* g, double* H, double* eps) {
    *eps = ee_mollifier_threshold(P3(prest, 0), P3(prest, 1), P3(prest, 2), P3(prest, 3));
    ee_mollifier(P3(p, 0), P3(p, 1), P3(p, 2), P3(p, 3), *eps, e, g, H);
}
void orc_barrier(double d, double dhat, double kappa, double* out3) {
    out3[0] = barrier(d, dhat, kappa);
    out3[1] = barrier_g(d, dhat, kappa);
    out3[2] = barrier_h(d, dhat, kappa);
}
void orc_sym_eig(int n, const double* A, double* w, double* V) { sym_eig(n, A, w, V); }
void orc_make_pd(int n, double* A) { make_pd(n, A); }
void orc_membrane(const double* x, const double* B, double vol, double lam, double mu, double* E, double* g, double* H) {
    *E = membrane_energy(P3(x, 0), P3(x, 1), P3(x, 2), B, vol, lam, mu);
    membrane_grad(P3(x, 0), P3(x, 1), P3(x, 2), B, vol, lam, mu, g);
    membrane_hess(P3(x, 0), P3(x, 1), P3(x, 2), B, vol, lam, mu, H);
}
void orc_dihedral(const double* x, double* th, double* g, double* H) {
    *th = dihedral_angle(P3(x, 0), P3(x, 1), P3(x, 2), P3(x, 3));
    dihedral_angle_grad(P3(x, 0), P3(x, 1), P3(x, 2), P3(x, 3), g);
    dihedral_angle_hess(P3(x, 0), P3(x, 1), P3(x, 2), P3(x, 3), H);
}

void orc_rod_spring(const double* x, double Er, double lbar, double* E, double* g, double* H) {
    *E = rod_spring_energy(P3(x, 0), P3(x, 1), Er, lbar);
    rod_spring_grad(P3(x, 0), P3(x, 1), Er, lbar, g);
    rod_spring_hess(P3(x, 0), P3(x, 1), Er, lbar, H);
}
void orc_rod_kappa2(const double* x, double* k2, double* g, double* H) {
    *k2 = rod_kappa2(P3(x, 0), P3(x, 1), P3(x, 2));
    rod_kappa2_grad(P3(x, 0), P3(x, 1), P3(x, 2), g);
    rod_kappa2_hess(P3(x, 0), P3(x, 1), P3(x, 2), H);
}

// ---------------------------------------------------------------- scenes
// params: rho, E, nu, thick, bendMult, dt, dHat, xi, PNTol, gx, gy, gz, kappaStiffMult, pcgTol
// iparams: collision, solver, pcgMaxIter, threads
// codim (may be null when nRod = nPart = 0): rodRho, rodE, rodBendMult, rodThick, partRho, partThick
// tets (may be null when nTet = 0): 4 per tet; tetParams: density, E, nu
void* orc_scene_create_full(int nV, const double* X0, int nT, const int* tri, int nDBC, const int* dbcIdx,
                            const double* dbcTarget, const double* params, const int* iparams, int nRod, const int* rods,
                            int nPart, const int* parts, const double* codim, int nTet, const int* tets, const double* tetParams) {
    Scene* S = new Scene();
    if (nTet) {
        S->tet.assign(tets, tets + 4 * (size_t)nTet);
        S->P.tetRho = tetParams[0]; S->P.tetE = tetParams[1]; S->P.tetNu = tetParams[2];
    }
    if (codim) {
        S->P.rodRho = codim[0]; S->P.rodE = codim[1]; S->P.rodBendMult = codim[2]; S->P.rodThick = codim[3];
        S->P.partRho = codim[4]; S->P.partThick = codim[5];
    }
    if (nRod) S->rod.assign(rods, rods + 2 * nRod);
    if (nPart) S->particle.assign(parts, parts + nPart);
    Params& P = S->P;
    P.rho = params[0]; P.E = params[1]; P.nu = params[2]; P.thick = params[3]; P.bendMult = params[4];
    P.dt = params[5]; P.dHat = params[6]; P.xi = params[7]; P.PNTol = params[8];
    P.grav[0] = params[9]; P.grav[1] = params[10]; P.grav[2] = params[11];
    P.kappaStiffMult = params[12]; P.pcgTol = params[13];
    P.collision = iparams[0]; P.solver = iparams[1]; P.pcgMaxIter = iparams[2]; P.threads = iparams[3] > 0 ? iparams[3] : 1;
    S->nV = nV;
    S->X0.assign(X0, X0 + 3 * nV);
    S->tri.assign(tri, tri + 3 * nT);
    S->dbcIdx.assign(dbcIdx, dbcIdx + nDBC);
    S->dbcTarget.assign(dbcTarget, dbcTarget + 3 * nDBC);
    scene_init(*S);
    return S;
}
void* orc_scene_create_codim(int nV, const double* X0, int nT, const int* tri, int nDBC, const int* dbcIdx,
                             const double* dbcTarget, const double* params, const int* iparams, int nRod, const int* rods,
                             int nPart, const int* parts, const double* codim) {
    return orc_scene_create_full(nV, X0, nT, tri, nDBC, dbcIdx, dbcTarget, params, iparams, nRod, rods, nPart, parts, codim, 0, nullptr,
                                 nullptr);
}
// 1 when the scene was set up (0: e.g. an inverted rest tet)
int orc_scene_ok(void* h) { return ((Scene*)h)->setupError.empty() ? 1 : 0; }
// ---- tets
int orc_tet_count(void* h) { return (int)((Scene*)h)->tet.size() / 4; }
// surface triangles of the tets (Find_Surface_TriMesh): returns the count, writes up to cap
int orc_tet_surface(void* h, int* out, int cap) {
    std::vector<int> tri;
    tet_surface(*(Scene*)h, tri);
    const int n = (int)tri.size() / 3;
    for (int i = 0; i < 3 * std::min(n, cap); ++i) out[i] = tri[i];
    return n;
}
// rest data: Dm^-1 (9 per tet, row-major), volume; lambda and mu
void orc_tet_arrays(void* h, double* IB, double* vol, double* lamMu) {
    Scene* S = (Scene*)h;
    if (IB) std::memcpy(IB, S->tetIB.data(), sizeof(double) * S->tetIB.size());
    if (vol) std::memcpy(vol, S->tetVol.data(), sizeof(double) * S->tetVol.size());
    if (lamMu) { lamMu[0] = S->tetLam; lamMu[1] = S->tetMu; }
}
// one tet's 12x12 Hessian with weight w (h^2 vol in the step)
void orc_tet_hess_local(void* h, const double* x, int t, double w, int project, double* H) {
    tet_hess_local(*(Scene*)h, x, t, w, project != 0, H);
}
void orc_nh_dPdF(const double* F, double lam, double mu, int project, double* D) { nh_dPdF(F, lam, mu, 1.0, project != 0, D); }
double orc_cubic_root(double a, double b, double c, double d) { return cubic_root(a, b, c, d, 1.e-6); }
double orc_inversion_free_stepsize(void* h, const double* x, const double* p, double alpha) {
    return inversion_free_stepsize(*(Scene*)h, x, p, alpha);
}
void* orc_scene_create(int nV, const double* X0, int nT, const int* tri, int nDBC, const int* dbcIdx,
                       const double* dbcTarget, const double* params, const int* iparams) {
    return orc_scene_create_codim(nV, X0, nT, tri, nDBC, dbcIdx, dbcTarget, params, iparams, 0, nullptr, 0, nullptr, nullptr);
}
// codim sizes: nRod, nRodHinge, nParticle, codimBN0, codimBN1
void orc_codim_sizes(void* h, int* out) {
    Scene* S = (Scene*)h;
    out[0] = (int)S->rod.size() / 2; out[1] = (int)S->rodHinge.size() / 3; out[2] = (int)S->particle.size();
    out[3] = S->codimBN0; out[4] = S->codimBN1;
}
void orc_codim_arrays(void* h, double* rodInfo, int* rodHinge, double* rodHingeInfo) {
    Scene* S = (Scene*)h;
    if (rodInfo) std::memcpy(rodInfo, S->rodInfo.data(), sizeof(double) * S->rodInfo.size());
    if (rodHinge) std::memcpy(rodHinge, S->rodHinge.data(), sizeof(int) * S->rodHinge.size());
    if (rodHingeInfo) std::memcpy(rodHingeInfo, S->rodHingeInfo.data(), sizeof(double) * S->rodHingeInfo.size());
}
void orc_scene_destroy(void* h) { delete (Scene*)h; }
void orc_set_threads(void* h, int t) { ((Scene*)h)->P.threads = t > 0 ? t : 1; }
void orc_set_solver(void* h, int solver, double tol) { ((Scene*)h)->P.solver = solver; ((Scene*)h)->P.pcgTol = tol; }
// the strain-limit stiffness kappa_s[0] as the steps have adapted it (IMPLICIT_EULER.h:630)
double orc_get_kappa_s(void* h) { return ((Scene*)h)->P.slKappa[0]; }
// k > 0: orc_step returns after k Newton iterations (positions left at that iterate)
void orc_set_max_newton(void* h, int k) { ((Scene*)h)->P.maxNewton = k > 0 ? k : 0; }
void orc_set_static_solve(void* h, int on) { ((Scene*)h)->P.staticSolve = on ? 1 : 0; }
// 0: Hessians are assembled from the unprojected local matrices (test switch)
void orc_set_project_pd(void* h, int on) { ((Scene*)h)->P.projectPD = on ? 1 : 0; }
// incremental potential before / after each Newton iteration's line search in the last step
int orc_energy_trace(void* h, double* before, double* after, int cap) {
    Scene* S = (Scene*)h;
    int n = (int)S->eBefore.size();
    for (int i = 0; i < n && i < cap; ++i) { before[i] = S->eBefore[i]; after[i] = S->eAfter[i]; }
    return n;
}
// sizes: nT, nBN, nBE, nBT, nHinge
void orc_scene_sizes(void* h, int* out) {
    Scene* S = (Scene*)h;
    out[0] = S->nT; out[1] = (int)S->bNode.size(); out[2] = (int)S->bEdge.size() / 2;
    out[3] = (int)S->bTri.size() / 3; out[4] = (int)S->hinge.size() / 4;
}
// scalars: kappa, kappaMax, dHat2, lambda, mu, kBend
void orc_scene_scalars(void* h, double* out) {
    Scene* S = (Scene*)h;
    out[0] = S->kappa; out[1] = S->kappaMax; out[2] = S->dHat2; out[3] = S->lam; out[4] = S->mu; out[5] = S->kBend;
}
void orc_scene_arrays(void* h, double* mass, double* Binv, double* vol, int* tri, int* hinge, double* hingeInfo,
                      int* bNode, int* bEdge) {
    Scene* S = (Scene*)h;
    if (mass) std::memcpy(mass, S->mass.data(), sizeof(double) * S->mass.size());
    if (Binv) std::memcpy(Binv, S->Binv.data(), sizeof(double) * S->Binv.size());
    if (vol) std::memcpy(vol, S->vol.data(), sizeof(double) * S->vol.size());
    if (tri) std::memcpy(tri, S->tri.data(), sizeof(int) * S->tri.size());
    if (hinge) std::memcpy(hinge, S->hinge.data(), sizeof(int) * S->hinge.size());
    if (hingeInfo) std::memcpy(hingeInfo, S->hingeInfo.data(), sizeof(double) * S->hingeInfo.size());
    if (bNode) std::memcpy(bNode, S->bNode.data(), sizeof(int) * S->bNode.size());
    if (bEdge) std::memcpy(bEdge, S->bEdge.data(), sizeof(int) * S->bEdge.size());
}
void orc_get_state(void* h, double* x, double* v) {
    Scene* S = (Scene*)h;
    if (x) std::memcpy(x, S->x.data(), sizeof(double) * S->x.size());
    if (v) std::memcpy(v, S->v.data(), sizeof(double) * S->v.size());
}
void orc_set_state(void* h, const double* x, const double* v) {
    Scene* S = (Scene*)h;
    if (x) std::memcpy(S->x.data(), x, sizeof(double) * S->x.size());
    if (v) std::memcpy(S->v.data(), v, sizeof(double) * S->v.size());
}
void orc_set_kappa(void* h, double k) { ((Scene*)h)->kappa = k; }

// constraint set at x: returns count, writes up to cap int4 entries
int orc_constraint_set(void* h, const double* x, int brute, int* out, int cap) {
    std::vector<Constraint> C;
    constraint_set(*(Scene*)h, x, C, brute != 0);
    int n = (int)C.size();
    for (int i = 0; i < n && i < cap; ++i) std::memcpy(out + 4 * i, C[i].c, 4 * sizeof(int));
    return n;
}
static std::vector<Constraint> toC(const int* c, int n) {
    std::vector<Constraint> C(n);
    for (int i = 0; i < n; ++i) std::memcpy(C[i].c, c + 4 * i, 4 * sizeof(int));
    return C;
}
// energies: [membrane, bending, inertia, barrier]; returns 1 if a constraint is non-positive
int orc_energy(void* h, const double* x, const double* xt, const int* c, int nC, double kappa, double* out4) {
    Scene& S = *(Scene*)h;
    double m, b;
    elastic_energy(S, x, &m, &b);
    out4[0] = m; out4[1] = b;
    out4[2] = xt ? inertia_energy(S, x, xt) : 0.0;
    int bad = 0;
    out4[3] = barrier_energy(S, x, toC(c, nC), kappa, &bad);
    return bad;
}
// gradient parts: g_el (elastic), g_bar (barrier)
void orc_gradient(void* h, const double* x, const int* c, int nC, double kappa, double* gel, double* gbar) {
    Scene& S = *(Scene*)h;
    if (gel) { std::memset(gel, 0, sizeof(double) * 3 * S.nV); elastic_grad(S, x, gel); }
    if (gbar) { std::memset(gbar, 0, sizeof(double) * 3 * S.nV); barrier_grad(S, x, toC(c, nC), kappa, gbar); }
}
// ---------------------------------------------------------------- Dirichlet motion
// DBC_MOTION entry over Dirichlet entries [begin, end) (BOUNDARY_CONDITION.h:10-16, 103)
void orc_dirichlet_add_motion(void* h, int begin, int end, const double* v, const double* center, const double* axis,
                              double angVelDeg) {
    Scene::DBCMotion M;
    M.begin = begin; M.end = end; M.angVelDeg = angVelDeg;
    for (int k = 0; k < 3; ++k) { M.v[k] = v[k]; M.center[k] = center[k]; M.axis[k] = axis[k]; }
    ((Scene*)h)->dbcMotion.push_back(M);
}
void orc_step_dirichlet(void* h) { step_dirichlet(*(Scene*)h); }
void orc_get_dirichlet_targets(void* h, double* t) {
    Scene* S = (Scene*)h;
    std::memcpy(t, S->dbcTarget.data(), sizeof(double) * S->dbcTarget.size());
}
void orc_set_dirichlet_targets(void* h, const double* t) {
    Scene* S = (Scene*)h;
    std::memcpy(S->dbcTarget.data(), t, sizeof(double) * S->dbcTarget.size());
}

// ---------------------------------------------------------------- strain limiting
// kappa_s, s, sHat as in the step call (index 0 upper, 1 lower)
void orc_set_strain_limit(void* h, const double* kappa, const double* s, const double* sHat) {
    Params& P = ((Scene*)h)->P;
    for (int i = 0; i < 2; ++i) { P.slKappa[i] = kappa[i]; P.slS[i] = s[i]; P.slSHat[i] = sHat[i]; }
}
// fibre model: k (4) = fiberStiffMult, lim (3) = inextLimit
void orc_set_fiber(void* h, const double* k, const double* lim) {
    Params& P = ((Scene*)h)->P;
    for (int i = 0; i < 4; ++i) P.fiberK[i] = k[i];
    for (int i = 0; i < 3; ++i) P.fiberLimit[i] = lim[i];
}
int orc_fiber_tri(void* h, const double* x, int t, int proj, double* E, double* g, double* H, double* E3) {
    return fiber_tri(*(Scene*)h, x, t, E, g, H, proj != 0, E3) ? 1 : 0;
}
// one triangle: returns 1 if valid; E, g (9), H (81, projected if proj), singular values (2)
int orc_strain_tri(void* h, const double* x, int t, int proj, double* E, double* g, double* H, double* sig) {
    return strain_limit_tri(*(Scene*)h, x, t, E, g, H, proj != 0, sig) ? 1 : 0;
}
// whole mesh: returns 1 if valid; E, gradient (3 nV, overwritten)
int orc_strain_energy_grad(void* h, const double* x, double* E, double* g) {
    Scene& S = *(Scene*)h;
    *E = 0;
    bool ok = strain_limit_energy(S, x, E);
    if (g) { std::memset(g, 0, sizeof(double) * 3 * S.nV); strain_limit_grad(S, x, g); }
    return ok ? 1 : 0;
}
// local 3x3 blocks of the projected Hessian; returns the count (9 per triangle)
int orc_strain_hess(void* h, const double* x, int* rows, int* cols, double* vals) {
    std::vector<Block> B;
    strain_limit_hess(*(Scene*)h, x, B);
    for (size_t i = 0; i < B.size(); ++i) {
        rows[i] = B[i].r; cols[i] = B[i].c;
        std::memcpy(vals + 9 * i, B[i].b, 9 * sizeof(double));
    }
    return (int)B.size();
}

// ---------------------------------------------------------------- friction
void orc_set_friction(void* h, double mu, double epsv2, int iterAmt) {
    Params& P = ((Scene*)h)->P;
    P.fricMu = mu; P.epsv2 = epsv2; P.fricIterAmt = iterAmt;
}
void orc_set_friction_components(void* h, int n, const int* range, const double* muComp) {
    Params& P = ((Scene*)h)->P;
    P.compNodeRange.assign(range, range + n);
    P.muComp.assign(muComp, muComp + (size_t)n * n);
}
static FricSet toF(const int* c, int nF, const double* info) {
    FricSet F;
    F.C = toC(c, nF);
    F.info.assign(info, info + 9 * nF);
    return F;
}
// friction stencils of constraint set c at x: returns count, writes up to cap entries
// (4 ints and 9 doubles each)
int orc_friction_basis(void* h, const double* x, const int* c, int nC, double kappa, int* outC, double* outInfo, int cap) {
    FricSet F;
    friction_basis(*(Scene*)h, x, toC(c, nC), kappa, F);
    int n = (int)F.C.size();
    for (int i = 0; i < n && i < cap; ++i) {
        std::memcpy(outC + 4 * i, F.C[i].c, 4 * sizeof(int));
        std::memcpy(outInfo + 9 * i, &F.info[9 * i], 9 * sizeof(double));
    }
    return n;
}
double orc_friction_energy(void* h, const double* x, const double* xn, const int* c, int nF, const double* info) {
    return friction_energy(*(Scene*)h, x, xn, toF(c, nF, info));
}
void orc_friction_grad(void* h, const double* x, const double* xn, const int* c, int nF, const double* info, double* g) {
    Scene& S = *(Scene*)h;
    std::memset(g, 0, sizeof(double) * 3 * S.nV);
    friction_grad(S, x, xn, toF(c, nF, info), g);
}
// local blocks (row vertex, col vertex, 9 doubles); returns the count (<= 16 nF)
int orc_friction_hess(void* h, const double* x, const double* xn, const int* c, int nF, const double* info, int* rows,
                      int* cols, double* vals) {
    std::vector<Block> B;
    friction_hess(*(Scene*)h, x, xn, toF(c, nF, info), B);
    for (size_t i = 0; i < B.size(); ++i) {
        rows[i] = B[i].r; cols[i] = B[i].c;
        std::memcpy(vals + 9 * i, B[i].b, 9 * sizeof(double));
    }
    return (int)B.size();
}

// assembled Hessian (elastic + barrier + M, DBC projected) in CSR; call with rp only to get nnz
static CSR g_lastA;
int orc_hessian(void* h, const double* x, const int* c, int nC, double kappa, int which) {
    Scene& S = *(Scene*)h;
    std::vector<Block> blocks;
    if (which & 1) {
        elastic_hess(S, x, blocks);
        strain_limit_hess(S, x, blocks);
        if (fiber_active(S)) fiber_hess(S, x, blocks);
    }
    if (which & 2) barrier_hess(S, x, toC(c, nC), kappa, blocks);
    assemble(S, blocks, g_lastA);
    return (int)g_lastA.val.size();
}
void orc_hessian_fetch(int* rp, int* ci, double* val) {
    std::memcpy(rp, g_lastA.rp.data(), sizeof(int) * g_lastA.rp.size());
    std::memcpy(ci, g_lastA.ci.data(), sizeof(int) * g_lastA.ci.size());
    std::memcpy(val, g_lastA.val.data(), sizeof(double) * g_lastA.val.size());
}
// local barrier Hessian blocks before assembly (for per-constraint parity)
int orc_barrier_blocks(void* h, const double* x, const int* c, int nC, double kappa, int* rc, double* blk, int cap) {
    Scene& S = *(Scene*)h;
    std::vector<Block> blocks;
    barrier_hess(S, x, toC(c, nC), kappa, blocks);
    int n = (int)blocks.size();
    for (int i = 0; i < n && i < cap; ++i) {
        rc[2 * i] = blocks[i].r; rc[2 * i + 1] = blocks[i].c;
        std::memcpy(blk + 9 * i, blocks[i].b, 9 * sizeof(double));
    }
    return n;
}
void orc_dist2_list(void* h, const double* x, const int* c, int nC, double* d2) {
    std::vector<double> d;
    constraint_dist2(*(Scene*)h, x, toC(c, nC), d);
    std::memcpy(d2, d.data(), sizeof(double) * d.size());
}
double orc_ccd(void* h, const double* x, const double* p, double alpha, double* clamped) {
    return ccd_stepsize(*(Scene*)h, x, p, alpha, false, clamped);
}
int orc_solve(int n, const int* rp, const int* ci, const double* val, const double* b, double* x, int solver, double tol,
              int* iters) {
    CSR A; A.n = n;
    A.rp.assign(rp, rp + n + 1);
    A.ci.assign(ci, ci + rp[n]);
    A.val.assign(val, val + rp[n]);
    if (solver == 0) return solve_direct(A, b, x);
    return solve_pcg(A, b, x, tol, 1000000, iters, 4);
}
// stats: newtonIters, nContacts, pcgIters, useGD, lsTrials, kappa, lastRes, E, tHess, tSolve, tCS, tCCD, tEnergy, tGrad, tTotal
int orc_step(void* h, double* stats) {
    StepStats st;
    int rc = advance_one_step(*(Scene*)h, st);
    if (stats) {
        const Scene& S = *(Scene*)h;
        double s[17] = {(double)st.newtonIters, (double)st.nContacts, (double)st.pcgIters, (double)st.useGDCount, (double)st.lsTrials,
                        st.kappa, st.lastRes, st.E, st.tHessian, st.tSolve, st.tCS, st.tCCD, st.tEnergy, st.tGrad, st.tTotal,
                        S.P.slKappa[0], (double)st.kappaSDoublings};
        std::memcpy(stats, s, sizeof(s));
    }
    return rc;
}

}  // extern "C"
