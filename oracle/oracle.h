// CPU FP64 oracle for the C-IPC contact-and-Newton hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain C++17 restatement of the reference
// algorithm (ipc-sim/Codim-IPC, Library/FEM/Shell/IMPLICIT_EULER.h and the files it
// calls).  It is the checker for the HIP path and the CPU baseline; nothing in the
// product links or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.
//
// Every function names the reference file:line it restates.  Small fixed-size
// linear algebra is written here (no Eigen).  Compiled with -ffp-contract=off so
// that classification predicates round exactly like the HIP kernels.
//
// Parity status of each part is listed in the header comment of the file that
// implements it.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>
#include <array>
#include <string>

namespace orc {

// ---------------------------------------------------------------- vec3
struct V3 { double x, y, z; };
inline V3 v3(const double* p) { return {p[0], p[1], p[2]}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double sq(V3 a) { return dot(a, a); }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---------------------------------------------------------------- distances
// squared distances and region classifiers (Math/Distance/*.h)
double dist2_pp(V3 a, V3 b);
double dist2_pe(V3 p, V3 e0, V3 e1);
double dist2_pt(V3 p, V3 t0, V3 t1, V3 t2);
double dist2_ee(V3 a0, V3 a1, V3 b0, V3 b1);
int pe_type(V3 p, V3 e0, V3 e1);
int pt_type(V3 p, V3 t0, V3 t1, V3 t2);
int ee_type(V3 a0, V3 a1, V3 b0, V3 b1);
double dist2_pt_unclassified(V3 p, V3 t0, V3 t1, V3 t2);
double dist2_ee_unclassified(V3 a0, V3 a1, V3 b0, V3 b1);
double dist2_pe_unclassified(V3 p, V3 e0, V3 e1);

// derivatives: g has 3*n entries, H is (3n)x(3n) row-major, n = 2,3,4 vertices
void dist2_pp_gh(V3 a, V3 b, double* g, double* H);
void dist2_pe_gh(V3 p, V3 e0, V3 e1, double* g, double* H);
void dist2_pt_gh(V3 p, V3 t0, V3 t1, V3 t2, double* g, double* H);
void dist2_ee_gh(V3 a0, V3 a1, V3 b0, V3 b1, double* g, double* H);

// edge-edge mollifier (Math/Distance/EDGE_EDGE_MOLLIFIER.h)
double ee_cross_norm2(V3 a0, V3 a1, V3 b0, V3 b1);
double ee_mollifier_threshold(V3 a0r, V3 a1r, V3 b0r, V3 b1r);
// e, 12-gradient, 12x12 Hessian (gradient/Hessian zero when c >= eps)
void ee_mollifier(V3 a0, V3 a1, V3 b0, V3 b1, double eps, double* e, double* g, double* H);

// log barrier, Math/BARRIER.h:10-62 (non-elastic branch)
double barrier(double d, double dhat, double kappa);
double barrier_g(double d, double dhat, double kappa);
double barrier_h(double d, double dhat, double kappa);

// symmetric eigen-decomposition (Householder tridiagonalisation + implicit QL)
// eigenvalues ascending in w, eigenvectors in columns of V (row-major n x n)
void sym_eig(int n, const double* A, double* w, double* V);
// Math/UTILS.h:10-27: clamp negative eigenvalues to zero
void make_pd(int n, double* A);

// ---------------------------------------------------------------- elasticity
// Neo-Hookean membrane, FEM/Shell/MEMBRANE.h:7-290 (useNH = true)
double membrane_energy(V3 x0, V3 x1, V3 x2, const double* B, double vol, double lam, double mu);
void membrane_grad(V3 x0, V3 x1, V3 x2, const double* B, double vol, double lam, double mu, double* g);
void membrane_hess(V3 x0, V3 x1, V3 x2, const double* B, double vol, double lam, double mu, double* H);

// dihedral angle, Math/DIHEDRAL_ANGLE.h:10-24 (angle and sign convention)
double dihedral_angle(V3 x0, V3 x1, V3 x2, V3 x3);
void dihedral_angle_grad(V3 x0, V3 x1, V3 x2, V3 x3, double* g);
void dihedral_angle_hess(V3 x0, V3 x1, V3 x2, V3 x3, double* H);

// ---------------------------------------------------------------- codimensional
// rod segment spring, FEM/Shell/Rod/ROD_STRETCHING.h:10-107 (without the h^2 pi t^2/4
// weight): E = Er / (2 lbar) (|x0 - x1| - lbar)^2, g 6, H 6x6 (not projected)
double rod_spring_energy(V3 x0, V3 x1, double Er, double lbar);
void rod_spring_grad(V3 x0, V3 x1, double Er, double lbar, double* g);
void rod_spring_hess(V3 x0, V3 x1, double Er, double lbar, double* H);
// squared discrete curvature of a rod hinge, FEM/Shell/Rod/ROD_BENDING.h:10-117:
// kappa = 2 (e0 x e1) / (|e0| |e1| + e0.e1), e0 = x1 - x0, e1 = x2 - x1; returns |kappa|^2,
// g 9, H 9x9 (not projected)
double rod_kappa2(V3 x0, V3 x1, V3 x2);
void rod_kappa2_grad(V3 x0, V3 x1, V3 x2, double* g);
void rod_kappa2_hess(V3 x0, V3 x1, V3 x2, double* H);

// ---------------------------------------------------------------- scene
struct Params {
    double rho = 200, E = 1e6, nu = 0.3, thick = 1e-3, bendMult = 1;
    double dt = 0.01, dHat = 1e-3, xi = 0, PNTol = 1e-2;
    double grav[3] = {0, -9.81, 0};
    double kappaStiffMult = 1;
    // rods (Initialize_Discrete_Rod) and particles (Initialize_Discrete_Particle)
    double rodRho = 0, rodE = 0, rodBendMult = 1, rodThick = 1e-3;
    double partRho = 0, partThick = 1e-3;
    // volumetric tets (add_object_3D + initialize_added_objects): density, E, nu
    double tetRho = 0, tetE = 0, tetNu = 0.3;
    int collision = 1;
    // lagged friction (Advance_One_Step_IE_Hinge's mu, epsv2, fricIterAmt)
    double fricMu = 0, epsv2 = 1e-6;
    int fricIterAmt = 1;
    // per-component friction (compNodeRange: exclusive end vertex of each component;
    // muComp: nComp x nComp coefficients, entry c0 + c1 nComp), FRICTION.h:114-171
    std::vector<int> compNodeRange;
    std::vector<double> muComp;
    // strain limiting (the step's kappa_s, s, sHat; upper limit index 0, lower 1)
    double slKappa[2] = {0, 0}, slS[2] = {1.01, 0}, slSHat[2] = {1, 1};
    // anisotropic (fibre) membrane: the step's fiberStiffMult (4) and inextLimit (3)
    double fiberK[4] = {0, 0, 0, 0}, fiberLimit[3] = {1, 1, 1};
    int solver = 0;            // 0 = direct (sparse Cholesky), 1 = block-Jacobi PCG
    double pcgTol = 1e-5;
    int pcgMaxIter = 100000;
    int threads = 1;
    int maxNewton = 0;         // > 0: return after this many Newton iterations (a timed prefix of a step)
    int projectPD = 1;         // 0: local Hessians assembled unprojected (finite-difference pins of the sums)
    int staticSolve = 0;       // the step's staticSolve argument
};

struct Scene;
// weight of the shell's membrane, bending, strain-limit and fibre terms: h^2, 1 in a static
// solve (staticSolve ? 1.0 : h, INC_POTENTIAL.h:90-94)
double shell_h2(const Scene& S);

struct Scene {
    Params P;
    int nV = 0, nT = 0;
    std::vector<double> X0;            // rest positions (3 nV) = nodeAttr.x0
    std::vector<double> x, v;          // state
    std::vector<int> tri;              // 3 nT (after degenerate filter)
    std::vector<char> isDBC;
    std::vector<int> dbcIdx;
    std::vector<double> dbcTarget;     // 3 per DBC entry
    // derived (DISCRETE_SHELL.h:218-357, 555-577; MESHIO.h:728-835)
    std::vector<double> mass;          // lumped, per vertex
    std::vector<double> Binv;          // inverse rest first fundamental form (3 per tri: 00, 01, 11)
    std::vector<double> vol;           // area * thickness
    double lam = 0, mu = 0, kBend = 0;
    std::vector<int> hinge;            // 4 per hinge
    std::vector<double> hingeInfo;     // theta_bar, |e_bar|, h_bar
    std::vector<int> bNode, bEdge, bTri;
    // codimensional primitives (DISCRETE_ROD.h:87-165, DISCRETE_PARTICLE.h:52-75)
    std::vector<int> rod;              // 2 per segment
    std::vector<double> rodInfo;       // E, rest length, thickness per segment
    std::vector<int> rodHinge;         // 3 per rod hinge (x0, x1 = shared node, x2)
    std::vector<double> rodHingeInfo;  // E * bend mult, rest |e0| + |e1|, thickness
    std::vector<int> particle;
    // tets (FEM/DEFORMATION_GRADIENT.h:39-77): 4 per tet, Dm^-1 row-major (9 per tet), rest volume
    std::vector<int> tet;
    std::vector<double> tetIB, tetVol;
    double tetLam = 0, tetMu = 0;
    int codimBN0 = 0, codimBN1 = 0;    // boundary-node slots where rod nodes / particles start
    double kappa = 0, kappaMax = 0, dHat2 = 0;
    double bboxDiag = 0;
    std::vector<double> slIB;          // strain limiting: B^-1 of the rest metric (00, 01, 11 per tri)
    std::vector<double> fiberIB;       // fibre model: B^-1 of the rest x-z coordinates (00, 01, 10, 11 per tri)
    // Dirichlet motion groups (FEM/BOUNDARY_CONDITION.h DBC_MOTION): entries [begin, end)
    struct DBCMotion { int begin, end; double v[3], center[3], axis[3], angVelDeg; };
    std::vector<DBCMotion> dbcMotion;
    // per Newton iteration of the last step: incremental potential before and after the
    // line search (the accepted iterate)
    std::vector<double> eBefore, eAfter;
    std::string setupError;            // non-empty when the scene could not be set up
};
// Step_Dirichlet (BOUNDARY_CONDITION.h:107-155): rotate then translate every group's targets
void step_dirichlet(Scene& S);

// constraint: 4 ints in the reference encoding (FEM/IPC.h:144-661)
struct Constraint { int c[4]; };
// local Hessian blocks, appended as (row vertex, col vertex, 3x3 block)
struct Block { int r, c; double b[9]; };

void scene_init(Scene& S);                       // A22
// rods and particles: mass, rod hinges, boundary primitives (called by scene_init)
void codim_init(Scene& S);
// rod springs added to *stretch, rod hinges to *bend; gradient accumulated into g;
// local blocks appended to out
void codim_energy(const Scene& S, const double* x, double* stretch, double* bend);
void codim_grad(const Scene& S, const double* x, double* g);
void codim_hess(const Scene& S, const double* x, std::vector<Block>& out);
// contact
void constraint_set(const Scene& S, const double* x, std::vector<Constraint>& out,
                    bool bruteForce = false);
double barrier_energy(const Scene& S, const double* x, const std::vector<Constraint>& C, double kappa,
                      int* bad = nullptr);
void barrier_grad(const Scene& S, const double* x, const std::vector<Constraint>& C, double kappa, double* g);
void barrier_hess(const Scene& S, const double* x, const std::vector<Constraint>& C, double kappa,
                  std::vector<Block>& out);
void constraint_dist2(const Scene& S, const double* x, const std::vector<Constraint>& C, std::vector<double>& d2);
// lagged friction (FEM/FRICTION.h): stencils, and per stencil 9 doubles: closest-point
// coordinates (2), tangent basis columns (2 x 3), normal force
struct FricSet {
    std::vector<Constraint> C;
    std::vector<double> info;
};
void friction_basis(const Scene& S, const double* x, const std::vector<Constraint>& C, double kappa, FricSet& F);
bool friction_by_component(const Scene& S);
double friction_energy(const Scene& S, const double* x, const double* xn, const FricSet& F);
void friction_grad(const Scene& S, const double* x, const double* xn, const FricSet& F, double* g);
void friction_hess(const Scene& S, const double* x, const double* xn, const FricSet& F, std::vector<Block>& out);
double ccd_stepsize(const Scene& S, const double* x, const double* p, double alpha, bool bruteForce = false,
                    double* clampedAlpha = nullptr);
// elastic + inertia
double elastic_energy(const Scene& S, const double* x, double* memb = nullptr, double* bend = nullptr);
void elastic_grad(const Scene& S, const double* x, double* g);
void elastic_hess(const Scene& S, const double* x, std::vector<Block>& out);
double inertia_energy(const Scene& S, const double* x, const double* xt);

// strain limiting (FEM/Shell/INEXT.h); energy adds to *E and returns false when a
// singular value reaches a limit
void strain_limit_setup(Scene& S);
bool strain_limit_energy(const Scene& S, const double* x, double* E);
void strain_limit_grad(const Scene& S, const double* x, double* g);
void strain_limit_hess(const Scene& S, const double* x, std::vector<Block>& out);
bool strain_limit_tri(const Scene& S, const double* x, int t, double* E, double* g9, double* H81, bool proj, double* sig);
void strain_max_sv(const Scene& S, const double* x, std::vector<double>& smax);

// ---------------------------------------------------------------- volumetric tets (tet.cpp)
bool tet_setup(Scene& S);                                   // rest basis, volume, lumped mass; false on an inverted rest tet
void tet_surface(const Scene& S, std::vector<int>& tri);    // Find_Surface_TriMesh
bool tets_valid(const Scene& S, const double* x);           // every det F > 0
double tet_energy(const Scene& S, const double* x);         // h^2 sum vol psi (valid states only)
void tet_grad(const Scene& S, const double* x, double* g);  // g +=
void tet_hess(const Scene& S, const double* x, std::vector<Block>& out);
void tet_hess_local(const Scene& S, const double* x, int t, double w, bool project, double* H144);
void nh_dPdF(const double* F, double lam, double mu, double w, bool project, double* D81);
double cubic_root(double a, double b, double c, double d, double tol);
double inversion_free_stepsize(const Scene& S, const double* x, const double* p, double alpha);
// anisotropic (fibre) membrane (FEM/Shell/ANISO_INEXT.h), replaces the Neo-Hookean membrane
// when fiberK[0] > 0 or fiberK[1] > 0; energy adds to *E and returns false when infeasible
void fiber_setup(Scene& S);
bool fiber_active(const Scene& S);
bool fiber_energy(const Scene& S, const double* x, double* E);
void fiber_grad(const Scene& S, const double* x, double* g);
void fiber_hess(const Scene& S, const double* x, std::vector<Block>& out);
bool fiber_tri(const Scene& S, const double* x, int t, double* E, double* g9, double* H81, bool proj, double* E3);

// assembled system in CSR (scalar) form
struct CSR { int n = 0; std::vector<int> rp, ci; std::vector<double> val; };
// + M, projection of the rows/columns flagged in proj (default: every Dirichlet node)
void assemble(const Scene& S, const std::vector<Block>& blocks, CSR& A, const std::vector<char>* proj = nullptr);
int solve_direct(const CSR& A, const double* b, double* x);                // 0 ok, 1 not SPD
int solve_pcg(const CSR& A, const double* b, double* x, double tol, int maxIter, int* iters, int threads = 1);

struct StepStats {
    int newtonIters = 0, nContacts = 0, pcgIters = 0, useGDCount = 0, lsTrials = 0;
    double kappa = 0, lastRes = 0, E = 0;
    double tHessian = 0, tSolve = 0, tCS = 0, tCCD = 0, tEnergy = 0, tGrad = 0, tTotal = 0;
    int kappaSDoublings = 0;
};
int advance_one_step(Scene& S, StepStats& st);

}  // namespace orc
