// Device-resident scene state and the host-side stages of one implicit-Euler step.
//
// Layout in HBM: per-vertex vectors (positions, velocities, gradients, Newton
// directions, Krylov vectors) are DOF-ordered double[3 nV] (x0 y0 z0 x1 ...), the order
// of the linear system and of the 3x3 BSR blocks.  Element and stencil attributes are
// structure-of-arrays (tri vertex k of all triangles contiguous, inverse rest metric
// component k contiguous, ...), so stencil kernels read them coalesced.
#pragma once
#include "common.hpp"
#include <array>
#include <memory>
#include <vector>
#include "../../include/cipc.h"

namespace cipc {

struct Params {
    double rho = 200, E = 1e6, nu = 0.3, thick = 1e-3, bendMult = 1;
    double dt = 0.01, dHat = 1e-3, xi = 0, PNTol = 1e-2;
    double grav[3] = {0, -9.81, 0};
    double kappaStiffMult = 1;
    double pcgTol = 1e-5;
    int pcgMaxIter = 100000;
    int collision = 1;
    int precision = 0;            // 0: FP64 PCG; 1: FP32 blocks + FP64 vectors; 2: all-FP32 PCG (1, 2 inside FP64 refinement)
    // 0: PCG, with the dense direct solve taking over a solve that reaches pcgMaxIter when
    // the system has at most kDirectMaxDof unknowns; 1: dense direct solve for every solve
    int linSolver = 0;
    // rods and particles
    double rodRho = 0, rodE = 0, rodBendMult = 1, rodThick = 1e-3;
    // volumetric tets (density, Young's modulus, Poisson ratio)
    double tetRho = 0, tetE = 0, tetNu = 0.3;
    // the step's staticSolve argument: shell terms weighted 1 instead of h^2, the body
    // force -x.(m g) instead of inertia, no velocity update (IMPLICIT_EULER.h:196-217,
    // 646-666, 893-899; INC_POTENTIAL.h:118-130, 233-249, 383-386)
    int staticSolve = 0;
    double partRho = 0, partThick = 1e-3;
    // lagged friction (Advance_One_Step_IE_Hinge's mu, epsv2, fricIterAmt)
    double fricMu = 0, epsv2 = 1e-6;
    int fricIterAmt = 1;
    // strain limiting (the step's kappa_s, s, sHat; index 0 upper limit, 1 lower)
    double slKappa[2] = {0, 0}, slS[2] = {1.01, 0}, slSHat[2] = {1, 1};
    // anisotropic (fibre) membrane: the step's fiberStiffMult and inextLimit
    double fiberK[4] = {0, 0, 0, 0}, fiberLimit[3] = {1, 1, 1};
};

// spatial grid used by the constraint-set broad phase and the CCD candidate search
// A primitive whose box covers more grid cells than this is not entered in the cell
// lists; queries test it directly from a compact list instead.
constexpr uint32_t kLargeCells = 512;

struct LearnedPrecond;

// Two-level correction of the learned preconditioner (coarse.hip): rigid-body modes
// (three translations, three rotations about the centroid) on aggregates of the elastic
// vertex graph, Galerkin coarse matrix P^T H P factorised densely, its inverse applied as
// two triangular products per PCG iteration.
constexpr int kCoarseMaxDof = 12288;      // largest coarse problem (6 per aggregate)
constexpr int kCoarseMaxNb = 256;         // neighbour aggregates of one aggregate (Galerkin row)
constexpr int kCoarseMaxVerts = 4096;     // vertices of one aggregate
struct CoarseSpace {
    bool built = false;                   // aggregates of the current elastic pattern
    bool ready = false;                   // factor valid for the current Hessian
    int nAgg = 0, nc = 0, npad = 0, nMembers = 0;
    int nnzbBuilt = -1, nVBuilt = -1;
    DevBuf<int> agg, aggPtr, aggVerts;    // per vertex aggregate (-1: not in the coarse space), CSR of members
    DevBuf<double> cent;                  // per aggregate: centroid xyz, 1 / rms radius
    DevBuf<double> d;                     // per vertex: (x - centroid) / rms radius
    DevBuf<double> dm;                    // the same per member, in aggVerts order (restriction reads it in sequence)
    DevBuf<double> A;                     // dense coarse matrix, then its Cholesky factor, then L^-1 (npad^2)
    DevBuf<double> W;                     // dense work: L^-1 (npad^2)
    DevBuf<double> diag0, tileInv;        // coarse diagonal before factorisation; inverses of the 64x64 diagonal tiles
    DevBuf<int> dropped;                  // coarse unknowns whose pivot vanished (degenerate modes)
    DevBuf<float> W32, WT32;              // L^-1 and its transpose in FP32, row-major npad^2
    DevBuf<uint16_t> W16, WT16;           // the same in BF16 (bf16 storage)
    bool bf16 = false;
    DevBuf<double> rc, wc, zc;            // coarse vectors
    DevBuf<int> flag;                     // Galerkin overflow flag
    int maxDof = 0;                       // largest coarse problem asked for (0: kCoarseMaxDof)
    int builtMaxDof = -1;
    int every = 1, sinceSetup = 0;        // set-up at a step's first solve and then every `every` solves
    bool stepStart = true;
    double lastSetupMs = 0;
    long setups = 0, failures = 0;        // set-ups, and those that found an aggregate with too many neighbours
};

// largest system (unknowns) of the dense direct solve: 24576^2 FP64 = 4.5 GiB of matrix
constexpr int kDirectMaxDof = 24576;

// an instantiated graph of one 16-iteration PCG chunk and the launch arguments it was
// captured with (replayed while they stay the same)
struct PcgGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<uintptr_t> key;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
    PcgGraph() = default;
    PcgGraph(const PcgGraph&) = delete;
    PcgGraph& operator=(const PcgGraph&) = delete;
    ~PcgGraph() { reset(); }
};

struct GridParams {
    double lb[3];
    double inv;        // 1 / cell size
    int dim[3];
};

// sparse symmetric matrix in 3x3 block CSR with structure-of-arrays block storage:
// val[k * nnzb + b] is entry k (row-major in the block) of block b
struct BSR {
    int nrow = 0, nnzb = 0;
    DevBuf<int> rowptr, col, rowOf;
    DevBuf<double> val;
};

// Dirichlet motion group (FEM/BOUNDARY_CONDITION.h DBC_MOTION): entries [begin, end)
struct DBCMotion {
    int begin, end;
    double v[3], center[3], axis[3], angVelDeg;
};

struct Ctx {
    hipStream_t s = nullptr;
    PrimWork pw;
    Params P;
    std::string lastError;

    int nV = 0, nT = 0, nH = 0, nBN = 0, nBE = 0, nDBC = 0;
    double kappa = 0, kappaMax = 0, dHat2 = 0, lam = 0, mu = 0, kBend = 0, cellSize = 0;

    // host mirrors of the topology
    std::vector<int> hTri, hHinge, hBNode, hBEdge, hDbcIdx;
    std::vector<double> hDbcTarget, hMass, hHingeInfo, hBinv, hVol, hSlIB, hFbIB;
    // codimensional primitives: rod segments are the last nRod boundary edges, rod nodes
    // occupy boundary-node slots [codimBN0, codimBN1), particles [codimBN1, nBN)
    int nRod = 0, nRH = 0, nPart = 0, codimBN0 = 0, codimBN1 = 0;
    std::vector<int> hRod, hRodHinge, hPart;
    std::vector<double> hRodInfo, hRodHingeInfo;
    DevBuf<int> rod;          // SoA [2][nRod]
    DevBuf<double> rodPar;    // SoA [3][nRod]: E, rest length, weight h^2 pi t^2 / 4
    DevBuf<int> rodHinge;     // SoA [3][nRH]
    DevBuf<double> rodHW;     // bending weight h^2 E t^4 pi / (64 lsum)
    DevBuf<int> partSlot;     // particles as point primitives (vertex ids, [1][nPart])
    // volumetric tets: SoA [4][nTet] vertices, Dm^-1 SoA [9][nTet] (row-major entries), rest
    // volumes; contact triangles (the tets' surface first, then the shell) SoA [3][nBT]
    int nTet = 0, nBT = 0;
    double tetLam = 0, tetMu = 0;
    std::vector<int> hTet, hBTri;
    DevBuf<int> tetIdx, btri;
    DevBuf<double> tetIB, tetVol;
    DevBuf<int> pLo;
    DevBuf<uint64_t> pKeys;
    DevBuf<uint32_t> pVals, pcellStart, pcellEnd;
    DevBuf<uint8_t> lgFlagP;
    DevBuf<int> lgListP;
    DevBuf<int> ccdPlo, ccdPhi;
    DevBuf<int> ccdTbox, ccdEbox, ccdPbox;   // the same voxel boxes as 32-byte records (lo xyz, hi xyz, 2 pad)
    DevBuf<int> ccdTrec, ccdErec;            // those records in cell order (third pad word = primitive)

    // device state
    DevBuf<double> x, xn, xt, v, xprev, X0, mass, p, g, rhs;
    DevBuf<uint8_t> dbc;
    DevBuf<int> tri;          // SoA [3][nT]
    DevBuf<double> Binv;      // SoA [3][nT]
    DevBuf<double> slIB;      // SoA [3][nT]: B^-1 of the rest metric (strain limiting)
    DevBuf<double> fbIB;      // SoA [4][nT]: B^-1 of the rest x-z coordinates (fibre model)
    DevBuf<double> vol;
    DevBuf<int> hinge;        // SoA [4][nH]
    DevBuf<double> hInfo;     // SoA [3][nH]
    DevBuf<int> bNode, bEdge; // bEdge SoA [2][nBE]

    // constraint set (reference int4 encoding, SoA [4][nC])
    int nC = 0, nCprev = 0;
    DevBuf<int> C, Cprev;
    DevBuf<double> d2prev;

    // scratch
    DevBuf<double> dscr, dscr2, red, scal;
    DevBuf<uint32_t> u32a, u32b, u32c, u32d, counters;
    DevBuf<uint64_t> k64a, k64b;
    DevBuf<int> i32a, i32b, i32c;
    DevBuf<uint32_t> cellStart, cellEnd, ecellStart, ecellEnd;
    uint32_t csCap = 1u << 16;
    DevBuf<uint64_t> csKey, csDKey;
    DevBuf<int> csC4;
    DevBuf<uint32_t> csPos, csIdx, csHead;
    DevBuf<int> triLo, edgeLo;
    DevBuf<int> ccdTlo, ccdThi, ccdElo, ccdEhi;
    // primitives spanning more than kLargeCells grid cells are kept out of the cell
    // lists and tested by every query directly (flag per primitive + compact list)
    DevBuf<uint8_t> lgFlagT, lgFlagE;
    DevBuf<int> lgListT, lgListE;

    // elastic Hessian: static BSR pattern + gather plan
    BSR Ael;
    DevBuf<int> planOff, planSlot;     // per BSR block: list of local block slots
    DevBuf<double> elLocal;            // local 3x3 blocks, [nSlots][9]
    DevBuf<double> hingeH;             // hinge Hessians before projection, [nH][12][12]
    int nElSlots = 0, nTriSlots = 0;
    DevBuf<int> gplanOff, gplanSlot;   // per vertex: gradient slots (tri 3 per, hinge 4 per)
    DevBuf<double> gLocal;             // local gradient 3-vectors, [nGSlots][3]
    int nGSlots = 0;

    // contact Hessian (rebuilt per Newton iteration)
    BSR Acon;
    DevBuf<double> conLocal;           // SoA [9][16 nC]
    DevBuf<double> conH;               // local barrier Hessians before projection, [nC][12][12]
    DevBuf<int> conNv;                 // vertices of each constraint's Hessian (0: not evaluated)
    DevBuf<int> conPlanSlot;           // sorted slot order
    DevBuf<double> conGLocal;          // SoA [3][4 nC]
    DevBuf<int> conGVert;              // vertex of each gradient slot (4 nC; -1 unused)

    // Dirichlet targets resident on the device, motion groups, and the penalty state of a
    // step whose Dirichlet motion the CCD blocked (dbcStiff > 0: only dbcFixed rows projected)
    bool dbcDevReady = false;
    DevBuf<int> dbcIdxD;
    DevBuf<double> dbcTgt;             // 3 per Dirichlet entry
    std::vector<DBCMotion> dbcMotion;
    DevBuf<uint8_t> dbcFixed;          // per vertex: Dirichlet with zero prescribed motion
    double dbcStiff = 0;

    // kappa_s rule: per triangle "largest singular value near the upper limit" at the
    // current and the previous Newton iterate (ping-pong)
    DevBuf<uint8_t> slNear0, slNear1;
    int slNearCur = 0;
    int kappaSDoublings = 0;           // in the last step

    // lagged friction stencils (FEM/FRICTION.h): SoA [4][nF] encoding, SoA [9][nF] data
    // (closest point 2, tangent basis columns 2 x 3, normal force)
    int nF = 0;
    int nFricComp = 0;                 // per-component coefficients (0: one mu for all)
    DevBuf<int> fricCompRange;         // exclusive end vertex of each component
    DevBuf<double> fricMuComp;         // nFricComp^2, entry c0 + c1 nFricComp
    bool fricOn = false;               // inside a step with mu > 0: potentials include friction
    DevBuf<int> fC;
    DevBuf<double> fI;
    DevBuf<double> fGLocal;            // SoA [3][4 nF]
    DevBuf<int> fGVert;

    // PCG
    DevBuf<double> pr, pz, pp, pq, pdinv, psol;
    // mixed forms: FP32 copies of the matrix values, FP32 Krylov vectors (fp32 form), FP64 residual
    DevBuf<float> valE32, valK32, fr, fz, fp, fq, fdinv, fx, fb;
    DevBuf<double> rres, xacc;
    int lastRefinements = 0;
    // learned preconditioner (created on first use); 0 = block Jacobi, 1 = learned
    std::shared_ptr<LearnedPrecond> lp;
    DevBuf<double> lpTmp;
    int precond = 0;
    int lpBudget = 3000;              // PCG iterations before the learned solve falls back
    int lastPrecondFallback = 0;
    long precondSolves = 0, precondFallbacks = 0;   // learned solves and how many fell back
    int lastLearnedRc = 0, lastLearnedIters = 0;     // outcome of the last learned attempt
    // the network runs at the first Newton solve of a step and then every lpEvery solves
    // (default: every solve -- in contact the Hessian changes enough between Newton
    // iterations that a fresh factor saves far more PCG iterations than the forward pass
    // costs); in between, the last factor is reused with the diagonal scaling of the new
    // Hessian
    int lpEvery = 1, lpSinceForward = 0;
    bool lpStepStart = true;
    // after a fallback the rest of the step's solves use block Jacobi; a learned attempt
    // is also cut off at 1.25x the iterations of the last block-Jacobi solve
    bool lpOffThisStep = false;
    int lastJacobiIters = 0;          // iterations of the last from-zero block-Jacobi solve
    bool solveWarm = false;           // inside the correction solve of a warm-started Newton solve
    int lastFallbackContinued = 0;    // the last fallback continued from the learned iterate
    // Newton solves after a line search that cut the step to alpha < 1 start from
    // (1 - alpha) p_prev, the quadratic model's prediction of the next direction
    bool warmStart = true;
    double lastWarmRatio = 1.0;       // ||b - H x0|| / ||b|| of the last warm-started solve
    DevBuf<double> warmR, warmX;
    DevBuf<double> fallbackX, fallbackR;   // learned iterate and its residual when block Jacobi takes over
    // two-level correction of the learned preconditioner (cipc_set_coarse)
    bool coarseOn = false;
    CoarseSpace cg;
    long fp64Retries = 0;                   // reduced-precision learned + coarse solves redone with FP64 blocks
    bool fp64ThisStep = false;              // a retry happened in this step: its remaining solves use FP64 blocks
    const double* solveX = nullptr;         // positions of the Newton iterate being solved (rotation modes)
    // dense direct solve: padded row-major matrix / factor, right-hand side, pivot flag
    DevBuf<double> denseA, denseY;
    DevBuf<int> denseBad;
    long directSolves = 0;

    // PCG chunks after the first replay a captured graph (cipc_set_graphs)
    bool useGraphs = true;
    PcgGraph pcgGraph;
    long pcgGraphCaptures = 0, pcgGraphLaunches = 0;

    // launch shapes and broad-phase cell size (set by the autotuner or the caller)
    int spmvGrid = 1024;               // workgroups of the PCG SpMV (multiple of 8)
    int spmvLanes = 16;                // lanes per block row of the PCG SpMV (4, 8 or 16)
    int spmvUnroll = 1;                // blocks per lane pass of the PCG SpMV (1 or 2)
    int lpGrid = 2048, lpLanes = 4;    // workgroups / lanes per row of the factor products
    double cellMult = 1.0;             // constraint-set grid cell = cellMult x mean rest edge length
    struct Tuner {
        bool pcg = false, cell = false;
        bool enabled = false;                      // PCG tuning requested (re-tunes allowed)
        int retunes = 0;
        long tunedNnzb = 0;                        // stored blocks of the matrix tuned on
        int pcgIdx = 0, cellIdx = 0, pcgPhase = 0;
        std::vector<std::array<int, 3>> spmvCand;  // phase 1: (spmvGrid, spmvLanes, spmvUnroll)
        std::vector<double> spmvMs;                // sampled SpMV time per shape
        std::vector<long> spmvN;
        long spmvTurn = 0;                         // chunk counter choosing the shape
        std::vector<std::array<int, 2>> pcgCand;   // phase 2: (lpGrid, lpLanes)
        std::vector<double> pcgCost, cellCand, cellCost;
    } tune;

    // host pinned scalars
    double* hScal = nullptr;           // 64 doubles
    uint32_t* hCnt = nullptr;          // 64 counters

    // statistics of the last step
    cipc_step_stats stats{};
    // per-step text logs of the reference's step (residual.txt, counter.txt, stretch.txt)
    // appended under this folder when it is set
    std::string outputFolder;

    // optional event timing of the dominant PCG kernel (elastic SpMV)
    int profiling = 0;
    int profPrecision = 0;                 // the solver form's precision when profiling started
    std::vector<hipEvent_t> evPool;
    long profLaunches = 0;
    double profMs = 0;
    double profBytes = 0, profFlops = 0;   // algorithmic totals over the timed launches
};

// ---- utilities (util.hip)
void fill_iota(uint32_t* a, long n, hipStream_t s);
void bbox_minmax(Ctx& c, const double* x, double* out6);   // min xyz, max xyz into device memory
void minmax6_final(const double* part, int nb, double* out6, hipStream_t s);   // reduce (lo, hi) records

// ---- stages (implemented in the .hip files)
void ctx_init_device(Ctx& c);                                      // upload + derived data
void constraint_set(Ctx& c, const double* x, DevBuf<int>& out, int& nOut);
double barrier_energy(Ctx& c, const double* x, const int* C, int nC, double kappa, int* bad);
// int4-row constraints, caller-supplied device partials / sum / flags; allocates nothing
void barrier_energy_rows(Ctx& c, const double* x, const int* Crows, int nC, double kappa, double* part, double* out, uint32_t* bad);
// g += barrier gradient; rows flagged in skip (projected out by the caller) are left alone
void barrier_gradient(Ctx& c, const double* x, const int* C, int nC, double kappa, double* g, const uint8_t* skip = nullptr);
// fills Acon; with fricXn the lagged friction blocks (displacement x - fricXn) are added
void barrier_hessian(Ctx& c, const double* x, const int* C, int nC, double kappa, const double* fricXn = nullptr);
void scatter_local_gradient(Ctx& c, const double* gl, const int* gv, long ns, double* g, const uint8_t* skip = nullptr);
// lagged friction (friction.hip): basis at x for constraint set C into c.fC / c.fI / c.nF
void friction_basis(Ctx& c, const double* x, const int* C, int nC, double kappa);
double friction_energy(Ctx& c, const double* x, const double* xn);
void friction_gradient(Ctx& c, const double* x, const double* xn, double* g, const uint8_t* skip = nullptr);   // g +=
void friction_hess_local(Ctx& c, const double* x, const double* xn, double* hl, long ns, long off);
void friction_hess_keys(Ctx& c, uint64_t* key);
// weight of the shell's membrane, bending, strain-limit and fibre terms: h^2, or 1 in a
// static solve (staticSolve ? 1.0 : h in the reference; rods and tets keep h^2)
inline double shell_h2(const Ctx& c) { return c.P.staticSolve ? 1.0 : c.P.dt * c.P.dt; }
// rows/columns projected out of the Newton system: every Dirichlet node, or only the
// fixed ones while the Dirichlet penalty is on
inline const uint8_t* proj_mask(const Ctx& c) { return c.dbcStiff ? c.dbcFixed.p : c.dbc.p; }
// moving Dirichlet nodes (dirichlet.hip)
void dirichlet_device_init(Ctx& c);
void step_dirichlet(Ctx& c);
double dbc_dist2(Ctx& c, const double* x);
double dbc_penalty_energy(Ctx& c, const double* x);
void dbc_penalty_grad(Ctx& c, const double* x, double* g);   // g +=
void dbc_penalty_hess(Ctx& c);                               // diagonal blocks of Ael +=
void dbc_fixed_mask(Ctx& c, const double* x);
// strain limiting (strain.hip): energy (+inf when a limit is reached), gradient and
// projected Hessian added to the membrane's local triangle slots
bool strain_limit_on(const Ctx& c);
double strain_limit_energy(Ctx& c, const double* x);
void strain_limit_grad_add(Ctx& c, const double* x, double* gl);
void strain_limit_hess_add(Ctx& c, const double* x, double* hl);
// kappa_s rule (IMPLICIT_EULER.h:606-645): records which triangles' largest singular value
// lies within 1e-4 (s - sHat) of the upper limit at x; true when one of them did at the
// previous call as well (first: no previous iterate in this step)
bool strain_limit_near_twice(Ctx& c, const double* x, bool first);
// anisotropic (fibre) membrane (strain.hip): replaces the membrane's energy and local slots
bool fiber_on(const Ctx& c);
double fiber_energy(Ctx& c, const double* x);               // +inf when infeasible
void fiber_grad_local(Ctx& c, const double* x, double* gl);
void fiber_hess_local(Ctx& c, const double* x, double* hl);
// volumetric tets (tet.hip): energy +inf when a tet is inverted
double tet_energy(Ctx& c, const double* x);
bool tets_valid(Ctx& c, const double* x);
void tet_grad_local(Ctx& c, const double* x, double* gl, long off);
void tet_hess_local(Ctx& c, const double* x, double* hl, long off);
double inversion_free_stepsize(Ctx& c, const double* x, const double* p, double alpha);
double min_dist2(Ctx& c, const double* x, const int* C, int nC, double* d2out);
double ccd_stepsize(Ctx& c, const double* x, const double* p, double alpha, double* clamped, int* zeroToi);
double elastic_energy(Ctx& c, const double* x, double* memb, double* bend);
double inertia_energy(Ctx& c, const double* x);
void elastic_gradient(Ctx& c, const double* x, double* g);        // g = elastic + inertia (overwrites)
void elastic_hessian(Ctx& c, const double* x);                    // fills Ael (incl. mass, DBC projection)
// dense FP64 Cholesky solve of (Ael + Acon) x = b (direct.hip); 0 ok, 1 non-positive
// pivot, 3 more than kDirectMaxDof unknowns
int direct_solve(Ctx& c, const double* b, double* x);
// tiled dense Cholesky in place of a padded row-major npad x npad matrix (lower factor;
// direct.hip).  bad: set on a non-positive pivot.  tileInv (npad / 64 tiles of 64 x 64):
// the diagonal tiles' inverses are kept there and the panels become products; nullptr:
// panels by substitution.  With diag0 / dropped: a pivot below tau x its original
// diagonal is dropped (pivot at that scale, zero column) and flagged instead.
void dense_cholesky(Ctx& c, double* A, int npad, int* bad, double* tileInv, const double* diag0 = nullptr, int* dropped = nullptr,
                    double tau = 0);
// W = L^-1 of a dense_cholesky factor whose diagonal-tile inverses are in tileInv
void dense_lower_inverse(Ctx& c, const double* L, double* W, int npad, const double* tileInv);
// two-level correction (coarse.hip)
bool coarse_prepare(Ctx& c, const double* x);            // aggregates (once), Galerkin matrix, factor; false: unavailable
template <class V>
void coarse_restrict_solve(Ctx& c, const V* r, const int* done);   // c.cg.zc = A_c^-1 P^T r
// Newton linear solve (PCG or the direct solve, as P.linSolver selects); 0 ok, 1 matrix
// not positive definite, 2 iteration limit, 3 too large for the direct solve
// x0: optional initial guess (nullptr: zero)
int pcg_solve(Ctx& c, const double* b, double* xsol, double tol, int maxIter, int* iters, double* relres, const double* x0 = nullptr);
// start the autotuner: bit 1 PCG launch shapes, bit 2 broad-phase cell size
void autotune_start(Ctx& c, int what);
int advance_one_step(Ctx& c);

}  // namespace cipc
