/*
 * cipc.h -- C ABI of the MI355X (gfx950) implementation of the C-IPC
 * contact-and-Newton step (ipc-sim/Codim-IPC, discrete-shell hinge path).
 *
 * Plain pointers and sizes only.  All arrays are caller-owned host memory;
 * per-vertex arrays are vertex-major xyz (double[3 n]).  Every entry point
 * returns CIPC_OK or a negative error code; cipc_last_error() describes the
 * last failure of a context.  A context owns one HIP stream and its device
 * state; it is not thread safe (one context per host thread).
 *
 * The step entry point replaces the reference's pybind11 binding
 *   FEM.DiscreteShell.Advance_One_Step_IE_Hinge
 *     = Advance_One_Step_IE_Discrete_Shell<double,3,KL=false,elasticIPC=false>
 *   (Library/FEM/Shell/DISCRETE_SHELL.h:1108, body IMPLICIT_EULER.h:151-902),
 * called from Python/Drivers/FEMDiscreteShellBase.py:353-361.  Scene set-up
 * replaces Initialize_Shell_Hinge_EIPC + Initialize_OIPC
 * (DISCRETE_SHELL.h:218-357, 555-577).  The component entry points expose the
 * reference functions the step is built from, for parity testing.
 */
#ifndef CIPC_H
#define CIPC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CIPC_OK 0
#define CIPC_E_INVALID (-1)      /* bad argument or size */
#define CIPC_E_HIP (-2)          /* HIP runtime failure */
#define CIPC_E_PENETRATION (-3)  /* a constraint distance <= offset (IPC.h:773-776 exit) */
#define CIPC_E_ZERO_TOI (-4)     /* CCD returned a zero step (IPC.h:2014-2032 exit) */
#define CIPC_E_UNSUPPORTED (-5)  /* configuration outside this build (e.g. blocked Dirichlet motion) */
#define CIPC_E_NO_CONVERGENCE (-6)
#define CIPC_E_CAPACITY (-7)     /* caller buffer too small; required size returned */

typedef struct cipc_ctx cipc_ctx;

/* Scene description: geometry, Dirichlet nodes and the parameters the reference
 * passes to Initialize_Shell_Hinge_EIPC / Initialize_OIPC / Advance_One_Step. */
typedef struct cipc_scene_desc {
    int32_t n_vertices;
    const double* x0;          /* rest = initial positions, 3 n_vertices */
    const double* v0;          /* initial velocities or NULL (zero) */
    int32_t n_triangles;
    const int32_t* triangles;  /* 3 n_triangles, consistently oriented */
    int32_t n_dbc;
    const int32_t* dbc_index;  /* Dirichlet vertex ids */
    const double* dbc_target;  /* 3 n_dbc target positions */
    double density, youngs, poisson, thickness, bending_stiff_mult;
    double dt, dhat, offset, newton_tol;
    double gravity[3];
    double kappa_stiff_mult;   /* Initialize_OIPC stiffMult */
    double pcg_tol;            /* relative residual of the Newton linear solve */
    int32_t pcg_max_iter;
    int32_t with_collision;
    /* codimensional primitives; zero counts for a pure shell scene.  Rod segments are
     * Initialize_Discrete_Rod (Rod/DISCRETE_ROD.h:87-165, per-segment springs, one
     * bending hinge per pair of segments sharing a node); particles are
     * Initialize_Discrete_Particle (DISCRETE_PARTICLE.h:52-75).  Both join the contact
     * primitives as in IMPLICIT_EULER.h:243-277 and the codimensional contact and CCD
     * branches of IPC.h:271-353, 2095-2162. */
    int32_t n_rod_segments;
    const int32_t* rod_segments;     /* 2 n_rod_segments vertex ids */
    double rod_density, rod_youngs, rod_bend_stiff_mult, rod_thickness;
    int32_t n_particles;
    const int32_t* particles;        /* n_particles vertex ids */
    double particle_density, particle_thickness;
    /* lagged friction: the step's mu, epsv2 and fricIterAmt arguments (IMPLICIT_EULER.h:
     * 167, FEM/FRICTION.h).  mu = 0 disables friction; with mu > 0, epsv2 > 0 is required
     * (the reference's default is 1e-6) and friction_iter_amt follows fricIterAmt: 1 =
     * basis taken at the step start only, k > 1 = re-lagged up to k - 1 times after
     * convergence, <= 0 = re-lagged until the updated Newton step is below tolerance. */
    double friction_mu, friction_epsv2;
    int32_t friction_iter_amt;
    /* isotropic strain limiting: the step's kappa_s, s, sHat arguments (FEM/Shell/INEXT.h,
     * index 0 = upper limit, 1 = lower limit).  kappa[0] = 0 disables it; with kappa[0] > 0
     * the singular values of each triangle's deformation gradient stay below s[0]
     * (barrier from shat[0]); kappa[1] > 0 adds the lower limit s[1] < shat[1]. */
    double strain_limit_kappa[2], strain_limit_s[2], strain_limit_shat[2];
    /* anisotropic (fibre) membrane: the step's fiberStiffMult and inextLimit arguments
     * (FEM/Shell/ANISO_INEXT.h).  With fiber_stiff_mult[0] > 0 or [1] > 0 it replaces the
     * Neo-Hookean membrane: log barriers on the Green strain components E00, E00 E11, E11,
     * E01 of the rest x-z frame, limits |E00| < fiber_limit[0], |E11| < [1], |E01| < [2]. */
    double fiber_stiff_mult[4], fiber_limit[3];
    /* volumetric tets inside the shell step (add_object_3D + initialize_added_objects of
     * FEMDiscreteShellBase.py, the step's tet / tetAttr / tetElasticity arguments):
     * Neo-Hookean (Physics/NEOHOOKEAN.h) with Young's modulus tet_youngs and Poisson ratio
     * tet_poisson, lumped mass tet_density vol / 4 per node (FEM/ELEM_TO_NODE.h:11-46);
     * their surface triangles (Find_Surface_TriMesh) join the contact primitives ahead of
     * the shell's.  Every rest tet must be positively oriented.  Zero count: no tets. */
    int32_t n_tets;
    const int32_t* tets;             /* 4 n_tets vertex ids */
    double tet_density, tet_youngs, tet_poisson;
} cipc_scene_desc;

typedef struct cipc_step_stats {
    int32_t newton_iters, contacts, pcg_iters, gd_fallbacks, ls_trials, status;
    double kappa, last_residual, energy;
    double t_total_ms, t_constraint_set_ms, t_ccd_ms, t_hessian_ms, t_solve_ms, t_energy_ms, t_gradient_ms;
    /* upper strain-limit stiffness kappa_s[0] after the step and how often the step doubled
     * it (IMPLICIT_EULER.h:606-645; the reference mutates its kappa_s argument in place) */
    double kappa_s;
    int32_t kappa_s_doublings;
} cipc_step_stats;

/* ---- lifetime ------------------------------------------------------------ */
/* Initialize_Shell_Hinge_EIPC + Initialize_OIPC (DISCRETE_SHELL.h:218-357, 555-577) */
int cipc_create(const cipc_scene_desc* desc, cipc_ctx** out);
void cipc_destroy(cipc_ctx* ctx);
const char* cipc_last_error(const cipc_ctx* ctx);
/* out[0..5] = n_vertices, n_triangles, n_boundary_nodes, n_boundary_edges, n_hinges, n_constraints */
int cipc_get_sizes(const cipc_ctx* ctx, int32_t* out);
/* out[0..5] = kappa, kappa_max, dhat^2, lambda, mu, hinge k (DISCRETE_SHELL.h:322-342, 568-574) */
int cipc_get_scalars(const cipc_ctx* ctx, double* out);
/* derived set-up arrays: lumped mass (n_vertices), hinge stencils (4 n_hinges) and
 * rest data (3 n_hinges: theta_bar, |e_bar|, h_bar), boundary edges (2 n_boundary_edges) */
int cipc_get_setup(const cipc_ctx* ctx, double* mass, int32_t* hinges, double* hinge_info, int32_t* bedges);
/* codimensional set-up (DISCRETE_ROD.h:87-165, IMPLICIT_EULER.h:243-277):
 * out[0..4] = n_rod_segments, n_rod_hinges, n_particles, first rod-node slot and first
 * particle slot among the boundary nodes */
int cipc_get_codim_sizes(const cipc_ctx* ctx, int32_t* out);
/* out[0] = n_tets, out[1] = contact triangles (the tets' surface, then the shell's) */
int cipc_get_tet_sizes(const cipc_ctx* ctx, int32_t* out);
/* Compute_Inversion_Free_StepSize (FEM/IPC.h:1785-1877): the largest step in (0, alpha]
 * along p (3 n_vertices) from x that keeps every tet's volume above 20% of its volume at
 * x (smallest positive root of the cubic det(x + t p) = 0.2 det(x), IPC.h:1751-1782).
 * The shell step itself keeps tets uninverted by halving its line search
 * (IMPLICIT_EULER.h:74-92), as the reference's shell step does. */
int cipc_inversion_free_stepsize(cipc_ctx* ctx, const double* x, const double* p, double alpha, double* out);
/* rod_info 3 per segment (E, rest length, thickness), rod_hinges 3 per hinge (x0, shared
 * node, x2), rod_hinge_info 3 per hinge (E * bend mult, rest |e0| + |e1|, thickness),
 * bnodes (n_boundary_nodes) */
int cipc_get_codim_setup(const cipc_ctx* ctx, double* rod_info, int32_t* rod_hinges, double* rod_hinge_info, int32_t* bnodes);
int cipc_set_state(cipc_ctx* ctx, const double* x, const double* v);
int cipc_get_state(const cipc_ctx* ctx, double* x, double* v);
int cipc_set_kappa(cipc_ctx* ctx, double kappa);

/* ---- the step ------------------------------------------------------------ */
/* Advance_One_Step_IE_Discrete_Shell<double,3,false,false> (IMPLICIT_EULER.h:151-902):
 * one implicit-Euler step; mutates positions, velocities and kappa. */
int cipc_step(cipc_ctx* ctx, cipc_step_stats* stats);
/* Many independent scenes stepped together (no counterpart in the reference, which
 * steps one scene per process): one cipc_step per context, run by `threads` host
 * threads (<= 0: one per hardware thread) taking contexts from a shared queue; each
 * context keeps its own HIP stream, so the scenes' kernels overlap on the GPU.  stats
 * and step_ms (may be NULL) receive each scene's statistics and step wall time; returns
 * the first error (every scene is still attempted).  Contexts must be distinct. */
int cipc_batch_step(cipc_ctx* const* ctxs, int32_t n, int32_t threads, cipc_step_stats* stats, double* step_ms);

/* ---- components at caller positions x (parity) ---------------------------- */
/* Compute_Constraint_Set (IPC.h:19-740): int4 encoding of the reference, order:
 * point-triangle by (boundary node, triangle), edge-edge family by (edge, edge),
 * deduplicated point-point/point-edge lexicographically with -multiplicity. */
int cipc_constraint_set(cipc_ctx* ctx, const double* x, int32_t* out, int32_t cap, int32_t* n);
/* Compute_Barrier (IPC.h:742-941) */
int cipc_barrier_energy(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, double* e);
/* No counterpart in the reference: the same result as cipc_barrier_energy for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  x, the constraint rows (int4 rows, reference
 * encoding), the per-block partials, the sum and the validity flags are staged in a
 * static 32 MiB device array per GPU; the call makes no hipMalloc / hipFree.  Inputs
 * that need more than 32 MiB (24 n_vertices + 16 n_cs + 8 ceil(n_cs / 256) bytes, each
 * rounded up to 256, + 512) return CIPC_E_CAPACITY.  Every context on one GPU shares that
 * GPU's array and the last call's contents stay in it: calls that overlap in time on one
 * GPU (two host threads, two contexts) race and must be serialised by the caller. */
int cipc_barrier_energy_noalloc(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, double* e);
/* Compute_Barrier_Gradient (IPC.h:943-1256); g overwritten (3 n_vertices) */
int cipc_barrier_gradient(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, double* g);
/* Compute_Membrane_Energy + Compute_Bending_Energy (MEMBRANE.h:7-50, BENDING.h:52-75): out[0..1] */
int cipc_elastic_energy(cipc_ctx* ctx, const double* x, double* out);
/* Compute_Membrane_Gradient + Compute_Bending_Gradient (MEMBRANE.h:52-119, BENDING.h:177-208) */
int cipc_elastic_gradient(cipc_ctx* ctx, const double* x, double* g);
/* Compute_IncPotential_Hessian (INC_POTENTIAL.h:277-395): which = 1 elastic (+mass,
 * Dirichlet projection), 2 barrier, 3 both.  Blocks as COO (duplicates to be summed):
 * brow/bcol block indices, bval 9 doubles (row-major) per block. */
int cipc_hessian(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, int32_t which,
                 int32_t* brow, int32_t* bcol, double* bval, int32_t cap, int32_t* nblocks);
/* Compute_Min_Dist2 (IPC.h:2246-2388): per-constraint d^2 (no offset) and min(d^2) - offset^2 */
int cipc_min_dist2(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double* d2, double* min_val);
/* Write_TriMesh_Obj (Utils/MESHIO.h:184-213): the current positions ("v %le %le %le") and
 * the shell triangles kept at set-up ("f i j k", 1-based) to an OBJ file. */
int cipc_write_obj(cipc_ctx* ctx, const char* path);
/* Moving Dirichlet nodes (FEM/BOUNDARY_CONDITION.h).  Targets live in the context and
 * start at cipc_scene_desc.dbc_target; a step moves each Dirichlet node to its target
 * when the CCD allows the whole motion, else it pulls the nodes there with the
 * mass-weighted penalty of IMPLICIT_EULER.h:384-401, 712-768 (raised until 99% of the way
 * is covered).  cipc_dirichlet_add_motion appends a DBC_MOTION group over Dirichlet
 * entries [begin, end) (Init_Dirichlet's range, BOUNDARY_CONDITION.h:21-104);
 * cipc_step_dirichlet is Step_Dirichlet (:107-155: rotation by ang_vel_deg * dt about
 * rot_axis through rot_center, then translation by v * dt); cipc_turn_dirichlet is
 * Turn_Dirichlet (:211-220); get/set read or replace the 3 n_dbc targets (Load_Dirichlet,
 * Reset_Dirichlet with the caller's arrays). */
int cipc_dirichlet_add_motion(cipc_ctx* ctx, int32_t begin, int32_t end, const double* v, const double* rot_center,
                              const double* rot_axis, double ang_vel_deg);
int cipc_step_dirichlet(cipc_ctx* ctx);
int cipc_turn_dirichlet(cipc_ctx* ctx);
int cipc_get_dirichlet_targets(cipc_ctx* ctx, double* targets);
int cipc_set_dirichlet_targets(cipc_ctx* ctx, const double* targets);
/* Strain-limit parameters for the following steps (kappa_s, s, sHat of
 * Advance_One_Step_IE_Hinge, IMPLICIT_EULER.h:162); same meaning as in cipc_scene_desc. */
/* The step's staticSolve argument (IMPLICIT_EULER.h:170): enable != 0 makes the following
 * steps static solves -- the shell's membrane, bending, strain-limit and fibre terms
 * weighted by 1 instead of h^2 (rods and tets keep h^2, as in the reference), the body
 * force potential -x.(M g) instead of inertia, no mass in the Hessian, the Newton residual
 * not divided by h, velocities left unchanged (IMPLICIT_EULER.h:196-217, 646-666, 893-899,
 * INC_POTENTIAL.h:118-130, 233-249, 383-386).  Lagged friction is refused in a static
 * solve (CIPC_E_UNSUPPORTED): the reference has no previous state for it there. */
int cipc_set_static_solve(cipc_ctx* ctx, int32_t enable);
/* The step's outputFolder argument (IMPLICIT_EULER.h:187): with a folder set (NULL or ""
 * unsets it) each step appends, as the reference does, one line per Newton iteration to
 * residual.txt ("PNIter avgResMag maxRes Eprev L2Norm", :648-670), "PNIter contacts" to
 * counter.txt (:868-875) and "maxs avgs minc avgc" of the triangles' in-plane stretch to
 * stretch.txt (:877-883, Compute_Max_And_Avg_Stretch, DISCRETE_SHELL.h:602-674). */
int cipc_set_output_folder(cipc_ctx* ctx, const char* folder);
int cipc_set_strain_limit(cipc_ctx* ctx, const double* kappa, const double* s, const double* shat);
/* kappa_s, s, sHat as the steps have left them: the step doubles kappa_s[0] (up to 1e5)
 * when a triangle stays at the limit, in place, as the reference does with its kappa_s
 * argument (IMPLICIT_EULER.h:606-645).  Null outputs are skipped. */
int cipc_get_strain_limit(const cipc_ctx* ctx, double* kappa, double* s, double* shat);
/* fiberStiffMult (4) and inextLimit (3) for the following steps; same meaning as in
 * cipc_scene_desc. */
int cipc_set_fiber(cipc_ctx* ctx, const double* stiff_mult, const double* limit);
/* Compute_Inextensibility_Energy (FEM/Shell/INEXT.h:120-178) at x: *valid = 0 when a
 * singular value reached a limit (then *e = 0).  Its gradient and Hessian are part of
 * cipc_elastic_gradient and cipc_hessian (which & 1) while strain limiting is on. */
int cipc_strain_limit_energy(cipc_ctx* ctx, const double* x, double* e, int32_t* valid);
/* Friction parameters for the following steps (the mu, epsv2, fricIterAmt arguments of
 * Advance_One_Step_IE_Hinge, IMPLICIT_EULER.h:167); same meaning as in cipc_scene_desc. */
int cipc_set_friction(cipc_ctx* ctx, double mu, double epsv2, int32_t iter_amt);
/* Per-component friction coefficients (the step's compNodeRange / muComp arguments,
 * Compute_Friction_Coef FRICTION.h:114-171): component i holds the vertices below
 * comp_node_range[i] (increasing, the last one >= n_vertices); a stencil's normal force is scaled by
 * mu_comp[c0 + c1 n_comp] for the components of its two primitives and mu is taken as 1.
 * n_comp = 0 returns to the single coefficient. */
int cipc_set_friction_components(cipc_ctx* ctx, int32_t n_comp, const int32_t* comp_node_range, const double* mu_comp);
/* Compute_Friction_Basis (FEM/FRICTION.h:17-112) at x for constraint set cs: the friction
 * stencils (cs without mollified entries, same order) are kept in the context for the
 * friction entry points below (a step replaces them) and returned: *n stencils,
 * out_cs 4 int32 each, out_info 9 doubles each (closest-point coordinates 2, tangent
 * basis columns 2 x 3, normal force).  CIPC_E_CAPACITY with *n set when n > cap. */
int cipc_friction_basis(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, int32_t* out_cs,
                        double* out_info, int32_t cap, int32_t* n);
/* Compute_Friction_Potential (FRICTION.h:173-243) of the kept stencils, displacement x - xn */
int cipc_friction_energy(cipc_ctx* ctx, const double* x, const double* xn, double* e);
/* Compute_Friction_Gradient (FRICTION.h:245-394); g overwritten (3 n_vertices) */
int cipc_friction_gradient(cipc_ctx* ctx, const double* x, const double* xn, double* g);
/* Compute_Friction_Hessian (FRICTION.h:396-661), blocks summed per (row, col) with
 * Dirichlet rows/columns dropped; COO layout as cipc_hessian */
int cipc_friction_hessian(cipc_ctx* ctx, const double* x, const double* xn, int32_t* brow, int32_t* bcol, double* bval,
                          int32_t cap, int32_t* nblocks);
/* Compute_Intersection_Free_StepSize (IPC.h:1879-2244) incl. the CCD hash step clamp
 * (SPATIAL_HASH.h:466-482); alpha_clamped = step after the clamp */
int cipc_ccd_stepsize(cipc_ctx* ctx, const double* x, const double* p, double alpha_in, double* alpha_out,
                      double* alpha_clamped);
/* Newton linear solve: assemble H(x, cs) (elastic + barrier + mass, Dirichlet
 * projected) and solve H sol = b by block-Jacobi PCG to ||r|| <= tol ||b||
 * (stands in for Solve_Direct, DIRECT_SOLVER.h:14-66). */
int cipc_linear_solve(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, const double* b,
                      double* sol, double tol, int32_t* iters, double* relres);

/* ---- solver precision (no counterpart in the reference, which factorises in FP64
 * with CHOLMOD, DIRECT_SOLVER.h:14-66) -----------------------------------------
 * mode 0 (fp64): FP64 block-Jacobi PCG.
 * mode 1 (mixed): FP32 Hessian blocks, FP64 Krylov vectors and SpMV accumulation.
 * mode 2 (fp32): FP32 blocks, FP32 Krylov vectors, FP32 SpMV accumulation.
 * Modes 1 and 2 run inside an FP64 iterative refinement that stops when the FP64
 * residual ||b - H x|| <= pcg_tol ||b||.  Affects cipc_step and cipc_linear_solve;
 * out (4 entries) of cipc_get_solver_info: out[0] = refinement sweeps of the last solve
 * (0 in mode 0), out[1], out[2] = PCG chunk graphs captured and replayed so far
 * (cipc_set_graphs), out[3] = dense direct solves so far (cipc_set_linear_solver). */
int cipc_set_precision(cipc_ctx* ctx, int32_t mode);
int cipc_get_solver_info(const cipc_ctx* ctx, int32_t* out);
/* Linear solver of the Newton systems (Solve_Direct, DIRECT_SOLVER.h:14-66: CHOLMOD
 * analyse + factorise + solve of every Newton matrix; a non-positive pivot returns
 * CHOLMOD_NOT_POSDEF and the step takes the gradient direction, IMPLICIT_EULER.h:515-554).
 * mode 0 (default): PCG as configured above; a solve that reaches pcg_max_iter is
 * redone by the dense direct solve when the system has at most 24576 unknowns.
 * mode 1: every Newton solve (cipc_step, cipc_linear_solve) is a dense FP64 Cholesky
 * factorisation and solve on the GPU; CIPC_E_UNSUPPORTED for scenes above 24576
 * unknowns (8192 vertices).  A non-positive pivot makes the step take the gradient
 * direction, as in the reference. */
int cipc_set_linear_solver(cipc_ctx* ctx, int32_t mode);
/* No reference counterpart.  enable != 0 (the default): every 16-iteration PCG chunk after
 * a solve's first runs as one replay of a captured HIP graph (its launches identical from
 * chunk to chunk; re-captured when the matrix, buffers or launch shapes change); 0: the
 * chunks' kernels are launched one by one.  Both give the same bits. */
int cipc_set_graphs(cipc_ctx* ctx, int32_t enable);

/* ---- learned preconditioner (no counterpart in the reference) ------------
 * A 4-layer message-passing network (hidden 64, BF16 weights, FP32 accumulation) over
 * the block graph of the Hessian predicts a block lower-triangular factor G on the
 * Hessian's sparsity pattern; the PCG uses M^-1 = S G^T G S with S = diag(H)^-1/2.
 * Training minimises the Kaporin condition number of G S H S G^T (Rademacher trace
 * estimate with n_probes probes, exact log-det).  mode 1 of cipc_set_preconditioner
 * uses it with `budget` PCG iterations, after which the solve restarts with block Jacobi
 * (mode 0, the default).  Weights: cipc_precond_num_params() FP32 values. */
int32_t cipc_precond_num_params(void);
int cipc_precond_init(cipc_ctx* ctx, uint64_t seed);
int cipc_precond_set_weights(cipc_ctx* ctx, const float* w, int32_t n);
int cipc_precond_get_weights(cipc_ctx* ctx, float* w, int32_t n);
int cipc_set_preconditioner(cipc_ctx* ctx, int32_t mode, int32_t budget);
/* run the network at the first Newton solve of each step and then every `every` solves
 * (>= 1; default 1: at every solve); the solves in between reuse its factor G with the
 * diagonal scaling S of their own Hessian */
int cipc_set_precond_reuse(cipc_ctx* ctx, int32_t every);
/* Two-level correction of the learned preconditioner (no counterpart in the reference,
 * which factorises every Newton matrix, DIRECT_SOLVER.h:14-66): every > 0 makes the
 * learned solves use M^-1 = F^T F + P X P^T, P = rigid-body modes (three translations,
 * three rotations about the centroid) on aggregates of the elastic vertex graph
 * (Dirichlet vertices and vertices without elastic neighbours excluded), X = W^T W with W
 * the inverse Cholesky factor of A_c = P^T H P rounded to FP32.  A_c is assembled and
 * factorised densely at the first solve of each step and then every `every` solves
 * (every = 0: off).  max_unknowns bounds the coarse problem (0: 12288; the aggregates of
 * a root-and-neighbours pass over the aggregate graph, merged pairwise while above the
 * bound); with an aggregate of more than 256 neighbours the solve runs without the
 * coarse level.  flags bit 0: W stored in BF16 instead of FP32 (half the bytes per
 * iteration; X stays positive semidefinite, its error grows with sqrt(cond(A_c))).
 * Affects the learned forms only (cipc_set_preconditioner mode 1); the stopping rule is
 * unchanged. */
int cipc_set_coarse(cipc_ctx* ctx, int32_t every, int32_t max_unknowns, int32_t flags);
/* out (7 entries): aggregates (-1 before the first build), coarse unknowns, coarse
 * set-ups so far, wall ms of the last set-up, 1 when the last set-up succeeded, set-ups
 * that found an aggregate with more than 256 neighbours (solved without the coarse level),
 * reduced-precision solves that failed with the coarse level and were redone with FP64
 * blocks (same preconditioner, from their iterate; block Jacobi only if that fails too) */
int cipc_get_coarse_info(const cipc_ctx* ctx, double* out);
/* Initial guess of the Newton solves (no counterpart in the reference, whose direct solve
 * needs none): enable != 0 (the default) starts a solve that follows a line search cut
 * to alpha < 1 from (1 - alpha) p_prev, the quadratic model's next direction, when that
 * guess leaves less than half of ||b|| as residual; the stopping rule
 * ||b - H p|| <= pcg_tol ||b|| is unchanged. */
int cipc_set_warm_start(cipc_ctx* ctx, int32_t enable);
/* assemble H(x, constraint set) as cipc_linear_solve does, then: loss and the flat
 * parameter gradient (grad may be NULL) / one Adam step at learning rate lr */
int cipc_precond_loss_grad(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, int32_t n_probes, uint64_t seed,
                           double* loss, float* grad);
int cipc_precond_train_step(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, int32_t n_probes, uint64_t seed,
                            double lr, double* loss);
/* the factor G at H(x, constraint set): 9 * nnzb floats (SoA [9][nnzb] over the elastic
 * block pattern, lower blocks, their transposes at the upper positions, G_ii on the
 * diagonal); use_mfma selects the BF16 matrix-core message layers (inference form) or
 * the FP32 ones (training form); *nnzb receives the pattern size when g is NULL */
int cipc_precond_factor(cipc_ctx* ctx, const double* x, const int32_t* cs, int32_t n_cs, double kappa, int32_t use_mfma, float* g,
                        int32_t* nnzb);
/* Adam moments and step count, so that training can move between scenes (contexts) */
int cipc_precond_get_adam(cipc_ctx* ctx, float* m, float* v, int32_t* t, int32_t n);
int cipc_precond_set_adam(cipc_ctx* ctx, const float* m, const float* v, int32_t t, int32_t n);
/* out[0] = 1 if the last solve fell back to block Jacobi; out[1], out[2] = learned
 * solves so far and how many of them fell back; out[3], out[4] = status (0 converged,
 * 1 non-positive curvature, 2 budget/refinement limit) and iterations of the last
 * learned attempt; out[5] = 1 if that fallback continued from the learned iterate (else
 * it restarted from zero); out[6] = iterations of the last from-zero block-Jacobi solve
 * (the cap of the learned budget: 1.25x this).  out holds 7 values. */
int cipc_precond_info(const cipc_ctx* ctx, int32_t* out);

/* ---- launch shapes and autotuning (no counterpart in the reference) -------
 * cipc_autotune(what): bit 1 -- the next Newton solves cycle twice through PCG launch
 * shapes and keep the fastest: first the SpMV shapes (grid 1024/2048 x {16 lanes per
 * block row x 1 block per lane pass, 8 x 1, 8 x 2}), timed by the SpMV's sampled dispatch time,
 * then, with the learned preconditioner, the factor-product shapes (grid 1024/2048 x
 * 2/4/8 lanes per row), timed per PCG iteration; the tuning reruns (at most twice) when
 * the number of stored Hessian blocks moves more than 5% from the tuned matrix's;
 * bit 2 -- the next ten constraint-set builds cycle twice through broad-phase cell
 * sizes of 0.5-2x the mean rest edge length and keep the fastest (the constraint set
 * does not depend on it).  cipc_set_tuning fixes the shapes and stops any tuning. */
typedef struct cipc_tuning {
    int32_t spmv_grid;      /* workgroups of the PCG SpMV, multiple of 8 */
    int32_t spmv_lanes;     /* lanes per block row: 4, 8 or 16 */
    int32_t spmv_unroll;    /* blocks per lane pass: 1 or 2 */
    int32_t factor_grid;    /* workgroups of the learned factor products, multiple of 8 */
    int32_t factor_lanes;   /* lanes per factor row: 2, 4, 8 or 16 */
    int32_t tuning_pcg;     /* out (get only): PCG tuning still running */
    int32_t tuning_cell;    /* out (get only): cell-size tuning still running */
    int32_t reserved;
    double cell_mult;       /* broad-phase cell = cell_mult x mean rest edge length, 0.125-16 */
} cipc_tuning;
int cipc_autotune(cipc_ctx* ctx, int32_t what);
int cipc_get_tuning(const cipc_ctx* ctx, cipc_tuning* out);
int cipc_set_tuning(cipc_ctx* ctx, const cipc_tuning* t);

/* ---- instrumentation (no counterpart in the reference) ------------------ */
/* Kernel timing of the dominant kernel of the step, the PCG SpMV (q = A p with
 * p = z + beta p_old formed on the fly), from the dispatch's own start/end
 * timestamps, sampled on every 8th launch (timing every launch slows the solve by about
 * 10%): enable != 0 starts recording (and resets the counters).
 * out[0..5] = launches timed, total ms, block rows, stored 3x3 blocks of the
 * elastic matrix, mean algorithmic bytes per launch (matrix values + column indices
 * + row pointers of the elastic and contact matrices, z and p_old read, p_new and q
 * written, each touched once), mean algorithmic flops per launch. */
int cipc_set_profiling(cipc_ctx* ctx, int32_t enable);
int cipc_get_profile(const cipc_ctx* ctx, double* out);

#ifdef __cplusplus
}
#endif
#endif /* CIPC_H */
