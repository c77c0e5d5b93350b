"""Host-side object over the C ABI: one simulated scene on one GPU.

``Simulator`` owns a ``cipc_ctx``; its methods map one-to-one onto the reference
functions the hot path is built from (names in the docstrings), taking and returning
numpy arrays (vertex-major xyz, float64).
"""
import ctypes

import numpy as np

from . import _lib


class CipcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("cipc error %d (%s): %s" % (code, _lib.ERRORS.get(code, "?"), msg))
        self.code = code


def _d(a):
    return None if a is None else a.ctypes.data_as(_lib.D)


def _i(a):
    return None if a is None else a.ctypes.data_as(_lib.I32)


class Simulator:
    PRECISION = {"fp64": 0, "mixed": 1, "fp32": 2}

    def __init__(self, sd, pcg_tol=1e-5, pcg_max_iter=100000, precision="fp64"):
        """Initialize_Shell_Hinge_EIPC + Initialize_OIPC (DISCRETE_SHELL.h:218-357, 555-577).

        precision: "fp64" (FP64 block-Jacobi PCG), "mixed" (FP32 Hessian blocks, FP64
        vectors) or "fp32" (FP32 blocks, vectors and SpMV accumulation); the last two run
        inside FP64 iterative refinement on the true residual."""
        self.lib = _lib.load()
        P = sd.params
        self._keep = [sd.X, sd.tris, sd.dbc_idx, sd.dbc_target]
        desc = _lib.SceneDesc()
        desc.n_vertices = sd.nV
        desc.x0 = _d(sd.X)
        desc.v0 = None
        desc.n_triangles = sd.nT
        desc.triangles = _i(sd.tris)
        desc.n_dbc = len(sd.dbc_idx)
        desc.dbc_index = _i(sd.dbc_idx)
        desc.dbc_target = _d(sd.dbc_target)
        desc.density = P["rho"]
        desc.youngs = P["E"]
        desc.poisson = P["nu"]
        desc.thickness = P["thick"]
        desc.bending_stiff_mult = P["bendMult"]
        desc.dt = P["dt"]
        desc.dhat = P["dHat"]
        desc.offset = P["xi"]
        desc.newton_tol = P["PNTol"]
        g = P.get("gravity", (0.0, -9.81, 0.0))
        desc.gravity[0], desc.gravity[1], desc.gravity[2] = g
        desc.kappa_stiff_mult = P.get("kappaStiffMult", 1.0)
        desc.pcg_tol = pcg_tol
        desc.pcg_max_iter = pcg_max_iter
        desc.with_collision = int(P.get("collision", 1))
        rods = getattr(sd, "rods", None)
        parts = getattr(sd, "particles", None)
        if rods is not None and len(rods):
            rods = np.ascontiguousarray(rods, dtype=np.int32)
            self._keep.append(rods)
            desc.n_rod_segments = len(rods)
            desc.rod_segments = _i(rods)
        if parts is not None and len(parts):
            parts = np.ascontiguousarray(parts, dtype=np.int32)
            self._keep.append(parts)
            desc.n_particles = len(parts)
            desc.particles = _i(parts)
        desc.rod_density = P.get("rodRho", 0.0)
        desc.rod_youngs = P.get("rodE", 0.0)
        desc.rod_bend_stiff_mult = P.get("rodBendMult", 1.0)
        desc.rod_thickness = P.get("rodThick", 1e-3)
        desc.particle_density = P.get("partRho", 0.0)
        desc.particle_thickness = P.get("partThick", 1e-3)
        tets = getattr(sd, "tets", None)
        if tets is not None and len(tets):
            tets = np.ascontiguousarray(tets, dtype=np.int32)
            self._keep.append(tets)
            desc.n_tets = len(tets)
            desc.tets = _i(tets)
            desc.tet_density = P.get("tetRho", 0.0)
            desc.tet_youngs = P.get("tetE", 0.0)
            desc.tet_poisson = P.get("tetNu", 0.3)
        desc.friction_mu = P.get("mu", 0.0)
        desc.friction_epsv2 = P.get("epsv2", 1e-6)
        desc.friction_iter_amt = int(P.get("fricIterAmt", 1))
        for i, (k, s_, sh) in enumerate(zip(P.get("kappa_s", (0.0, 0.0)), P.get("s", (1.01, 0.0)), P.get("sHat", (1.0, 1.0)))):
            desc.strain_limit_kappa[i], desc.strain_limit_s[i], desc.strain_limit_shat[i] = k, s_, sh
        for i, k in enumerate(P.get("fiberStiffMult", (0.0, 0.0, 0.0, 0.0))):
            desc.fiber_stiff_mult[i] = k
        for i, l in enumerate(P.get("inextLimit", (1.0, 1.0, 1.0))):
            desc.fiber_limit[i] = l
        h = ctypes.c_void_p()
        rc = self.lib.cipc_create(ctypes.byref(desc), ctypes.byref(h))
        if rc != 0:
            raise CipcError(rc, "cipc_create failed")
        self.h = h
        if P.get("compNodeRange") is not None:
            self.set_friction_components(P["compNodeRange"], P["muComp"])
        self.nV = sd.nV
        self.nDBC = len(sd.dbc_idx)
        s = np.zeros(6, dtype=np.int32)
        self.lib.cipc_get_sizes(self.h, _i(s))
        self.nT, self.nBN, self.nBE, self.nH = int(s[1]), int(s[2]), int(s[3]), int(s[4])
        self.lib.cipc_get_codim_sizes(self.h, _i(s))
        self.nRod, self.nRodHinge, self.nParticle, self.codimBN0, self.codimBN1 = [int(v) for v in s[:5]]
        if precision != "fp64":
            self.set_precision(precision)

    def close(self):
        if getattr(self, "h", None):
            self.lib.cipc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise CipcError(rc, self.lib.cipc_last_error(self.h).decode())

    # ------------------------------------------------------------ state
    def scalars(self):
        out = np.zeros(6)
        self._check(self.lib.cipc_get_scalars(self.h, _d(out)))
        return dict(kappa=out[0], kappaMax=out[1], dHat2=out[2], lam=out[3], mu=out[4], kBend=out[5])

    def setup_arrays(self):
        mass = np.zeros(self.nV)
        hinge = np.zeros(4 * self.nH, dtype=np.int32)
        hinfo = np.zeros(3 * self.nH)
        bedge = np.zeros(2 * self.nBE, dtype=np.int32)
        self._check(self.lib.cipc_get_setup(self.h, _d(mass), _i(hinge), _d(hinfo), _i(bedge)))
        return dict(mass=mass, hinge=hinge.reshape(-1, 4), hinge_info=hinfo.reshape(-1, 3), bedge=bedge.reshape(-1, 2))

    def autotune(self, pcg=True, cell=True):
        """Tune the PCG launch shapes over the next Newton solves and the broad-phase cell
        size over the next constraint-set builds (cipc_autotune)."""
        self._check(self.lib.cipc_autotune(self.h, (1 if pcg else 0) | (2 if cell else 0)))

    def tuning(self):
        t = _lib.Tuning()
        self._check(self.lib.cipc_get_tuning(self.h, ctypes.byref(t)))
        return dict(spmv_grid=t.spmv_grid, spmv_lanes=t.spmv_lanes, spmv_unroll=t.spmv_unroll, factor_grid=t.factor_grid,
                    factor_lanes=t.factor_lanes, cell_mult=t.cell_mult, tuning_pcg=bool(t.tuning_pcg),
                    tuning_cell=bool(t.tuning_cell))

    def set_tuning(self, spmv_grid=1024, spmv_lanes=16, spmv_unroll=1, factor_grid=2048, factor_lanes=4, cell_mult=1.0):
        t = _lib.Tuning(spmv_grid=spmv_grid, spmv_lanes=spmv_lanes, spmv_unroll=spmv_unroll, factor_grid=factor_grid,
                        factor_lanes=factor_lanes, cell_mult=cell_mult)
        self._check(self.lib.cipc_set_tuning(self.h, ctypes.byref(t)))

    def tet_sizes(self):
        """(tets, contact triangles: the tets' surface then the shell's)"""
        s = np.zeros(2, dtype=np.int32)
        self._check(self.lib.cipc_get_tet_sizes(self.h, _i(s)))
        return int(s[0]), int(s[1])

    def inversion_free_stepsize(self, x, p, alpha=1.0):
        """Compute_Inversion_Free_StepSize (IPC.h:1785-1877)"""
        out = np.zeros(1)
        self._check(self.lib.cipc_inversion_free_stepsize(self.h, _d(self._x(x)), _d(self._x(p)), float(alpha), _d(out)))
        return out[0]

    def codim_setup(self):
        """Rod segment data, rod hinges and the boundary-node list (DISCRETE_ROD.h:87-165,
        IMPLICIT_EULER.h:243-277)."""
        ri = np.zeros(3 * self.nRod)
        rh = np.zeros(3 * self.nRodHinge, dtype=np.int32)
        rhi = np.zeros(3 * self.nRodHinge)
        bn = np.zeros(self.nBN, dtype=np.int32)
        self._check(self.lib.cipc_get_codim_setup(self.h, _d(ri), _i(rh), _d(rhi), _i(bn)))
        return dict(rod_info=ri.reshape(-1, 3), rod_hinge=rh.reshape(-1, 3), rod_hinge_info=rhi.reshape(-1, 3), bnode=bn)

    def get_state(self):
        x = np.zeros(3 * self.nV)
        v = np.zeros(3 * self.nV)
        self._check(self.lib.cipc_get_state(self.h, _d(x), _d(v)))
        return x, v

    def set_state(self, x=None, v=None):
        xx = None if x is None else np.ascontiguousarray(x, dtype=np.float64).ravel()
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64).ravel()
        self._check(self.lib.cipc_set_state(self.h, _d(xx), _d(vv)))

    def set_warm_start(self, enable=True):
        """Seed the Newton solves that follow a cut line search with (1 - alpha) p_prev
        (cipc_set_warm_start; on by default)."""
        self._check(self.lib.cipc_set_warm_start(self.h, 1 if enable else 0))

    def set_kappa(self, k):
        self._check(self.lib.cipc_set_kappa(self.h, float(k)))

    # ------------------------------------------------------------ step
    def step(self):
        """Advance_One_Step_IE_Discrete_Shell<double,3,false,false> (IMPLICIT_EULER.h:151-902)."""
        st = _lib.StepStats()
        rc = self.lib.cipc_step(self.h, ctypes.byref(st))
        self._check(rc)
        return st.as_dict()

    def set_precision(self, precision):
        self._check(self.lib.cipc_set_precision(self.h, self.PRECISION[precision]))

    def solver_info(self):
        out = np.zeros(4, np.int32)
        self._check(self.lib.cipc_get_solver_info(self.h, _i(out)))
        return dict(refinements=int(out[0]), graph_captures=int(out[1]), graph_launches=int(out[2]),
                    direct_solves=int(out[3]))

    def set_coarse(self, enable=True, every=1, max_unknowns=0, bf16=False):
        """two-level correction of the learned preconditioner: rigid-body modes on
        aggregates of the elastic vertex graph, Galerkin coarse matrix set up at a step's
        first solve and then every `every` solves, at most `max_unknowns` coarse unknowns
        (0: the library's bound), inverse factor in FP32 or BF16 (cipc_set_coarse; learned
        forms only)"""
        self._check(self.lib.cipc_set_coarse(self.h, int(every) if enable else 0, int(max_unknowns), 1 if bf16 else 0))

    def coarse_info(self):
        out = np.zeros(7)
        self._check(self.lib.cipc_get_coarse_info(self.h, _d(out)))
        return dict(aggregates=int(out[0]), unknowns=int(out[1]), setups=int(out[2]), last_setup_ms=float(out[3]),
                    ready=bool(out[4]), failures=int(out[5]), fp64_retries=int(out[6]))

    LINEAR_SOLVER = {"pcg": 0, "direct": 1}

    def set_linear_solver(self, solver):
        """'pcg' (default; a solve that reaches pcg_max_iter is redone by the dense direct
        solve when the system is small enough) or 'direct' (dense FP64 Cholesky for every
        Newton solve, as the reference's Solve_Direct; at most 24576 unknowns)
        (cipc_set_linear_solver)"""
        self._check(self.lib.cipc_set_linear_solver(self.h, self.LINEAR_SOLVER[solver]))

    def set_output_folder(self, folder):
        """the step's outputFolder: residual.txt, counter.txt, stretch.txt appended per step
        (cipc_set_output_folder; None: off)"""
        self._check(self.lib.cipc_set_output_folder(self.h, None if folder is None else str(folder).encode()))

    def set_static_solve(self, enable=True):
        """the step's staticSolve argument (cipc_set_static_solve)"""
        self._check(self.lib.cipc_set_static_solve(self.h, 1 if enable else 0))

    def set_graphs(self, enable=True):
        """replay captured HIP graphs of the PCG chunks (cipc_set_graphs; on by default)"""
        self._check(self.lib.cipc_set_graphs(self.h, 1 if enable else 0))

    # ------------------------------------------------------------ learned preconditioner
    def precond_init(self, seed=1):
        self._check(self.lib.cipc_precond_init(self.h, ctypes.c_uint64(seed)))

    def precond_weights(self):
        n = self.lib.cipc_precond_num_params()
        w = np.zeros(n, np.float32)
        self._check(self.lib.cipc_precond_get_weights(self.h, w.ctypes.data_as(_lib.F32), n))
        return w

    def precond_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.lib.cipc_precond_set_weights(self.h, w.ctypes.data_as(_lib.F32), len(w)))

    def set_preconditioner(self, mode, budget=0, reuse=None):
        """mode "jacobi" (block Jacobi) or "learned" (falls back to block Jacobi after `budget`
        iterations).  reuse: run the network at the first Newton solve of a step and then
        every `reuse` solves (the others rescale its last factor)."""
        self._check(self.lib.cipc_set_preconditioner(self.h, {"jacobi": 0, "learned": 1}[mode], int(budget)))
        if reuse is not None:
            self._check(self.lib.cipc_set_precond_reuse(self.h, int(reuse)))

    def precond_loss_grad(self, x, C, kappa, n_probes=8, seed=1, want_grad=True):
        x, C = self._x(x), self._cs(C)
        loss = np.zeros(1)
        g = np.zeros(self.lib.cipc_precond_num_params(), np.float32) if want_grad else None
        self._check(self.lib.cipc_precond_loss_grad(self.h, _d(x), _i(C), len(C), float(kappa), int(n_probes), ctypes.c_uint64(seed),
                                                    _d(loss), None if g is None else g.ctypes.data_as(_lib.F32)))
        return loss[0], g

    def precond_train_step(self, x, C, kappa, lr=1e-3, n_probes=8, seed=1):
        x, C = self._x(x), self._cs(C)
        loss = np.zeros(1)
        self._check(self.lib.cipc_precond_train_step(self.h, _d(x), _i(C), len(C), float(kappa), int(n_probes), ctypes.c_uint64(seed),
                                                     float(lr), _d(loss)))
        return loss[0]

    def precond_factor(self, x, C, kappa, mfma=True):
        """factor G (SoA [9][nnzb] float32) at H(x, C); mfma selects the BF16 matrix-core layers"""
        x, C = self._x(x), self._cs(C)
        nnzb = ctypes.c_int32(0)
        self._check(self.lib.cipc_precond_factor(self.h, _d(x), _i(C), len(C), float(kappa), 1 if mfma else 0, None, ctypes.byref(nnzb)))
        g = np.zeros(9 * nnzb.value, np.float32)
        self._check(self.lib.cipc_precond_factor(self.h, _d(x), _i(C), len(C), float(kappa), 1 if mfma else 0, g.ctypes.data_as(_lib.F32),
                                                 ctypes.byref(nnzb)))
        return g.reshape(9, -1)

    def precond_get_state(self):
        """(weights, Adam m, Adam v, Adam step) -- to continue training in another context"""
        n = self.lib.cipc_precond_num_params()
        m, v = np.zeros(n, np.float32), np.zeros(n, np.float32)
        t = ctypes.c_int32(0)
        self._check(self.lib.cipc_precond_get_adam(self.h, m.ctypes.data_as(_lib.F32), v.ctypes.data_as(_lib.F32), ctypes.byref(t), n))
        return self.precond_weights(), m, v, t.value

    def precond_set_state(self, state):
        w, m, v, t = state
        self.precond_set_weights(w)
        m, v = np.ascontiguousarray(m, np.float32), np.ascontiguousarray(v, np.float32)
        self._check(self.lib.cipc_precond_set_adam(self.h, m.ctypes.data_as(_lib.F32), v.ctypes.data_as(_lib.F32), int(t), len(m)))

    def precond_info(self):
        out = np.zeros(8, np.int32)
        self._check(self.lib.cipc_precond_info(self.h, _i(out)))
        return dict(fallback=int(out[0]), solves=int(out[1]), fallbacks=int(out[2]), last_status=int(out[3]), last_iters=int(out[4]),
                    continued=int(out[5]), last_jacobi_iters=int(out[6]))

    def set_profiling(self, enable=True):
        self._check(self.lib.cipc_set_profiling(self.h, 1 if enable else 0))

    def profile(self):
        """Dispatch timing of the PCG SpMV (the step's dominant kernel) since set_profiling."""
        out = np.zeros(6)
        self._check(self.lib.cipc_get_profile(self.h, _d(out)))
        return dict(launches=int(out[0]), total_ms=out[1], block_rows=int(out[2]), nnzb=int(out[3]),
                    bytes_per_launch=out[4], flops_per_launch=out[5])

    # ------------------------------------------------------------ components
    @staticmethod
    def _x(x):
        return np.ascontiguousarray(x, dtype=np.float64).ravel()

    @staticmethod
    def _cs(C):
        return np.ascontiguousarray(C, dtype=np.int32).reshape(-1, 4)

    def constraint_set(self, x):
        """Compute_Constraint_Set (IPC.h:19-740)."""
        x = self._x(x)
        n = ctypes.c_int32(0)
        cap = 1 << 16
        while True:
            out = np.zeros(4 * cap, dtype=np.int32)
            rc = self.lib.cipc_constraint_set(self.h, _d(x), _i(out), cap, ctypes.byref(n))
            if rc == -7:
                cap = n.value
                continue
            self._check(rc)
            return out[:4 * n.value].reshape(-1, 4)

    def barrier_energy(self, x, C, kappa):
        """Compute_Barrier (IPC.h:742-941)."""
        x, C = self._x(x), self._cs(C)
        e = np.zeros(1)
        self._check(self.lib.cipc_barrier_energy(self.h, _d(x), _i(C), len(C), float(kappa), _d(e)))
        return e[0]

    def barrier_energy_noalloc(self, x, C, kappa):
        """Compute_Barrier (IPC.h:742-941) through the entry point that stages everything in
        the library's static per-GPU scratch array (no device allocation; calls that overlap
        on one GPU race)."""
        x, C = self._x(x), self._cs(C)
        e = np.zeros(1)
        self._check(self.lib.cipc_barrier_energy_noalloc(self.h, _d(x), _i(C), len(C), float(kappa), _d(e)))
        return e[0]

    def barrier_gradient(self, x, C, kappa):
        """Compute_Barrier_Gradient (IPC.h:943-1256)."""
        x, C = self._x(x), self._cs(C)
        g = np.zeros(3 * self.nV)
        self._check(self.lib.cipc_barrier_gradient(self.h, _d(x), _i(C), len(C), float(kappa), _d(g)))
        return g

    def elastic_energy(self, x):
        """Compute_Membrane_Energy + Compute_Bending_Energy: (membrane, bending)."""
        out = np.zeros(2)
        self._check(self.lib.cipc_elastic_energy(self.h, _d(self._x(x)), _d(out)))
        return out

    def elastic_gradient(self, x):
        g = np.zeros(3 * self.nV)
        self._check(self.lib.cipc_elastic_gradient(self.h, _d(self._x(x)), _d(g)))
        return g

    def hessian(self, x, C, kappa, which=3):
        """Compute_IncPotential_Hessian (INC_POTENTIAL.h:277-395) as a scipy CSR matrix."""
        import scipy.sparse as sp
        x, C = self._x(x), self._cs(C)
        nb = ctypes.c_int32(0)
        rc = self.lib.cipc_hessian(self.h, _d(x), _i(C), len(C), float(kappa), which, None, None, None, 0, ctypes.byref(nb))
        if rc not in (0, -7):
            self._check(rc)
        cap = nb.value
        br = np.zeros(max(cap, 1), dtype=np.int32)
        bc = np.zeros(max(cap, 1), dtype=np.int32)
        bv = np.zeros((max(cap, 1), 9))
        self._check(self.lib.cipc_hessian(self.h, _d(x), _i(C), len(C), float(kappa), which, _i(br), _i(bc), _d(bv), cap,
                                          ctypes.byref(nb)))
        n = nb.value
        rows = (3 * br[:n, None] + np.array([0, 0, 0, 1, 1, 1, 2, 2, 2])[None, :]).ravel()
        cols = (3 * bc[:n, None] + np.array([0, 1, 2, 0, 1, 2, 0, 1, 2])[None, :]).ravel()
        return sp.csr_matrix((bv[:n].ravel(), (rows, cols)), shape=(3 * self.nV, 3 * self.nV))

    def min_dist2(self, x, C):
        """Compute_Min_Dist2 (IPC.h:2246-2388): (per-constraint d^2, min d^2 - offset^2)."""
        x, C = self._x(x), self._cs(C)
        d2 = np.zeros(max(len(C), 1))
        m = np.zeros(1)
        self._check(self.lib.cipc_min_dist2(self.h, _d(x), _i(C), len(C), _d(d2), _d(m)))
        return d2[:len(C)], m[0]

    def write_obj(self, path):
        """Write_TriMesh_Obj (MESHIO.h:184-213) of the current positions."""
        self._check(self.lib.cipc_write_obj(self.h, str(path).encode()))

    # ---- moving Dirichlet nodes (FEM/BOUNDARY_CONDITION.h)
    def dirichlet_add_motion(self, begin, end, v=(0, 0, 0), center=(0, 0, 0), axis=(0, 1, 0), ang_vel_deg=0.0):
        """A DBC_MOTION group over Dirichlet entries [begin, end)."""
        arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (v, center, axis)]
        self._check(self.lib.cipc_dirichlet_add_motion(self.h, int(begin), int(end), _d(arr[0]), _d(arr[1]), _d(arr[2]),
                                                       float(ang_vel_deg)))

    def step_dirichlet(self):
        """Step_Dirichlet (BOUNDARY_CONDITION.h:107-155)."""
        self._check(self.lib.cipc_step_dirichlet(self.h))

    def turn_dirichlet(self):
        self._check(self.lib.cipc_turn_dirichlet(self.h))

    def dirichlet_targets(self):
        t = np.zeros(3 * self.nDBC)
        self._check(self.lib.cipc_get_dirichlet_targets(self.h, _d(t)))
        return t.reshape(-1, 3)

    def set_dirichlet_targets(self, t):
        self._check(self.lib.cipc_set_dirichlet_targets(self.h, _d(np.ascontiguousarray(t, dtype=np.float64).ravel())))

    # ---- strain limiting (FEM/Shell/INEXT.h)
    def set_strain_limit(self, kappa, s=(1.01, 0.0), sHat=(1.0, 1.0)):
        """The step's kappa_s, s, sHat arguments (index 0 upper limit, 1 lower)."""
        k, s_, sh = (np.ascontiguousarray(v, dtype=np.float64) for v in (kappa, s, sHat))
        self._check(self.lib.cipc_set_strain_limit(self.h, _d(k), _d(s_), _d(sh)))

    def strain_limit(self):
        """(kappa_s, s, sHat) as the steps have left them (kappa_s[0] doubles in place at the
        limit, IMPLICIT_EULER.h:606-645)"""
        k, s_, sh = np.zeros(2), np.zeros(2), np.zeros(2)
        self._check(self.lib.cipc_get_strain_limit(self.h, _d(k), _d(s_), _d(sh)))
        return k, s_, sh

    def strain_limit_energy(self, x):
        """Compute_Inextensibility_Energy (INEXT.h:120-178): (valid, energy)."""
        e = np.zeros(1)
        ok = np.zeros(1, dtype=np.int32)
        self._check(self.lib.cipc_strain_limit_energy(self.h, _d(self._x(x)), _d(e), _i(ok)))
        return bool(ok[0]), e[0]

    # ---- lagged friction (FEM/FRICTION.h)
    def set_friction(self, mu, epsv2=1e-6, iter_amt=1):
        """The step's mu, epsv2 and fricIterAmt arguments (IMPLICIT_EULER.h:167)."""
        self._check(self.lib.cipc_set_friction(self.h, float(mu), float(epsv2), int(iter_amt)))

    def set_friction_components(self, comp_node_range, mu_comp):
        """compNodeRange / muComp (Compute_Friction_Coef, FRICTION.h:114-171)."""
        r = np.ascontiguousarray(comp_node_range, dtype=np.int32)
        m = np.ascontiguousarray(mu_comp, dtype=np.float64).ravel()
        if len(m) != len(r) * len(r):
            raise ValueError("mu_comp needs %d coefficients (n_components^2), got %d" % (len(r) * len(r), len(m)))
        self._check(self.lib.cipc_set_friction_components(self.h, len(r), _i(r), _d(m)))

    def friction_basis(self, x, C, kappa):
        """Compute_Friction_Basis (FRICTION.h:17-112): (stencils (n, 4), info (n, 9)); the
        stencils stay in the context for friction_energy / _gradient / _hessian."""
        x, C = self._x(x), self._cs(C)
        cap = max(len(C), 1)
        oc = np.zeros(4 * cap, dtype=np.int32)
        oi = np.zeros(9 * cap)
        n = ctypes.c_int32(0)
        self._check(self.lib.cipc_friction_basis(self.h, _d(x), _i(C), len(C), float(kappa), _i(oc), _d(oi), cap, ctypes.byref(n)))
        return oc[:4 * n.value].reshape(-1, 4), oi[:9 * n.value].reshape(-1, 9)

    def friction_energy(self, x, xn):
        """Compute_Friction_Potential (FRICTION.h:173-243) of the kept stencils."""
        e = np.zeros(1)
        self._check(self.lib.cipc_friction_energy(self.h, _d(self._x(x)), _d(self._x(xn)), _d(e)))
        return e[0]

    def friction_gradient(self, x, xn):
        """Compute_Friction_Gradient (FRICTION.h:245-394)."""
        g = np.zeros(3 * self.nV)
        self._check(self.lib.cipc_friction_gradient(self.h, _d(self._x(x)), _d(self._x(xn)), _d(g)))
        return g

    def friction_hessian(self, x, xn):
        """Compute_Friction_Hessian (FRICTION.h:396-661) as a scipy CSR matrix (Dirichlet rows/columns dropped)."""
        import scipy.sparse as sp
        x, xn = self._x(x), self._x(xn)
        nb = ctypes.c_int32(0)
        rc = self.lib.cipc_friction_hessian(self.h, _d(x), _d(xn), None, None, None, 0, ctypes.byref(nb))
        if rc not in (0, -7):
            self._check(rc)
        cap = max(nb.value, 1)
        br = np.zeros(cap, dtype=np.int32)
        bc = np.zeros(cap, dtype=np.int32)
        bv = np.zeros((cap, 9))
        self._check(self.lib.cipc_friction_hessian(self.h, _d(x), _d(xn), _i(br), _i(bc), _d(bv), cap, ctypes.byref(nb)))
        n = nb.value
        rows = (3 * br[:n, None] + np.array([0, 0, 0, 1, 1, 1, 2, 2, 2])[None, :]).ravel()
        cols = (3 * bc[:n, None] + np.array([0, 1, 2, 0, 1, 2, 0, 1, 2])[None, :]).ravel()
        return sp.csr_matrix((bv[:n].ravel(), (rows, cols)), shape=(3 * self.nV, 3 * self.nV))

    def ccd_stepsize(self, x, p, alpha=1.0):
        """Compute_Intersection_Free_StepSize (IPC.h:1879-2244): (alpha, alpha after the hash clamp)."""
        a = np.zeros(1)
        cl = np.zeros(1)
        self._check(self.lib.cipc_ccd_stepsize(self.h, _d(self._x(x)), _d(self._x(p)), float(alpha), _d(a), _d(cl)))
        return a[0], cl[0]

    def linear_solve(self, x, C, kappa, b, tol=1e-5):
        """Assemble H(x, C) and solve H s = b by block-Jacobi PCG (stands in for Solve_Direct)."""
        x, C, b = self._x(x), self._cs(C), self._x(b)
        sol = np.zeros(3 * self.nV)
        it = ctypes.c_int32(0)
        rr = np.zeros(1)
        self._check(self.lib.cipc_linear_solve(self.h, _d(x), _i(C), len(C), float(kappa), _d(b), _d(sol), float(tol),
                                               ctypes.byref(it), _d(rr)))
        return sol, it.value, rr[0]
