"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker -- never by the product path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


def _dp(a):
    return None if a is None else a.ctypes.data_as(D)


def _ip(a):
    return None if a is None else a.ctypes.data_as(I)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.orc_dist2.restype = ctypes.c_double
        L.orc_dist2_unclassified.restype = ctypes.c_double
        L.orc_scene_create.restype = ctypes.c_void_p
        L.orc_scene_create_codim.restype = ctypes.c_void_p
        L.orc_scene_create_full.restype = ctypes.c_void_p
        L.orc_scene_ok.argtypes = [ctypes.c_void_p]
        L.orc_tet_count.argtypes = [ctypes.c_void_p]
        L.orc_tet_surface.argtypes = [ctypes.c_void_p, I, ctypes.c_int]
        L.orc_tet_arrays.argtypes = [ctypes.c_void_p, D, D, D]
        L.orc_tet_hess_local.argtypes = [ctypes.c_void_p, D, ctypes.c_int, ctypes.c_double, ctypes.c_int, D]
        L.orc_nh_dPdF.argtypes = [D, ctypes.c_double, ctypes.c_double, ctypes.c_int, D]
        L.orc_cubic_root.argtypes = [ctypes.c_double] * 4
        L.orc_cubic_root.restype = ctypes.c_double
        L.orc_inversion_free_stepsize.argtypes = [ctypes.c_void_p, D, D, ctypes.c_double]
        L.orc_inversion_free_stepsize.restype = ctypes.c_double
        L.orc_ccd.restype = ctypes.c_double
        for name in ("orc_scene_destroy", "orc_scene_sizes", "orc_scene_scalars", "orc_scene_arrays", "orc_get_state",
                     "orc_set_state", "orc_set_kappa", "orc_constraint_set", "orc_energy", "orc_gradient", "orc_hessian",
                     "orc_barrier_blocks", "orc_dist2_list", "orc_ccd", "orc_step", "orc_set_threads", "orc_set_solver"):
            getattr(L, name).argtypes = None
        L.orc_set_kappa.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_ccd.argtypes = [ctypes.c_void_p, D, D, ctypes.c_double, D]
        L.orc_set_solver.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.orc_set_max_newton.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_get_kappa_s.argtypes = [ctypes.c_void_p]
        L.orc_get_kappa_s.restype = ctypes.c_double
        L.orc_set_project_pd.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_static_solve.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_energy_trace.argtypes = [ctypes.c_void_p, D, D, ctypes.c_int]
        L.orc_barrier.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, D]
        L.orc_energy.argtypes = [ctypes.c_void_p, D, D, I, ctypes.c_int, ctypes.c_double, D]
        L.orc_gradient.argtypes = [ctypes.c_void_p, D, I, ctypes.c_int, ctypes.c_double, D, D]
        L.orc_hessian.argtypes = [ctypes.c_void_p, D, I, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.orc_barrier_blocks.argtypes = [ctypes.c_void_p, D, I, ctypes.c_int, ctypes.c_double, I, D, ctypes.c_int]
        L.orc_membrane.argtypes = [D, D, ctypes.c_double, ctypes.c_double, ctypes.c_double, D, D, D]
        L.orc_solve.argtypes = [ctypes.c_int, I, I, D, D, D, ctypes.c_int, ctypes.c_double, I]
        L.orc_set_friction.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_set_friction_components.argtypes = [ctypes.c_void_p, ctypes.c_int, I, D]
        L.orc_set_strain_limit.argtypes = [ctypes.c_void_p, D, D, D]
        L.orc_set_fiber.argtypes = [ctypes.c_void_p, D, D]
        L.orc_fiber_tri.argtypes = [ctypes.c_void_p, D, ctypes.c_int, ctypes.c_int, D, D, D, D]
        L.orc_dirichlet_add_motion.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, D, D, D, ctypes.c_double]
        L.orc_step_dirichlet.argtypes = [ctypes.c_void_p]
        L.orc_get_dirichlet_targets.argtypes = [ctypes.c_void_p, D]
        L.orc_set_dirichlet_targets.argtypes = [ctypes.c_void_p, D]
        L.orc_strain_tri.argtypes = [ctypes.c_void_p, D, ctypes.c_int, ctypes.c_int, D, D, D, D]
        L.orc_strain_energy_grad.argtypes = [ctypes.c_void_p, D, D, D]
        L.orc_strain_hess.argtypes = [ctypes.c_void_p, D, I, I, D]
        L.orc_friction_basis.argtypes = [ctypes.c_void_p, D, I, ctypes.c_int, ctypes.c_double, I, D, ctypes.c_int]
        L.orc_friction_energy.restype = ctypes.c_double
        L.orc_friction_energy.argtypes = [ctypes.c_void_p, D, D, I, ctypes.c_int, D]
        L.orc_friction_grad.argtypes = [ctypes.c_void_p, D, D, I, ctypes.c_int, D, D]
        L.orc_friction_hess.argtypes = [ctypes.c_void_p, D, D, I, ctypes.c_int, D, I, I, D]
        _LIB = L
    return _LIB


# ------------------------------------------------------------------ primitives
KIND = {"pp": 0, "pe": 1, "pt": 2, "ee": 3}


def dist2(kind, pts):
    p = np.ascontiguousarray(pts, dtype=np.float64)
    return lib().orc_dist2(KIND[kind], _dp(p))


def dist2_gh(kind, pts):
    p = np.ascontiguousarray(pts, dtype=np.float64)
    n = p.size
    g = np.zeros(n)
    H = np.zeros((n, n))
    lib().orc_dist2_gh(KIND[kind], _dp(p), _dp(g), _dp(H))
    return g, H


def dtype(kind, pts):
    p = np.ascontiguousarray(pts, dtype=np.float64)
    return lib().orc_type(KIND[kind], _dp(p))


def dist2_unclassified(kind, pts):
    p = np.ascontiguousarray(pts, dtype=np.float64)
    return lib().orc_dist2_unclassified(KIND[kind], _dp(p))


def mollifier(pts, rest):
    p = np.ascontiguousarray(pts, dtype=np.float64)
    r = np.ascontiguousarray(rest, dtype=np.float64)
    e = np.zeros(1)
    eps = np.zeros(1)
    g = np.zeros(12)
    H = np.zeros((12, 12))
    lib().orc_mollifier(_dp(p), _dp(r), _dp(e), _dp(g), _dp(H), _dp(eps))
    return e[0], g, H, eps[0]


def barrier(d, dhat, kappa):
    out = np.zeros(3)
    lib().orc_barrier(d, dhat, kappa, _dp(out))
    return out


def sym_eig(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n))
    lib().orc_sym_eig(n, _dp(A), _dp(w), _dp(V))
    return w, V


def make_pd(A):
    A = np.array(A, dtype=np.float64, copy=True)
    lib().orc_make_pd(A.shape[0], _dp(A))
    return A


def membrane(x, B, vol, lam, mu):
    x = np.ascontiguousarray(x, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    E = np.zeros(1)
    g = np.zeros(9)
    H = np.zeros((9, 9))
    lib().orc_membrane(_dp(x), _dp(B), vol, lam, mu, _dp(E), _dp(g), _dp(H))
    return E[0], g, H


def dihedral(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    th = np.zeros(1)
    g = np.zeros(12)
    H = np.zeros((12, 12))
    lib().orc_dihedral(_dp(x), _dp(th), _dp(g), _dp(H))
    return th[0], g, H


def rod_spring(x, Er, lbar):
    x = np.ascontiguousarray(x, dtype=np.float64)
    E = np.zeros(1)
    g = np.zeros(6)
    H = np.zeros((6, 6))
    lib().orc_rod_spring(_dp(x), ctypes.c_double(Er), ctypes.c_double(lbar), _dp(E), _dp(g), _dp(H))
    return E[0], g, H


def rod_kappa2(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    k2 = np.zeros(1)
    g = np.zeros(9)
    H = np.zeros((9, 9))
    lib().orc_rod_kappa2(_dp(x), _dp(k2), _dp(g), _dp(H))
    return k2[0], g, H


def nh_dPdF(F, lam, mu, project=True):
    """d2 psi / dF2 of the Neo-Hookean energy (9x9, vec(F) column-major), projected
    block-wise as NEOHOOKEAN.h:192-300 does"""
    Fr = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(3, 3)).ravel()
    D = np.zeros(81)
    lib().orc_nh_dPdF(_dp(Fr), float(lam), float(mu), int(project), _dp(D))
    return D.reshape(9, 9)


def cubic_root(a, b, c, d):
    """smallest positive real root of a t^3 + b t^2 + c t + d (IPC.h:1751-1782), < 0 if none"""
    return lib().orc_cubic_root(a, b, c, d)


def solve(A_csr, b, solver=0, tol=1e-5):
    import scipy.sparse as sp
    A = sp.csr_matrix(A_csr)
    A.sort_indices()
    rp = A.indptr.astype(np.int32)
    ci = A.indices.astype(np.int32)
    val = A.data.astype(np.float64)
    x = np.zeros(A.shape[0])
    it = np.zeros(1, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.float64)
    rc = lib().orc_solve(A.shape[0], _ip(rp), _ip(ci), _dp(val), _dp(b), _dp(x), solver, tol, _ip(it))
    return rc, x, int(it[0])


# ------------------------------------------------------------------ scenes
class OracleScene:
    """A scene held by the oracle; mirrors the state kept by the HIP context."""

    def __init__(self, sd, solver=0, pcg_tol=1e-5, threads=1):
        P = sd.params
        g = P.get("gravity", (0.0, -9.81, 0.0))
        self._params = np.array([P["rho"], P["E"], P["nu"], P["thick"], P["bendMult"], P["dt"], P["dHat"], P["xi"],
                                 P["PNTol"], g[0], g[1], g[2], P.get("kappaStiffMult", 1.0), pcg_tol], dtype=np.float64)
        self._iparams = np.array([P.get("collision", 1), solver, 1000000, threads], dtype=np.int32)
        L = lib()
        self.sd = sd
        rods = getattr(sd, "rods", None)
        parts = getattr(sd, "particles", None)
        rods = np.zeros((0, 2), np.int32) if rods is None else np.ascontiguousarray(rods, dtype=np.int32)
        parts = np.zeros(0, np.int32) if parts is None else np.ascontiguousarray(parts, dtype=np.int32)
        self._codim = np.array([P.get("rodRho", 0.0), P.get("rodE", 0.0), P.get("rodBendMult", 1.0), P.get("rodThick", 1e-3),
                                P.get("partRho", 0.0), P.get("partThick", 1e-3)], dtype=np.float64)
        self._rods, self._parts = rods, parts
        tets = getattr(sd, "tets", None)
        tets = np.zeros((0, 4), np.int32) if tets is None else np.ascontiguousarray(tets, dtype=np.int32)
        self._tets = tets
        self._tetp = np.array([P.get("tetRho", 0.0), P.get("tetE", 0.0), P.get("tetNu", 0.3)], dtype=np.float64)
        self.h = ctypes.c_void_p(L.orc_scene_create_full(
            ctypes.c_int(sd.nV), _dp(sd.X), ctypes.c_int(sd.nT), _ip(sd.tris), ctypes.c_int(len(sd.dbc_idx)),
            _ip(sd.dbc_idx), _dp(sd.dbc_target), _dp(self._params), _ip(self._iparams), ctypes.c_int(len(rods)), _ip(rods),
            ctypes.c_int(len(parts)), _ip(parts), _dp(self._codim), ctypes.c_int(len(tets)), _ip(tets), _dp(self._tetp)))
        if not L.orc_scene_ok(self.h):
            raise ValueError("oracle scene set-up failed (an inverted rest tet)")
        s = np.zeros(5, dtype=np.int32)
        L.orc_scene_sizes(self.h, _ip(s))
        self.nT, self.nBN, self.nBE, self.nBT, self.nH = [int(v) for v in s]
        L.orc_codim_sizes(self.h, _ip(s))
        self.nRod, self.nRodHinge, self.nParticle, self.codimBN0, self.codimBN1 = [int(v) for v in s]
        self.nV = sd.nV
        L.orc_set_friction(self.h, float(P.get("mu", 0.0)), float(P.get("epsv2", 1e-6)), int(P.get("fricIterAmt", 1)))
        if P.get("compNodeRange") is not None:
            self._comp = (np.ascontiguousarray(P["compNodeRange"], dtype=np.int32),
                          np.ascontiguousarray(P["muComp"], dtype=np.float64).ravel())
            L.orc_set_friction_components(self.h, len(self._comp[0]), _ip(self._comp[0]), _dp(self._comp[1]))
        self._sl = [np.array(P.get(k, d), dtype=np.float64) for k, d in
                    (("kappa_s", (0.0, 0.0)), ("s", (1.01, 0.0)), ("sHat", (1.0, 1.0)))]
        L.orc_set_strain_limit(self.h, _dp(self._sl[0]), _dp(self._sl[1]), _dp(self._sl[2]))
        self._fb = [np.array(P.get("fiberStiffMult", (0.0, 0.0, 0.0, 0.0)), dtype=np.float64),
                    np.array(P.get("inextLimit", (1.0, 1.0, 1.0)), dtype=np.float64)]
        L.orc_set_fiber(self.h, _dp(self._fb[0]), _dp(self._fb[1]))

    def __del__(self):
        try:
            lib().orc_scene_destroy(self.h)
        except Exception:
            pass

    def set_threads(self, t):
        lib().orc_set_threads(self.h, ctypes.c_int(t))

    def set_solver(self, solver, tol=1e-5):
        lib().orc_set_solver(self.h, solver, tol)

    # ---- volumetric tets
    def tet_surface(self):
        """Find_Surface_TriMesh (MESHIO.h:486-544) of the scene's tets, (n, 3)"""
        n = lib().orc_tet_surface(self.h, None, 0)
        out = np.zeros(3 * max(n, 1), dtype=np.int32)
        lib().orc_tet_surface(self.h, _ip(out), n)
        return out[:3 * n].reshape(-1, 3)

    def tet_arrays(self):
        n = lib().orc_tet_count(self.h)
        IB, vol, lm = np.zeros(9 * n), np.zeros(n), np.zeros(2)
        lib().orc_tet_arrays(self.h, _dp(IB), _dp(vol), _dp(lm))
        return dict(IB=IB.reshape(-1, 3, 3), vol=vol, lam=lm[0], mu=lm[1])

    def tet_hess_local(self, x, t, w=1.0, project=True):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        H = np.zeros(144)
        lib().orc_tet_hess_local(self.h, _dp(x), int(t), float(w), int(project), _dp(H))
        return H.reshape(12, 12)

    def inversion_free_stepsize(self, x, p, alpha=1.0):
        """Compute_Inversion_Free_StepSize (IPC.h:1785-1877)"""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        p = np.ascontiguousarray(p, dtype=np.float64).ravel()
        return lib().orc_inversion_free_stepsize(self.h, _dp(x), _dp(p), float(alpha))

    def set_strain_limit(self, kappa, s, sHat):
        """the step's kappa_s, s, sHat (kappa_s[0] is then adapted in place by the steps)"""
        self._sl = [np.ascontiguousarray(v, dtype=np.float64) for v in (kappa, s, sHat)]
        lib().orc_set_strain_limit(self.h, _dp(self._sl[0]), _dp(self._sl[1]), _dp(self._sl[2]))

    def kappa_s(self):
        return lib().orc_get_kappa_s(self.h)

    def set_static_solve(self, on):
        """the step's staticSolve argument"""
        lib().orc_set_static_solve(self.h, 1 if on else 0)

    def set_project_pd(self, on):
        """False: Hessians are assembled from unprojected local matrices (finite-difference pins)"""
        lib().orc_set_project_pd(self.h, 1 if on else 0)

    def energy_trace(self):
        """(before, after): incremental potential before and after each Newton iteration's line
        search in the last step"""
        b, a = np.zeros(4096), np.zeros(4096)
        n = lib().orc_energy_trace(self.h, _dp(b), _dp(a), 4096)
        return b[:n], a[:n]

    def set_max_newton(self, k):
        """k > 0: step() returns after k Newton iterations (a timed prefix of a step); 0: off"""
        lib().orc_set_max_newton(self.h, int(k))

    def scalars(self):
        out = np.zeros(6)
        lib().orc_scene_scalars(self.h, _dp(out))
        return dict(kappa=out[0], kappaMax=out[1], dHat2=out[2], lam=out[3], mu=out[4], kBend=out[5])

    def arrays(self):
        mass = np.zeros(self.nV)
        Binv = np.zeros(3 * self.nT)
        vol = np.zeros(self.nT)
        tri = np.zeros(3 * self.nT, dtype=np.int32)
        hinge = np.zeros(4 * self.nH, dtype=np.int32)
        hinfo = np.zeros(3 * self.nH)
        bnode = np.zeros(self.nBN, dtype=np.int32)
        bedge = np.zeros(2 * self.nBE, dtype=np.int32)
        lib().orc_scene_arrays(self.h, _dp(mass), _dp(Binv), _dp(vol), _ip(tri), _ip(hinge), _dp(hinfo), _ip(bnode),
                               _ip(bedge))
        return dict(mass=mass, Binv=Binv.reshape(-1, 3), vol=vol, tri=tri.reshape(-1, 3), hinge=hinge.reshape(-1, 4),
                    hinge_info=hinfo.reshape(-1, 3), bnode=bnode, bedge=bedge.reshape(-1, 2))

    def codim_arrays(self):
        ri = np.zeros(3 * self.nRod)
        rh = np.zeros(3 * self.nRodHinge, dtype=np.int32)
        rhi = np.zeros(3 * self.nRodHinge)
        lib().orc_codim_arrays(self.h, _dp(ri), _ip(rh), _dp(rhi))
        return dict(rod_info=ri.reshape(-1, 3), rod_hinge=rh.reshape(-1, 3), rod_hinge_info=rhi.reshape(-1, 3))

    def get_state(self):
        x = np.zeros(3 * self.nV)
        v = np.zeros(3 * self.nV)
        lib().orc_get_state(self.h, _dp(x), _dp(v))
        return x, v

    def set_state(self, x=None, v=None):
        xx = None if x is None else np.ascontiguousarray(x, dtype=np.float64).ravel()
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64).ravel()
        lib().orc_set_state(self.h, _dp(xx), _dp(vv))

    def set_kappa(self, k):
        lib().orc_set_kappa(self.h, k)

    def constraint_set(self, x, brute=False):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        cap = 1 << 16
        while True:
            out = np.zeros(4 * cap, dtype=np.int32)
            n = lib().orc_constraint_set(self.h, _dp(x), int(brute), _ip(out), cap)
            if n <= cap:
                return out[:4 * n].reshape(-1, 4)
            cap = n

    def energy(self, x, xt, C, kappa):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        xt = None if xt is None else np.ascontiguousarray(xt, dtype=np.float64).ravel()
        C = np.ascontiguousarray(C, dtype=np.int32).reshape(-1, 4)
        out = np.zeros(4)
        bad = lib().orc_energy(self.h, _dp(x), _dp(xt), _ip(C), len(C), kappa, _dp(out))
        return out, bad

    def gradient(self, x, C, kappa):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        C = np.ascontiguousarray(C, dtype=np.int32).reshape(-1, 4)
        ge = np.zeros(3 * self.nV)
        gb = np.zeros(3 * self.nV)
        lib().orc_gradient(self.h, _dp(x), _ip(C), len(C), kappa, _dp(ge), _dp(gb))
        return ge, gb

    def hessian(self, x, C, kappa, which=3):
        import scipy.sparse as sp
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        C = np.ascontiguousarray(C, dtype=np.int32).reshape(-1, 4)
        nnz = lib().orc_hessian(self.h, _dp(x), _ip(C), len(C), kappa, which)
        n = 3 * self.nV
        rp = np.zeros(n + 1, dtype=np.int32)
        ci = np.zeros(nnz, dtype=np.int32)
        val = np.zeros(nnz)
        lib().orc_hessian_fetch(_ip(rp), _ip(ci), _dp(val))
        return sp.csr_matrix((val, ci, rp), shape=(n, n))

    def barrier_blocks(self, x, C, kappa):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        C = np.ascontiguousarray(C, dtype=np.int32).reshape(-1, 4)
        cap = 16 * len(C) + 1
        rc = np.zeros(2 * cap, dtype=np.int32)
        blk = np.zeros(9 * cap)
        n = lib().orc_barrier_blocks(self.h, _dp(x), _ip(C), len(C), kappa, _ip(rc), _dp(blk), cap)
        return rc[:2 * n].reshape(-1, 2), blk[:9 * n].reshape(-1, 3, 3)

    # ---- Dirichlet motion (FEM/BOUNDARY_CONDITION.h)
    def dirichlet_add_motion(self, begin, end, v=(0, 0, 0), center=(0, 0, 0), axis=(0, 1, 0), ang_vel_deg=0.0):
        arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (v, center, axis)]
        lib().orc_dirichlet_add_motion(self.h, int(begin), int(end), _dp(arr[0]), _dp(arr[1]), _dp(arr[2]), float(ang_vel_deg))

    def step_dirichlet(self):
        lib().orc_step_dirichlet(self.h)

    def dirichlet_targets(self):
        t = np.zeros(3 * len(self.sd.dbc_idx))
        lib().orc_get_dirichlet_targets(self.h, _dp(t))
        return t.reshape(-1, 3)

    def set_dirichlet_targets(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64).ravel()
        lib().orc_set_dirichlet_targets(self.h, _dp(t))

    # ---- strain limiting (FEM/Shell/INEXT.h)
    def strain_tri(self, x, t, proj=True):
        """(valid, E, g (9), H (9, 9), singular values (2)) of triangle t."""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        E, g, H, sg = np.zeros(1), np.zeros(9), np.zeros(81), np.zeros(2)
        ok = lib().orc_strain_tri(self.h, _dp(x), int(t), int(proj), _dp(E), _dp(g), _dp(H), _dp(sg))
        return bool(ok), E[0], g, H.reshape(9, 9), sg

    def fiber_tri(self, x, t, proj=True):
        """(feasible, E, g (9), H (9, 9), Green strain (E00, E11, E01)) of triangle t."""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        E, g, H, e3 = np.zeros(1), np.zeros(9), np.zeros(81), np.zeros(3)
        ok = lib().orc_fiber_tri(self.h, _dp(x), int(t), int(proj), _dp(E), _dp(g), _dp(H), _dp(e3))
        return bool(ok), E[0], g, H.reshape(9, 9), e3

    def strain_energy_grad(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        E, g = np.zeros(1), np.zeros(3 * self.nV)
        ok = lib().orc_strain_energy_grad(self.h, _dp(x), _dp(E), _dp(g))
        return bool(ok), E[0], g

    def strain_hessian(self, x):
        import scipy.sparse as sp
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        cap = 9 * self.nT + 1
        r, c, v = np.zeros(cap, dtype=np.int32), np.zeros(cap, dtype=np.int32), np.zeros(9 * cap)
        nb = lib().orc_strain_hess(self.h, _dp(x), _ip(r), _ip(c), _dp(v))
        rows = (3 * r[:nb, None] + np.array([0, 0, 0, 1, 1, 1, 2, 2, 2])[None, :]).ravel()
        cols = (3 * c[:nb, None] + np.array([0, 1, 2, 0, 1, 2, 0, 1, 2])[None, :]).ravel()
        n = 3 * self.nV
        return sp.csr_matrix((v[:9 * nb], (rows, cols)), shape=(n, n))

    # ---- lagged friction (FEM/FRICTION.h)
    def friction_basis(self, x, C, kappa):
        """Compute_Friction_Basis: (stencils int (n, 4), info (n, 9): closest point (2), tangent
        basis columns (2 x 3), normal force)."""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        C = np.ascontiguousarray(C, dtype=np.int32).reshape(-1, 4)
        cap = max(len(C), 1)
        oc = np.zeros(4 * cap, dtype=np.int32)
        oi = np.zeros(9 * cap)
        n = lib().orc_friction_basis(self.h, _dp(x), _ip(C), len(C), kappa, _ip(oc), _dp(oi), cap)
        return oc[:4 * n].reshape(-1, 4), oi[:9 * n].reshape(-1, 9)

    @staticmethod
    def _f(x, xn, FC, FI):
        return (np.ascontiguousarray(x, dtype=np.float64).ravel(), np.ascontiguousarray(xn, dtype=np.float64).ravel(),
                np.ascontiguousarray(FC, dtype=np.int32).reshape(-1, 4), np.ascontiguousarray(FI, dtype=np.float64).reshape(-1, 9))

    def friction_energy(self, x, xn, FC, FI):
        x, xn, FC, FI = self._f(x, xn, FC, FI)
        return lib().orc_friction_energy(self.h, _dp(x), _dp(xn), _ip(FC), len(FC), _dp(FI))

    def friction_gradient(self, x, xn, FC, FI):
        x, xn, FC, FI = self._f(x, xn, FC, FI)
        g = np.zeros(3 * self.nV)
        lib().orc_friction_grad(self.h, _dp(x), _dp(xn), _ip(FC), len(FC), _dp(FI), _dp(g))
        return g

    def friction_hessian(self, x, xn, FC, FI):
        """Friction Hessian as a scipy CSR matrix (local blocks summed; no mass, no Dirichlet projection)."""
        import scipy.sparse as sp
        x, xn, FC, FI = self._f(x, xn, FC, FI)
        cap = 16 * len(FC) + 1
        r = np.zeros(cap, dtype=np.int32)
        c = np.zeros(cap, dtype=np.int32)
        v = np.zeros(9 * cap)
        nb = lib().orc_friction_hess(self.h, _dp(x), _dp(xn), _ip(FC), len(FC), _dp(FI), _ip(r), _ip(c), _dp(v))
        rows = (3 * r[:nb, None] + np.array([0, 0, 0, 1, 1, 1, 2, 2, 2])[None, :]).ravel()
        cols = (3 * c[:nb, None] + np.array([0, 1, 2, 0, 1, 2, 0, 1, 2])[None, :]).ravel()
        n = 3 * self.nV
        return sp.csr_matrix((v[:9 * nb], (rows, cols)), shape=(n, n))

    def dist2_list(self, x, C):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        C = np.ascontiguousarray(C, dtype=np.int32).reshape(-1, 4)
        d2 = np.zeros(len(C))
        lib().orc_dist2_list(self.h, _dp(x), _ip(C), len(C), _dp(d2))
        return d2

    def ccd(self, x, p, alpha=1.0):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        p = np.ascontiguousarray(p, dtype=np.float64).ravel()
        cl = np.zeros(1)
        a = lib().orc_ccd(self.h, _dp(x), _dp(p), alpha, _dp(cl))
        return a, cl[0]

    STAT_NAMES = ("newton_iters", "contacts", "pcg_iters", "gd_fallbacks", "ls_trials", "kappa", "last_res", "E",
                  "t_hessian", "t_solve", "t_constraint_set", "t_ccd", "t_energy", "t_grad", "t_total", "kappa_s",
                  "kappa_s_doublings")

    def step(self):
        st = np.zeros(17)
        rc = lib().orc_step(self.h, _dp(st))
        return rc, dict(zip(self.STAT_NAMES, st.tolist()))
