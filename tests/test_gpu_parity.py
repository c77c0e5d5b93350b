"""HIP path vs the CPU FP64 oracle, through the C ABI (include/cipc.h).

Bars (stated per test): integer/index results bit-exact (constraint sets, set-up
topology); distances bit-exact (same IEEE operation order, no contraction); energies,
gradients and Hessians to a relative tolerance (different summation order, different
eigensolver in the PSD projection); CCD step never later than the oracle's; PCG
solutions to the stated relative residual of the oracle's own assembled matrix.
"""
import numpy as np
import pytest

import oracle as O
from bars import trajectory_bar
from cipc import Simulator, scenes

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


SCENES = {
    # every constraint kind active in the first state, all primitives grid-listed
    "contact_stack": scenes.scene_contact_stack,
    # ground triangles/edges cover thousands of cells: large-primitive path of both broad phases
    "cloth_on_ground": scenes.scene_cloth_on_ground,
    # rods and particles on the sheet (offset 1 mm): codimensional point-edge / point-point
    # entries, rod springs and hinges, particle CCD
    "codim_contact": scenes.scene_codim_contact,
}


@pytest.fixture(scope="module", params=sorted(SCENES))
def stack(request, require_gpu):
    sd = SCENES[request.param]()
    gpu = Simulator(sd)
    orc = O.OracleScene(sd, threads=4)
    # states: the initial one (every constraint kind active) and two after oracle steps
    states = []
    x, _ = orc.get_state()
    states.append(x.copy())
    for _ in range(2):
        rc, _ = orc.step()
        assert rc > 0
        states.append(orc.get_state()[0].copy())
    return sd, gpu, orc, states


def test_setup_matches_oracle(stack):
    sd, gpu, orc, _ = stack
    a, g = orc.arrays(), gpu.setup_arrays()
    assert np.array_equal(g["hinge"], a["hinge"])
    assert np.array_equal(g["bedge"], a["bedge"])
    assert np.array_equal(g["mass"], a["mass"])
    assert np.array_equal(g["hinge_info"], a["hinge_info"])
    so, sg = orc.scalars(), gpu.scalars()
    for k in ("kappa", "kappaMax", "dHat2", "lam", "mu", "kBend"):
        assert sg[k] == so[k], k
    assert (gpu.nRod, gpu.nRodHinge, gpu.nParticle, gpu.codimBN0, gpu.codimBN1) == \
        (orc.nRod, orc.nRodHinge, orc.nParticle, orc.codimBN0, orc.codimBN1)
    cg, co = gpu.codim_setup(), orc.codim_arrays()
    assert np.array_equal(cg["bnode"], a["bnode"])
    for k in ("rod_info", "rod_hinge", "rod_hinge_info"):
        assert np.array_equal(cg[k], co[k]), k


def test_constraint_set_bit_exact(stack):
    sd, gpu, orc, states = stack
    for x in states:
        Cg = gpu.constraint_set(x)
        Co = orc.constraint_set(x)
        Cb = orc.constraint_set(x, brute=True)
        assert np.array_equal(Co, Cb)
        assert Cg.shape == Co.shape
        assert np.array_equal(Cg, Co)
    assert len(gpu.constraint_set(states[0])) > 200


def test_min_dist2_bit_exact(stack):
    sd, gpu, orc, states = stack
    x = states[0]
    C = orc.constraint_set(x)
    d2g, mg = gpu.min_dist2(x, C)
    d2o = orc.dist2_list(x, C)
    assert np.array_equal(d2g, d2o)
    assert mg == d2o.min() - sd.params["xi"] ** 2


def test_barrier_energy_gradient(stack):
    sd, gpu, orc, states = stack
    kappa = orc.scalars()["kappa"]
    for x in states:
        C = orc.constraint_set(x)
        eo, bad = orc.energy(x, None, C, kappa)
        assert bad == 0
        eg = gpu.barrier_energy(x, C, kappa)
        assert abs(eg - eo[3]) <= 1e-11 * abs(eo[3])          # summation order only
        _, gbo = orc.gradient(x, C, kappa)
        gbg = gpu.barrier_gradient(x, C, kappa)
        assert rel(gbg, gbo) <= 1e-11


def test_barrier_hessian(stack):
    """Barrier blocks after PSD projection (Jacobi on the GPU, Householder-QL in the oracle)."""
    sd, gpu, orc, states = stack
    kappa = orc.scalars()["kappa"]
    none = np.zeros((0, 4), np.int32)
    for x in states[:2]:
        C = orc.constraint_set(x)
        # the oracle always adds the mass diagonal and projects Dirichlet rows; remove that part
        D = orc.hessian(x, C, kappa, which=2) - orc.hessian(x, none, kappa, which=0)
        Hg = gpu.hessian(x, C, kappa, which=2)
        assert rel(Hg.toarray(), D.toarray()) <= 1e-9


def test_elastic_energy_gradient_hessian(stack):
    sd, gpu, orc, states = stack
    x = states[2]
    eo, _ = orc.energy(x, None, np.zeros((0, 4), np.int32), 1.0)
    eg = gpu.elastic_energy(x)
    # near the rest shape each triangle's energy is a difference of terms of size
    # h^2 (lam + mu) vol, so the sum carries cancellation error of a few ulps of
    # h^2 (lam + mu) sum(vol): bar = 1e-8 relative + 4 eps of that scale
    sc = orc.scalars()
    scale = sd.params["dt"] ** 2 * (sc["lam"] + sc["mu"]) * float(np.sum(orc.arrays()["vol"]))
    assert abs(eg[0] - eo[0]) <= 1e-8 * abs(eo[0]) + 4 * np.finfo(float).eps * scale
    assert abs(eg[1] - eo[1]) <= 1e-8 * abs(eo[1]) + 1e-30
    geo, _ = orc.gradient(x, np.zeros((0, 4), np.int32), 1.0)
    geg = gpu.elastic_gradient(x)
    assert rel(geg, geo) <= 1e-10
    Ho = orc.hessian(x, np.zeros((0, 4), np.int32), 1.0, which=1)
    Hg = gpu.hessian(x, np.zeros((0, 4), np.int32), 1.0, which=1)
    assert rel(Hg.toarray(), Ho.toarray()) <= 1e-9


def test_ccd_matches_oracle(stack):
    sd, gpu, orc, states = stack
    rng = np.random.default_rng(7)
    x = states[0]
    n = x.size
    dbc = np.zeros(sd.nV, bool)
    dbc[sd.dbc_idx] = True
    for scale in (1e-4, 1e-3, 1e-2):
        p = rng.standard_normal(n) * scale
        p.reshape(-1, 3)[dbc] = 0
        ao, clo = orc.ccd(x, p, 1.0)
        ag, clg = gpu.ccd_stepsize(x, p, 1.0)
        assert clg == clo
        assert ag <= ao                        # never later than the FP64 oracle
        assert ag >= ao * (1 - 1e-9)


def test_linear_solve_residual(stack):
    sd, gpu, orc, states = stack
    import scipy.sparse.linalg as spla
    kappa = orc.scalars()["kappa"]
    x = states[0]
    C = orc.constraint_set(x)
    H = orc.hessian(x, C, kappa, which=3)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(x.size)
    b.reshape(-1, 3)[sd.dbc_idx] = 0
    sol, it, rr = gpu.linear_solve(x, C, kappa, b, tol=1e-8)
    assert rr <= 1e-8
    assert np.linalg.norm(H @ sol - b) <= 2e-8 * np.linalg.norm(b)
    ref = spla.spsolve(H.tocsc(), b)
    assert rel(sol, ref) <= 1e-5


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_linear_solve_mixed_refinement(stack, precision):
    """FP32 blocks (and, in the fp32 form, FP32 vectors) inside FP64 iterative refinement:
    the FP64 residual of the oracle's own assembled matrix meets the requested 1e-8, which
    takes more than one sweep in the fp32 form (its inner floor is 1e-6)."""
    sd, gpu, orc, states = stack
    kappa = orc.scalars()["kappa"]
    x = states[0]
    C = orc.constraint_set(x)
    H = orc.hessian(x, C, kappa, which=3)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(x.size)
    b.reshape(-1, 3)[sd.dbc_idx] = 0
    gpu.set_precision(precision)
    try:
        sol, it, rr = gpu.linear_solve(x, C, kappa, b, tol=1e-8)
        sweeps = gpu.solver_info()["refinements"]
    finally:
        gpu.set_precision("fp64")
    assert rr <= 1e-8
    assert np.linalg.norm(H @ sol - b) <= 2e-8 * np.linalg.norm(b)
    assert sweeps >= (2 if precision == "fp32" else 1)


def test_direct_solve(stack):
    """Dense FP64 Cholesky of the Newton matrix on the GPU (linear solver 'direct', the
    reference's Solve_Direct).  On the library's own assembled matrix the residual is of
    the order of LAPACK's dense solve of that matrix (at most 10x it: a backward-stable
    solve's round-off) and the solution equals scipy's sparse LU solution to 1e-9; on the
    oracle's assembled matrix (equal to 1e-9, different PSD eigensolvers) the residual meets
    the 2e-8 bar of the PCG test."""
    sd, gpu, orc, states = stack
    import scipy.sparse.linalg as spla
    kappa = orc.scalars()["kappa"]
    x = states[1]
    C = orc.constraint_set(x)
    H = orc.hessian(x, C, kappa, which=3)
    rng = np.random.default_rng(7)
    b = rng.standard_normal(x.size)
    b.reshape(-1, 3)[sd.dbc_idx] = 0
    before = gpu.solver_info()["direct_solves"]
    gpu.set_linear_solver("direct")
    try:
        sol, it, rr = gpu.linear_solve(x, C, kappa, b, tol=1e-8)
    finally:
        gpu.set_linear_solver("pcg")
    assert gpu.solver_info()["direct_solves"] == before + 1
    assert it == 0
    assert rr <= 1e-10
    Hg = gpu.hessian(x, C, kappa, which=3)
    lapack = np.linalg.norm(Hg @ np.linalg.solve(Hg.toarray(), b) - b)
    assert np.linalg.norm(Hg @ sol - b) <= 10 * lapack + 1e-14 * np.linalg.norm(b)
    assert rel(sol, spla.spsolve(Hg.tocsc(), b)) <= 1e-9
    assert np.linalg.norm(H @ sol - b) <= 2e-8 * np.linalg.norm(b)


def test_direct_solve_takes_over_at_iteration_limit(require_gpu):
    """A PCG solve that reaches pcg_max_iter is redone by the direct solve (systems of at
    most 24576 unknowns), so the solve still meets its tolerance."""
    sd = scenes.scene_contact_stack()
    gpu = Simulator(sd, pcg_max_iter=3)
    orc = O.OracleScene(sd, threads=4)
    kappa = orc.scalars()["kappa"]
    x = orc.get_state()[0]
    C = orc.constraint_set(x)
    H = orc.hessian(x, C, kappa, which=3)
    b = np.random.default_rng(9).standard_normal(x.size)
    b.reshape(-1, 3)[sd.dbc_idx] = 0
    sol, it, rr = gpu.linear_solve(x, C, kappa, b, tol=1e-8)
    assert it == 3
    assert gpu.solver_info()["direct_solves"] == 1
    assert rr <= 1e-10
    Hg = gpu.hessian(x, C, kappa, which=3)
    lapack = np.linalg.norm(Hg @ np.linalg.solve(Hg.toarray(), b) - b)
    assert np.linalg.norm(Hg @ sol - b) <= 10 * lapack + 1e-14 * np.linalg.norm(b)
    assert np.linalg.norm(H @ sol - b) <= 2e-8 * np.linalg.norm(b)


def test_direct_solve_size_limit(require_gpu):
    """The dense direct solve is refused above 24576 unknowns (8192 vertices)."""
    sd = scenes.scene_cloth_on_ground(n=100)
    assert 3 * sd.X.shape[0] > 24576
    gpu = Simulator(sd)
    with pytest.raises(Exception) as e:
        gpu.set_linear_solver("direct")
    assert getattr(e.value, "code", None) == -5
    small = Simulator(scenes.scene_contact_stack())
    small.set_linear_solver("direct")
    small.set_linear_solver("pcg")


@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp32", "direct"])
def test_step_trajectory(require_gpu, precision):
    """Five steps of the contact stack, tight Newton tolerance; FP64 oracle.  Before each
    step the oracle takes the GPU's state (positions, velocities, barrier stiffness), so
    every step is compared from the same input (bars.trajectory_bar: 1e-6 of the step's
    displacement or the Newton stopping resolution).  The mixed solver refines to the same
    1e-10 linear residual, and the direct solve (dense Cholesky) solves exactly, so the
    same bar applies."""
    sd = scenes.scene_contact_stack()
    sd.params["PNTol"] = 1e-6 * sd.bbox_diag()
    gpu = Simulator(sd, pcg_tol=1e-10, precision="fp64" if precision == "direct" else precision)
    if precision == "direct":
        gpu.set_linear_solver("direct")
    orc = O.OracleScene(sd, threads=4)
    for s in range(5):
        x0, v0 = gpu.get_state()
        orc.set_state(x0, v0)
        orc.set_kappa(gpu.scalars()["kappa"])
        st = gpu.step()
        rc, so = orc.step()
        assert rc > 0 and st["status"] == 0
        xg, vg = gpu.get_state()
        xo, vo = orc.get_state()
        disp = np.abs(xo - x0).max()
        assert np.abs(xg - xo).max() <= trajectory_bar(sd, disp), (s, st["newton_iters"], so)
        d2, m = gpu.min_dist2(xg, gpu.constraint_set(xg)) if len(gpu.constraint_set(xg)) else (None, 1.0)
        assert m > 0
    if precision == "direct":
        assert gpu.solver_info()["direct_solves"] > 0 and st["pcg_iters"] == 0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_codim_step_trajectory(require_gpu, precision):
    """Six steps of the falling rods / particles / sheet mix (offset 1 mm) against the
    FP64 oracle, tight Newton tolerance, each from the GPU's state as in
    test_step_trajectory; every state intersection free."""
    sd = scenes.scene_codim_small()
    sd.params["PNTol"] = 1e-6 * sd.bbox_diag()
    gpu = Simulator(sd, pcg_tol=1e-10, precision=precision)
    orc = O.OracleScene(sd, threads=4)
    for s in range(6):
        x0, v0 = gpu.get_state()
        orc.set_state(x0, v0)
        orc.set_kappa(gpu.scalars()["kappa"])
        st = gpu.step()
        rc, so = orc.step()
        assert rc > 0 and st["status"] == 0
        xg, _ = gpu.get_state()
        xo, _ = orc.get_state()
        disp = np.abs(xo - x0).max()
        assert np.abs(xg - xo).max() <= trajectory_bar(sd, disp), (s, st["newton_iters"], so)
        C = gpu.constraint_set(xg)
        if len(C):
            assert gpu.min_dist2(xg, C)[1] > 0          # min d^2 - offset^2
    assert st["contacts"] > 0


def test_c3_one_step(require_gpu):
    """The 324k-vertex codimensional mix (C1 + 20,000 rods + 100,000 particles): one step in
    FP64 and in the FP32 form at a tight Newton / linear tolerance; they agree to 1e-6 of
    the step's displacement."""
    sd = scenes.scene_c3()
    sd.params["PNTol"] = 1e-6 * sd.bbox_diag()
    out = {}
    for prec in ("fp64", "fp32"):
        gpu = Simulator(sd, pcg_tol=1e-10, precision=prec)
        st = gpu.step()
        assert st["status"] == 0
        x, _ = gpu.get_state()
        out[prec] = x
        print(prec, {k: st[k] for k in ("newton_iters", "pcg_iters", "contacts", "t_total_ms")}, flush=True)
        gpu.close()
    disp = np.abs(out["fp64"] - sd.X.ravel()).max()
    assert disp > 0
    assert np.abs(out["fp32"] - out["fp64"]).max() <= 1e-6 * disp + 1e-9


def test_constraint_entries_outside_vertex_range_rejected(require_gpu):
    """a constraint entry outside [-n_vertices, n_vertices) would be read out of bounds on the
    device: every entry point taking a constraint set returns CIPC_E_INVALID instead"""
    from cipc.sim import CipcError
    sd = scenes.scene_contact_stack(n=6)
    gpu = Simulator(sd)
    x, _ = gpu.get_state()
    for bad in ([-1, 0, 1, sd.nV], [-sd.nV - 1, 0, 1, 2]):
        C = np.array([bad], dtype=np.int32)
        for call in (lambda: gpu.barrier_energy(x, C, 1.0), lambda: gpu.barrier_gradient(x, C, 1.0),
                     lambda: gpu.min_dist2(x, C), lambda: gpu.hessian(x, C, 1.0)):
            with pytest.raises(CipcError) as e:
                call()
            assert e.value.code == -1
    gpu.close()
