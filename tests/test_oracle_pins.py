"""Independent pins of oracle functions that per-stencil checks leave open (CPU only):

* hinge bending: closed-form energy of a folded two-triangle hinge, rest data by hand,
  central differences energy -> gradient -> unprojected Hessian, and the projected block
  against numpy's eigh projection of the finite-difference Hessian;
* the membrane and the whole elastic sum over a small curved sheet by central differences;
* the barrier summed over a whole constraint set (point-triangle, edge-edge, mollified
  edge-edge / point-edge, point-edge / point-point with multiplicity) by directional
  central differences of energy -> gradient -> unprojected Hessian;
* the step: the end state of a converged contact step is a stationary point of the
  incremental potential (its Newton decrement, by an independent scipy solve, is below
  the stopping tolerance) and the potential never rises across accepted iterates;
* the shell lumped mass against rho t area / 3 per incident triangle.

The reference ships no values for any of these (BENDING.h:502-600 and MEMBRANE.h:293-386
only print finite-difference errors), so the mathematics is the pin.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from cipc import scenes


def hinge_scene(fold_rest_deg=0.0, bend_mult=1.0):
    """two triangles sharing the edge v0-v1 on the x axis; the apex of triangle (v1, v0, v3)
    is turned about the edge by fold_rest_deg in the rest shape"""
    a = math.radians(fold_rest_deg)
    X = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.4, 0.0, 0.8],
                  [0.6, 0.7 * math.sin(a), -0.7 * math.cos(a)]])
    T = np.array([[0, 1, 2], [1, 0, 3]], dtype=np.int32)
    params = dict(scenes.DEFAULT_PARAMS)
    params.update(bendMult=bend_mult, collision=0, PNTol=1e-3)
    return scenes.SceneData(X, T, np.zeros(0, np.int32), np.zeros((0, 3)), params, "hinge")


def fold(X, deg):
    """turn vertex 3 about the x axis (the hinge edge) by deg"""
    a = math.radians(deg)
    x = X.copy()
    y, z = x[3, 1], x[3, 2]
    x[3, 1], x[3, 2] = math.cos(a) * y - math.sin(a) * z, math.sin(a) * y + math.cos(a) * z
    return x


def k_bend(P):
    """bendingStiffMult * E t^3 / (24 (1 - nu^2)), DISCRETE_SHELL.h:336-342"""
    return P["bendMult"] * P["E"] * P["thick"] ** 3 / (24.0 * (1.0 - P["nu"] ** 2))


def energies(orc, x):
    e, _ = orc.energy(x, None, np.zeros((0, 4), np.int32), 0.0)
    return e   # membrane, bending, inertia, barrier


def elastic_grad(orc, x):
    return orc.gradient(x, np.zeros((0, 4), np.int32), 0.0)[0]


def elastic_hess(orc, sd, x, project):
    """assembled elastic Hessian without the lumped mass the assembly adds; Dirichlet rows
    and columns (projected to the identity by the assembly) are zeroed"""
    orc.set_project_pd(project)
    H = orc.hessian(x, np.zeros((0, 4), np.int32), 0.0, which=1).toarray()
    orc.set_project_pd(True)
    m = orc.arrays()["mass"]
    H = H - np.diag(np.repeat(m, 3))
    fixed = np.repeat(np.isin(np.arange(sd.nV), sd.dbc_idx), 3)
    H[fixed, :] = 0
    H[:, fixed] = 0
    return H


def free_jac(sd, J):
    """a finite-difference Jacobian with the Dirichlet rows and columns zeroed"""
    fixed = np.repeat(np.isin(np.arange(sd.nV), sd.dbc_idx), 3)
    J = J.copy()
    J[fixed, :] = 0
    J[:, fixed] = 0
    return J


def fd_grad(f, x, eps):
    g = np.zeros_like(x)
    for i in range(len(x)):
        xp, xm = x.copy(), x.copy()
        xp[i] += eps
        xm[i] -= eps
        g[i] = (f(xp) - f(xm)) / (2 * eps)
    return g


def fd_jac(f, x, eps):
    cols = []
    for i in range(len(x)):
        xp, xm = x.copy(), x.copy()
        xp[i] += eps
        xm[i] -= eps
        cols.append((f(xp) - f(xm)) / (2 * eps))
    return np.stack(cols, 1)


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def psd_projection(H):
    w, V = np.linalg.eigh(0.5 * (H + H.T))
    return (V * np.maximum(w, 0.0)) @ V.T


# ----------------------------------------------------------------------- hinge bending
@pytest.mark.parametrize("rest_deg,deg", [(0.0, 30.0), (0.0, -55.0), (40.0, 70.0), (40.0, 10.0)])
def test_hinge_bending_closed_form(rest_deg, deg):
    """E = h^2 k (theta - theta_bar)^2 |e_bar| / h_bar (BENDING.h:52-75) with the rest data
    of DISCRETE_SHELL.h:169-213 computed by hand: |e_bar| = 1, h_bar = (|n1| + |n2|) /
    (6 |e_bar|) = (0.8 + 0.7) / 6 = 0.25 (|n| = twice the triangle area); turning one
    triangle about the shared edge is an isometry, so the membrane energy stays zero"""
    sd = hinge_scene(rest_deg)
    orc = O.OracleScene(sd)
    ar = orc.arrays()
    assert ar["hinge"].shape == (1, 4)
    th_bar, e_bar, h_bar = ar["hinge_info"][0]
    # the angle comes from acos of the normals' cosine (DIHEDRAL_ANGLE.h:10-24), which near
    # 0 and pi resolves angles only to about sqrt(2 eps) = 2.1e-8
    assert abs(abs(th_bar) - math.radians(rest_deg)) < 3e-8
    assert abs(e_bar - 1.0) < 1e-15 and abs(h_bar - 0.25) < 1e-15
    k = k_bend(sd.params)
    assert abs(orc.scalars()["kBend"] - k) <= 1e-15 * k
    x = fold(sd.X, deg - rest_deg)
    memb, bend = energies(orc, x.ravel())[:2]
    h2 = sd.params["dt"] ** 2
    exp = h2 * k * math.radians(deg - rest_deg) ** 2 * 1.0 / 0.25
    assert abs(bend - exp) <= 2e-7 * exp
    mscale = h2 * orc.scalars()["mu"] * orc.arrays()["vol"].sum()
    assert abs(memb) <= 1e-15 * mscale    # zero up to rounding of the membrane's own scale
    # rigid motions leave it unchanged
    R = np.linalg.qr(np.random.default_rng(3).normal(size=(3, 3)))[0]
    bend2 = energies(orc, (x @ R.T + np.array([0.3, -2.0, 1.0])).ravel())[1]
    assert abs(bend2 - exp) <= 2e-7 * exp


def test_hinge_bending_derivatives():
    """bending gradient (BENDING.h:177-208) = central difference of the bending energy;
    unprojected bending Hessian (BENDING.h:441-496) = central difference of the gradient;
    the projected block = numpy eigh projection of that finite-difference Hessian.
    The bending part is isolated by subtracting a bendMult = 0 copy of the scene; bendMult
    = 1e4 (a linear factor) lifts it above the rounding of the membrane terms."""
    rng = np.random.default_rng(7)
    for rest_deg in (0.0, 35.0):
        sdA, sdB = hinge_scene(rest_deg, 1e4), hinge_scene(rest_deg, 0.0)
        A, B = O.OracleScene(sdA), O.OracleScene(sdB)
        for trial in range(3):
            x = (fold(sdA.X, rng.uniform(-60, 60)) + rng.uniform(-0.05, 0.05, sdA.X.shape)).ravel()
            bend = lambda y: energies(A, y)[1]
            gb = elastic_grad(A, x) - elastic_grad(B, x)
            assert energies(B, x)[1] == 0.0
            assert relerr(gb, fd_grad(bend, x, 1e-6)) < 1e-7
            Hb = elastic_hess(A, sdA, x, False) - elastic_hess(B, sdB, x, False)
            Hfd = fd_jac(lambda y: elastic_grad(A, y) - elastic_grad(B, y), x, 1e-6)
            assert relerr(Hb, Hfd) < 1e-6
            assert np.allclose(Hb, Hb.T, rtol=0, atol=1e-12 * np.abs(Hb).max())
            Hp = elastic_hess(A, sdA, x, True) - elastic_hess(B, sdB, x, True)
            assert relerr(Hp, psd_projection(Hfd)) < 1e-6


def curved_sheet(n=6, bend_mult=1.0):
    """an n x n sheet with a curved rest shape and two Dirichlet corners"""
    X, T = scenes.cloth_grid(n, 0.0)
    X = X.copy()
    X[:, 1] = 0.15 * np.sin(2.0 * X[:, 0]) * np.cos(3.0 * X[:, 2])
    dbc = np.array([0, n - 1], dtype=np.int32)
    params = dict(scenes.DEFAULT_PARAMS)
    params.update(bendMult=bend_mult, collision=0, PNTol=1e-3)
    return scenes.SceneData(X, T, dbc, X[dbc].copy(), params, "curved")


def test_elastic_sums_by_finite_differences():
    """membrane + hinge bending summed over a curved 6 x 6 sheet (every interior hinge, a
    triangle per slot, MEMBRANE.h:7-290 and BENDING.h): gradient = central difference of
    the energy and unprojected Hessian = central difference of the gradient, for the
    membrane alone (bendMult = 0) and for the bending part alone (difference of the two);
    bendMult = 1e6 makes bending dominate so the sum is checked as a whole as well"""
    rng = np.random.default_rng(11)
    sd0, sd1, sdB = curved_sheet(6, 0.0), curved_sheet(6, 1.0), curved_sheet(6, 1e6)
    o0, o1, oB = (O.OracleScene(s) for s in (sd0, sd1, sdB))
    assert o1.arrays()["hinge"].shape[0] == 3 * 5 * 5 - 2 * 5   # interior edges of the 6 x 6 grid
    x = (sd0.X + rng.uniform(-0.01, 0.01, sd0.X.shape)).ravel()
    # membrane
    memb = lambda y: energies(o0, y)[0]
    g0 = elastic_grad(o0, x)
    assert relerr(g0, fd_grad(memb, x, 1e-7)) < 1e-6
    H0 = elastic_hess(o0, sd0, x, False)
    assert relerr(H0, free_jac(sd0, fd_jac(lambda y: elastic_grad(o0, y), x, 1e-7))) < 1e-6
    # bending alone
    bend = lambda y: energies(o1, y)[1]
    gb = elastic_grad(o1, x) - g0
    assert relerr(gb, fd_grad(bend, x, 1e-6)) < 1e-6
    Hb = elastic_hess(o1, sd1, x, False) - H0
    assert relerr(Hb, free_jac(sd0, fd_jac(lambda y: elastic_grad(o1, y) - elastic_grad(o0, y), x, 1e-6))) < 1e-5
    # bending-dominated whole
    tot = lambda y: energies(oB, y)[:2].sum()
    gB = elastic_grad(oB, x)
    assert relerr(gB, fd_grad(tot, x, 1e-7)) < 1e-6
    assert relerr(elastic_hess(oB, sdB, x, False), free_jac(sdB, fd_jac(lambda y: elastic_grad(oB, y), x, 1e-7))) < 1e-6


# ----------------------------------------------------------------------- barrier over a set
def test_barrier_set_by_directional_differences():
    """barrier energy, gradient and unprojected Hessian summed over the whole constraint set
    of the contact stack (IPC.h:742-1731, multiplicities and mollifiers included): along
    random directions, d/de E(x + e v) = g . v and d/de g(x + e v) = H v (central
    differences, the set held fixed as in the reference's line search)"""
    sd = scenes.scene_contact_stack()
    orc = O.OracleScene(sd, threads=4)
    x, _ = orc.get_state()
    C = orc.constraint_set(x)
    kappa = orc.scalars()["kappa"]
    pt = (C[:, 0] < 0) & (C[:, 3] >= 0)
    pe_pp_mult = (C[:, 0] < 0) & (C[:, 3] < -1)
    ee = (C[:, 0] >= 0) & (C[:, 2] >= 0) & (C[:, 3] >= 0)
    ee_moll = (C[:, 0] >= 0) & (C[:, 2] < 0)
    pe_moll = (C[:, 0] >= 0) & (C[:, 2] >= 0) & (C[:, 3] < 0)
    assert pt.sum() and pe_pp_mult.sum() and ee.sum() and ee_moll.sum() and pe_moll.sum()
    E = lambda y: orc.energy(y, None, C, kappa)[0][3]
    G = lambda y: orc.gradient(y, C, kappa)[1]
    g = G(x)
    orc.set_project_pd(False)
    H = orc.hessian(x, C, kappa, which=2)
    orc.set_project_pd(True)
    m = orc.arrays()["mass"]
    H = H - sp.diags(np.repeat(m, 3))
    # rows/columns of Dirichlet nodes are projected out by the assembly
    free = np.repeat(~np.isin(np.arange(sd.nV), sd.dbc_idx), 3)
    rng = np.random.default_rng(5)
    d2min = orc.dist2_list(x, C).min()
    for trial in range(6):
        v = rng.normal(size=x.size) * free
        v *= math.sqrt(d2min) / np.abs(v).max()      # directions on the scale of the closest gap
        eps = 1e-5
        dE = (E(x + eps * v) - E(x - eps * v)) / (2 * eps)
        assert abs(dE - g @ v) <= 1e-6 * np.abs(g).max() * np.abs(v).sum() + 1e-7 * abs(g @ v)
        dG = (G(x + eps * v) - G(x - eps * v)) / (2 * eps)
        Hv = H @ v
        assert relerr(Hv[free], dG[free]) < 1e-5
    # whole set, single direction per vertex (catches a term dropped on one constraint kind)
    for kind in (pt, pe_pp_mult, ee, ee_moll, pe_moll):
        Ck = C[kind]
        Ek = lambda y: orc.energy(y, None, Ck, kappa)[0][3]
        gk = orc.gradient(x, Ck, kappa)[1]
        v = rng.normal(size=x.size) * free
        v *= math.sqrt(d2min) / np.abs(v).max()
        eps = 1e-5
        dE = (Ek(x + eps * v) - Ek(x - eps * v)) / (2 * eps)
        assert abs(dE - gk @ v) <= 1e-6 * abs(gk @ v) + 1e-9 * np.abs(gk).max()


# ----------------------------------------------------------------------- the step
def newton_decrement(orc, sd, x, xt, kappa):
    """rms Newton step / h at x for the incremental potential (elastic + inertia + barrier),
    Dirichlet rows removed, solved by scipy's sparse LU on the oracle's projected Hessian"""
    C = orc.constraint_set(x)
    ge, gb = orc.gradient(x, C, kappa)
    m = orc.arrays()["mass"]
    g = ge + gb + np.repeat(m, 3) * (x - xt)
    free = np.repeat(~np.isin(np.arange(sd.nV), sd.dbc_idx), 3)
    H = orc.hessian(x, C, kappa, which=3).tocsc()
    Hf = H[free][:, free]
    p = spla.spsolve(Hf, -g[free])
    h = sd.params["dt"]
    return math.sqrt((p * p).sum() / (sd.nV - len(sd.dbc_idx))) / h


def test_contact_step_ends_stationary_and_energy_never_rises():
    """after a converged contact step (direct solve), the Newton decrement of the
    incremental potential at the end state is below the stopping tolerance PNTol
    (IMPLICIT_EULER.h:647-710, 866) while at the start it is far above it; every accepted
    line search lowered the potential (IMPLICIT_EULER.h:9-149)"""
    sd = scenes.scene_contact_stack()
    sd.params["PNTol"] = 1e-4 * sd.bbox_diag()     # 100x tighter than the scenes' default
    orc = O.OracleScene(sd, solver=0, threads=4)
    h = sd.params["dt"]
    g = np.array(sd.params["gravity"])
    for step in range(2):
        x0, v0 = orc.get_state()
        m = orc.arrays()["mass"]
        xt = x0 + h * v0 + (h * h * np.where(m[:, None] > 0, g[None, :], 0.0)).ravel()
        k0 = orc.scalars()["kappa"]
        before = newton_decrement(orc, sd, x0, xt, k0)
        rc, st = orc.step()
        assert rc >= 3
        x1, _ = orc.get_state()
        after = newton_decrement(orc, sd, x1, xt, st["kappa"])
        tol = sd.params["PNTol"]
        assert before > 100 * tol, (before, tol)
        assert after <= tol, (after, tol)
        eb, ea = orc.energy_trace()
        assert len(eb) == st["newton_iters"]
        assert (ea <= eb).all()


# ----------------------------------------------------------------------- lumped mass
def test_shell_lumped_mass_closed_form():
    """DISCRETE_SHELL.h:279-299: each triangle adds rho t area / 3 to its three nodes; on an
    unjittered 9 x 9 grid over 1 m^2 every cloth triangle has area (1/8)^2 / 2, the cloth's
    total is rho t (1 m^2), the ground's rho t (2 m)^2; the sphere's is rho t times its
    polyhedral area"""
    sd = scenes.cloth_on_sphere(sheets=1, n=9, jitter=0.0, sphere_level=2, name="mass")
    orc = O.OracleScene(sd)
    m = orc.arrays()["mass"]
    rho, t = sd.params["rho"], sd.params["thick"]
    nc = sd.n_cloth
    inc = np.bincount(sd.tris[(sd.tris < nc).all(1)].ravel(), minlength=nc)   # cloth triangles per node
    assert np.allclose(m[:nc], rho * t * inc * (1.0 / 8) ** 2 / 2 / 3, rtol=1e-13, atol=0)
    assert abs(m[:nc].sum() - rho * t * 1.0) < 1e-14
    ng = sd.nV - 4
    assert abs(m[ng:].sum() - rho * t * 4.0) < 1e-13
    Xs = sd.X[nc:ng]
    Ts = sd.tris[(sd.tris >= nc).all(1) & (sd.tris < ng).all(1)] - nc
    area = 0.5 * np.linalg.norm(np.cross(Xs[Ts[:, 1]] - Xs[Ts[:, 0]], Xs[Ts[:, 2]] - Xs[Ts[:, 0]]), axis=1).sum()
    assert abs(m[nc:ng].sum() - rho * t * area) < 1e-13 * rho * t * area
    # the polyhedral sphere area approaches 4 pi r^2 from below
    assert 0.95 * 4 * math.pi * 0.2 ** 2 < area < 4 * math.pi * 0.2 ** 2
