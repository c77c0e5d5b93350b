"""Pins for the oracle's lagged friction (FEM/FRICTION.h, FEM/FRICTION_UTILS.h), CPU only.

* stencils: the constraint set without its mollified entries, in the same order;
  closest points satisfy the orthogonality conditions of a closest pair; tangent bases
  are orthonormal and orthogonal to the contact normal; the normal force equals the norm
  of the barrier gradient on the stencil's first primitive (point, or edge a).
* smoothing: f0 hand values on both sides of eps_v h.
* derivatives: gradient = central difference of the energy, Hessian = central difference
  of the gradient, in the static and the sliding regimes.
* dynamics: a sheet sliding on the ground loses horizontal momentum mu / h * sum(lambda)
  per step in the sliding direction (Newton stationarity with the closed-form force).
"""
import numpy as np
import pytest

import oracle as O
from cipc import scenes


def fric_scene(mu=0.4, n=9, gap=5e-4, v0=(1.0, 0.0, 0.0), fric_iter=1):
    """An n x n sheet (0.3 m) resting `gap` above a 2-triangle ground, far from the
    ground's diagonal and boundary, moving with velocity v0; ground Dirichlet."""
    X, T = scenes.cloth_grid(n, 0.0)
    X = 0.3 * X + np.array([0.45, gap, -0.45])
    Xg = np.array([[-1.0, 0.0, -1.0], [1.0, 0.0, -1.0], [1.0, 0.0, 1.0], [-1.0, 0.0, 1.0]])
    Tg = np.array([[0, 2, 1], [0, 3, 2]], dtype=np.int32) + len(X)
    Xa = np.concatenate([X, Xg])
    Ta = np.concatenate([T, Tg]).astype(np.int32)
    dbc = np.arange(len(X), len(Xa), dtype=np.int32)
    params = dict(scenes.DEFAULT_PARAMS, mu=mu, epsv2=1e-6, fricIterAmt=fric_iter)
    sd = scenes.SceneData(Xa, Ta, dbc, Xa[dbc].copy(), params, "sheet_on_ground")
    sd.params["PNTol"] = 1e-6 * sd.bbox_diag()
    sd.n_cloth = len(X)
    sd.v0 = np.zeros_like(Xa)
    sd.v0[: len(X)] = v0
    return sd


@pytest.fixture(scope="module")
def stack():
    sd = scenes.scene_contact_stack()
    sd.params.update(mu=0.2, epsv2=1e-6)
    orc = O.OracleScene(sd, threads=4)
    x, _ = orc.get_state()
    C = orc.constraint_set(x)
    kappa = orc.scalars()["kappa"]
    FC, FI = orc.friction_basis(x, C, kappa)
    return sd, orc, x, C, kappa, FC, FI


def _moll(c):
    return c[0] >= 0 and (c[2] < 0 or c[3] < 0)


def test_friction_stencils_and_bases(stack):
    sd, orc, x, C, kappa, FC, FI = stack
    keep = np.array([not _moll(c) for c in C])
    assert keep.sum() < len(C)                       # the stack has mollified entries
    assert np.array_equal(FC, C[keep])
    X = x.reshape(-1, 3)
    kinds = set()
    for c, f in zip(FC, FI):
        b0, b1 = f[2:5], f[5:8]
        assert abs(b0 @ b0 - 1) < 1e-12 and abs(b1 @ b1 - 1) < 1e-12 and abs(b0 @ b1) < 1e-12
        if c[0] >= 0:
            a0, a1, q0, q1 = X[c[0]], X[c[1]], X[c[2]], X[c[3]]
            pa = a0 + f[0] * (a1 - a0)
            pb = q0 + f[1] * (q1 - q0)
            d = pa - pb
            sc = np.linalg.norm(a1 - a0) * max(np.linalg.norm(d), 1e-12)
            assert abs(d @ (a1 - a0)) <= 1e-9 * sc and abs(d @ (q1 - q0)) <= 1e-9 * sc
            nrm = np.cross(a1 - a0, q1 - q0)
            kinds.add("ee")
        else:
            p = X[-c[0] - 1]
            if c[2] < 0:
                d = X[c[1]] - p
                nrm = d
                kinds.add("pp")
            elif c[3] < 0:
                e0, e1 = X[c[1]], X[c[2]]
                q = e0 + f[0] * (e1 - e0)
                assert abs((p - q) @ (e1 - e0)) <= 1e-9 * np.linalg.norm(e1 - e0) * np.linalg.norm(p - q)
                nrm = p - q
                kinds.add("pe")
            else:
                t0, t1, t2 = X[c[1]], X[c[2]], X[c[3]]
                q = t0 + f[0] * (t1 - t0) + f[1] * (t2 - t0)
                r = p - q
                for e in (t1 - t0, t2 - t0):
                    assert abs(r @ e) <= 1e-9 * np.linalg.norm(e) * np.linalg.norm(r)
                nrm = np.cross(t1 - t0, t2 - t0)
                kinds.add("pt")
        nrm = nrm / np.linalg.norm(nrm)
        assert abs(b0 @ nrm) < 1e-9 and abs(b1 @ nrm) < 1e-9
    assert {"pt", "ee"} <= kinds


def test_normal_force_is_barrier_force(stack):
    """lambda = -b'(d^2) 2 d = |d b / d x| on the stencil's point (or on edge a as a whole)."""
    sd, orc, x, C, kappa, FC, FI = stack
    for c, f in list(zip(FC, FI))[:: max(1, len(FC) // 60)]:
        _, gb = orc.gradient(x, c[None, :], kappa)
        G = gb.reshape(-1, 3)
        if c[0] >= 0:
            force = G[c[0]] + G[c[1]]
            mult = 1
        else:
            force = G[-c[0] - 1]
            mult = -c[3] if c[3] < 0 else 1
        assert abs(np.linalg.norm(force) / mult - f[8]) <= 1e-9 * f[8]
        assert f[8] > 0


def f0_ref(x, ev):
    return x if x >= ev else x * x * (-x / 3 + ev) / ev ** 2 + ev / 3


def test_smoothing_values():
    """f0 is C1 at eps: value and slope match |u| there; f0(0) = eps / 3 (FRICTION_UTILS.h:9-19),
    checked through the energy of a single point-triangle stencil."""
    sd = fric_scene(mu=1.0)
    orc = O.OracleScene(sd)
    x, _ = orc.get_state()
    C = orc.constraint_set(x)
    FC, FI = orc.friction_basis(x, C, orc.scalars()["kappa"])
    k = int(np.where(FC[:, 3] >= 0)[0][0])         # a point-triangle stencil
    FC1, FI1 = FC[k:k + 1], FI[k:k + 1]
    lam = FI1[0, 8]
    p = -FC1[0, 0] - 1
    ev = np.sqrt(1e-6) * sd.params["dt"]
    for s in (0.0, 0.3, 0.999, 1.0, 1.5, 40.0):
        xx = x.copy()
        xx[3 * p] += s * ev                          # tangential (x) displacement of the point
        e = orc.friction_energy(xx, x, FC1, FI1)
        assert abs(e - lam * f0_ref(s * ev, ev)) <= 1e-9 * lam * max(ev, s * ev)   # x + d - x rounding
    e0 = orc.friction_energy(x, x, FC1, FI1)
    assert abs(e0 - lam * ev / 3) <= 1e-12 * lam * ev


@pytest.mark.parametrize("scale", [0.2, 3.0, 50.0])
def test_friction_derivatives_fd(stack, scale):
    """scale = |relative displacement| / eps_v h: static, mixed and sliding regimes."""
    sd, orc, x, C, kappa, FC, FI = stack
    ev = np.sqrt(1e-6) * sd.params["dt"]
    rng = np.random.default_rng(int(scale * 10))
    xn = x.copy()
    xx = xn + rng.standard_normal(x.size) * ev * scale
    g = orc.friction_gradient(xx, xn, FC, FI)
    H = orc.friction_hessian(xx, xn, FC, FI)
    for _ in range(3):
        v = rng.standard_normal(x.size)
        t = 1e-4 * ev
        ep = orc.friction_energy(xx + t * v, xn, FC, FI)
        em = orc.friction_energy(xx - t * v, xn, FC, FI)
        fd = (ep - em) / (2 * t)
        assert abs(fd - g @ v) <= 1e-6 * (np.abs(g).sum() * np.abs(v).max())
        gp = orc.friction_gradient(xx + t * v, xn, FC, FI)
        gm = orc.friction_gradient(xx - t * v, xn, FC, FI)
        hv = H @ v
        assert np.abs((gp - gm) / (2 * t) - hv).max() <= 1e-5 * np.abs(hv).max()
    # symmetric and positive semi-definite
    Hd = H.toarray()
    assert np.abs(Hd - Hd.T).max() <= 1e-12 * np.abs(Hd).max()
    idx = np.unique(np.concatenate([[3 * v + k for k in range(3)] for v in np.unique(np.abs(FC.clip(min=0)))]))
    w = np.linalg.eigvalsh(Hd[np.ix_(idx, idx)])
    assert w.min() >= -1e-10 * w.max()


def test_sliding_sheet_momentum():
    """One step of a sheet sliding at 1 m/s on the ground (mu = 0.4, every contact sliding):
    sum_a m_a (v1 - v0)_a = -(mu / h) sum_k lambda_k B_k u_k / |u_k|, and the elastic and
    barrier forces carry no horizontal resultant."""
    mu = 0.4
    # gap 0.95 d_hat: the lagged normal force is small enough that the sheet keeps sliding
    sd = fric_scene(mu=mu, gap=9.5e-4, v0=(3.0, 0.0, 0.0))
    orc = O.OracleScene(sd)
    orc.set_state(None, sd.v0.ravel())
    x0, v0 = orc.get_state()
    C = orc.constraint_set(x0)
    assert len(C) > 0 and all(c[0] < 0 and c[3] >= 0 for c in C)   # point-triangle only
    FC, FI = orc.friction_basis(x0, C, orc.scalars()["kappa"])
    rc, _ = orc.step()
    assert rc > 0
    x1, v1 = orc.get_state()
    m = orc.arrays()["mass"]
    nc = sd.n_cloth
    dP = (m[:nc, None] * (v1 - v0).reshape(-1, 3)[:nc]).sum(axis=0)
    h = sd.params["dt"]
    ev = np.sqrt(1e-6) * h
    F = np.zeros(3)
    for c, f in zip(FC, FI):
        p = -c[0] - 1
        B = np.stack([f[2:5], f[5:8]], axis=1)
        u = B.T @ (x1.reshape(-1, 3)[p] - x0.reshape(-1, 3)[p])
        assert np.linalg.norm(u) >= ev                # sliding
        F += mu * f[8] * (B @ (u / np.linalg.norm(u)))
    pred = -F / h
    assert pred[0] < 0
    assert abs(dP[0] - pred[0]) <= 1e-4 * abs(pred[0])
    assert abs(dP[2] - pred[2]) <= 1e-4 * abs(pred[0])
    # friction slowed the sheet: the same step without friction keeps v_x
    sd0 = fric_scene(mu=0.0, gap=9.5e-4, v0=(3.0, 0.0, 0.0))
    o0 = O.OracleScene(sd0)
    o0.set_state(None, sd0.v0.ravel())
    o0.step()
    _, v1f = o0.get_state()
    assert v1.reshape(-1, 3)[:nc, 0].mean() < v1f.reshape(-1, 3)[:nc, 0].mean() - 0.5 * mu * 9.81 * h * 0.1


def test_friction_relag_iterations():
    """fricIterAmt = 0 re-lags the friction at every converged iterate until the updated
    Newton step is below tolerance (IMPLICIT_EULER.h:770-859): the end state is then
    stationary for the potential whose friction basis and normal forces are taken at that
    state, which the single lag (fricIterAmt = 1) is not."""
    res = {}
    for it in (1, 0):
        sd = fric_scene(mu=0.4, fric_iter=it)
        orc = O.OracleScene(sd)
        orc.set_state(None, sd.v0.ravel())
        x0, v0 = orc.get_state()
        rc, _ = orc.step()
        assert rc > 0
        x, _ = orc.get_state()
        C = orc.constraint_set(x)
        assert orc.dist2_list(x, C).min() > 0
        kappa = orc.scalars()["kappa"]
        FC, FI = orc.friction_basis(x, C, kappa)
        gel, gb = orc.gradient(x, C, kappa)
        m = orc.arrays()["mass"]
        h = sd.params["dt"]
        xt = x0.reshape(-1, 3) + h * v0.reshape(-1, 3) + h * h * np.array(sd.params["gravity"])
        gi = (m[:, None] * (x.reshape(-1, 3) - xt)).ravel()
        gf = orc.friction_gradient(x, x0, FC, FI)
        tot = (gel + gb + gi + gf).reshape(-1, 3)[: sd.n_cloth]
        res[it] = (np.abs(tot).max(), np.abs(gb).max())
    assert res[0][0] <= 1e-6 * res[0][1]
    assert res[1][0] >= 1e3 * res[0][0]


def test_friction_components():
    """Compute_Friction_Coef (FRICTION.h:114-171): each stencil's normal force is scaled by
    the coefficient of its two primitives' components and mu is taken as 1; a uniform
    coefficient matrix reproduces the single-mu potential."""
    sd = scenes.scene_contact_stack()
    nv, nS = sd.nV, 144
    ranges = [nS, 2 * nS, 3 * nS, nv]
    mu = np.array([[0.1, 0.2, 0.3, 0.4], [0.2, 0.5, 0.6, 0.7], [0.3, 0.6, 0.8, 0.9], [0.4, 0.7, 0.9, 1.0]])
    base = O.OracleScene(sd)
    x, _ = base.get_state()
    C = base.constraint_set(x)
    kappa = base.scalars()["kappa"]
    F0, I0 = base.friction_basis(x, C, kappa)
    comp = lambda v: int(np.searchsorted(ranges, v, side="right"))
    for m in (mu, np.full((4, 4), 0.3)):
        sd2 = scenes.scene_contact_stack()
        sd2.params.update(compNodeRange=ranges, muComp=m.T.ravel())   # entry c0 + c1 n
        orc = O.OracleScene(sd2)
        F1, I1 = orc.friction_basis(x, C, kappa)
        assert np.array_equal(F0, F1)
        for c, a, b in zip(F0, I0, I1):
            pa, pb = (c[0], c[2]) if c[0] >= 0 else (-c[0] - 1, c[1])
            assert abs(b[8] - m[comp(pa), comp(pb)] * a[8]) <= 1e-15 * a[8]
    # uniform 0.3 matrix == mu 0.3
    sd3 = scenes.scene_contact_stack()
    sd3.params.update(mu=0.3)
    o3 = O.OracleScene(sd3)
    xx = x + np.random.default_rng(1).standard_normal(x.size) * 1e-6
    F3, I3 = o3.friction_basis(x, C, kappa)
    assert abs(orc.friction_energy(xx, x, F1, I1) - o3.friction_energy(xx, x, F3, I3)) <= 1e-14 * o3.friction_energy(xx, x, F3, I3)
