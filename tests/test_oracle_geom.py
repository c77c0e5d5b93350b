"""Pins for the oracle's per-stencil geometry (CPU only).

The reference ships no tests or golden values for these functions (its Check_* helpers
print finite-difference errors without expected values), so each function is pinned by
what the mathematics fixes: closed forms and brute-force minimisation for distances and
region types, float64 autograd of the plain energy formula plus central finite
differences for every gradient and Hessian, numpy.linalg.eigh fixtures for the
eigensolver behind the PSD projection.
"""
import os

import numpy as np
import pytest
import torch

import oracle as O

torch.set_default_dtype(torch.float64)
RNG = np.random.default_rng(1234)
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def jac_hess(f, x):
    xt = torch.tensor(x, requires_grad=True)
    g = torch.autograd.functional.jacobian(f, xt).numpy()
    H = torch.autograd.functional.hessian(f, xt).numpy()
    return g, H


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ----------------------------------------------------------------------- formulas (torch)
def t_pp(x):
    return ((x[0:3] - x[3:6]) ** 2).sum()


def t_pe(x):
    p, e0, e1 = x[0:3], x[3:6], x[6:9]
    c = torch.linalg.cross(e0 - p, e1 - p)
    return (c @ c) / ((e1 - e0) @ (e1 - e0))


def t_pt(x):
    p, t0, t1, t2 = x[0:3], x[3:6], x[6:9], x[9:12]
    b = torch.linalg.cross(t1 - t0, t2 - t0)
    a = (p - t0) @ b
    return a * a / (b @ b)


def t_ee(x):
    a0, a1, b0, b1 = x[0:3], x[3:6], x[6:9], x[9:12]
    b = torch.linalg.cross(a1 - a0, b1 - b0)
    a = (b0 - a0) @ b
    return a * a / (b @ b)


@pytest.mark.parametrize("kind,f,n", [("pp", t_pp, 6), ("pe", t_pe, 9), ("pt", t_pt, 12), ("ee", t_ee, 12)])
def test_distance_derivatives_vs_autograd(kind, f, n):
    for _ in range(20):
        x = RNG.standard_normal(n)
        d = O.dist2(kind, x.reshape(-1, 3))
        assert abs(d - f(torch.tensor(x)).item()) <= 1e-13 * max(abs(d), 1e-300)
        g, H = O.dist2_gh(kind, x.reshape(-1, 3))
        gt, Ht = jac_hess(f, x)
        assert relerr(g, gt) < 1e-12
        assert relerr(H, Ht) < 1e-11
        assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()


def test_distance_gradient_central_difference():
    """Check_*-style finite differences (MEMBRANE.h:293-386 methodology, central, eps 1e-6)."""
    for kind, n in (("pe", 9), ("pt", 12), ("ee", 12)):
        x = RNG.standard_normal(n)
        g, H = O.dist2_gh(kind, x.reshape(-1, 3))
        eps = 1e-6
        gfd = np.zeros(n)
        Hfd = np.zeros((n, n))
        for i in range(n):
            xp, xm = x.copy(), x.copy()
            xp[i] += eps
            xm[i] -= eps
            gfd[i] = (O.dist2(kind, xp.reshape(-1, 3)) - O.dist2(kind, xm.reshape(-1, 3))) / (2 * eps)
            Hfd[:, i] = (O.dist2_gh(kind, xp.reshape(-1, 3))[0] - O.dist2_gh(kind, xm.reshape(-1, 3))[0]) / (2 * eps)
        assert relerr(gfd, g) < 1e-7
        assert relerr(Hfd, H) < 1e-6


# ----------------------------------------------------------------------- region types vs brute force
def closest_on_triangle(p, a, b, c, m=301):
    # brute force over a barycentric lattice, then refine on the best edge/interior
    best = (np.inf, None)
    u = np.linspace(0, 1, m)
    U, V = np.meshgrid(u, u)
    mask = U + V <= 1
    P = a[None, :] + U[mask][:, None] * (b - a) + V[mask][:, None] * (c - a)
    d = ((P - p) ** 2).sum(1)
    i = int(np.argmin(d))
    return d[i], U[mask][i], V[mask][i]


def pt_region_exact(p, a, b, c):
    """Closed-form closest-point region of a triangle (Ericson, Real-Time Collision Detection 5.1.5)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return 0
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return 1
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return 3
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return 2
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return 5
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return 4
    return 6


def test_pt_type_and_distance_vs_exact():
    agree = 0
    for _ in range(2000):
        pts = RNG.standard_normal((4, 3))
        t = O.dtype("pt", pts)
        r = pt_region_exact(*pts)
        agree += (t == r)
        d = O.dist2_unclassified("pt", pts)
        # exact closest distance from the region's closed form
        p, a, b, c = pts
        if r == 6:
            n = np.cross(b - a, c - a)
            de = ((p - a) @ n) ** 2 / (n @ n)
        elif r in (0, 1, 2):
            de = ((p - (a, b, c)[r]) ** 2).sum()
        else:
            e0, e1 = {3: (a, b), 4: (b, c), 5: (c, a)}[r]
            de = (np.cross(e0 - p, e1 - p) ** 2).sum() / ((e1 - e0) ** 2).sum()
        assert abs(d - de) <= 1e-12 * max(de, 1e-300) or t != r
    assert agree == 2000


def test_pt_distance_brute_force_small():
    for _ in range(30):
        pts = RNG.standard_normal((4, 3))
        d = O.dist2_unclassified("pt", pts)
        db, _, _ = closest_on_triangle(*pts)
        assert d <= db + 1e-12 and db - d <= 2e-3 * max(db, 1e-3)


def seg_seg_dist2_brute(a0, a1, b0, b1, m=2001):
    s = np.linspace(0, 1, m)
    P = a0[None] + s[:, None] * (a1 - a0)
    best = np.inf
    for t in np.linspace(0, 1, m):
        Q = b0 + t * (b1 - b0)
        best = min(best, ((P - Q) ** 2).sum(1).min())
    return best


def test_ee_unclassified_distance_vs_brute_force():
    for _ in range(15):
        pts = RNG.standard_normal((4, 3))
        d = O.dist2_unclassified("ee", pts)
        db = seg_seg_dist2_brute(*pts, m=801)
        assert d <= db + 1e-12 and db - d <= 2e-5   # sampling grid 1/800 per edge


def test_ee_parallel_fallback_types():
    # exactly parallel edges: the near-parallel branch (DISTANCE_TYPE.h:120) picks a PE case
    a0, a1 = np.array([0.0, 0, 0]), np.array([1.0, 0, 0])
    b0, b1 = np.array([0.25, 1e-3, 0]), np.array([1.5, 1e-3, 0])
    t = O.dtype("ee", np.array([a0, a1, b0, b1]))
    assert t in (2, 5, 6, 7)
    d = O.dist2_unclassified("ee", np.array([a0, a1, b0, b1]))
    assert abs(d - 1e-6) <= 1e-15


def test_pe_type_and_unclassified():
    for _ in range(200):
        p, e0, e1 = RNG.standard_normal((3, 3))
        r = (e1 - e0) @ (p - e0) / ((e1 - e0) @ (e1 - e0))
        t = O.dtype("pe", np.array([p, e0, e1]))
        assert t == (0 if r < 0 else (1 if r > 1 else 2))
        d = O.dist2_unclassified("pe", np.array([p, e0, e1]))
        q = e0 + min(max(r, 0), 1) * (e1 - e0)
        assert abs(d - ((p - q) ** 2).sum()) <= 1e-12 * max(d, 1e-300)


# ----------------------------------------------------------------------- mollifier & barrier
def test_mollifier_vs_autograd():
    for _ in range(10):
        rest = RNG.standard_normal((4, 3))
        # nearly parallel current edges so that c < eps
        x = rest.copy()
        x[3] = x[2] + (x[1] - x[0]) * 0.7 + 1e-3 * RNG.standard_normal(3)
        e, g, H, eps = O.mollifier(x, rest)

        def f(z):
            c = torch.linalg.cross(z[3:6] - z[0:3], z[9:12] - z[6:9])
            cn = c @ c
            return (-cn / eps + 2.0) * cn / eps

        xt = torch.tensor(x.ravel())
        assert abs(e - f(xt).item()) <= 1e-12
        gt, Ht = jac_hess(f, x.ravel())
        assert relerr(g, gt) < 1e-10 and relerr(H, Ht) < 1e-10
    # far from parallel: e = 1 and zero derivatives
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 1, 1.0]])
    e, g, H, eps = O.mollifier(x, x)
    assert e == 1.0 and not g.any() and not H.any()


def test_barrier_closed_forms():
    dhat, kappa = 1e-6, 3.0
    for d in (1e-9, 1e-8, 3e-7, 9.9e-7):
        b, bg, bh = O.barrier(d, dhat, kappa)
        dt = torch.tensor(d, requires_grad=True)
        bt = -kappa * (dt - dhat) ** 2 * torch.log(dt / dhat)
        g1, = torch.autograd.grad(bt, dt, create_graph=True)
        g2, = torch.autograd.grad(g1, dt)
        assert abs(b - bt.item()) <= 1e-14 * abs(b)
        assert abs(bg - g1.item()) <= 1e-12 * abs(bg)
        assert abs(bh - g2.item()) <= 1e-12 * abs(bh)
    # b(dhat) = b'(dhat) = 0, b'' > 0 below dhat
    assert O.barrier(dhat, dhat, kappa)[0] == 0 and O.barrier(dhat, dhat, kappa)[1] == 0


def test_kappa_formula_hand_value():
    """Initialize_EIPC<double,false> (DISCRETE_SHELL.h:568-574) for one hand-computed scene."""
    import sys
    from cipc import scenes
    sd = scenes.scene_small(n=3, sheets=1, sphere_level=0)
    o = O.OracleScene(sd)
    s = o.scalars()
    mass = o.arrays()["mass"]
    d2 = sd.params["dHat"] ** 2
    d = 1e-16
    Hb = (np.log(d / d2) * -2.0 - (d - d2) * 4.0 / d) + 1.0 / (d * d) * (d - d2) ** 2
    kappa = 1e11 * mass.mean() * 3 / (4e-16 * Hb)
    assert abs(s["kappa"] - kappa) <= 1e-12 * kappa
    assert s["kappaMax"] == 100 * s["kappa"]
    assert sys is not None


# ----------------------------------------------------------------------- eigen / PSD
def test_sym_eig_golden():
    f = np.load(os.path.join(GOLD, "eig_psd.npz"))
    for A, w_ref, P_ref in zip(f["A"], f["w"], f["P"]):
        w, V = O.sym_eig(A)
        assert np.abs(w - w_ref).max() <= 1e-12 * np.abs(w_ref).max()
        assert np.abs(V @ np.diag(w) @ V.T - A).max() <= 1e-12 * np.abs(A).max()
        P = O.make_pd(A)
        assert np.abs(P - P_ref).max() <= 1e-12 * np.abs(A).max()


def test_make_pd_properties():
    for n in (6, 9, 12):
        M = RNG.standard_normal((n, n))
        S = M @ M.T + np.eye(n)
        assert np.array_equal(O.make_pd(S), S)     # positive definite: unchanged
        A = M + M.T
        P = O.make_pd(A)
        assert np.linalg.eigvalsh(P).min() >= -1e-12 * np.abs(A).max()
        w, V = np.linalg.eigh(A)
        assert np.abs(P - (V * np.maximum(w, 0)) @ V.T).max() <= 1e-12 * np.abs(A).max()


# ----------------------------------------------------------------------- elasticity
def t_memb(x, B, vol, lam, mu):
    x0, x1, x2 = x[0:3], x[3:6], x[6:9]
    D0, D1 = x1 - x0, x2 - x0
    A = torch.stack([torch.stack([D0 @ D0, D0 @ D1]), torch.stack([D0 @ D1, D1 @ D1])])
    Bm = torch.tensor([[B[0], B[1]], [B[1], B[2]]])
    lnJ = torch.log(torch.det(A) * torch.det(Bm)) / 2
    return vol * (mu / 2 * (torch.trace(Bm @ A) - 2 - 2 * lnJ) + lam / 2 * lnJ * lnJ)


def test_membrane_vs_autograd():
    for _ in range(10):
        X = RNG.standard_normal(9)
        x = X + 0.2 * RNG.standard_normal(9)
        D0, D1 = X[3:6] - X[0:3], X[6:9] - X[0:3]
        Bi = np.linalg.inv(np.array([[D0 @ D0, D0 @ D1], [D0 @ D1, D1 @ D1]]))
        B = np.array([Bi[0, 0], Bi[0, 1], Bi[1, 1]])
        E, g, H = O.membrane(x.reshape(3, 3), B, 0.37, 3.3e5, 3.8e5)
        f = lambda z: t_memb(z, B, 0.37, 3.3e5, 3.8e5)  # noqa: E731
        assert abs(E - f(torch.tensor(x)).item()) <= 1e-12 * abs(E)
        gt, Ht = jac_hess(f, x)
        assert relerr(g, gt) < 1e-11 and relerr(H, Ht) < 1e-11
    # rest state: zero energy and zero gradient
    E, g, H = O.membrane(X.reshape(3, 3), B, 0.37, 3.3e5, 3.8e5)
    assert abs(E) < 1e-9 and np.abs(g).max() < 1e-6


def t_dihedral(x):
    x0, x1, x2, x3 = x[0:3], x[3:6], x[6:9], x[9:12]
    n1 = torch.linalg.cross(x1 - x0, x2 - x0)
    n2 = torch.linalg.cross(x2 - x3, x1 - x3)
    e = x1 - x2
    return torch.atan2(torch.linalg.cross(n2, n1) @ e / torch.sqrt(e @ e), n1 @ n2)


def test_dihedral_vs_autograd_and_convention():
    for _ in range(20):
        x = RNG.standard_normal(12)
        th, g, H = O.dihedral(x.reshape(4, 3))
        assert abs(th - t_dihedral(torch.tensor(x)).item()) <= 1e-12
        gt, Ht = jac_hess(t_dihedral, x)
        assert relerr(g, gt) < 1e-11 and relerr(H, Ht) < 1e-10
    # flat, consistently oriented hinge: angle 0; folding the wing up gives a signed angle
    flat = np.array([[0.5, 0, -1.0], [0, 0, 0], [1.0, 0, 0], [0.5, 0, 1.0]])
    assert abs(O.dihedral(flat)[0]) < 1e-15
    up = flat.copy()
    up[3] = [0.5, 1.0, 1.0]
    down = flat.copy()
    down[3] = [0.5, -1.0, 1.0]
    a_up, a_dn = O.dihedral(up)[0], O.dihedral(down)[0]
    assert abs(abs(a_up) - np.pi / 4) < 1e-12 and abs(a_up + a_dn) < 1e-12
