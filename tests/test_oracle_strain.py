"""Pins for the oracle's strain limiting (FEM/Shell/INEXT.h), CPU only.

* closed forms: a triangle stretched uniformly by l (both singular values l) or along
  one edge direction (singular values l and 1) has energy h^2 vol kappa (2 or 1) f(l)
  with f(s) = -(sHat - s)^2 ln((slim - s)/(slim - sHat)) / (slim - sHat)^2;
  invalid at and above the limit; no energy below sHat.
* derivatives: gradient = central difference of the energy; the unprojected Hessian =
  central difference of the gradient (one and two active singular values, lower limit);
  the projected Hessian is symmetric PSD and equals the unprojected one when that is PSD.
* step: a sheet stretched by gravity between pinned corners stays below the limit, and
  without the limit it stretches beyond it.
"""
import numpy as np
import pytest

import oracle as O
from cipc import scenes


def tri_scene(kappa=(1e3, 0.0), s=(1.01, 0.0), sHat=(1.0, 1.0)):
    X = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.03, 0.0, 0.08]])
    T = np.array([[0, 1, 2]], dtype=np.int32)     # local vertex order = global order
    params = dict(scenes.DEFAULT_PARAMS, PNTol=1e-3, collision=0, kappa_s=kappa, s=s, sHat=sHat)
    sd = scenes.SceneData(X, T, np.zeros(0, np.int32), np.zeros((0, 3)), params, "one_triangle")
    return sd, O.OracleScene(sd)


def f_sl(s, slim, shat):
    return -(shat - s) ** 2 * np.log((slim - s) / (slim - shat)) / (slim - shat) ** 2


def test_closed_forms():
    sd, orc = tri_scene()
    X = sd.X
    h2vol = sd.params["dt"] ** 2 * orc.arrays()["vol"][0]
    for l in (0.99, 1.0, 1.003, 1.009, 1.0099):
        ok, E, _, _, sg = orc.strain_tri((X * l).ravel(), 0)
        assert ok and np.allclose(sg, [l, l], rtol=1e-12)
        ref = 2 * 1e3 * h2vol * f_sl(l, 1.01, 1.0) if l > 1.0 else 0.0
        assert abs(E - ref) <= 1e-9 * max(abs(ref), 1e-30)
    # stretch along x only: F has singular values l (x) and 1
    for l in (1.002, 1.008):
        Xs = X * np.array([l, 1.0, 1.0])
        ok, E, _, _, sg = orc.strain_tri(Xs.ravel(), 0)
        assert ok and np.allclose(sorted(sg), [1.0, l], rtol=1e-12)
        assert abs(E - 1e3 * h2vol * f_sl(l, 1.01, 1.0)) <= 1e-9 * E
    for l in (1.01, 1.02):
        ok, *_ = orc.strain_tri((X * l).ravel(), 0)
        assert not ok
    ok, E, g = orc.strain_energy_grad((X * 1.02).ravel())
    assert not ok


@pytest.mark.parametrize("case", ["uniform", "uniaxial", "sheared", "lower"])
def test_derivatives_fd(case):
    kw = dict(kappa=(1e3, 1e3), s=(1.01, 0.9), sHat=(1.0, 0.95)) if case == "lower" else {}
    sd, orc = tri_scene(**kw)
    rng = np.random.default_rng(5)
    X = sd.X.copy()
    if case == "uniform":
        X = X * 1.005
    elif case == "uniaxial":
        X = X * np.array([1.006, 1.0, 1.0])
    elif case == "sheared":
        X = X @ np.array([[1.004, 0.0, 0.003], [0.0, 1.0, 0.0], [0.002, 0.0, 1.002]])
    else:
        X = X * np.array([0.93, 1.0, 1.003])
    X = X + rng.standard_normal(X.shape) * 1e-5
    x = X.ravel()
    ok, E, g, H, sg = orc.strain_tri(x, 0, proj=False)
    assert ok and E > 0
    for _ in range(4):
        v = rng.standard_normal(9)
        t = 1e-7
        Ep = orc.strain_tri(x + t * v, 0)[1]
        Em = orc.strain_tri(x - t * v, 0)[1]
        assert abs((Ep - Em) / (2 * t) - g @ v) <= 1e-6 * np.abs(g).sum() * np.abs(v).max()
        gp = orc.strain_tri(x + t * v, 0, proj=False)[2]
        gm = orc.strain_tri(x - t * v, 0, proj=False)[2]
        hv = H @ v
        assert np.abs((gp - gm) / (2 * t) - hv).max() <= 1e-5 * np.abs(H).max() * np.abs(v).max()
    Hp = orc.strain_tri(x, 0, proj=True)[3]
    assert np.abs(Hp - Hp.T).max() <= 1e-12 * np.abs(Hp).max()
    w = np.linalg.eigvalsh(Hp)
    assert w.min() >= -1e-10 * w.max()
    wu = np.linalg.eigvalsh(0.5 * (H + H.T))
    if wu.min() >= 0:
        assert np.abs(Hp - H).max() <= 1e-9 * np.abs(H).max()
    # the projection keeps the PSD part: Hp - H is PSD
    assert np.linalg.eigvalsh(0.5 * (Hp - H + (Hp - H).T)).min() >= -1e-9 * np.abs(H).max()


def hammock(kappa):
    """A 9x9 sheet (0.3 m) pinned at its four corners, no contact: gravity stretches it."""
    X, T = scenes.cloth_grid(9, 0.0)
    X = 0.3 * X
    corners = np.array([0, 8, 72, 80], dtype=np.int32)
    params = dict(scenes.DEFAULT_PARAMS, collision=0, E=2e3, gravity=(0.0, -30.0, 0.0),
                  kappa_s=(kappa, 0.0), s=(1.01, 0.0), sHat=(1.0, 1.0))
    sd = scenes.SceneData(X, T, corners, X[corners].copy(), params, "hammock")
    sd.params["PNTol"] = 1e-6 * sd.bbox_diag()
    return sd


def max_sigma(orc, x, nT):
    return max(orc.strain_tri(x, t)[4].max() for t in range(nT))


def test_step_keeps_strain_below_limit():
    out = {}
    for kappa in (0.0, 1e3):
        sd = hammock(kappa)
        orc = O.OracleScene(sd)
        for _ in range(8):
            rc, _ = orc.step()
            assert rc > 0
        x, _ = orc.get_state()
        sdl = hammock(1e3)
        probe = O.OracleScene(sdl)
        out[kappa] = max_sigma(probe, x, sd.nT)
    assert out[0.0] > 1.01          # the soft sheet stretches past the limit without it
    assert 1.0 < out[1e3] < 1.01    # and stays inside it with it


# ---------------------------------------------------------------- anisotropic (fibre) model
def fiber_tri_scene(k=(1e3, 5e2, 1e3, 2e2), lim=(0.05, 0.05, 0.05)):
    sd, _ = tri_scene()
    sd.params.update(fiberStiffMult=k, inextLimit=lim)
    return sd, O.OracleScene(sd)


def eta(s, m):
    return -m * np.log((m - s) / m)


def test_fiber_closed_forms():
    k, lim = (1e3, 5e2, 1e3, 2e2), (0.05, 0.04, 0.03)
    sd, orc = fiber_tri_scene(k, lim)
    w = sd.params["dt"] ** 2 * orc.arrays()["vol"][0]
    X = sd.X
    for l in (1.01, 1.03):                       # stretch along x: E00 = (l^2 - 1)/2
        ok, E, _, _, e3 = orc.fiber_tri((X * np.array([l, 1.0, 1.0])).ravel(), 0)
        E00 = 0.5 * (l * l - 1)
        assert ok and np.allclose(e3, [E00, 0, 0], atol=1e-15)
        assert abs(E - w * k[0] / 2 * eta(E00 ** 2, lim[0] ** 2)) <= 1e-12 * E
    for gmm in (0.02, 0.05):                     # shear x += g z: E11 = g^2/2, E01 = g/2
        Xs = X.copy()
        Xs[:, 0] += gmm * X[:, 2]
        ok, E, _, _, e3 = orc.fiber_tri(Xs.ravel(), 0)
        assert ok and np.allclose(e3, [0, gmm * gmm / 2, gmm / 2], atol=1e-15)
        ref = w * (k[2] / 2 * eta((gmm * gmm / 2) ** 2, lim[1] ** 2) + k[3] * eta((gmm / 2) ** 2, lim[2] ** 2))
        assert abs(E - ref) <= 1e-12 * ref
    assert not orc.fiber_tri((X * np.array([1.06, 1.0, 1.0])).ravel(), 0)[0]   # E00 = 0.0618 >= 0.05
    ok, E, *_ = orc.fiber_tri(X.ravel(), 0)
    assert ok and E == 0.0


@pytest.mark.parametrize("case", ["stretch", "shear", "mixed"])
def test_fiber_derivatives_fd(case):
    sd, orc = fiber_tri_scene()
    rng = np.random.default_rng(7)
    X = sd.X.copy()
    if case == "stretch":
        X = X * np.array([1.02, 1.0, 0.99])
    elif case == "shear":
        X[:, 0] += 0.03 * X[:, 2]
    else:
        X = X @ np.array([[1.02, 0.01, 0.015], [0.0, 1.0, 0.0], [-0.01, 0.02, 0.985]])
    X = X + rng.standard_normal(X.shape) * 1e-5
    x = X.ravel()
    ok, E, g, H, _ = orc.fiber_tri(x, 0, proj=False)
    assert ok and E > 0
    for _ in range(4):
        v = rng.standard_normal(9)
        t = 1e-7
        fd = (orc.fiber_tri(x + t * v, 0)[1] - orc.fiber_tri(x - t * v, 0)[1]) / (2 * t)
        assert abs(fd - g @ v) <= 1e-6 * np.abs(g).sum() * np.abs(v).max()
        gd = (orc.fiber_tri(x + t * v, 0, proj=False)[2] - orc.fiber_tri(x - t * v, 0, proj=False)[2]) / (2 * t)
        assert np.abs(gd - H @ v).max() <= 1e-5 * np.abs(H).max() * np.abs(v).max()
    Hp = orc.fiber_tri(x, 0, proj=True)[3]
    assert np.linalg.eigvalsh(0.5 * (Hp + Hp.T)).min() >= -1e-10 * np.abs(Hp).max()


def test_fiber_step_stays_feasible():
    """the pinned sheet under gravity with the fibre model (limits 0.03): every step ends
    with each Green strain component inside its limit, and the limit is reached closely"""
    sd = hammock(0.0)
    sd.params.update(fiberStiffMult=(5e3, 1e3, 5e3, 5e3), inextLimit=(0.03, 0.03, 0.03))
    orc = O.OracleScene(sd)
    for _ in range(6):
        rc, _ = orc.step()
        assert rc > 0
    x, _ = orc.get_state()
    e = np.array([orc.fiber_tri(x, t)[4] for t in range(sd.nT)])
    assert 0.02 < np.abs(e).max() < 0.03


def test_kappa_s_doubles_when_a_triangle_stays_at_the_limit():
    """IMPLICIT_EULER.h:606-645: with a weak barrier (kappa_s = 1) gravity drives the pinned
    sheet's corner triangles to within 1e-4 (s - sHat) of the limit in successive Newton
    iterations of the fourth step, and kappa_s doubles there (three times, to 8) and keeps
    the new value in later steps; with kappa_s = 1e3 the limit is never approached that
    closely and it never changes; every state stays inside the limit"""
    for k0, expect in ((1.0, [1.0, 1.0, 1.0, 8.0, 8.0]), (1e3, [1e3] * 5)):
        sd = hammock(k0)
        orc = O.OracleScene(sd)
        got = []
        for _ in range(5):
            rc, st = orc.step()
            assert rc > 0
            got.append(st["kappa_s"])
            assert st["kappa_s"] == k0 * 2.0 ** round(np.log2(st["kappa_s"] / k0))
            x, _ = orc.get_state()
            assert max_sigma(orc, x, sd.nT) < 1.01
        assert got == expect, got
        assert orc.kappa_s() == expect[-1]
