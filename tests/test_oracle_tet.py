"""Pins for the oracle's volumetric tets inside the shell step (CPU only).

* set-up: rest volumes, lumped mass (rho vol / 4 per node) and the Lame parameters;
  the surface of a cube of Kuhn tets has 12 n^2 outward triangles and is the set of faces
  that belong to one tet only;
* Neo-Hookean energy of uniform and uniaxial stretch in closed form, h^2 vol weighted;
* gradient = central difference of the energy; the first Piola-Kirchhoff stress read off
  the gradient of a unit tet, its central difference = the unprojected dP/dF; the projected
  dP/dF = numpy's eigh projection of that (the energy is isotropic, so the block-wise
  clamp of NEOHOOKEAN.h:192-300 is the exact eigen-clamp); the element Hessian = D^T dPdF D;
* the inversion-free step size against numpy's roots of det(x + t p) = 0.2 det(x);
* steps of a cloth draped on a soft tet cube on the ground: no tet inverts, every state
  intersection free, the end state stationary for the incremental potential.
"""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
from cipc import scenes

NONE = np.zeros((0, 4), np.int32)


def one_tet(E=1e5, nu=0.4, rho=1000.0, X=None):
    X = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]) if X is None else np.asarray(X, float)
    params = dict(scenes.DEFAULT_PARAMS, collision=0, PNTol=1e-3, tetRho=rho, tetE=E, tetNu=nu)
    sd = scenes.SceneData(X, np.zeros((0, 3), np.int32), np.zeros(0, np.int32), np.zeros((0, 3)), params, "one_tet")
    sd.tets = np.array([[0, 1, 2, 3]], dtype=np.int32)
    return sd, O.OracleScene(sd)


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def test_setup_mass_volume_surface():
    X, T = scenes.tet_box((0, 0, 0), (0.2, 0.1, 0.3), 3)
    params = dict(scenes.DEFAULT_PARAMS, collision=0, PNTol=1e-3, tetRho=500.0, tetE=1e5, tetNu=0.3)
    sd = scenes.SceneData(X, np.zeros((0, 3), np.int32), np.zeros(0, np.int32), np.zeros((0, 3)), params, "box")
    sd.tets = T
    orc = O.OracleScene(sd)
    ta = orc.tet_arrays()
    assert np.allclose(ta["vol"], 0.2 * 0.1 * 0.3 / (27 * 6), rtol=1e-12)
    lam, mu = lame(1e5, 0.3)
    assert abs(ta["lam"] - lam) <= 1e-12 * lam and abs(ta["mu"] - mu) <= 1e-12 * mu
    m = orc.arrays()["mass"]
    inc = np.bincount(T.ravel(), minlength=sd.nV)
    assert np.allclose(m, 500.0 * ta["vol"][0] / 4 * inc, rtol=1e-12)
    assert abs(m.sum() - 500.0 * 0.2 * 0.1 * 0.3) <= 1e-12
    S = orc.tet_surface()
    # 2 triangles per boundary square: 2 (3x3) per face, 6 faces
    assert len(S) == 2 * 9 * 6
    # the faces that belong to exactly one tet, oriented outward
    faces = {}
    for t in T:
        for f in ((t[0], t[2], t[1]), (t[0], t[3], t[2]), (t[0], t[1], t[3]), (t[1], t[2], t[3])):
            key = tuple(sorted(f))
            faces[key] = faces.get(key, 0) + 1
    assert sorted(tuple(sorted(f)) for f in S) == sorted(k for k, c in faces.items() if c == 1)
    ctr = X.mean(0)
    a, b, c = X[S[:, 0]], X[S[:, 1]], X[S[:, 2]]
    nrm = np.cross(b - a, c - a)
    assert (np.einsum("ij,ij->i", nrm, (a + b + c) / 3 - ctr) > 0).all()
    # boundary primitives of a tet-only scene: the surface's nodes, edges and triangles
    assert orc.nBT == len(S)
    assert orc.nBN == len(np.unique(S))


@pytest.mark.parametrize("kind,s", [("uniform", 1.07), ("uniform", 0.9), ("uniaxial", 1.2), ("uniaxial", 0.8)])
def test_energy_closed_form(kind, s):
    E, nu = 1e5, 0.4
    sd, orc = one_tet(E, nu)
    lam, mu = lame(E, nu)
    F = np.diag([s, s, s]) if kind == "uniform" else np.diag([s, 1.0, 1.0])
    x = (sd.X @ F.T).ravel()
    e = orc.energy(x, None, NONE, 0.0)[0]
    h2 = sd.params["dt"] ** 2
    lj = 3 * math.log(s) if kind == "uniform" else math.log(s)
    f2 = 3 * s * s if kind == "uniform" else s * s + 2
    exp = h2 / 6 * (mu / 2 * (f2 - 3) - mu * lj + lam / 2 * lj * lj)
    assert abs(e[0] - exp) <= 1e-12 * exp
    # a rotation changes nothing
    R = np.linalg.qr(np.random.default_rng(1).normal(size=(3, 3)))[0]
    R *= np.sign(np.linalg.det(R))
    e2 = orc.energy((x.reshape(-1, 3) @ R.T).ravel(), None, NONE, 0.0)[0]
    assert abs(e2[0] - exp) <= 1e-11 * exp


def stress(orc, sd, F):
    """first Piola-Kirchhoff stress of the unit tet at F from the nodal gradient:
    g_a = h^2 vol P e_a (a = 1, 2, 3)"""
    g = orc.gradient((sd.X @ F.T).ravel(), NONE, 0.0)[0].reshape(4, 3)
    return g[1:].T / (sd.params["dt"] ** 2 / 6)


def test_gradient_stress_and_stress_derivative():
    E, nu = 1e5, 0.4
    sd, orc = one_tet(E, nu)
    lam, mu = lame(E, nu)
    rng = np.random.default_rng(3)
    for trial in range(4):
        F = np.eye(3) + rng.uniform(-0.25, 0.25, (3, 3))
        assert np.linalg.det(F) > 0.3
        # gradient vs central differences of the energy
        x = (sd.X @ F.T).ravel()
        f = lambda y: orc.energy(y, None, NONE, 0.0)[0][0]
        g = orc.gradient(x, NONE, 0.0)[0]
        fd = np.array([(f(x + 1e-7 * e) - f(x - 1e-7 * e)) / 2e-7 for e in np.eye(12)])
        assert np.abs(g - fd).max() <= 1e-6 * np.abs(g).max()
        # P against the closed form mu (F - F^-T) + lambda ln J F^-T
        P = stress(orc, sd, F)
        Fit = np.linalg.inv(F).T
        assert np.abs(P - (mu * (F - Fit) + lam * math.log(np.linalg.det(F)) * Fit)).max() <= 1e-9 * np.abs(P).max()
        # dP/dF (vec column-major) by central differences of P
        D = np.zeros((9, 9))
        for q in range(9):
            dF = np.zeros(9)
            dF[q] = 1e-6
            dFm = dF.reshape(3, 3, order="F")
            D[:, q] = ((stress(orc, sd, F + dFm) - stress(orc, sd, F - dFm)) / 2e-6).ravel(order="F")
        raw = O.nh_dPdF(F, lam, mu, project=False)
        assert np.abs(raw - D).max() <= 1e-5 * np.abs(D).max()
        w, V = np.linalg.eigh(0.5 * (D + D.T))
        proj = (V * np.maximum(w, 0)) @ V.T
        assert np.abs(O.nh_dPdF(F, lam, mu, project=True) - proj).max() <= 1e-5 * np.abs(proj).max()


def test_element_hessian_from_stress_derivative():
    """12x12 = D^T dPdF D with dF_ij/dx_a,k = delta_ik c_a(j), c_a = row a-1 of Dm^-1 and
    c_0 = -sum; unprojected it equals the central difference of the gradient"""
    rng = np.random.default_rng(5)
    X = np.array([[0.0, 0, 0], [0.9, 0.1, 0], [0.2, 1.1, 0.1], [0.1, 0.2, 0.8]])
    sd, orc = one_tet(X=X)
    ta = orc.tet_arrays()
    B = ta["IB"][0]
    c = np.vstack([-B.sum(0), B])
    Dx = np.zeros((9, 12))
    for a in range(4):
        for k in range(3):
            for j in range(3):
                Dx[k + 3 * j, 3 * a + k] = c[a, j]
    x = (X + rng.uniform(-0.1, 0.1, X.shape)).ravel()
    Fm = (x.reshape(4, 3)[1:] - x.reshape(4, 3)[0]).T @ B
    lam, mu = ta["lam"], ta["mu"]
    for project in (False, True):
        H = orc.tet_hess_local(x, 0, 1.0, project)
        ref = Dx.T @ O.nh_dPdF(Fm, lam, mu, project) @ Dx
        assert np.abs(H - ref).max() <= 1e-9 * np.abs(ref).max()
    Hraw = orc.tet_hess_local(x, 0, 1.0, False)
    w = sd.params["dt"] ** 2 * ta["vol"][0]
    gfun = lambda y: orc.gradient(y, NONE, 0.0)[0] / w
    fd = np.stack([(gfun(x + 1e-6 * e) - gfun(x - 1e-6 * e)) / 2e-6 for e in np.eye(12)], 1)
    assert np.abs(Hraw - fd).max() <= 1e-5 * np.abs(fd).max()
    Hp = orc.tet_hess_local(x, 0, 1.0, True)
    assert np.linalg.eigvalsh(0.5 * (Hp + Hp.T)).min() >= -1e-9 * np.abs(Hp).max()


def test_cubic_root_and_inversion_free_stepsize():
    rng = np.random.default_rng(9)
    for _ in range(200):
        a, b, c, d = rng.normal(size=4)
        r = np.roots([a, b, c, d])
        real = [z.real for z in r if abs(z.imag) < 1e-9 and z.real > 0]
        got = O.cubic_root(a, b, c, d)
        if real:
            assert abs(got - min(real)) <= 1e-9 * max(1.0, min(real))
        elif got > 0:   # a root the Cardano form reads as real within its 1e-6 imaginary tolerance
            assert min(abs(z.imag) for z in r) < 1e-5
    X = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    sd, orc = one_tet(X=X)
    hits = 0
    for _ in range(100):
        p = rng.normal(size=12)
        x = X.ravel()
        def det_at(t):
            y = (x + t * p).reshape(4, 3)
            return np.linalg.det((y[1:] - y[0]).T)
        # det(x + t p) - 0.2 det(x): cubic through 4 samples
        ts = np.array([0.0, 1.0, 2.0, 3.0])
        coef = np.polyfit(ts, [det_at(t) - 0.2 * det_at(0.0) for t in ts], 3)
        r = [z.real for z in np.roots(coef) if abs(z.imag) < 1e-9 and z.real > 0]
        alpha = orc.inversion_free_stepsize(x, p, 1.0)
        if r and min(r) < 1.0:
            hits += 1
            assert abs(alpha - min(r)) <= 1e-7 * min(r)
            assert det_at(alpha) > 0
        else:
            assert alpha == 1.0 or alpha > 0
    assert hits > 20


def test_tet_cloth_steps():
    """cloth on a soft tet cube on the ground: no tet inverts, every state intersection
    free, each accepted line search lowers the potential, the step ends stationary"""
    sd = scenes.scene_tet_cloth()
    sd.params["PNTol"] = 1e-4 * sd.bbox_diag()
    orc = O.OracleScene(sd, solver=0, threads=4)
    h = sd.params["dt"]
    g = np.array(sd.params["gravity"])
    contacts = []
    for s in range(4):
        x0, v0 = orc.get_state()
        rc, st = orc.step()
        assert rc >= 3, (s, rc)
        x, _ = orc.get_state()
        C = orc.constraint_set(x)
        contacts.append(len(C))
        assert orc.dist2_list(x, C).min() > 0
        Xt = x.reshape(-1, 3)[sd.tets]
        assert (np.linalg.det(np.stack([Xt[:, 1] - Xt[:, 0], Xt[:, 2] - Xt[:, 0], Xt[:, 3] - Xt[:, 0]], -1)) > 0).all()
        eb, ea = orc.energy_trace()
        assert (ea <= eb).all()
    assert min(contacts) > 100
    # stationarity of the last step's end state
    m = orc.arrays()["mass"]
    xt = x0 + h * v0 + (h * h * np.where(m[:, None] > 0, g[None, :], 0.0)).ravel()
    kappa = st["kappa"]
    C = orc.constraint_set(x)
    ge, gb = orc.gradient(x, C, kappa)
    grad = ge + gb + np.repeat(m, 3) * (x - xt)
    free = np.repeat(~np.isin(np.arange(sd.nV), sd.dbc_idx), 3)
    H = orc.hessian(x, C, kappa, which=3).tocsc()
    p = spla.spsolve(H[free][:, free], -grad[free])
    assert math.sqrt((p * p).sum() / (sd.nV - len(sd.dbc_idx))) / h <= sd.params["PNTol"]
