"""Pins for the oracle's contact stages (CPU only): constraint set vs brute force and
the reference predicates, CCD conservativeness on closed-form cases and on scenes
(every accepted state free of edge-triangle intersections), solver accuracy vs scipy."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
from cipc import scenes


@pytest.fixture(scope="module")
def stack():
    sd = scenes.scene_contact_stack()
    orc = O.OracleScene(sd, threads=4)
    return sd, orc


def test_hashed_set_equals_brute_force(stack):
    sd, orc = stack
    x, _ = orc.get_state()
    rng = np.random.default_rng(0)
    for trial in range(3):
        xx = x.copy()
        if trial:
            xx[: 3 * sd.n_cloth] += rng.uniform(-2e-4, 2e-4, 3 * sd.n_cloth)
        C = orc.constraint_set(xx)
        assert np.array_equal(C, orc.constraint_set(xx, brute=True))
        assert len(C) > 500


def test_constraint_predicates(stack):
    """every active stencil is closer than (dhat + offset); encoding invariants of IPC.h:144-661"""
    sd, orc = stack
    x, _ = orc.get_state()
    C = orc.constraint_set(x)
    d2 = orc.dist2_list(x, C)
    dh = sd.params["dHat"] + sd.params["xi"]
    assert (d2 < dh * dh).all() and (d2 > 0).all()
    pt = C[(C[:, 0] < 0) & (C[:, 3] >= 0)]
    dup = C[(C[:, 0] < 0) & (C[:, 3] < 0)]
    assert len(pt) and len(dup)
    assert (dup[:, 3] <= -1).all()
    # deduplicated entries are unique and lexicographically sorted (std::map order)
    keys = [tuple(r[:3]) for r in dup]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    # point-triangle entries ordered by point then triangle
    assert (np.diff(-pt[:, 0] - 1) >= 0).all()


def point_plane_toi(p, dp, z):
    return (p[1] - z) / -dp[1]


def test_ccd_point_triangle_closed_form():
    """a point falling onto a static triangle: the CCD step never exceeds the exact time of
    impact (conservative) nor the hash step clamp, and stays within the additive-CCD gap"""
    # the falling point is a vertex of a far-away static sliver triangle so that it is a
    # boundary node of the contact primitives
    X = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.2, 0.1, 0.2], [5.0, 5.0, 5.0], [5.1, 5.0, 5.0]])
    T = np.array([[0, 2, 1], [3, 4, 5]], dtype=np.int32)
    fixed = np.array([0, 1, 2, 4, 5], dtype=np.int32)
    sd = scenes.SceneData(X, T, fixed, X[fixed].copy(), dict(scenes.DEFAULT_PARAMS, PNTol=1e-3))
    orc = O.OracleScene(sd)
    x = X.ravel().copy()
    for vy in (-0.05, -0.2, -1.0):
        p = np.zeros_like(x)
        p[3 * 3 + 1] = vy
        a, cl = orc.ccd(x, p, 1.0)
        toi = 0.1 / -vy
        assert a <= min(cl, toi)
        if toi < cl:
            assert a >= 0.5 * toi


@pytest.mark.parametrize("layout", ["crossing", "parallel"])
def test_ccd_edge_edge_closed_form(layout):
    """A moving edge approaching a static edge at speed v from distance d0 (offset 0):
    exact time of impact d0 / v.  "crossing": perpendicular edges whose closest points are
    both interior (edge-edge, every vertex at least 1 m from the other triangle);
    "parallel": the moving edge lies above the static one and parallel to it (the
    near-parallel fallback of the edge-edge distance and the point-edge case)."""
    d0 = 0.01
    A = [[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, -1.0, 0.0]]
    if layout == "crossing":
        B = [[0.0, d0, -1.0], [0.0, d0, 1.0], [0.0, d0 + 1.0, 0.0]]
    else:
        B = [[-0.5, d0, 0.0], [0.5, d0, 0.0], [0.0, d0 + 1.0, 1.0]]
    X = np.array(A + B)
    T = np.array([[0, 1, 2], [3, 4, 5]], dtype=np.int32)
    fixed = np.array([0, 1, 2], dtype=np.int32)
    sd = scenes.SceneData(X, T, fixed, X[fixed].copy(), dict(scenes.DEFAULT_PARAMS, PNTol=1e-3, xi=0.0))
    orc = O.OracleScene(sd)
    x = X.ravel().copy()
    for v in (0.005, 0.02, 0.2):
        p = np.zeros_like(x).reshape(-1, 3)
        p[3:, 1] = -v
        a, cl = orc.ccd(x, p.ravel(), 1.0)
        toi = d0 / v
        assert a <= min(cl, toi)
        if toi < cl:
            assert a >= 0.85 * toi
        else:
            assert a >= 0.85 * cl


def seg_tri_intersect(e0, e1, t0, t1, t2):
    M = np.column_stack([t1 - t0, t2 - t0, e0 - e1])
    det = np.linalg.det(M)
    if abs(det) < 1e-300:
        return False
    u, v, s = np.linalg.solve(M, e0 - t0)
    return u >= 0 and v >= 0 and u + v <= 1 and 0 <= s <= 1


def test_ccd_states_intersection_free():
    """two sheets driven through each other by a prescribed direction: the CCD-limited
    step never produces an edge-triangle crossing (IPC.h:2390-2441 check)"""
    sd = scenes.scene_small(n=5, sheets=2, y0=0.3, dy=0.004, twist_deg=20.0)
    orc = O.OracleScene(sd)
    x, _ = orc.get_state()
    n_cloth = sd.n_cloth
    p = np.zeros_like(x)
    p.reshape(-1, 3)[: n_cloth // 2, 1] = 0.05   # lower sheet moves up through the upper
    a, cl = orc.ccd(x, p, 1.0)
    assert 0 < a < 0.08
    xa = x + a * p
    X = xa.reshape(-1, 3)
    tris = sd.tris[: 2 * 2 * (5 - 1) ** 2]
    arr = orc.arrays()
    for e in arr["bedge"]:
        if e[0] >= n_cloth or e[1] >= n_cloth:
            continue
        for t in tris:
            if e[0] in t or e[1] in t:
                continue
            assert not seg_tri_intersect(X[e[0]], X[e[1]], X[t[0]], X[t[1]], X[t[2]])


def test_direct_solve_vs_scipy(stack):
    sd, orc = stack
    x, _ = orc.get_state()
    C = orc.constraint_set(x)
    kappa = orc.scalars()["kappa"]
    H = orc.hessian(x, C, kappa, which=3)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(H.shape[0])
    b.reshape(-1, 3)[sd.dbc_idx] = 0
    rc, sol, _ = O.solve(H, b, solver=0)
    assert rc == 0
    assert np.linalg.norm(H @ sol - b) <= 1e-12 * np.linalg.norm(b)
    ref = spla.spsolve(H.tocsc(), b)
    assert np.abs(sol - ref).max() <= 1e-10 * np.abs(ref).max()
    rc, solp, it = O.solve(H, b, solver=1, tol=1e-8)
    assert rc == 0 and np.linalg.norm(H @ solp - b) <= 1e-8 * np.linalg.norm(b)
    # an indefinite matrix is reported, not factorised (CHOLMOD_NOT_POSDEF path)
    rc, _, _ = O.solve(-H, b, solver=0)
    assert rc == 1
