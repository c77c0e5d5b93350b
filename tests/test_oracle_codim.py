"""Oracle pins for the codimensional primitives: rod springs and rod hinges (derivatives
against torch autograd of the reference's literal formulas, Rod/ROD_STRETCHING.h:22-24 and
Rod/ROD_BENDING.h:24-30), rod / particle set-up (DISCRETE_ROD.h:87-165,
DISCRETE_PARTICLE.h:52-75, IMPLICIT_EULER.h:243-277), the codimensional contact branches
(IPC.h:271-353: hashed = brute force, closed-form cases, multiplicity) and the particle CCD
(IPC.h:2095-2162: never later than the exact time of impact of closed-form motions)."""
import math

import numpy as np
import pytest
import torch

import oracle as O
from cipc import scenes


def _spring_ref(x, Er, lbar):
    x0, x1 = x[:3], x[3:]
    return lbar * Er / 2 * ((x0 - x1).norm() / lbar - 1) ** 2


def _kappa2_ref(x):
    x0, x1, x2 = x[:3], x[3:6], x[6:]
    e0, e1 = x1 - x0, x2 - x1
    k = torch.linalg.cross(e0, e1) * 2 / (torch.sqrt((e0 @ e0) * (e1 @ e1)) + e0 @ e1)
    return k @ k


def _autograd(f, x):
    t = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    e = f(t)
    g = torch.autograd.grad(e, t, create_graph=True)[0]
    H = torch.stack([torch.autograd.grad(g[i], t, retain_graph=True)[0] for i in range(len(x))])
    return float(e), g.detach().numpy(), H.detach().numpy()


@pytest.mark.parametrize("stretch", [0.7, 1.0 + 1e-6, 1.3])
def test_rod_spring_derivatives(stretch):
    rng = np.random.default_rng(int(stretch * 100))
    d = rng.standard_normal(3)
    d /= np.linalg.norm(d)
    lbar = 0.01
    x = np.r_[rng.standard_normal(3) * 0.1, np.zeros(3)]
    x[3:] = x[:3] + stretch * lbar * d
    E, g, H = O.rod_spring(x, 1e7, lbar)
    Er, gr, Hr = _autograd(lambda t: _spring_ref(t, 1e7, lbar), x)
    assert abs(E - Er) <= 1e-12 * max(abs(Er), 1e-300) + 1e-300
    assert np.abs(g - gr).max() <= 1e-9 * np.abs(gr).max() + 1e-12
    assert np.abs(H - Hr).max() <= 1e-9 * np.abs(Hr).max()


@pytest.mark.parametrize("angle_deg", [0.5, 20.0, 90.0, 150.0])
def test_rod_kappa2_derivatives(angle_deg):
    """|kappa|^2 = 4 tan^2(theta / 2) for turning angle theta; derivatives vs autograd of the
    cross-product form and central differences of the gradient."""
    rng = np.random.default_rng(int(angle_deg * 10))
    a = math.radians(angle_deg)
    l0, l1 = 0.01, 0.013
    x = np.array([-l0, 0, 0, 0, 0, 0, l1 * math.cos(a), l1 * math.sin(a), 0.0])
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    x = (x.reshape(3, 3) @ Q.T + rng.standard_normal(3) * 0.1).ravel()
    k2, g, H = O.rod_kappa2(x)
    assert abs(k2 - 4 * math.tan(a / 2) ** 2) <= 1e-10 * max(1.0, 4 * math.tan(a / 2) ** 2)
    kr, gr, Hr = _autograd(_kappa2_ref, x)
    assert np.abs(g - gr).max() <= 1e-8 * np.abs(gr).max()
    assert np.abs(H - Hr).max() <= 1e-7 * np.abs(Hr).max()
    h = 1e-7
    for i in range(9):
        e = np.zeros(9)
        e[i] = h
        fd = (O.rod_kappa2(x + e)[1] - O.rod_kappa2(x - e)[1]) / (2 * h)
        assert np.abs(fd - H[i]).max() <= 1e-4 * np.abs(H).max()


def test_codim_setup():
    sd = scenes.scene_codim_small()
    o = O.OracleScene(sd)
    nR, nP = len(sd.rods), len(sd.particles)
    assert (o.nRod, o.nParticle) == (nR, nP)
    # 6 rods of 5 segments: 4 interior nodes each -> 24 hinges
    assert o.nRodHinge == 24
    a = o.arrays()
    c = o.codim_arrays()
    P = sd.params
    X = sd.X
    lens = np.linalg.norm(X[sd.rods[:, 0]] - X[sd.rods[:, 1]], axis=1)
    assert np.allclose(c["rod_info"][:, 1], lens, rtol=0, atol=0)
    # mass: shell + pi t^2/4 rho l per segment + 4/3 pi t^3/8 rho per particle
    shell = O.OracleScene(_shell_only(sd))
    m_shell = shell.arrays()["mass"]
    rod_m = lens * math.pi * P["rodThick"] ** 2 / 4 * P["rodRho"]
    part_m = 4.0 / 3.0 * math.pi * P["partThick"] ** 3 / 8 * P["partRho"]
    assert abs(a["mass"].sum() - (m_shell.sum() + rod_m.sum() + nP * part_m)) <= 1e-12 * a["mass"].sum()
    assert abs(part_m - scenes.PARTICLE_MASS) <= 1e-15
    # boundary nodes: surface nodes, rod nodes ascending, particles in order
    bn = a["bnode"]
    rod_nodes = np.unique(sd.rods)
    assert np.array_equal(bn[o.codimBN0:o.codimBN1], rod_nodes)
    assert np.array_equal(bn[o.codimBN1:], sd.particles)
    assert np.array_equal(a["bedge"][-nR:], sd.rods)
    # hinge stencils: (neighbour, node, neighbour) in ascending node order
    rh = c["rod_hinge"]
    assert np.all(np.diff(rh[:, 1]) > 0)
    for h in rh:
        assert abs(h[0] - h[1]) == 1 and abs(h[2] - h[1]) == 1


def _shell_only(sd):
    s = scenes.SceneData(sd.X, sd.tris, sd.dbc_idx, sd.dbc_target, sd.params, "shell")
    return s


def test_codim_constraint_set_hash_equals_brute():
    sd = scenes.scene_codim_contact()
    o = O.OracleScene(sd, threads=4)
    x, _ = o.get_state()
    C = o.constraint_set(x)
    assert np.array_equal(C, o.constraint_set(x, brute=True))
    rods = set(map(tuple, sd.rods.tolist()))
    parts = set(sd.particles.tolist())
    pe = [c for c in C if c[0] < 0 and c[2] >= 0 and c[3] < 0 and (int(c[1]), int(c[2])) in rods]
    pp = [c for c in C if c[0] < 0 and c[2] == -1 and (-int(c[0]) - 1) in parts and int(c[1]) in parts]
    assert len(pe) > 10 and len(pp) > 5


def _bare_codim(rod_nodes, particles, xi=1e-3):
    """ground only (Dirichlet) + the given rods and particles, no gravity."""
    Xg, Tg = scenes.ground(1.0, y=-0.5)
    base = scenes.SceneData(Xg, Tg, np.arange(4, dtype=np.int32), Xg.copy(), dict(scenes.DEFAULT_PARAMS), "bare")
    base.params["gravity"] = (0.0, 0.0, 0.0)
    base.n_cloth = 0
    return scenes.with_codim(base, rod_nodes, particles, xi=xi)


def test_codim_constraint_closed_forms():
    """On a rod bent into a roof (peak at its middle node), a particle 1.5 mm off the middle
    of the first segment gives one point-edge entry; one 1.5 mm above the peak lies beyond
    the ends of both segments there and gives a point-point entry of multiplicity 2 (each
    segment reports it, IPC.h:599-661); two particles 1.5 mm apart give one point-point
    entry (offset 1 mm: activation distance 2 mm)."""
    rod = np.array([[-0.02, -0.004, 0], [-0.01, -0.002, 0], [0, 0, 0], [0.01, -0.002, 0], [0.02, -0.004, 0]])
    n = np.array([-0.002, 0.01, 0]) / math.hypot(0.002, 0.01)
    parts = np.array([np.array([-0.015, -0.003, 0]) + 0.0015 * n, [0.0, 0.0015, 0.0], [0.1, 0.1, 0.1], [0.1015, 0.1, 0.1]])
    sd = _bare_codim([rod], parts)
    o = O.OracleScene(sd)
    x, _ = o.get_state()
    C = {tuple(int(v) for v in c) for c in o.constraint_set(x)}
    r0 = 4                       # first rod node id (after the 4 ground vertices)
    p0, p1, p2, p3 = 9, 10, 11, 12
    assert (-p0 - 1, r0, r0 + 1, -1) in C           # point-edge, segment 0
    assert (-p1 - 1, r0 + 2, -1, -2) in C           # point-point with the middle node, multiplicity 2
    assert (-p2 - 1, p3, -1, -1) in C               # particle pair
    assert len(C) == 3


def test_codim_ccd_closed_forms():
    xi = 1e-3
    d0, v = 0.004, 0.01
    rod = scenes.make_rod((-0.02, 0, 0), (0.02, 0, 0), 1)
    parts = np.array([[0.003, d0, 0.0], [0.3, 0.2, 0.3], [0.3 + d0, 0.2, 0.3]])
    sd = _bare_codim([rod], parts, xi=xi)
    o = O.OracleScene(sd)
    x, _ = o.get_state()
    # particle 0 falls onto the segment: exact time (d0 - xi) / v
    p = np.zeros_like(x).reshape(-1, 3)
    p[6] = (0, -v, 0)
    a, cl = o.ccd(x, p.ravel(), 1.0)
    t = (d0 - xi) / v
    assert cl == 1.0 and 0.85 * t <= a <= t
    # particles 1 and 2 approach head-on: exact time (d0 - xi) / (2 v)
    p = np.zeros_like(x).reshape(-1, 3)
    p[7] = (v, 0, 0)
    p[8] = (-v, 0, 0)
    a, cl = o.ccd(x, p.ravel(), 1.0)
    t = (d0 - xi) / (2 * v)
    assert 0.85 * t <= a <= t


def test_codim_steps_intersection_free():
    sd = scenes.scene_codim_small()
    o = O.OracleScene(sd, threads=4)
    xi2 = sd.params["xi"] ** 2
    for _ in range(6):
        rc, st = o.step()
        assert rc > 0
        x, _ = o.get_state()
        C = o.constraint_set(x)
        if len(C):
            assert o.dist2_list(x, C).min() > xi2
    # the rods and particles have come down onto the cloth
    assert st["contacts"] > 0


def test_c3_scene_layout():
    sd = scenes.scene_c3()
    assert sd.nV == 324247 and len(sd.rods) == 200000 and len(sd.particles) == 100000
    assert scenes.min_codim_separation(sd) > sd.params["xi"] + sd.params["dHat"]
    cloth_top = sd.X[:sd.n_cloth, 1].max()
    codim = np.unique(np.r_[sd.rods.ravel(), sd.particles])
    assert sd.X[codim, 1].min() - cloth_top > 0.04
