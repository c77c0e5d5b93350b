"""Synthetic scenes for the contact-and-Newton step.

The reference drives its FEMShell scenes from OBJ meshes and Python scene scripts
(Projects/FEMShell/*.py, e.g. 18_sphere_on_cloth_stack.py); the benchmark scenes here
are generated instead, from recorded seeds, so that every consumer (the HIP path,
the CPU oracle, the benchmark) sees identical bits:

* ``cloth_grid``: an n x n regular grid over [-0.5, 0.5]^2 in the xz plane, two
  triangles per quad with a consistent +y normal (same topology family as
  Projects/FEMShell/input/square41.obj);
* ``icosphere``: the level-k subdivided icosahedron (level 4: 2,562 vertices,
  5,120 faces);
* ``ground``: two triangles spanning [-1, 1]^2 at y = 0;
* vertex jitter: uniform(-1e-4, 1e-4) m per coordinate from a splitmix64 stream.

``scene_c1`` and ``scene_c2`` build the two cloth-on-sphere workloads (41x41 cloth,
and four 201x201 sheets 5 mm apart); sphere and ground vertices are Dirichlet-fixed.
"""
import math

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64_uniform(seed, count):
    """``count`` doubles in [0, 1) from a splitmix64 stream (53-bit mantissa)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & MASK64) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def cloth_grid(n, y=0.5):
    idx = np.arange(n)
    xs = -0.5 + idx / (n - 1)
    X = np.zeros((n * n, 3))
    for j in range(n):
        X[j * n:(j + 1) * n, 0] = xs
        X[j * n:(j + 1) * n, 2] = -0.5 + j / (n - 1)
    X[:, 1] = y
    tris = []
    for j in range(n - 1):
        for i in range(n - 1):
            v00 = i + n * j
            v10 = v00 + 1
            v01 = v00 + n
            v11 = v01 + 1
            tris.append((v00, v11, v10))
            tris.append((v00, v01, v11))
    return X, np.array(tris, dtype=np.int32)


def icosphere(level=4, radius=0.2, center=(0.0, 0.0, 0.0)):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
             (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
             (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
             (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
             (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for (a, b, c) in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    X = np.array(verts) * radius + np.asarray(center, dtype=np.float64)
    return X, np.array(faces, dtype=np.int32)


def ground(half=1.0, y=0.0):
    X = np.array([[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]], dtype=np.float64)
    T = np.array([[0, 2, 1], [0, 3, 2]], dtype=np.int32)
    return X, T


class SceneData:
    """Arrays and parameters of one scene (all float64 / int32, vertex-major xyz)."""

    def __init__(self, X, tris, dbc_idx, dbc_target, params, name=""):
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.tris = np.ascontiguousarray(tris, dtype=np.int32)
        self.dbc_idx = np.ascontiguousarray(dbc_idx, dtype=np.int32)
        self.dbc_target = np.ascontiguousarray(dbc_target, dtype=np.float64)
        self.params = dict(params)
        self.name = name
        self.rods = None        # (n, 2) int32 rod segments or None
        self.particles = None   # (n,) int32 particle vertex ids or None
        self.tets = None        # (n, 4) int32 volumetric tets or None (params tetRho, tetE, tetNu)

    @property
    def nV(self):
        return self.X.shape[0]

    @property
    def nT(self):
        return self.tris.shape[0]

    def bbox_diag(self):
        return float(np.linalg.norm(self.X.max(axis=0) - self.X.min(axis=0)))


DEFAULT_PARAMS = dict(rho=200.0, E=1e6, nu=0.3, thick=1e-3, bendMult=1.0, dt=0.01, dHat=1e-3, xi=0.0,
                      gravity=(0.0, -9.81, 0.0), kappaStiffMult=1.0, collision=1)


def cloth_on_sphere(sheets=1, n=41, y0=0.5, dy=0.005, seed=1, sphere_level=4, sphere_radius=0.2,
                    ground_half=1.0, jitter=1e-4, name="", twist_deg=0.0, twists=None):
    parts_X, parts_T = [], []
    nv = 0
    for s in range(sheets):
        X, T = cloth_grid(n, y0 + s * dy)
        tw = twists[s] if twists is not None else twist_deg * s
        if tw:
            a = math.radians(tw)
            c, sn = math.cos(a), math.sin(a)
            X = X.copy()
            X[:, 0], X[:, 2] = c * X[:, 0] - sn * X[:, 2], sn * X[:, 0] + c * X[:, 2]
        parts_X.append(X)
        parts_T.append(T + nv)
        nv += X.shape[0]
    n_cloth = nv
    Xc = np.concatenate(parts_X)
    if jitter:
        u = splitmix64_uniform(seed, Xc.size).reshape(Xc.shape)
        Xc = Xc + (2.0 * u - 1.0) * jitter
    Xs, Ts = icosphere(sphere_level, sphere_radius)
    Xg, Tg = ground(ground_half)
    X = np.concatenate([Xc, Xs, Xg])
    T = np.concatenate(parts_T + [Ts + n_cloth, Tg + n_cloth + Xs.shape[0]])
    dbc = np.arange(n_cloth, X.shape[0], dtype=np.int32)
    params = dict(DEFAULT_PARAMS)
    sd = SceneData(X, T, dbc, X[dbc].copy(), params, name)
    sd.params["PNTol"] = 1e-2 * sd.bbox_diag()
    sd.n_cloth = n_cloth
    return sd


def scene_c1(seed=1):
    """41x41 cloth (1,681 v / 3,200 tris) over the icosphere and ground, 4,247 vertices."""
    return cloth_on_sphere(sheets=1, n=41, seed=seed, name="c1_cloth41_sphere")


def scene_c2(seed=2):
    """Four 201x201 sheets 5 mm apart (161,604 cloth v / 320,000 tris) over the icosphere and ground."""
    return cloth_on_sphere(sheets=4, n=201, seed=seed, name="c2_cloth4x201_sphere")


def scene_small(n=9, sheets=2, seed=3, y0=0.22, dy=0.004, twist_deg=0.0, sphere_level=2):
    """A few-hundred-vertex cloth stack just above the sphere: contact within a few steps."""
    return cloth_on_sphere(sheets=sheets, n=n, y0=y0, dy=dy, seed=seed, sphere_level=sphere_level, name="small_stack",
                           twist_deg=twist_deg)


def scene_contact_stack(n=12, seed=5):
    """Three sheets 0.8 mm apart (inside d_hat = 1 mm) resting 0.8 mm above the sphere top; the two
    lower sheets are aligned (parallel edges: mollified edge-edge / point-edge / point-point) and the top
    one is twisted by 10 degrees (general point-triangle and edge-edge), so every constraint kind is active
    in the first state."""
    sd = cloth_on_sphere(sheets=3, n=n, y0=0.2008, dy=0.0008, seed=seed, sphere_level=3, name="contact_stack",
                         twists=[0.0, 0.0, 10.0])
    return sd


def scene_cloth_on_ground(n=40, gap=6e-4, seed=9, twist_deg=0.0, ground_half=1.0):
    """One n x n sheet `gap` above the 2-triangle ground (no sphere).  The ground triangles
    and edges cover thousands of grid cells, so they take the large-primitive path of the
    broad phases; the sheet's diagonals run parallel to the ground diagonal, so mollified
    edge-edge entries appear next to point-triangle and edge-edge ones."""
    X, T = cloth_grid(n, gap)
    if twist_deg:
        a = math.radians(twist_deg)
        c, sn = math.cos(a), math.sin(a)
        X[:, 0], X[:, 2] = c * X[:, 0] - sn * X[:, 2], sn * X[:, 0] + c * X[:, 2]
    u = splitmix64_uniform(seed, X.size).reshape(X.shape)
    X = X + (2.0 * u - 1.0) * np.array([1e-4, 1e-4 * min(1.0, gap / 1e-3) * 0.5, 1e-4])
    n_cloth = X.shape[0]
    Xg, Tg = ground(ground_half)
    Xa = np.concatenate([X, Xg])
    Ta = np.concatenate([T, Tg + n_cloth])
    dbc = np.arange(n_cloth, Xa.shape[0], dtype=np.int32)
    sd = SceneData(Xa, Ta, dbc, Xa[dbc].copy(), dict(DEFAULT_PARAMS), "cloth_on_ground")
    sd.params["PNTol"] = 1e-2 * sd.bbox_diag()
    sd.n_cloth = n_cloth
    return sd


# ---------------------------------------------------------------- codimensional scenes
ROD_PARAMS = dict(rodRho=1200.0, rodE=1e7, rodBendMult=1.0, rodThick=1e-3)
PARTICLE_MASS = 1e-5
PARTICLE_PARAMS = dict(partThick=1e-3, partRho=PARTICLE_MASS / (math.pi * 1e-3 ** 3 / 6))


def make_rod(p0, p1, nseg):
    """Make_Rod (Rod/DISCRETE_ROD.h:6-28) between two points: nseg + 1 nodes, nseg segments."""
    t = np.linspace(0.0, 1.0, nseg + 1)[:, None]
    return (1 - t) * np.asarray(p0, float) + t * np.asarray(p1, float)


def with_codim(sd, rod_nodes, particles, xi=1e-3, name=None, rod_params=None, particle_params=None):
    """Append rods (list of node arrays, each a polyline) and particles ((n, 3) array) to a
    scene: initialize_rod / initialize_particle of Python/Drivers/FEMDiscreteShellBase.py, offset xi."""
    X = [sd.X]
    nv = sd.nV
    segs = []
    for nodes in rod_nodes:
        k = len(nodes)
        X.append(np.asarray(nodes, float))
        segs.append(np.stack([np.arange(nv, nv + k - 1), np.arange(nv + 1, nv + k)], 1))
        nv += k
    parts = np.arange(nv, nv + len(particles), dtype=np.int32)
    X.append(np.asarray(particles, float).reshape(-1, 3))
    out = SceneData(np.concatenate(X), sd.tris, sd.dbc_idx, sd.dbc_target, sd.params, name or sd.name)
    out.n_cloth = getattr(sd, "n_cloth", 0)
    out.rods = np.concatenate(segs).astype(np.int32) if segs else None
    out.particles = parts if len(parts) else None
    out.params.update(ROD_PARAMS if rod_params is None else rod_params)
    out.params.update(PARTICLE_PARAMS if particle_params is None else particle_params)
    out.params["xi"] = xi
    out.params["PNTol"] = 1e-2 * out.bbox_diag()
    return out


def scene_codim_contact(seed=11):
    """Every codimensional contact kind active in the first state (offset xi = 1 mm, so the
    activation distance is 2 mm): a 12x12 sheet 1.5 mm above the sphere top, four rods of
    eight segments lying 1.5 mm above the sheet and 1.5 mm apart, particles 1.5 mm above
    the rods (point-edge) in pairs 1.5 mm apart (point-point)."""
    base = cloth_on_sphere(sheets=1, n=12, y0=0.2025, seed=seed, sphere_level=3, name="codim_contact")
    base.X[:base.n_cloth, 0] *= 0.1
    base.X[:base.n_cloth, 2] *= 0.1
    rng = np.random.default_rng(seed)
    rods = []
    for k in range(4):
        z = -0.00225 + 0.0015 * k
        y = 0.2040
        rods.append(make_rod((-0.03, y, z), (0.03, y, z), 8) + rng.uniform(-5e-5, 5e-5, 3))
    parts = []
    for k in range(4):
        z = -0.00225 + 0.0015 * k
        for xk in (-0.012, 0.009):
            parts.append((xk, 0.2055, z))
            parts.append((xk + 0.0015, 0.2055, z))
    parts = np.array(parts) + rng.uniform(-5e-5, 5e-5, (len(parts), 3))
    return with_codim(base, rods, parts, name="codim_contact")


def scene_codim_small(seed=13):
    """A falling codimensional mix: a 12x12 sheet 3 mm above the sphere, 6 rods of 5 segments
    4 mm above it, 30 particles 4 mm above the rods."""
    base = cloth_on_sphere(sheets=1, n=12, y0=0.203, seed=seed, sphere_level=3, name="codim_small")
    base.X[:base.n_cloth, 0] *= 0.2
    base.X[:base.n_cloth, 2] *= 0.2
    rng = np.random.default_rng(seed)
    rods = [make_rod((-0.04, 0.207, -0.025 + 0.01 * k), (0.04, 0.207, -0.025 + 0.01 * k), 5) + rng.uniform(-1e-4, 1e-4, 3)
            for k in range(6)]
    g = np.stack(np.meshgrid(np.linspace(-0.04, 0.04, 6), [0.211], np.linspace(-0.02, 0.02, 5), indexing="ij"), -1).reshape(-1, 3)
    parts = g + rng.uniform(-2e-4, 2e-4, g.shape)
    return with_codim(base, rods, parts, name="codim_small")


def scene_c3(seed=3, n_rods=20000, n_seg=10, n_particles=100000):
    """C1 (41x41 cloth over the sphere and ground) + 20,000 rods (10 segments, 0.1 m, E 1e7)
    + 100,000 particles (1e-5 kg) in a 0.3 m cube above the cloth, offset xi = 1 mm.
    Rods run along x in two columns (x in [-0.145, -0.045] and [0.045, 0.145]) on a
    100 x 100 (y, z) lattice of 3 mm pitch; particles fill the 9 cm gap between the
    columns on a jittered 4.2 x 4.3 x 4.3 mm lattice.  Seeded jitter; the initial
    separation of every pair of codimensional nodes / segments exceeds xi + d_hat = 2 mm
    (checked by cipc.scenes.min_codim_separation)."""
    base = scene_c1(seed=1)
    rng = np.random.default_rng(seed)
    lines = n_rods // 2
    m = int(math.ceil(math.sqrt(lines)))
    pitch = 0.3 / m
    rods = []
    for i in range(lines):
        iy, iz = divmod(i, m)
        y = 0.55 + (iy + 0.5) * pitch
        z = -0.15 + (iz + 0.5) * pitch
        for x0 in (-0.145, 0.045):
            j = rng.uniform(-2e-4, 2e-4, 3)
            rods.append(make_rod((x0, y, z), (x0 + 0.1, y, z), n_seg) + j)
    nx, ny, nz = 20, 71, 71
    gx = np.linspace(-0.04, 0.04, nx)
    gy = 0.55 + (np.arange(ny) + 0.5) * (0.3 / ny)
    gz = -0.15 + (np.arange(nz) + 0.5) * (0.3 / nz)
    g = np.stack(np.meshgrid(gx, gy, gz, indexing="ij"), -1).reshape(-1, 3)[:n_particles]
    parts = g + rng.uniform(-4e-4, 4e-4, g.shape)
    return with_codim(base, rods, parts, name="c3_codim_mix")


def min_codim_separation(sd):
    """Smallest distance between a particle and any other vertex and between nodes of
    different rods (a lower bound check of the initial separation)."""
    from scipy.spatial import cKDTree
    n0 = sd.nV - (0 if sd.particles is None else len(sd.particles))
    tree = cKDTree(sd.X)
    best = np.inf
    if sd.particles is not None:
        d, _ = tree.query(sd.X[sd.particles], k=2)
        best = min(best, float(d[:, 1].min()))
    if sd.rods is not None:
        rod_of = -np.ones(sd.nV, dtype=np.int64)
        rid = np.cumsum(np.r_[1, (sd.rods[1:, 0] != sd.rods[:-1, 1]).astype(int)]) - 1
        rod_of[sd.rods[:, 0]] = rid
        rod_of[sd.rods[:, 1]] = rid
        nodes = np.nonzero(rod_of >= 0)[0]
        t2 = cKDTree(sd.X[nodes])
        pairs = t2.query_pairs(0.0025, output_type="ndarray")
        if len(pairs):
            other = rod_of[nodes[pairs[:, 0]]] != rod_of[nodes[pairs[:, 1]]]
            if other.any():
                dd = np.linalg.norm(sd.X[nodes[pairs[other, 0]]] - sd.X[nodes[pairs[other, 1]]], axis=1)
                best = min(best, float(dd.min()))
    return best


def scene_c5(seed=5, count=256, tri_min=1e3, tri_max=1e5):
    """`count` independent scenes: an N x N sheet (1 m square) whose triangle count 2 (N-1)^2
    is drawn log-uniform in [tri_min, tri_max], its own icosphere (radius uniform in
    [0.1, 0.3] m, level 3) and ground, dropped from a height uniform in [0.2, 0.8] m
    above the sphere top.  Returns a list of SceneData."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        tris = 10 ** rng.uniform(np.log10(tri_min), np.log10(tri_max))
        n = max(3, int(round(math.sqrt(tris / 2.0))) + 1)
        r = float(rng.uniform(0.1, 0.3))
        y0 = r + float(rng.uniform(0.2, 0.8))
        sd = cloth_on_sphere(sheets=1, n=n, y0=y0, seed=10000 + k, sphere_level=3, sphere_radius=r, name="c5_%03d" % k)
        out.append(sd)
    return out


# ---------------------------------------------------------------- volumetric tets
def tet_box(lo, hi, n):
    """A box [lo, hi] cut into n^3 cubes of six tets each (the Kuhn subdivision along the
    main diagonal: conforming across cubes), every tet positively oriented.
    Returns (X (m, 3), tets (6 n^3, 4))."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    g = np.arange(n + 1)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    X = lo + (hi - lo) * np.stack([I, J, K], -1).reshape(-1, 3) / n
    vid = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for pm in perms:
                    c = [i, j, k]
                    t = [vid(*c)]
                    for ax in pm:
                        c[ax] += 1
                        t.append(vid(*c))
                    a, b, cc, d = (X[v] for v in t)
                    if np.linalg.det(np.stack([b - a, cc - a, d - a], 1)) < 0:
                        t[1], t[2] = t[2], t[1]
                    tets.append(t)
    return X, np.array(tets, dtype=np.int32)


TET_PARAMS = dict(tetRho=1000.0, tetE=1e5, tetNu=0.4)


def with_tets(sd, X, tets, name=None, tet_params=None):
    """Append a tet mesh (add_object_3D + initialize_added_objects) to a scene."""
    out = SceneData(np.concatenate([sd.X, X]), sd.tris, sd.dbc_idx, sd.dbc_target, sd.params, name or sd.name)
    out.n_cloth = getattr(sd, "n_cloth", 0)
    out.rods, out.particles = sd.rods, sd.particles
    out.tets = np.asarray(tets, dtype=np.int32) + sd.nV
    out.params.update(TET_PARAMS if tet_params is None else tet_params)
    out.params["PNTol"] = 1e-2 * out.bbox_diag()
    return out


def scene_tet_cloth(n=12, cells=4, gap=8e-4, seed=21):
    """A soft tet cube (0.2 m, 6 cells^3 tets, rho 1000, E 1e5, nu 0.4) resting on the ground
    (gap above it) with an n x n sheet (0.4 m) gap above the cube's top: cloth-tet contact,
    tet-ground contact and the sheet draping over the cube's edges."""
    X, T = cloth_grid(n, 0.0)
    X = 0.4 * X
    X[:, 1] = 0.2 + 2 * gap
    u = splitmix64_uniform(seed, X.size).reshape(X.shape)
    X = X + (2.0 * u - 1.0) * np.array([1e-4, 0.2 * gap, 1e-4])
    n_cloth = X.shape[0]
    Xg, Tg = ground(1.0)
    Xa = np.concatenate([X, Xg])
    Ta = np.concatenate([T, Tg + n_cloth])
    dbc = np.arange(n_cloth, Xa.shape[0], dtype=np.int32)
    base = SceneData(Xa, Ta, dbc, Xa[dbc].copy(), dict(DEFAULT_PARAMS), "tet_cloth")
    base.n_cloth = n_cloth
    Xt, Tt = tet_box((-0.1, gap, -0.1), (0.1, 0.2 + gap, 0.1), cells)
    return with_tets(base, Xt, Tt, name="tet_cloth")
