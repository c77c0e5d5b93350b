"""CPU checks of the boundary and host logic: libcipc.so loads and exports every symbol
include/cipc.h declares; the scene generators reproduce the workload sizes; the seeded
jitter stream is stable."""
import ctypes
import os
import re

import numpy as np

import cipc
from cipc import _lib, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "cipc.h")).read()
    return sorted(set(re.findall(r"\b(cipc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header():
    assert os.path.exists(_lib.LIB_PATH), "libcipc.so not built"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.PROTOTYPES), "python prototypes out of sync with include/cipc.h"


def test_invalid_arguments_rejected_without_gpu():
    lib = cipc.load()
    # a NULL description and a bad triangle index are rejected before any device call
    h = ctypes.c_void_p()
    assert lib.cipc_create(None, ctypes.byref(h)) == -1
    sd = scenes.scene_small(n=3, sheets=1, sphere_level=0)
    desc = _lib.SceneDesc()
    desc.n_vertices = sd.nV
    desc.x0 = sd.X.ctypes.data_as(_lib.D)
    bad = sd.tris.copy()
    bad[0, 0] = sd.nV + 5
    desc.n_triangles = sd.nT
    desc.triangles = bad.ctypes.data_as(_lib.I32)
    assert lib.cipc_create(ctypes.byref(desc), ctypes.byref(h)) == -1


def test_scene_sizes():
    c1 = scenes.scene_c1()
    assert (c1.nV, c1.nT, c1.n_cloth) == (4247, 8322, 1681)
    c2 = scenes.scene_c2()
    assert (c2.nV, c2.nT, c2.n_cloth) == (164170, 325122, 161604)
    ys = c2.X[: c2.n_cloth, 1].reshape(4, -1)
    assert np.allclose(ys.mean(axis=1), [0.5, 0.505, 0.51, 0.515], atol=1e-5)
    assert np.abs(c1.X[: c1.n_cloth] - scenes.cloth_grid(41)[0]).max() <= 1e-4


def test_splitmix64_known_values():
    # splitmix64 reference outputs for seed 0 (first three 64-bit words >> 11, scaled)
    u = scenes.splitmix64_uniform(0, 3)
    words = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert np.array_equal(u, np.array([(w >> 11) * 2.0 ** -53 for w in words]))


def test_feature_arguments_rejected_without_gpu():
    """friction, strain-limit, fibre and Dirichlet entry points validate before any device
    call: a NULL context, and scene descriptions with inconsistent limits"""
    lib = cipc.load()
    d3 = (ctypes.c_double * 3)(0, 0, 0)
    d2 = (ctypes.c_double * 2)(1e3, 0)
    assert lib.cipc_set_friction(None, 0.1, 1e-6, 1) == -1
    assert lib.cipc_set_strain_limit(None, d2, d2, d2) == -1
    assert lib.cipc_set_fiber(None, d3, d3) == -1
    assert lib.cipc_dirichlet_add_motion(None, 0, 1, d3, d3, d3, 0.0) == -1
    assert lib.cipc_step_dirichlet(None) == -1
    assert lib.cipc_set_friction_components(None, 0, None, None) == -1
    sd = scenes.scene_small(n=3, sheets=1, sphere_level=0)
    h = ctypes.c_void_p()
    for field, value in (("friction_mu", -0.1), ("strain_limit", ((1e3, 0.0), (1.0, 0.0), (1.01, 1.0))),
                         ("fiber", ((1e3, 0.0, 0.0, 0.0), (0.0, 0.1, 0.1)))):
        desc = _lib.SceneDesc()
        desc.n_vertices = sd.nV
        desc.x0 = sd.X.ctypes.data_as(_lib.D)
        desc.n_triangles = sd.nT
        desc.triangles = sd.tris.ctypes.data_as(_lib.I32)
        if field == "friction_mu":
            desc.friction_mu = value                  # negative coefficient
        elif field == "strain_limit":
            for i in range(2):                        # limit s below sHat
                desc.strain_limit_kappa[i], desc.strain_limit_s[i], desc.strain_limit_shat[i] = value[0][i], value[1][i], value[2][i]
        else:
            for i in range(4):
                desc.fiber_stiff_mult[i] = value[0][i]
            for i in range(3):                        # a zero limit with the model on
                desc.fiber_limit[i] = value[1][i]
        assert lib.cipc_create(ctypes.byref(desc), ctypes.byref(h)) == -1, field


def test_friction_component_table_size_checked():
    """set_friction_components rejects a coefficient table that is not n_components^2
    before anything reaches the library"""
    import pytest

    class NoLib:
        h = None
        lib = None
    with pytest.raises(ValueError):
        cipc.Simulator.set_friction_components(NoLib(), [10, 20], [0.1, 0.2, 0.3])
