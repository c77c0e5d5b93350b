"""Learned preconditioner on the GPU vs its CPU reference (tests/learned_ref.py).

Bars: Kaporin loss within 1e-4 (relative + absolute 1e-4) and every parameter tensor's
gradient within 1e-3 relative (Frobenius) of the float64 CPU reference on the same
BF16-rounded weights and the same probe vectors (the GPU keeps activations in FP32);
PCG with the learned preconditioner meets the FP64 residual bar of the oracle's own
assembled matrix, and after training it needs fewer iterations than block Jacobi.
"""
import numpy as np
import pytest

import learned_ref as R
import oracle as O
from cipc import Simulator, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(require_gpu):
    sd = scenes.scene_contact_stack()
    gpu = Simulator(sd)
    orc = O.OracleScene(sd, threads=4)
    states = [orc.get_state()[0].copy()]
    for _ in range(3):
        rc, _ = orc.step()
        assert rc > 0
        states.append(orc.get_state()[0].copy())
    return sd, gpu, orc, states


def random_weights(gpu, seed, scale=0.02):
    """library initialisation, then the decoder moved off zero so every layer gets a gradient"""
    gpu.precond_init(seed)
    w = gpu.precond_weights()
    rng = np.random.default_rng(seed)
    L, _ = R.layout()
    for name in ("wo", "bo", "wd", "bd"):
        o, n = L[name]
        w[o:o + n] = (rng.standard_normal(n) * scale).astype(np.float32)
    gpu.precond_set_weights(w)
    return w


@pytest.mark.parametrize("state", [0, 2])
def test_loss_and_gradient_match_cpu(setup, state):
    sd, gpu, orc, states = setup
    x = states[state]
    C = orc.constraint_set(x)
    assert len(C) > 0
    kappa = orc.scalars()["kappa"]
    w = random_weights(gpu, 3 + state)
    loss_g, grad_g = gpu.precond_loss_grad(x, C, kappa, n_probes=4, seed=11)
    H = orc.hessian(x, C, kappa, which=3)
    a = orc.arrays()
    dbc = np.zeros(sd.nV, bool)
    dbc[sd.dbc_idx] = True
    loss_c, grad_c = R.loss_and_grad(H, a["mass"], dbc, a["tri"], a["hinge"], w, 4, 11, R.contact_row_counts(sd.nV, C, dbc))
    assert abs(loss_g - loss_c) <= 1e-4 * (1 + abs(loss_c))
    L, _ = R.layout()
    for name, (o, n) in L.items():
        gc, gg = grad_c[o:o + n], grad_g[o:o + n].astype(np.float64)
        assert np.linalg.norm(gg - gc) <= 1e-3 * np.linalg.norm(gc) + 1e-9, name


def test_mfma_forward_matches_fp32_forward(setup):
    """inference form (BF16 activations into the matrix cores, FP32 accumulation) against
    the FP32 training form of the same network: BF16 keeps 8 mantissa bits, so the factor
    agrees to a few 1e-3 relative (bar 2e-2)"""
    sd, gpu, orc, states = setup
    x = states[2]
    C = orc.constraint_set(x)
    kappa = orc.scalars()["kappa"]
    random_weights(gpu, 9, scale=0.1)
    g32 = gpu.precond_factor(x, C, kappa, mfma=False)
    g16 = gpu.precond_factor(x, C, kappa, mfma=True)
    assert np.isfinite(g16).all()
    assert np.array_equal(g16 == 0, g32 == 0)
    err = np.linalg.norm(g16 - g32) / np.linalg.norm(g32 - np.round(g32 * 0))
    assert err <= 2e-2, err
    # the deviation from identity is what the network contributes; compare it too
    dev32 = g32 - (np.abs(g32 - 1) < 0.3) * 1.0
    assert np.linalg.norm(g16 - g32) <= 5e-2 * np.linalg.norm(dev32)


def test_untrained_learned_pcg_is_point_jacobi_and_converges(setup):
    sd, gpu, orc, states = setup
    x = states[1]
    C = orc.constraint_set(x)
    kappa = orc.scalars()["kappa"]
    H = orc.hessian(x, C, kappa, which=3)
    b = np.random.default_rng(2).standard_normal(x.size)
    b.reshape(-1, 3)[sd.dbc_idx] = 0
    gpu.precond_init(1)          # decoder at zero: G = I
    gpu.set_preconditioner("learned", budget=20000)
    try:
        sol, it, rr = gpu.linear_solve(x, C, kappa, b, tol=1e-8)
        assert gpu.precond_info()["fallback"] == 0
    finally:
        gpu.set_preconditioner("jacobi")
    assert rr <= 1e-8
    assert np.linalg.norm(H @ sol - b) <= 2e-8 * np.linalg.norm(b)


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_learned_refinement_forms(setup, precision):
    """the learned preconditioner inside the FP64 refinement of the reduced-precision
    forms meets the FP64 residual bar of the oracle's matrix"""
    sd, gpu, orc, states = setup
    x = states[1]
    C = orc.constraint_set(x)
    kappa = orc.scalars()["kappa"]
    H = orc.hessian(x, C, kappa, which=3)
    b = np.random.default_rng(6).standard_normal(x.size)
    b.reshape(-1, 3)[sd.dbc_idx] = 0
    random_weights(gpu, 13, scale=1e-3)   # near point Jacobi: the test is about the plumbing
    gpu.set_precision(precision)
    gpu.set_preconditioner("learned", budget=20000)
    try:
        sol, it, rr = gpu.linear_solve(x, C, kappa, b, tol=1e-8)
        info = gpu.precond_info()
        print(precision, "iterations", it, "relres", rr, info, gpu.solver_info())
        assert info["fallback"] == 0, info
    finally:
        gpu.set_preconditioner("jacobi")
        gpu.set_precision("fp64")
    assert rr <= 1e-8
    assert np.linalg.norm(H @ sol - b) <= 2e-8 * np.linalg.norm(b)


def test_training_cuts_pcg_iterations(setup):
    """train on three states, evaluate on a fourth: fewer PCG iterations than block Jacobi"""
    sd, gpu, orc, states = setup
    kappa = orc.scalars()["kappa"]
    sets = [orc.constraint_set(x) for x in states]
    gpu.precond_init(7)
    losses = []
    for it in range(240):
        k = it % 3
        losses.append(gpu.precond_train_step(states[k], sets[k], kappa, lr=2e-3, n_probes=4, seed=100 + it))
    assert np.mean(losses[-30:]) < np.mean(losses[:30])
    x, C = states[3], sets[3]
    H = orc.hessian(x, C, kappa, which=3)
    b = np.random.default_rng(4).standard_normal(x.size)
    b.reshape(-1, 3)[sd.dbc_idx] = 0
    gpu.set_preconditioner("jacobi")
    _, it_j, _ = gpu.linear_solve(x, C, kappa, b, tol=1e-6)
    gpu.set_preconditioner("learned", budget=20000)
    try:
        sol, it_l, rr = gpu.linear_solve(x, C, kappa, b, tol=1e-6)
        assert gpu.precond_info()["fallback"] == 0
    finally:
        gpu.set_preconditioner("jacobi")
    assert rr <= 1e-6
    assert np.linalg.norm(H @ sol - b) <= 2e-6 * np.linalg.norm(b)
    print("block Jacobi %d iterations, learned %d" % (it_j, it_l))
    assert it_l < 0.8 * it_j


def test_learned_fallback_continues_from_learned_iterate(setup):
    """a budget too small for the (untrained, point-Jacobi) learned solve: block Jacobi
    takes over from the learned iterate and the FP64 residual bar still holds"""
    sd, gpu, orc, states = setup
    x = states[1]
    C = orc.constraint_set(x)
    kappa = orc.scalars()["kappa"]
    H = orc.hessian(x, C, kappa, which=3)
    b = np.random.default_rng(4).standard_normal(x.size)
    b.reshape(-1, 3)[sd.dbc_idx] = 0
    gpu.precond_init(1)
    gpu.set_preconditioner("learned", budget=30)
    try:
        sol, it, rr = gpu.linear_solve(x, C, kappa, b, tol=1e-8)
        info = gpu.precond_info()
        assert info["fallback"] == 1
        assert info["continued"] == 1          # block Jacobi went on from the learned iterate
    finally:
        gpu.set_preconditioner("jacobi")
    assert it > 30
    assert rr <= 1e-8
    assert np.linalg.norm(H @ sol - b) <= 2e-8 * np.linalg.norm(b)
