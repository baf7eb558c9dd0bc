"""Learned preconditioner kernels (k_mpnn.hip) against the CPU oracle of the
network (oracle/pympnn.py), and the flexible-GMRES plumbing.

Inputs: the committed network on what the solver gives it, the block
diagonal before factorisation, the preconditioned residual and the scaled
solution of a GMRES solve on the 8^3 graded box (features of order one, no
saturated activations).  Bars: the forward correction z - v and the
gradients agree with the oracle's bf16 model to 1e-2 normwise (bf16
roundings of intermediate values can land on different sides of a tie when
float sums are taken in another order; the oracle's own gradients are
pinned by finite differences in float64).  A zero network makes FGMRES
identical to block-diagonal GMRES.
"""
import os

import numpy as np
import pytest

import pympnn as M
import pyoracle
from phasta_amd import mesh, solver

pytestmark = pytest.mark.gpu
TOL_NET = 1e-2
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "phasta_amd", "weights", "mpnn_box64.npz")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def net_case():
    """The committed network and the pair of a block-diagonal solve."""
    case = mesh.box_case(8, mach=0.3, kspace=10, grade=1.1)
    print(f"[mpnn] box8 graded: nshg={case.nshg}", flush=True)
    g = M.periodic_graph(case)
    w = np.load(WEIGHTS)
    p = solver.mpnn_unflatten(w["weights"])
    dscale = np.asarray(w["dscale"], dtype=np.float32)
    orc = pyoracle.Oracle(case)
    y, ac = orc.itr_predict_bc(case.yold)
    gpu = solver.GpuSolver(case)
    gpu.mpnn_init(dscale)
    gpu.mpnn_set_weights(w["weights"])
    gpu.mpnn_enable(use_pc=0, keep_pair=True)
    gpu.solgmrs(y, ac)
    bd, v, t = gpu.mpnn_last_pair()
    gpu.mpnn_enable(use_pc=0, keep_pair=False)
    yield case, g, p, bd, v, t, dscale, gpu
    gpu.close()


@pytest.fixture(scope="module")
def random_case():
    """A fresh network with random biases and random inputs (the optimiser
    test starts from a loss well above its minimum)."""
    case = mesh.box_case(4, mach=0.3, kspace=10)
    print(f"[mpnn] box4: nshg={case.nshg}", flush=True)
    g = M.periodic_graph(case)
    rng = np.random.default_rng(11)
    p = M.init_params(5)
    for k in p:
        if k.startswith("b"):
            p[k] = (0.1 * rng.standard_normal(p[k].shape)).astype(np.float32)
    n = case.nshg
    bd = np.asfortranarray(rng.standard_normal((n, 5, 5)))
    v = np.asfortranarray(rng.standard_normal((n, 5)))
    t = np.asfortranarray(v + 0.3 * rng.standard_normal((n, 5)))
    dscale = np.full(25, 2.0, dtype=np.float32)
    gpu = solver.GpuSolver(case)
    gpu.mpnn_init(dscale)
    gpu.mpnn_set_weights(solver.mpnn_flatten(p))
    yield case, g, p, bd, v, t, dscale, gpu
    gpu.close()


def test_graph_and_forward_match_oracle(net_case):
    case, g, p, bd, v, t, dscale, gpu = net_case
    z = gpu.mpnn_apply(bd, v)
    x, s = M.features(bd, v, dscale, g.active)
    y, _ = M.forward(p, g, x)
    z_ref = v + y / s
    z_ref[~g.active] = v[~g.active]
    e = rel(z - v, z_ref - v)
    print(f"[mpnn] forward rel err {e:.2e}", flush=True)
    assert e < TOL_NET
    assert np.array_equal(z[~g.active], v[~g.active])


def test_gradients_match_oracle(net_case):
    case, g, p, bd, v, t, dscale, gpu = net_case
    loss, grad = gpu.mpnn_train_step(bd, v, t, lr=0.0, want_grad=True)
    x, s = M.features(bd, v, dscale, g.active)
    ystar = ((t - v) * s).astype(np.float32)
    loss_ref, gr, _ = M.loss_and_grad(p, g, x, ystar)
    gg = solver.mpnn_unflatten(grad)
    print(f"[mpnn] loss gpu {loss:.6e} oracle {loss_ref:.6e}", flush=True)
    assert abs(loss - loss_ref) <= TOL_NET * loss_ref
    for k, _ in solver.MPNN_SHAPES:
        e = rel(gg[k], gr[k])
        print(f"[mpnn] grad {k} rel err {e:.2e}", flush=True)
        assert e < TOL_NET, k


def test_adam_steps_reduce_the_loss(random_case):
    case, g, p, bd, v, t, dscale, gpu = random_case
    losses = [gpu.mpnn_train_step(bd, v, t, lr=2e-3) for _ in range(30)]
    print(f"[mpnn] loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert losses[-1] < 0.7 * losses[0]
    gpu.mpnn_set_weights(solver.mpnn_flatten(p))  # leave the fixture as it was


def test_zero_network_fgmres_equals_block_diagonal_gmres():
    case = mesh.box_case(4, mach=0.3, kspace=10)
    orc = pyoracle.Oracle(case)
    y, ac = orc.itr_predict_bc(case.yold)
    ref = orc.solgmrs(y, ac)
    gpu = solver.GpuSolver(case)
    try:
        gpu.mpnn_init(np.ones(25, dtype=np.float32))
        gpu.mpnn_set_weights(np.zeros(solver.lib().phg_mpnn_param_count(), np.float32))
        gpu.mpnn_enable(use_pc=10, keep_pair=True)  # every direction of the cycle
        out = gpu.solgmrs(y, ac)
        bd, b, z = gpu.mpnn_last_pair()
    finally:
        gpu.close()
    assert out["iKs"] == ref["iKs"]
    assert out["pc_fallback"] == 0
    assert rel(out["Dy"], ref["Dy"]) < 1e-12
    # the kept pair: b is the preconditioned residual, z solves the scaled system
    assert rel(b, ref["res"]) < 1e-12
    assert np.linalg.norm(z) > 0



def test_bad_network_direction_is_rejected():
    """Block-diagonal GMRES is the yardstick of the network's first
    direction: a network whose direction leaves a larger one-step residual
    than the block-diagonal one is set aside for the cycle (pc_fallback 3),
    and the solve is then the block-diagonal solve of the oracle."""
    case = mesh.box_case(4, mach=0.3, kspace=10)
    orc = pyoracle.Oracle(case)
    y, ac = orc.itr_predict_bc(case.yold)
    ref = orc.solgmrs(y, ac)
    p = M.init_params(9)
    p["Wout"] = (10.0 * np.random.default_rng(4).standard_normal(p["Wout"].shape)).astype(np.float32)
    gpu = solver.GpuSolver(case)
    try:
        gpu.mpnn_init(np.ones(25, dtype=np.float32))
        gpu.mpnn_set_weights(solver.mpnn_flatten(p))
        gpu.mpnn_enable(use_pc=1)
        out = gpu.solgmrs(y, ac)
    finally:
        gpu.close()
    assert out["pc_fallback"] == 3
    assert out["iKs"] == ref["iKs"]
    assert rel(out["Dy"], ref["Dy"]) < 1e-12

def test_rejected_network_is_set_aside_for_the_time_step():
    """Resident time stepping: a network that loses the acceptance test in
    a time step's first Newton iteration (pc_fallback 3) is not run again
    in that step (4) and is tried again after the next predictor; the
    steps are the oracle's block-diagonal ones."""
    case = mesh.box_case(4, mach=0.3, kspace=10)
    orc = pyoracle.Oracle(case)
    acold = np.zeros((case.nshg, 5), order="F")
    yo_ref, aco_ref, hist = orc.time_steps(case.yold, acold, steps=2, newton=2)
    p = M.init_params(9)
    p["Wout"] = (10.0 * np.random.default_rng(4).standard_normal(p["Wout"].shape)).astype(np.float32)
    gpu = solver.GpuSolver(case)
    try:
        gpu.mpnn_init(np.ones(25, dtype=np.float32))
        gpu.mpnn_set_weights(solver.mpnn_flatten(p))
        gpu.mpnn_enable(use_pc=1)
        gpu.upload_state(case.yold, acold)
        its = []
        for _ in range(2):
            gpu.predict()
            its += [gpu.newton_iteration() for _ in range(2)]
            gpu.update()
        yo, _ = gpu.download_old()
    finally:
        gpu.close()
    assert [s["pc_fallback"] for s in its] == [3, 4, 3, 4]
    assert [s["iKs"] for s in its] == [h["iKs"] for h in hist]
    assert rel(yo, yo_ref) < 1e-12


def test_residual_loss_matches_host_and_training_reduces_it():
    """The residual loss ||K z(v) - v||^2 / ||v||^2 of the first Krylov vector
    v = b / ||b||, formed on the device, equals the same quantity from the
    host-array entry points; its gradient matches the oracle's backward pass;
    Adam steps on it reduce it."""
    case = mesh.box_case(8, mach=0.3, kspace=10, grade=1.1)
    orc = pyoracle.Oracle(case)
    y, ac = orc.itr_predict_bc(case.yold)
    w = np.load(WEIGHTS)
    p = solver.mpnn_unflatten(w["weights"])
    dscale = np.asarray(w["dscale"], dtype=np.float32)
    gpu = solver.GpuSolver(case)
    try:
        gpu.mpnn_init(dscale)
        gpu.mpnn_set_weights(w["weights"])
        gpu.mpnn_enable(use_pc=0, keep_pair=True)
        out = gpu.solgmrs(y, ac, want_lhs=True)      # lhsK: the scaled matrix K
        bd, b, _ = gpu.mpnn_last_pair()
        # loss and gradient on the solve's first Krylov vector, before the
        # host-array entry points below overwrite the context's Krylov basis
        loss_dev, grad = gpu.mpnn_train_krylov(1, lr=0.0, want_grad=True)
        v = b / np.linalg.norm(b)
        z = gpu.mpnn_apply(bd, v)
        kz = gpu.bc3per(gpu.sparse_ap(out["lhsK"], z))  # SparseAp, then bc3per
        loss_host = np.sum((kz - v) ** 2)
        print(f"[mpnn] residual loss device {loss_dev:.6e} host {loss_host:.6e}", flush=True)
        assert abs(loss_dev - loss_host) <= 1e-6 * loss_host
        # its gradient: the oracle's backward pass from dL/dy = -2 K^T (v - K z) / s
        g = M.periodic_graph(case)
        x, s = M.features(bd, v, dscale, g.active)
        dz = gpu.sparse_ap_t(out["lhsK"], v - kz)      # K^T (v - K z)
        dy = np.where(g.active[:, None], -2.0 * dz / s, 0.0)
        y_o, _ = M.forward(p, g, x)
        na = int(g.active.sum())
        ystar = (y_o - 0.5 * na * dy).astype(np.float32)   # 2 (y - y*) / na = dy
        _, gr, _ = M.loss_and_grad(p, g, x, ystar)
        gg = solver.mpnn_unflatten(grad)
        for k, _ in solver.MPNN_SHAPES:
            e = rel(gg[k], gr[k])
            print(f"[mpnn] residual-loss grad {k} rel err {e:.2e}", flush=True)
            assert e < TOL_NET, k
        # training on the Krylov vectors of this solve, at the step size the
        # committed network was trained with
        gpu.solgmrs(y, ac)
        l0 = gpu.mpnn_train_krylov(out["iKs"] + 1, lr=0.0)
        for _ in range(20):
            gpu.mpnn_train_krylov(out["iKs"] + 1, lr=2e-5)
        l1 = gpu.mpnn_train_krylov(out["iKs"] + 1, lr=0.0)
        print(f"[mpnn] residual training {l0:.4e} -> {l1:.4e}", flush=True)
        assert l1 < l0
    finally:
        gpu.close()


def test_trained_network_against_block_diagonal_gmres():
    """What the committed network (phasta_amd/weights/mpnn_box64.npz) gains
    over the method without it, on a 32^3 mesh seed and states it was not
    trained on: Krylov iterations and solve time of flexible GMRES with the
    network on the first direction of each cycle against block-diagonal
    GMRES, both to the 1e-3 tolerance, and the residual loss ||K z - v||^2
    of the network against z = v on the Krylov vectors of the plain solve.
    The numbers are printed (the gain may be negative: the network trained
    on 64^3 boxes has not beaten block-diagonal GMRES here).  Asserted: both
    solves converge to the same tolerance (the learned one tested on its true
    residual), and what the acceptance test guarantees: a network whose first
    direction is rejected (pc_fallback 3) costs no Krylov iteration over
    block-diagonal GMRES, and an accepted one at most one cycle of Kspace
    directions (the cycle is discarded when it does not reduce the true
    residual)."""
    import time
    from phasta_amd import config
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    w = np.load(os.path.join(root, "phasta_amd", "weights", "mpnn_box64.npz"))
    case = mesh.box_case(32, mach=0.5, jitter=0.1, kspace=20,
                         mesh_seed=config.MESH_SEED + 4242)
    gpu = solver.GpuSolver(case)
    rows = []
    try:
        gpu.mpnn_init(w["dscale"])
        gpu.mpnn_set_weights(w["weights"])
        zero = np.zeros_like(w["weights"])
        for mach, seed in ((0.25, 101), (0.55, 202), (0.75, 303)):
            gpu.set_params(**config.solver_params(32, mach, 20))
            state = mesh.box_state(case, mach, seed)
            out = {}
            for arm, use_pc in (("block", 0), ("learned", 1)):
                gpu.mpnn_enable(use_pc=use_pc, keep_pair=(arm == "block"))
                gpu.upload_state(state)
                gpu.predict()
                gpu.sync()
                t0 = time.perf_counter()
                st = gpu.newton_iteration()
                gpu.sync()
                out[arm] = (st, 1e3 * (time.perf_counter() - t0))
                if arm == "block":
                    loss_net = gpu.mpnn_train_krylov(st["iKs"] + 1, 0.0)
                    gpu.mpnn_set_weights(zero)
                    loss_id = gpu.mpnn_train_krylov(st["iKs"] + 1, 0.0)
                    gpu.mpnn_set_weights(w["weights"])
            (sb, tb), (sl, tl) = out["block"], out["learned"]
            print(f"[ablation] mach {mach}: block-diagonal {sb['iKs']} its {tb:.2f} ms; "
                  f"learned {sl['iKs']} its {tl:.2f} ms (fallback {sl['pc_fallback']}); "
                  f"residual loss network {loss_net:.4g} vs z = v {loss_id:.4g}", flush=True)
            rows.append((sb, sl))
    finally:
        gpu.close()
    for sb, sl in rows:
        assert sb["converged"] and sl["converged"]
        if sl["pc_fallback"] == 3:
            assert sl["iKs"] == sb["iKs"]
        else:
            assert sl["iKs"] <= sb["iKs"] + 20
