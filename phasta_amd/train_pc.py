"""Harvest training pairs with the plain solver and train the learned
preconditioner on the GPU with the library's own forward/backward kernels.

Pairs: for each of `meshes` 64^3 boxes (jitter ~ U[0, 0.15] h, seeded) and
`pairs // meshes` states on it (Mach ~ U[0.1, 0.8], seeded field), the FP64
solver runs a predictor and Newton iterations with block-diagonal GMRES
(Kspace 40, tolerance 1e-4) and keeps, per Newton iteration, the scaled
right-hand side b = L^-1 res, the BDiag before factorisation and the scaled
solution z.  The network first learns z from (BDiag, b); then it is tuned on
the residual ||K z(v) - v||^2 / ||v||^2 of the Krylov vectors v of fresh
block-diagonal solves, the inputs it meets inside flexible GMRES.

    python -m phasta_amd.train_pc --out phasta_amd/weights/mpnn_box64.npz
    python -m phasta_amd.train_pc --eval phasta_amd/weights/mpnn_box64.npz
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from phasta_amd import config, mesh, solver  # noqa: E402

HARVEST_K = 40
HARVEST_ETOL = 1e-4


def log(msg):
    print(f"[train_pc {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def step_params(N, mach, kspace, grade, etol=1e-3):
    """Solver scalars of a box: CFL 1 on h = 1/N, or CFL 3 on the first wall
    cell of a graded box (the rule of the graded workloads)."""
    return config.solver_params(N, mach, kspace, etol=etol, grade=grade,
                                cfl=3.0 if grade > 1.0 else 1.0)


def harvest(N, meshes, pairs, newton, seed, grade_max=1.0):
    rng = np.random.default_rng(seed)
    per_mesh = pairs // meshes
    ctxs, slots = [], []
    for m in range(meshes):
        jitter = float(rng.uniform(0.0, 0.15))
        grade = float(rng.uniform(1.0, grade_max)) if grade_max > 1.0 else 1.0
        case = mesh.box_case(N, mach=0.5, jitter=jitter, grade=grade, kspace=HARVEST_K,
                             etol=HARVEST_ETOL, mesh_seed=config.MESH_SEED + 1000 + m,
                             cfl=3.0 if grade > 1.0 else 1.0)
        case.grade = grade
        gpu = solver.GpuSolver(case)
        gpu.mpnn_init(np.ones(25, dtype=np.float32))
        if ctxs:
            gpu.mpnn_share(ctxs[0][1])
        gpu.mpnn_dataset(per_mesh)
        gpu.mpnn_enable(use_pc=False, keep_pair=True)
        k = 0
        iters = []
        while k < per_mesh:
            mach = float(rng.uniform(0.1, 0.8))
            fseed = int(rng.integers(1, 2 ** 31))
            gpu.set_params(**step_params(N, mach, HARVEST_K, grade, etol=HARVEST_ETOL))
            gpu.upload_state(mesh.box_state(case, mach, fseed))
            gpu.predict()
            for _ in range(newton):
                if k >= per_mesh:
                    break
                st = gpu.newton_iteration()
                gpu.mpnn_store_pair(k)
                iters.append(st["iKs"])
                slots.append((m, k))
                k += 1
        gpu.sync()
        ctxs.append((case, gpu))
        log(f"mesh {m}: jitter {jitter:.3f}, grade {grade:.3f}, {per_mesh} pairs, Krylov its "
            f"mean {np.mean(iters):.1f} max {max(iters)}")
    return ctxs, slots


def train(ctxs, slots, epochs, lr, seed):
    rng = np.random.default_rng(seed + 1)
    total = epochs * len(slots)
    step = 0
    history = []
    for ep in range(epochs):
        order = rng.permutation(len(slots))
        losses = []
        t0 = time.perf_counter()
        for idx in order:
            m, k = slots[idx]
            lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
            losses.append(ctxs[m][1].mpnn_train_pair(k, lr_t))
            step += 1
        dt = time.perf_counter() - t0
        history.append(float(np.mean(losses)))
        log(f"epoch {ep}: mean loss {history[-1]:.5f} ({dt:.1f} s, "
            f"{1e3 * dt / len(slots):.2f} ms per pair)")
    return history


def train_residual(ctxs, N, rounds, states, inner, lr, seed, newton=1):
    """Fine-tuning on the residual loss ||K z(v) - v||^2 / ||v||^2 over the
    Krylov vectors of fresh block-diagonal solves (the directions flexible
    GMRES hands to the network), K the scaled matrix of each solve; the
    first `newton` Newton iterations of every state are used."""
    rng = np.random.default_rng(seed + 3)
    hist = []
    for r in range(rounds):
        losses = []
        t0 = time.perf_counter()
        for m, (case, gpu) in enumerate(ctxs):
            gpu.mpnn_enable(use_pc=0, keep_pair=True)
            for _ in range(states):
                mach = float(rng.uniform(0.1, 0.8))
                fseed = int(rng.integers(1, 2 ** 31))
                gpu.set_params(**step_params(N, mach, 20, case.grade))
                gpu.upload_state(mesh.box_state(case, mach, fseed))
                gpu.predict()
                lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * r / max(1, rounds)))
                for _ in range(newton):
                    st = gpu.newton_iteration()
                    for _ in range(inner):
                        losses.append(gpu.mpnn_train_krylov(st["iKs"] + 1, lr_t))
        hist.append(float(np.mean(losses)))
        log(f"residual round {r}: mean loss {hist[-1]:.5f} ({time.perf_counter() - t0:.1f} s)")
    return hist


def evaluate(N, weights, seed, states=4, grade=1.0, newton=3):
    """Plain block-diagonal GMRES vs FGMRES with the network on the first
    direction of every cycle, on a mesh and states not used in training
    (Kspace 20, tolerance 1e-3, `newton` Newton iterations per state)."""
    rng = np.random.default_rng(seed + 7)
    case = mesh.box_case(N, mach=0.5, jitter=0.1, grade=grade, kspace=20,
                         mesh_seed=config.MESH_SEED + 9999, cfl=3.0 if grade > 1.0 else 1.0)
    gpu = solver.GpuSolver(case)
    gpu.mpnn_init(np.ones(25, dtype=np.float32))
    gpu.mpnn_set_weights(weights)
    zero = np.zeros_like(weights)
    rows = []
    for _ in range(states):
        mach = float(rng.uniform(0.1, 0.8))
        fseed = int(rng.integers(1, 2 ** 31))
        gpu.set_params(**step_params(N, mach, 20, grade))
        res = {}
        for arm, dirs in (("plain", 0), ("first", 1)):
            gpu.mpnn_enable(use_pc=dirs, keep_pair=(arm == "plain"))
            gpu.upload_state(mesh.box_state(case, mach, fseed))
            gpu.predict()
            its, fb, ms, conv = [], [], 0.0, True
            for k in range(newton):
                st = gpu.newton_iteration()
                its.append(st["iKs"])
                fb.append(st["pc_fallback"])
                ms += st["ms_solve"]
                conv = conv and bool(st["converged"])
                if arm == "plain" and k == 0:
                    # residual loss on this solve's Krylov vectors: network vs z = v
                    loss_net = gpu.mpnn_train_krylov(st["iKs"] + 1, 0.0)
                    gpu.mpnn_set_weights(zero)
                    loss_id = gpu.mpnn_train_krylov(st["iKs"] + 1, 0.0)
                    gpu.mpnn_set_weights(weights)
            res[arm] = (its, fb, ms, conv)
        rows.append((mach, res, loss_net, loss_id))
        log(f"eval mach {mach:.2f}: residual loss {loss_net:.4g} (z = v: {loss_id:.4g}); " +
            ", ".join(f"{a} its {r[0]} fallback {r[1]} ({r[2]:.2f} ms solve, converged {r[3]})"
                      for a, r in res.items()))
    gpu.close()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=64)
    ap.add_argument("--meshes", type=int, default=8)
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--newton", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--mode", choices=["supervised", "residual", "both"], default="both",
                    help="supervised on (b, z) pairs, residual loss on Krylov vectors, "
                         "or the first then the second")
    ap.add_argument("--rounds", type=int, default=6)
    ap.add_argument("--states", type=int, default=8)
    ap.add_argument("--inner", type=int, default=2)
    ap.add_argument("--res-newton", type=int, default=1,
                    help="Newton iterations per state in the residual phase")
    ap.add_argument("--res-lr", type=float, default=2e-5,
                    help="Adam step of the residual-loss phase (the loss is stiff: "
                         "K amplifies the network's output error)")
    ap.add_argument("--grade-max", type=float, default=1.0,
                    help="y-grading ratio of the training meshes ~ U[1, grade-max]")
    ap.add_argument("--eval-N", type=int, default=None, help="box size of the evaluation")
    ap.add_argument("--eval-grade", type=float, default=1.0, help="y-grading of the evaluation box")
    ap.add_argument("--out", help="weights file to write")
    ap.add_argument("--eval", help="only evaluate the weights in this file")
    args = ap.parse_args()
    eval_n = args.eval_N or args.N
    if args.eval:
        evaluate(eval_n, np.load(args.eval)["weights"], args.seed, grade=args.eval_grade)
        return
    if not args.out:
        ap.error("--out is required for training")
    t0 = time.perf_counter()
    ctxs, slots = harvest(args.N, args.meshes, args.pairs, args.newton, args.seed,
                          args.grade_max)
    log(f"harvested {len(slots)} pairs in {time.perf_counter() - t0:.1f} s")
    owner = ctxs[0][1]
    owner.mpnn_set_weights(solver.mpnn_flatten(init_params(args.seed)))
    hist = []
    if args.mode in ("supervised", "both"):
        hist = train(ctxs, slots, args.epochs, args.lr, args.seed)
    rhist = []
    if args.mode in ("residual", "both"):
        rhist = train_residual(ctxs, args.N, args.rounds, args.states, args.inner, args.res_lr,
                               args.seed, args.res_newton)
    w = owner.mpnn_get_weights()
    meta = dict(N=args.N, meshes=args.meshes, pairs=len(slots), newton=args.newton,
                epochs=args.epochs, lr=args.lr, seed=args.seed, loss=hist,
                mode=args.mode, residual_loss=rhist, rounds=args.rounds, states=args.states,
                inner=args.inner, res_lr=args.res_lr, res_newton=args.res_newton,
                grade_max=args.grade_max,
                harvest_kspace=HARVEST_K, harvest_etol=HARVEST_ETOL,
                seconds=time.perf_counter() - t0)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.savez(args.out, weights=w, dscale=np.ones(25, dtype=np.float32),
             meta=json.dumps(meta))
    log(f"wrote {args.out}: {json.dumps(meta)}")
    evaluate(eval_n, w, args.seed, grade=args.eval_grade)
    # shared parameters belong to ctxs[0]: release the others first
    for _, g in ctxs[1:]:
        g.close()
    owner.close()


def init_params(seed):
    """Initial weights (same scheme as the CPU oracle's init_params)."""
    rng = np.random.default_rng(seed)
    p = {"W0": rng.standard_normal((32, 64)) * np.sqrt(2.0 / 32), "b0": np.zeros(64)}
    for l in (1, 2, 3):
        p[f"W{l}"] = rng.standard_normal((320, 64)) * (0.5 / np.sqrt(320))
        p[f"b{l}"] = np.zeros(64)
    p["Wout"] = np.zeros((64, 5))
    p["bout"] = np.zeros(5)
    return p


if __name__ == "__main__":
    main()
