#!/usr/bin/env python
"""Benchmark: time per Newton iteration of the compressible hot path.

A step is one Newton iteration on the device-resident state: ElmGMRs
(interior + wall assembly, BCs) + SolGMRs (block LU, scaling, GMRES to the
1e-3 tolerance) + itrCorrect/itrBC, i.e. the unit the reference runs once per
"0 1" pair of its step sequence (itrdrv.f:427-599).  Every `newton` steps the
predictor restarts the time step from the same yold, so each timed step
repeats a real Newton iteration of the workload.

The default workload is the 128^3 y-graded box the metric is quoted on.
The arms run on the same state: plain FP64 block-diagonal GMRES ("fp64"),
the mixed-precision form (float tangent and matrix, double Krylov/MGS/Givens
and refinement residual) with block-diagonal GMRES ("mixed"), and the same
with the learned preconditioner on the first direction of every flexible
GMRES cycle ("learned", weights phasta_amd/weights/mpnn_box64.npz); a
fourth arm runs the matrix-free solver SolMFG in FP64 ("mfg": no lhsK, each
Krylov product one modified-residual pass over the elements).
value = interior tets per second of the learned arm, the form the metric
names (nel / time per Newton iteration); every arm's time, its ratio to
FP64 and its update's error against FP64 are reported beside it.  Inputs
are resident in HBM before the timed region.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("time/Newton iter to 1e-3 tol: FP32+learned PC vs plain FP64; "
          "elements/s")
HBM_PEAK_GBS = 8000.0
# MI355X dense peaks (TFLOP/s): fp32 157.3 (packed / matrix), fp64 78.6
PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}
# e3 at four points, residual + tangent, one element: FP operations with no
# structurally zero operand (oracle/count_flops.cpp, tests/test_flop_count.py)
E3_FLOPS = 38268
WEIGHTS = os.path.join(ROOT, "phasta_amd", "weights", "mpnn_box64.npz")


def spmv_bytes(nnz, nshg, s_a):
    """Algorithmic bytes of one block-CSR matvec: blocks + column indices +
    row pointers + one read and one write of the double 5-vector."""
    return nnz * (25 * s_a + 4) + 4 * (nshg + 1) + 2 * nshg * 5 * 8


def measured_traffic(kernel, case):
    """HBM bytes per launch of `kernel` on `case` from the committed rocprof
    summary (calibrated FETCH_SIZE + WRITE_SIZE), or None."""
    path = os.path.join(ROOT, "profiles", f"{case}_newton_iteration.md")
    if not os.path.exists(path):
        return None
    key = f"traffic {kernel} {case}:"
    for ln in open(path):
        if ln.startswith(key):
            return int(ln[len(key):].strip())
    return None


def host_cpu():
    """CPU model and the cores this process may run on."""
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def cpu_newton(case, threads):
    """One oracle Newton iteration (ElmGMRs + SolGMRs + itrCorrect/itrBC)
    of `case` on `threads` OpenMP threads; returns seconds and the update."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    used = pyoracle.set_threads(threads)
    orc = pyoracle.Oracle(case)
    acold = np.zeros((case.nshg, 5), order="F")
    y, ac = orc.itr_predict_bc(case.yold, acold)
    t0 = time.perf_counter()
    out = orc.solgmrs(y, ac)
    orc.itr_correct_bc(y, ac, case.yold, acold, out["Dy"])
    dt = time.perf_counter() - t0
    return dt, used, out


def run_arm(case, precision, newton, steps, warmup, weights=None, idiff=0, mfg=False):
    """Warm-up, then `steps` timed Newton iterations; returns timings, the
    per-step statistics and the update of the first Newton iteration.
    mfg: matrix-free GMRES (SolMFG, lhs = 0), its interval eGMRES computed by
    itrFDI at the first Newton iteration only (the reference recomputes it
    once every 20 time steps, solmfg.f:163) and held by the context."""
    import numpy as np
    from phasta_amd import solver
    prm = dict(case.params, idiff=idiff)
    if mfg:
        prm.update(lhs=0, iprec=1)
    gpu = solver.GpuSolver(case, params=prm, precision=precision)
    if weights is not None:
        w = np.load(weights)
        gpu.mpnn_init(w["dscale"])
        gpu.mpnn_set_weights(w["weights"])
        gpu.mpnn_enable(use_pc=1)
    acold = np.zeros((case.nshg, 5), order="F")
    gpu.upload_state(case.yold, acold)
    gpu.predict()
    first = gpu.newton_iteration_mfg(egmres=0.0, fdi=True) if mfg else gpu.newton_iteration()
    _, _, dy_first = gpu.download_state()
    gpu.upload_state(case.yold, acold)
    it = [0]

    def step():
        if it[0] % newton == 0:
            gpu.predict()
        st = gpu.newton_iteration_mfg() if mfg else gpu.newton_iteration()
        it[0] += 1
        return st

    for _ in range(warmup):
        step()
    gpu.sync()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(step())
    gpu.sync()
    t1 = time.perf_counter()
    ms_copy, by_copy = gpu.bench_copy(5)
    # the matvec on the arm's resident matrix, 20 launches back to back
    # between two HIP events on the library stream: the kernel's own time,
    # without the launch gaps that the per-launch events inside GMRES see
    # (the host waits for each Hessenberg column before the next launch)
    ms_mv, by_mv = gpu.bench_spmv(20) if not mfg else (float("nan"), float("nan"))
    info = gpu.info()
    gpu.close()
    return dict(ms=(t1 - t0) * 1e3 / steps, stats=stats, first=first, dy=dy_first,
                ms_copy=ms_copy, by_copy=by_copy, ms_mv=ms_mv, by_mv=by_mv, info=info)


def pc_accepted(stats):
    """Fraction of Newton iterations whose GMRES used the network's
    direction (pc_fallback 0: its cycle converged, 1: its cycle was
    discarded, 2: finished by block-diagonal cycles; 3 and 4: the network's
    direction lost the acceptance test or was set aside)."""
    return sum(1 for s in stats if s["pc_fallback"] in (0, 1, 2)) / max(1, len(stats))


def breakdown(arm, steps, nel):
    st = arm["stats"]
    ms_asm = sum(s["ms_assembly"] for s in st) / steps
    return {"ms_per_step": round(arm["ms"], 3),
            "ms_assembly_per_step": round(ms_asm, 3),
            "ms_spmv_per_step": round(sum(s["ms_spmv"] for s in st) / steps, 3),
            "ms_mgs_per_step": round(sum(s["ms_mgs"] for s in st) / steps, 3),
            "ms_point_pass": round(sum(s["ms_point"] for s in st) / steps, 3),
            "ms_block_pass": round(sum(s["ms_blocks"] for s in st) / steps, 3),
            "krylov_iters": [s["iKs"] for s in st],
            "gmres_cycles": [s["refine_cycles"] for s in st],
            "pc_fallback": [s["pc_fallback"] for s in st],
            "converged": [bool(s["converged"]) for s in st],
            "assembly_elements_per_s": round(nel / (ms_asm / 1e3), 1),
            "device_bytes": arm["info"]["dev_bytes"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--case", default="box128")
    ap.add_argument("--arms", default="fp64,mixed,learned,mfg",
                    help="comma-separated subset of fp64, mixed, learned, mfg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--idiff", type=int, default=0, choices=(0, 1, 3),
                    help="global reconstruction of the diffusive flux (0: off, the "
                         "workload's flag set)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; sharding over several "
                          "GPUs is a design, not code", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    from phasta_amd import config, mesh, solver

    spec = dict(config.CASES[args.case])
    case = mesh.named_case(args.case)
    nel = case.nel
    arms = [a for a in ("fp64", "mixed", "learned", "mfg") if a in args.arms.split(",")]
    if not arms:
        ap.error("--arms names none of fp64, mixed, learned, mfg")
    res = {}
    for a in arms:
        prec = {"fp64": solver.PHG_FP64, "mfg": solver.PHG_FP64_MATRIX_FREE}.get(a, solver.PHG_MIXED)
        wts = WEIGHTS if a == "learned" else None
        print(f"[bench] {args.case} arm {a}", file=sys.stderr, flush=True)
        res[a] = run_arm(case, prec, spec["newton"], args.steps, args.warmup, wts, args.idiff,
                         mfg=a == "mfg")
    head_name = "learned" if "learned" in res else ("mixed" if "mixed" in res else "fp64")
    if head_name not in res:
        ap.error("the mfg arm is reported beside one of fp64, mixed, learned")
    head = res[head_name]
    value = nel / (head["ms"] / 1e3)

    # live assembly timing of the headline arm (HIP events on the library
    # stream around the point pass and the block pass): FLOP rate of the
    # frozen e3 count against the chip's peak in the arm's tangent type
    st = head["stats"]
    ms_pt = sum(s["ms_point"] for s in st) / len(st)
    ms_bl = sum(s["ms_blocks"] for s in st) / len(st)
    asm_tf = E3_FLOPS * nel / ((ms_pt + ms_bl) * 1e-3) / 1e12
    tdt = "f64" if head_name == "fp64" else "f32"
    kasm = (f"k_asm_qp<{'double' if tdt == 'f64' else 'float'}, false, false, false> + "
            f"k_asm_pairs<{'double, double' if tdt == 'f64' else 'float, float'}>")
    roof_asm = {"kernel": f"{kasm} (ElmGMRs interior assembly: e3 residual + tangent, "
                          f"res/BDiag/lhsK written once)",
                "bound": "mfma", "compute": f"VALU {tdt} (no MFMA in this kernel; the peak "
                "is the chip's dense fp32/fp64 figure, packed FP32 is off in this build)",
                "achieved": round(asm_tf, 2), "peak": PEAK_TFLOPS[tdt], "unit": "TFLOP/s",
                "frac": round(asm_tf / PEAK_TFLOPS[tdt], 4),
                "traffic": measured_traffic(kasm, args.case),
                "flop_per_element": E3_FLOPS, "elements_per_launch": nel,
                "ms_point_pass": round(ms_pt, 4), "ms_block_pass": round(ms_bl, 4)}
    # matvec of the headline arm: HIP events on the library stream, 20
    # back-to-back launches on the arm's resident matrix after the timed
    # steps (the kernel's time); the average of the per-launch events inside
    # the timed GMRES solves (which include each launch's gap) beside it
    ms_spmv = sum(s["ms_spmv"] for s in st)
    n_spmv = sum(s["n_spmv"] for s in st)
    s_a = 8 if head_name == "fp64" else 4
    by = spmv_bytes(case.nnz_tot, case.nshg, s_a)
    assert abs(by - head["by_mv"]) < 1.0, (by, head["by_mv"])
    spmv_avg = head["ms_mv"]
    spmv_gbs = by / (spmv_avg * 1e-3) / 1e9
    kname = "k_spmv<double, double>" if head_name == "fp64" else "k_spmv<float, double>"
    roof_mv = {"kernel": f"{kname} (block-CSR matvec, SparseAp)", "bound": "hbm",
                "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(spmv_gbs / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(kname, args.case),
                "bytes_per_launch": by, "ms_per_launch": round(spmv_avg, 5),
                "ms_per_launch_in_solves": round(ms_spmv / max(1, n_spmv), 5)}

    if head_name == "learned":
        acc = pc_accepted(head["stats"])
        pc_label = ("learned" if acc > 0 else
                    "block-diagonal (the network's direction was rejected in every timed "
                    "Newton iteration)")
    else:
        pc_label = "block-diagonal"
    desc = (f"{args.case}: {spec['N']}^3-hex unit box, Kuhn 6-tet split ({nel} tets, "
            f"{case.nshg} nodes, {case.nnz_tot} 5x5 blocks), Mach {spec['mach']}, jitter "
            f"{spec['jitter']}h, y-grading {spec['grade']}, periodic x/z, no-slip walls y; "
            f"one Newton iteration = ElmGMRs + GMRES(K={spec['kspace']}) to 1e-3 + itrCorrect")
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "elements/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(head["ms"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head_name == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": desc, "N": spec["N"], "nel": nel, "nshg": case.nshg,
                   "nnz_blocks": case.nnz_tot, "newton_per_time_step": spec["newton"],
                   "arm": head_name, "pc": pc_label, "idiff": args.idiff},
        "roofline": roof_asm,
        "roofline_matvec": roof_mv,
        "arms": {a: breakdown(res[a], args.steps, nel) for a in res},
        "copy_GBs": round(head["by_copy"] / (head["ms_copy"] * 1e-3) / 1e9, 1),
    }
    if "learned" in res:
        line["arms"]["learned"]["pc_accepted_fraction"] = round(
            pc_accepted(res["learned"]["stats"]), 3)
    if "fp64" in res:
        d64 = res["fp64"]["dy"]
        for a in res:
            if a == "fp64":
                continue
            # time ratio, and the first Newton update against the FP64 arm's
            line["arms"][a]["speedup_vs_fp64"] = round(res["fp64"]["ms"] / res[a]["ms"], 3)
            line["arms"][a]["vs_fp64"] = {
                "Dy_rel_err": float(np.linalg.norm(res[a]["dy"] - d64) / np.linalg.norm(d64)),
                "iKs": [res[a]["first"]["iKs"], res["fp64"]["first"]["iKs"]],
                "converged": [bool(res[a]["first"]["converged"]),
                              bool(res["fp64"]["first"]["converged"])]}
        if head_name != "fp64":
            line["speedup_vs_fp64"] = line["arms"][head_name]["speedup_vs_fp64"]
    if not args.no_cpu:
        # the C oracle (FP64, the reference's operation order) on the host,
        # on a bounded sample of the workload: one Newton iteration of the
        # 64^3 box (the same per-element and per-row work, 1/8 of the
        # elements), on all the cores this process may use and on one core
        model, ncores = host_cpu()
        threads = int(os.environ.get("OMP_NUM_THREADS", ncores))
        small = mesh.named_case("box64")
        dt_all, used, orc64 = cpu_newton(small, threads)
        dt_one, _, _ = cpu_newton(small, 1)
        line["cpu_baseline"] = {
            "value": round(small.nel / dt_all, 1), "unit": "elements/s", "cores": used,
            "kind": "port", "cpu": model, "cores_available": ncores,
            "sample": f"one FP64 Newton iteration (ElmGMRs + SolGMRs + itrCorrect, "
                      f"{orc64['iKs']} Krylov iterations) of the 64^3 box ({small.nel} tets, "
                      f"the workload's per-element work on 1/8 of its elements) by the C "
                      f"oracle on {used} OpenMP threads: {dt_all:.1f} s",
            "one_core": {"value": round(small.nel / dt_one, 1), "unit": "elements/s",
                         "seconds": round(dt_one, 2)},
            "gpu_vs_all_cores": round(value / (small.nel / dt_all), 1)}
        # the same Newton iteration on the GPU (FP64) against the oracle's
        gpu = solver.GpuSolver(small)
        gpu.upload_state(small.yold, np.zeros((small.nshg, 5), order="F"))
        gpu.predict()
        st = gpu.newton_iteration()
        _, _, dy = gpu.download_state()
        gpu.close()
        d_ref = orc64["Dy"]
        line["oracle_parity"] = {
            "case": "box64", "Dy_rel_err": float(np.linalg.norm(dy - d_ref) / np.linalg.norm(d_ref)),
            "iKs": [st["iKs"], orc64["iKs"]]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
