#!/usr/bin/env python
"""Benchmark: time per Newton iteration of the compressible hot path.

A step is one Newton iteration on the device-resident state: ElmGMRs
(interior + wall assembly, BCs) + SolGMRs (block LU, scaling, block-diagonal
GMRES to the 1e-3 tolerance) + itrCorrect/itrBC, i.e. the unit the reference
runs once per "0 1" pair of its step sequence (itrdrv.f:427-599).  Every
`newton` steps the predictor restarts the time step from the same yold, so
each timed step repeats a real Newton iteration of the workload.

value = interior tets assembled and solved per second (nel / time per
Newton iteration).  Inputs are resident in HBM before the timed region.
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
FP64_PEAK_TFLOPS = 78.6


def spmv_bytes(nnz, nshg, s_a):
    """Algorithmic bytes of one block-CSR matvec: blocks + column indices +
    row pointers + one read and one write of the 5-vector."""
    return nnz * (25 * s_a + 4) + 4 * (nshg + 1) + 2 * nshg * 5 * 8


def cpu_baseline(case, newton_index=0):
    """One oracle Newton iteration (1 core) of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    orc = pyoracle.Oracle(case)
    acold = np.zeros((case.nshg, 5), order="F")
    y, ac = orc.itr_predict_bc(case.yold, acold)
    t0 = time.perf_counter()
    out = orc.solgmrs(y, ac)
    y1, _ = orc.itr_correct_bc(y, ac, case.yold, acold, out["Dy"])
    dt = time.perf_counter() - t0
    return dt, out, y1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--case", default="box64")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None,
                    help="override the case's precision")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; sharding over several "
                          "GPUs is a design, not code", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    from phasta_amd import config, mesh, solver

    spec = dict(config.CASES[args.case])
    if args.precision:
        spec["precision"] = args.precision
    case = mesh.named_case(args.case)
    newton = spec["newton"]
    precision = solver.PHG_MIXED if spec["precision"] == "mixed" else solver.PHG_FP64
    gpu = solver.GpuSolver(case, precision=precision)
    acold = np.zeros((case.nshg, 5), order="F")
    gpu.upload_state(case.yold, acold)
    info = gpu.info()

    it = [0]
    stats = []

    def step():
        if it[0] % newton == 0:
            gpu.predict()
        st = gpu.newton_iteration()
        it[0] += 1
        return st

    for _ in range(args.warmup):
        step()
    gpu.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(step())
    gpu.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    nel = case.nel
    value = nel / (ms_per_step / 1e3)

    # live per-kernel HIP-event times from the timed steps
    ms_spmv = sum(s["ms_spmv"] for s in stats)
    n_spmv = sum(s["n_spmv"] for s in stats)
    ms_mgs = sum(s["ms_mgs"] for s in stats)
    ms_asm = sum(s["ms_assembly"] for s in stats)
    s_a = 4 if precision == solver.PHG_MIXED else 8
    by = spmv_bytes(case.nnz_tot, case.nshg, s_a)
    spmv_avg = ms_spmv / max(1, n_spmv)
    spmv_gbs = by / (spmv_avg * 1e-3) / 1e9
    roofline = {"kernel": "k_spmv (block-CSR matvec, SparseAp)", "bound": "hbm",
                "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(spmv_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": by, "ms_per_launch": round(spmv_avg, 5)}

    # a plain copy of known size, after the timed region: the achievable
    # stream bandwidth on this box and the calibration kernel of the
    # profiler's memory counters
    ms_copy, by_copy = gpu.bench_copy(5)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "elements/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == solver.PHG_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{args.case}: {spec['N']}^3-hex unit box, Kuhn 6-tet split "
                               f"({nel} tets, {case.nshg} nodes, {case.nnz_tot} 5x5 blocks), "
                               f"Mach {spec['mach']}, jitter {spec['jitter']}h, grade "
                               f"{spec['grade']}, periodic x/z, no-slip walls y; one Newton "
                               f"iteration = ElmGMRs + block-diagonal GMRES(K={spec['kspace']}) "
                               f"to 1e-3 + itrCorrect",
                   "N": spec["N"], "nel": nel, "nshg": case.nshg, "nnz_blocks": case.nnz_tot,
                   "newton_per_time_step": newton, "precision": spec["precision"],
                   "device_bytes": info["dev_bytes"]},
        "roofline": roofline,
        "breakdown": {"ms_assembly_per_step": round(ms_asm / args.steps, 3),
                      "ms_spmv_per_step": round(ms_spmv / args.steps, 3),
                      "ms_mgs_per_step": round(ms_mgs / args.steps, 3),
                      "krylov_iters": [s["iKs"] for s in stats],
                      "assembly_elements_per_s": round(nel / (ms_asm / args.steps / 1e3), 1),
                      "copy_GBs": round(by_copy / (ms_copy * 1e-3) / 1e9, 1),
                      "copy_bytes": by_copy},
    }
    if not args.no_cpu:
        dt, ref, _ = cpu_baseline(case)
        line["cpu_baseline"] = {"value": round(nel / dt, 1), "unit": "elements/s", "cores": 1,
                                "kind": "port",
                                "sample": f"one Newton iteration of {args.case} by the C "
                                          f"oracle on 1 host core ({dt:.1f} s)"}
        # parity of the first Newton iteration against the oracle
        gpu.upload_state(case.yold, acold)
        gpu.predict()
        g = gpu.newton_iteration()
        _, _, dy = gpu.download_state()
        rel = float(np.linalg.norm(dy - ref["Dy"]) / np.linalg.norm(ref["Dy"]))
        line["parity"] = {"Dy_rel_err_vs_oracle": rel, "iKs_gpu": g["iKs"],
                          "iKs_oracle": ref["iKs"]}
    gpu.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
