"""Benchmark: FD solves/sec at 128^3 nodes (rel. err <= 1e-6 vs FP64).

Workload: finite-difference implicit interpolation on a 128^3-node unit-box
grid (127 cells per axis), 5000 value and 1000 normal constraints of
f = z + 0.1 sin(2 pi x) cos(2 pi y), regularisation 0.1, reference default
weights (npw = 10, borders 1, ridge 1e-8, Jacobi column scaling).  One step =
one complete solve from constraint arrays already resident in HBM: point
location + row assembly + transpose/diagonal/rhs + PCG to the tolerance that
meets rel. L2 error <= 1e-6 against a tightly converged FP64 solution.

Forms timed: 'plain' = FP64 Jacobi-PCG (one system); 'mixed' = FP32 stencil
and vectors with FP64 dots, FP64 residual recomputation and outer refinement.
``value`` is the best form that meets the accuracy bar.

Prints ONE JSON line (see the round contract) with ``roofline`` for the
dominant kernel (the regulariser stencil, HIP-event timed) and
``cpu_baseline`` (the CPU oracle: the reference's scipy CG path, restated).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "FD solves/sec at 128³ (rel.err≤1e-6 vs FP64) batched vs plain; % HBM roofline"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=127)
    ap.add_argument("--values", type=int, default=5000)
    ap.add_argument("--normals", type=int, default=1000)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-hbm-sweep", action="store_true")
    ap.add_argument("--forms", default="plain,mixed")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the batched and "
              f"single-system solvers here use one MI355X. Multi-GPU sharding is design-only.",
              flush=True)
        sys.exit(2)
    import torch

    from loopstructural_amd import _lib, synthetic
    from loopstructural_amd import fd_system as fs

    dev = _lib.device()
    cells = args.cells
    grid = _lib.make_grid((0.0, 0.0, 0.0), (1.0 / cells,) * 3, (cells + 1,) * 3)
    n = (cells + 1) ** 3
    values, normals = synthetic.constraints(args.values, args.normals, 1005)
    vals_d = torch.from_numpy(values).to(dev)
    nrms_d = torch.from_numpy(normals).to(dev)
    w = {"dxx": 0.1, "dyy": 0.1, "dzz": 0.1, "dxy": 0.1, "dxz": 0.1, "dyz": 0.1,
         "dx": 1.0, "dy": 1.0, "dz": 1.0}
    system = fs.FdSystem(grid, fs.reg_struct(w, (cells + 1,) * 3, 1e-8))

    def assemble():
        blocks = []
        for pts, kind, wcol, ws in ((nrms_d, _lib.ROWS_NORMAL, 6, 10.0 / 3),
                                    (vals_d, _lib.ROWS_VALUE, 4, 1.0)):
            _, ins, _, _ = fs.locate(grid, pts[:, :3].contiguous(), want_w=False)
            blocks.append(fs.build_rows(grid, kind, pts[ins].contiguous(), wcol, ws, n))
        system.set_rows(fs.concat_rows(blocks, n))

    log(f"workload 128^3-class: cells={cells} nodes={n} values={args.values} "
        f"normals={args.normals}")
    assemble()
    t = time.perf_counter()
    c_ref, info_ref = system.solve(rtol=1e-13, precision="fp64", check_every=256)
    torch.cuda.synchronize()
    log(f"FP64 reference solve (rtol 1e-13): {info_ref['iterations']} it, "
        f"{time.perf_counter() - t:.2f}s, converged={info_ref['converged']}")
    ref_norm = float(torch.linalg.vector_norm(c_ref))

    forms = {}
    for form in args.forms.split(","):
        prec = "fp64" if form == "plain" else "fp32"
        kw = dict(rtol=args.rtol, precision=prec, check_every=128)
        if prec == "fp32":
            kw["inner_rtol"] = 1e-3

        def step():
            assemble()
            return system.solve(**kw)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            c, info = step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        err = float(torch.linalg.vector_norm(c - c_ref)) / ref_norm
        forms[form] = {"ms_per_solve": dt * 1e3, "solves_per_s": 1.0 / dt,
                       "iterations": info["iterations"], "refinement_sweeps": info["outer"],
                       "rel_err_vs_fp64": err, "meets_1e-6": err <= 1e-6,
                       "dtype": "f64" if prec == "fp64" else "f32+f64"}
        log(f"{form}: {dt * 1e3:.1f} ms/solve, it={info['iterations']} outer={info['outer']} "
            f"relerr={err:.2e}")

    ok = {k: v for k, v in forms.items() if v["meets_1e-6"]}
    best = max(ok, key=lambda k: ok[k]["solves_per_s"]) if ok else "plain"
    prec = "fp64" if best == "plain" else "fp32"
    sb = 8 if prec == "fp64" else 4
    # roofline of the dominant kernel: regulariser stencil, HIP events on the
    # launch stream; algorithmic bytes = n * (read x + write y)
    avg_ms = system.time_stencil(prec, iters=50)
    bytes_alg = n * 2 * sb
    achieved = bytes_alg / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": f"k_stencil<{'double' if prec == 'fp64' else 'float'}>",
                "avg_us": round(avg_ms * 1e3, 2), "algorithmic_bytes": bytes_alg,
                "note": "128^3 working set is Infinity-Cache (MALL) resident",
                "autotune": system.stencil_info(prec)}
    log(f"stencil {prec}: {avg_ms * 1e3:.2f} us, {achieved:.0f} GB/s")
    hbm = None
    if not args.no_hbm_sweep:
        big = 511
        g2 = _lib.make_grid((0.0, 0.0, 0.0), (1.0 / big,) * 3, (big + 1,) * 3)
        s2 = fs.FdSystem(g2, fs.reg_struct(w, (big + 1,) * 3, 1e-8))
        n2 = (big + 1) ** 3
        res = {}
        for p2, sb2 in (("fp32", 4), ("fp64", 8)):
            ms2 = s2.time_stencil(p2, iters=20)
            gbs = n2 * 2 * sb2 / (ms2 * 1e-3) / 1e9
            res[p2] = {"avg_us": round(ms2 * 1e3, 1), "achieved_GBs": round(gbs, 1),
                       "frac": round(gbs / HBM_PEAK_GBS, 4), "autotune": s2.stencil_info(p2)}
            log(f"512^3 stencil {p2}: {ms2 * 1e3:.1f} us, {gbs:.0f} GB/s")
        s2.close()
        del s2
        torch.cuda.empty_cache()
        hbm = {"grid": "512^3 nodes, HBM-resident (x+y > 256 MB)", **res}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(cells, values, normals, forms["plain"]["iterations"]
                           if "plain" in forms else info_ref["iterations"])

    out = {
        "metric": METRIC,
        "value": round(forms[best]["solves_per_s"], 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(forms[best]["ms_per_solve"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": forms[best]["dtype"],
        "data": "synthetic",
        "config": {"workload": f"FD interpolation solve, {cells + 1}^3 nodes unit box, "
                               f"{args.values} value + {args.normals} normal constraints, "
                               f"regularisation 0.1, rel.err<=1e-6 vs FP64",
                   "form": best, "rtol": args.rtol, "parallelism": "single GPU"},
        "forms": forms,
        "fp64_reference_iterations": info_ref["iterations"],
        "roofline": roofline,
        "roofline_hbm_512": hbm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(cells, values, normals, gpu_iterations):
    """The reference's CPU path (oracle restatement: explicit scipy assembly,
    ridge + column scaling, scipy cg on A.T@A) on a bounded sample: full
    assembly/preprocess at the same size + 50 CG iterations, extrapolated to
    the GPU's iteration count."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from scipy.sparse import linalg as spla

    from oracle import fd as ofd

    og = ofd.Grid((0, 0, 0), (cells,) * 3, (1.0 / cells,) * 3)
    t0 = time.perf_counter()
    sysm = ofd.assemble_fd(og, {"value": values.copy(), "normal": normals.copy()},
                           regularisation=0.1)
    A, b = sysm.build_matrix()
    A, b, S = ofd.preprocess(A, b)
    N = A.T @ A
    rhs = A.T @ b
    t_setup = time.perf_counter() - t0
    k = 50
    t1 = time.perf_counter()
    spla.cg(N, rhs, rtol=1e-300, atol=0.0, maxiter=k)
    t_it = (time.perf_counter() - t1) / k
    t_solve = t_setup + t_it * gpu_iterations
    log(f"cpu oracle: setup {t_setup:.1f}s, {t_it * 1e3:.1f} ms/it -> {t_solve:.0f} s/solve")
    return {"value": round(1.0 / t_solve, 6), "unit": "solves/s", "cores": 1,
            "kind": "port",
            "sample": f"{cells + 1}^3 nodes: full scipy assembly+preprocess+A^T A "
                      f"({t_setup:.1f}s) + {k} CG iterations ({t_it * 1e3:.1f} ms/it), "
                      f"extrapolated to {gpu_iterations} iterations (GPU FP64 count)",
            "host_cpus_visible": len(os.sched_getaffinity(0))}


if __name__ == "__main__":
    main()
