"""Benchmark: FD solves/sec at 128^3 nodes (rel. err <= 1e-6 vs FP64).

Workload: finite-difference implicit interpolation on a 128^3-node unit-box
grid (127 cells per axis), 5000 value and 1000 normal constraints of
f = z + 0.1 sin(2 pi x) cos(2 pi y), regularisation 0.1, reference default
weights (npw = 10, borders 1, ridge 1e-8, Jacobi column scaling).  One step =
one complete solve from constraint arrays already resident in HBM: point
location + row assembly + transpose/diagonal/rhs + PCG to the tolerance that
meets rel. L2 error <= 1e-6 against a tightly converged FP64 solution.

Forms timed: 'plain' = FP64 Jacobi-PCG (one system, the reference's column
scaling); 'mixed' = FP32 stencil and vectors with FP64 dots, FP64 residual
recomputation and outer refinement; 'mg' = geometric-multigrid preconditioned
flexible CG (FP32 V-cycle, FP64 outer iteration and stopping test);
'batched' = B perturbed realisations of the same workload (value noise 0.02,
normals tilted 5 degrees at the same locations: one matrix, B right-hand
sides) solved together by the multi-rhs multigrid (every launch serves all B
lanes), each realisation checked against its own FP64 solution.  ``value`` is
the best form that meets the accuracy bar.

Also reported (not ``value``): the 256-system uneven-grid ensemble
(32^3..96^3 nodes), the P1 solve on 128^3 cells, FD and P1 evaluation
throughput, the adjoint gradient cost, and the stencil at 512^3 (HBM-resident).

Prints ONE JSON line with ``roofline`` for the dominant kernel (the regulariser
stencil, HIP-event timed) and ``cpu_baseline`` (the CPU oracle: the
reference's scipy CG path, restated).
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
W = {"dxx": 0.1, "dyy": 0.1, "dzz": 0.1, "dxy": 0.1, "dxz": 0.1, "dyz": 0.1,
     "dx": 1.0, "dy": 1.0, "dz": 1.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=127)
    ap.add_argument("--values", type=int, default=5000)
    ap.add_argument("--normals", type=int, default=1000)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--batch", type=int, default=64, help="realisations in the batched form")
    ap.add_argument("--ensemble", type=int, default=256, help="systems in the uneven ensemble")
    ap.add_argument("--forms", default="plain,mixed,mg,batched")
    ap.add_argument("--ref-solver", default="jacobi", choices=("jacobi", "mg", "none"),
                    help="FP64 reference solves (rtol 1e-13): Jacobi PCG (default), multigrid "
                         "(faster), or none (kernel traces of the timed work only: accuracy is "
                         "not checked and the line reports no value)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-hbm-sweep", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip ensemble, P1, evaluation and adjoint measurements")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run a 512 MiB copy, the stencil at 128^3 and 512^3 and the "
                         "256-system ensemble stencil on both tile widths and z-chunks (the workload of the "
                         "FETCH_SIZE / WRITE_SIZE counter runs)")
    ap.add_argument("--p1-probe", action="store_true",
                    help="only time the P1 stencil at 129^3 and 257^3 nodes (every tuned form, "
                         "the workload of the P1 kernel traces)")
    ap.add_argument("--ensemble-probe", action="store_true",
                    help="only prepare and solve the 256-system ensemble by batched multigrid "
                         "(the workload of its kernel trace)")
    ap.add_argument("--eval-probe", action="store_true",
                    help="only run the FD evaluation measurement on a 128^3 grid with seeded "
                         "coefficients (the workload of its kernel trace and counter runs)")
    ap.add_argument("--workers", type=int, default=12,
                    help="host threads of the ensemble drivers")
    return ap.parse_args()


def traffic_probe():
    """Copy kernel of known size (counter calibration), then the regulariser
    stencil at 128^3 and 512^3 in FP32 and FP64 (autotuned z-chunk)."""
    import torch

    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    x = torch.rand(512 ** 3, device=_lib.device(), dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)  # 512 MiB read + 512 MiB written
    torch.cuda.synchronize()
    out = {}
    for n1 in (128, 512):
        g = _lib.make_grid((0, 0, 0), (1 / (n1 - 1),) * 3, (n1,) * 3)
        s = fs.FdSystem(g, fs.reg_struct(W, (n1,) * 3, 1e-8))
        for prec in ("fp32", "fp64"):
            ms = s.time_stencil(prec, iters=3)
            out[f"{n1}:{prec}"] = {"avg_us": round(ms * 1e3, 2), **s.stencil_info(prec)}
        s.close()
    # the 256-system ensemble (32^3..96^3 nodes), FP32 stencil in each form
    from loopstructural_amd import synthetic

    ens = synthetic.ensemble(256, 1004)
    grids = [_lib.make_grid((0, 0, 0), (1.0 / (k - 1),) * 3, (k,) * 3) for k, _, _ in ens]
    tot = sum(k ** 3 for k, _, _ in ens)
    for tile, zc in ((64, 16), (32, 16), (64, 32), (32, 32)):
        os.environ["LSAMD_BATCH_TILE"], os.environ["LSAMD_BATCH_ZC"] = str(tile), str(zc)
        bs = fs.BatchSystem(grids, fs.reg_struct(W, (32, 32, 32), 1e-8))
        bs.solve(rtol=1e-5, maxiter=1, precision="fp32", check_every=1)  # fixes the form
        ms = bs.time_stencil("fp32", iters=3)
        out[f"ensemble:fp32:{tile}:z{zc}"] = {"avg_us": round(ms * 1e3, 2), "dof": tot,
                                              "GBs": round(tot * 8 / (ms * 1e-3) / 1e9, 1)}
        bs.close()
        del os.environ["LSAMD_BATCH_TILE"], os.environ["LSAMD_BATCH_ZC"]
    print(json.dumps({"traffic_probe": out}), flush=True)


def p1_probe():
    """The P1 edge-jump stencil at 129^3 and 257^3 nodes in FP64 and FP32:
    the autotuner times every tile form, then 20 launches of the chosen one."""
    import torch

    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    out = {}
    for cells in (128, 256):
        g = _lib.make_grid((0, 0, 0), (1 / cells,) * 3, (cells + 1,) * 3)
        s = fs.P1System(g, 0.1, 1e-8)
        n = (cells + 1) ** 3
        for prec, sb in (("fp64", 8), ("fp32", 4)):
            ms = s.time_stencil(prec, iters=20)
            gbs = 2 * sb * n / (ms * 1e-3) / 1e9
            out[f"{cells + 1}:{prec}"] = {"avg_us": round(ms * 1e3, 2), "GBs": round(gbs, 1),
                                         "frac": round(gbs / HBM_PEAK_GBS, 4)}
        s.close()
    torch.cuda.synchronize()
    print(json.dumps({"p1_probe": out}), flush=True)


def eval_probe():
    """FD evaluation (values + gradients) on the 128^3 grid of the headline
    solve, seeded random coefficients: 1e7 uniform random points and the
    216^3 linspace points, three timed calls each after one warm call.  A
    512 MiB device copy first calibrates the counter runs."""
    import torch

    from loopstructural_amd import _lib
    from loopstructural_amd import fd_system as fs

    x = torch.rand(512 ** 3, device=_lib.device(), dtype=torch.float32)
    torch.empty_like(x).copy_(x)
    del x
    n1 = 128
    grid = _lib.make_grid((0, 0, 0), (1 / (n1 - 1),) * 3, (n1,) * 3)
    g = torch.Generator(device="cuda").manual_seed(3)
    c = torch.rand(n1 ** 3, dtype=torch.float64, device="cuda", generator=g)
    res = fd_evaluation(torch, _lib, fs, grid, c, n1 ** 3)
    print(json.dumps({"eval_probe": res}), flush=True)


def ensemble_probe(args):
    """The 256-system ensemble (32^3..96^3 nodes) by batched multigrid: grid
    hierarchies, then the timed rows, row-dependent hierarchy parts and
    batched solve (as in the bench's ensemble measurement)."""
    import torch

    from loopstructural_amd import _lib, synthetic
    from loopstructural_amd import fd_system as fs

    dev = _lib.device()
    ens = synthetic.ensemble(args.ensemble, 1004)
    grids, ns = [], []
    for nnode, _, _ in ens:
        grids.append(_lib.make_grid((0, 0, 0), (1.0 / (nnode - 1),) * 3, (nnode,) * 3))
        ns.append(nnode ** 3)
    reg = fs.reg_struct(W, (32, 32, 32), 1e-8)
    vd = [(torch.from_numpy(v).to(dev), torch.from_numpy(m).to(dev)) for _, v, m in ens]
    blocks = [fd_rows(fs, _lib, g, v, m, k) for g, (v, m), k in zip(grids, vd, ns)]
    bm = fs.BatchedMultigrid(grids, reg, workers=args.workers)
    t0 = time.perf_counter()
    bm.prepare(blocks)
    torch.cuda.synchronize()
    t_prep = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, infos = bm.solve(blocks, rtol=args.rtol)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out = {"systems": len(ens), "workers": args.workers, "setup_seconds": round(t_prep, 3),
           "seconds": round(dt, 3), "solves_per_s": round(len(ens) / dt, 2),
           "iterations_max": max(i["iterations"] for i in infos),
           "converged": sum(i["converged"] for i in infos),
           "phase_seconds": {k: round(v, 3) for k, v in bm.last_timing.items()}}
    bm.close()
    print(json.dumps({"ensemble_probe": out}), flush=True)


def stencil_traffic(nn, prec, zc):
    """HBM bytes per dispatch of the stencil from the committed counter
    summary (profiles/stencil_traffic.json), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "stencil_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic"].get(f"{nn}:{prec}:{zc}")
    except (OSError, ValueError, KeyError):
        return None


def batched_traffic(key):
    """HBM bytes per dispatch of a batched-form kernel from the committed
    counter summary (profiles/fd_batched_traffic.json), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "fd_batched_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic"].get(key)
    except (OSError, ValueError, KeyError):
        return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def fd_rows(fs, _lib, grid, vals_d, nrms_d, n):
    blocks = []
    for pts, kind, wcol, ws in ((nrms_d, _lib.ROWS_NORMAL, 6, 10.0 / 3),
                                (vals_d, _lib.ROWS_VALUE, 4, 1.0)):
        _, ins, _, _ = fs.locate(grid, pts[:, :3].contiguous(), want_w=False)
        blocks.append(fs.build_rows(grid, kind, pts[ins].contiguous(), wcol, ws, n))
    return fs.concat_rows(blocks, n)


def main():
    args = parse()
    if args.traffic_probe:
        traffic_probe()
        return
    if args.p1_probe:
        p1_probe()
        return
    if args.ensemble_probe:
        ensemble_probe(args)
        return
    if args.eval_probe:
        eval_probe()
        return
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
    nn = cells + 1
    grid = _lib.make_grid((0.0, 0.0, 0.0), (1.0 / cells,) * 3, (nn,) * 3)
    n = nn ** 3
    values, normals = synthetic.constraints(args.values, args.normals, 1005)
    vals_d = torch.from_numpy(values).to(dev)
    nrms_d = torch.from_numpy(normals).to(dev)
    reg = fs.reg_struct(W, (nn,) * 3, 1e-8)
    system = fs.FdSystem(grid, reg)

    def assemble():
        system.set_rows(fd_rows(fs, _lib, grid, vals_d, nrms_d, n))

    log(f"workload: {nn}^3 nodes, values={args.values} normals={args.normals}")
    assemble()
    c_ref, info_ref, ref_norm = None, {"iterations": None}, None
    if args.ref_solver != "none":
        t = time.perf_counter()
        c_ref, info_ref = system.solve(rtol=1e-13, precision="fp64", check_every=256)
        torch.cuda.synchronize()
        log(f"FP64 reference solve (rtol 1e-13): {info_ref['iterations']} it, "
            f"{time.perf_counter() - t:.2f}s, converged={info_ref['converged']}")
        ref_norm = float(torch.linalg.vector_norm(c_ref))

    forms = {}
    want = args.forms.split(",")
    for form in [f for f in want if f in ("plain", "mixed", "mg")]:
        prec = "fp64" if form in ("plain", "mg") else "fp32"
        kw = dict(rtol=args.rtol, precision=prec, check_every=128)
        if form == "mixed":
            kw["inner_rtol"] = 1e-3
        if form == "mg":
            kw.update(multigrid=True, check_every=8)

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
        err = (float(torch.linalg.vector_norm(c - c_ref)) / ref_norm if c_ref is not None
               else float("nan"))
        forms[form] = {"ms_per_solve": dt * 1e3, "solves_per_s": 1.0 / dt,
                       "iterations": info["iterations"], "refinement_sweeps": info["outer"],
                       "rel_err_vs_fp64": err, "meets_1e-6": err <= 1e-6,
                       "dtype": "f64" if form == "plain" else "f32+f64"}
        log(f"{form}: {dt * 1e3:.1f} ms/solve, it={info['iterations']} outer={info['outer']} "
            f"relerr={err:.2e}")

    if "batched" in want and args.batch > 0:
        forms["batched"] = batched_form(args, torch, _lib, fs, synthetic, grid, reg, values,
                                        normals, n, system)

    ok = {k: v for k, v in forms.items() if v["meets_1e-6"]}
    if args.ref_solver == "none":  # trace runs: accuracy unchecked, no value reported
        ok = {}
        best = max(forms, key=lambda k: forms[k]["solves_per_s"])
    else:
        best = max(ok, key=lambda k: ok[k]["solves_per_s"]) if ok else "plain"
    prec = "fp64" if forms[best]["dtype"] == "f64" else "fp32"
    sb = 8 if prec == "fp64" else 4
    # roofline of the dominant kernel: regulariser stencil of the single-system
    # plan, HIP events on the launch stream; algorithmic bytes = n*(read x + write y)
    avg_ms = system.time_stencil(prec, iters=50)
    bytes_alg = n * 2 * sb
    achieved = bytes_alg / (avg_ms * 1e-3) / 1e9
    info_st = system.stencil_info(prec)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": stencil_traffic(nn, prec, info_st["zc"]),
                "traffic_source": "profiles/stencil_traffic.json (FETCH_SIZE x2 + WRITE_SIZE, "
                                  "copy-kernel calibrated)",
                "kernel": f"regulariser stencil ({prec}): k_b_tiles<{prec}, zc>",
                "avg_us": round(avg_ms * 1e3, 2), "algorithmic_bytes": bytes_alg,
                "profile": "profiles/stencil_roofline_kernel_trace.md (rocprofv3 per-dispatch "
                           "durations of the same launches)",
                "note": f"{nn}^3 working set is Infinity-Cache (MALL) resident; the HBM-bound "
                        f"figure is roofline_hbm_512",
                "autotune": info_st}
    log(f"stencil {prec}: {avg_ms * 1e3:.2f} us, {achieved:.0f} GB/s")
    roofline_128 = roofline
    rb = forms.get("batched", {}).get("roofline_batched")
    if best == "batched" and rb and "cheb" in rb:
        # the batched form's largest kernel share: the fine-level Chebyshev step
        c = rb["cheb"]
        roofline = {"bound": "hbm", "achieved": c["achieved"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": c["frac"], "traffic": c["traffic"],
                    "traffic_source": "profiles/fd_batched_traffic.json (FETCH_SIZE x2 + "
                                      "WRITE_SIZE of the same timed launches, copy-kernel "
                                      "calibrated; profiles/fd_batched_traffic_pmc.md)",
                    "kernel": c["kernel"], "avg_us": c["avg_us"],
                    "algorithmic_bytes": c["algorithmic_bytes"],
                    "profile": "profiles/fd_batched_128_kernel_trace.md (rocprofv3 per-dispatch "
                               "durations by kernel, grid and multigrid level, and of these "
                               "timed launches)",
                    "note": "largest kernel share of the batched form's time; the other "
                            "fine-level kernels are in roofline_batched, the single-system "
                            "128^3 stencil in roofline_stencil_128"}
    hbm = None if args.no_hbm_sweep else hbm_sweep(torch, _lib, fs)

    extras = {}
    if not args.no_extras and c_ref is not None:
        extras["fd_evaluation"] = fd_evaluation(torch, _lib, fs, grid, c_ref, n)
        extras["p1"] = p1_c3(torch, _lib, fs, synthetic)
        extras["adjoint"] = adjoint_c5(torch, synthetic)
        extras["ensemble"] = ensemble_c4(args, torch, _lib, fs, synthetic)

    cpu = None
    if not args.no_cpu_baseline and info_ref["iterations"] is not None:
        cpu = cpu_baseline(cells, values, normals, forms["plain"]["iterations"]
                           if "plain" in forms else info_ref["iterations"])
        if not args.no_extras:
            extras["c1_phases"] = c1_phases()

    out = {
        "metric": METRIC,
        "value": round(forms[best]["solves_per_s"], 4) if ok else None,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(forms[best].get("ms_per_batch", forms[best]["ms_per_solve"]), 3),
        "solves_per_step": forms[best].get("batch", 1),
        "ms_per_solve": round(forms[best]["ms_per_solve"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": forms[best]["dtype"],
        "data": "synthetic",
        "config": {"workload": f"FD interpolation solve, {nn}^3 nodes unit box, "
                               f"{args.values} value + {args.normals} normal constraints, "
                               f"regularisation 0.1, rel.err<=1e-6 vs FP64",
                   "form": best, "rtol": args.rtol, "parallelism": "single GPU"},
        "forms": forms,
        "batched_vs_plain": (round(forms["batched"]["solves_per_s"] /
                                   forms["plain"]["solves_per_s"], 3)
                             if "batched" in forms and "plain" in forms else None),
        "fp64_reference_iterations": info_ref["iterations"],
        "roofline": roofline,
        "roofline_batched": forms.get("batched", {}).get("roofline_batched"),
        "roofline_stencil_128": roofline_128,
        "roofline_hbm_512": hbm,
        "extras": extras,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def batched_form(args, torch, _lib, fs, synthetic, grid, reg, values, normals, n, single):
    """B perturbed realisations of the 128^3 workload (value noise and tilted
    normals at the same locations: one matrix, B right-hand sides) in one
    multi-rhs multigrid solve: rows of every realisation, the matrix from the
    first, the right-hand side of each from its row values, then all B
    lanes per launch sequence."""
    B = args.batch
    dev = _lib.device()
    reals = [synthetic.perturb(values, normals, 2000 + b) for b in range(B)]
    vd = [(torch.from_numpy(v).to(dev), torch.from_numpy(m).to(dev)) for v, m in reals]
    # FP64 references (single-system solves, rtol 1e-13) of 16 realisations
    # spread over the batch (every one up to 16)
    ref_idx = sorted(set(np.linspace(0, B - 1, min(B, 16)).astype(int).tolist()))
    refs = []
    rows0 = None
    for b, (v, m) in enumerate(vd):
        rows = fd_rows(fs, _lib, grid, v, m, n)
        if rows0 is None:
            rows0 = rows
        elif not (torch.equal(rows.cols, rows0.cols) and torch.equal(rows.coef, rows0.coef)
                  and torch.equal(rows.w, rows0.w)):
            raise RuntimeError("realisations do not share one matrix")
        if args.ref_solver == "none" or b not in ref_idx:
            continue
        single.set_rows(rows)
        if args.ref_solver == "mg":
            c, _ = single.solve(rtol=1e-13, multigrid=True, check_every=8)
        else:
            c, _ = single.solve(rtol=1e-13, precision="fp64", check_every=256)
        refs.append(c.clone())
    sysb = fs.FdSystem(grid, reg)

    def step():
        rows = [fd_rows(fs, _lib, grid, v, m, n) for v, m in vd]
        sysb.set_rows(rows[0])
        rhs = torch.stack([sysb.rhs_for(r.b) for r in rows])
        return sysb.solve_multi(rhs, rtol=args.rtol, check_every=8)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        c, infos = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    errs = [float(torch.linalg.vector_norm(c[b] - r) / torch.linalg.vector_norm(r))
            for b, r in zip(ref_idx, refs)] or [float("nan")]
    res = {"ms_per_solve": dt * 1e3 / B, "solves_per_s": B / dt, "batch": B,
           "ms_per_batch": dt * 1e3, "method": "multi-rhs multigrid, all lanes per launch",
           "iterations_max": max(i["iterations"] for i in infos), "refinement_sweeps_max": 0,
           "rel_err_vs_fp64_max": max(errs), "rel_err_checked": len(refs),
           "meets_1e-6": max(errs) <= 1e-6, "dtype": "f32+f64"}
    log(f"batched x{B}: {dt * 1e3:.0f} ms/batch = {B / dt:.2f} solves/s, max relerr "
        f"{max(errs):.2e}")
    if B >= 24:
        # fine-level kernels of the widest lane group (64 lanes from 48
        # realisations, else 32), HIP-event timed on the launch stream (after
        # the timed region: the group's work is reused)
        lanes = 64 if B >= 48 else 32
        res["roofline_batched"] = {}
        for kname, kfull in (("cheb", f"k_mr_cheb_all<4> (Chebyshev step over every fine "
                                      f"node, {lanes} lanes)"),
                             ("update", f"k_mr_upd<{lanes}> (FP64 x, FP32 r update, {lanes} "
                                        f"lanes)"),
                             ("stencil", f"k_b_tiles<float, zc> over {lanes} lanes "
                                         f"(regulariser)")):
            ms, nbytes = sysb.time_multi(kname, iters=20)
            ach = nbytes / (ms * 1e-3) / 1e9
            res["roofline_batched"][kname] = {
                "kernel": kfull, "bound": "hbm", "avg_us": round(ms * 1e3, 2),
                "algorithmic_bytes": int(nbytes), "achieved": round(ach, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "lanes": lanes, "traffic": batched_traffic(f"{kname}:{lanes}")}
            log(f"batched {kname}: {ms * 1e3:.1f} us, {ach:.0f} GB/s")
    sysb.close()
    del sysb
    torch.cuda.empty_cache()
    return res


def hbm_sweep(torch, _lib, fs):
    big = 511
    g2 = _lib.make_grid((0.0, 0.0, 0.0), (1.0 / big,) * 3, (big + 1,) * 3)
    s2 = fs.FdSystem(g2, fs.reg_struct(W, (big + 1,) * 3, 1e-8))
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
    # what a plain device copy of the same size (x read, y written) reaches:
    # the practical ceiling for a kernel that streams x in and y out
    x = torch.rand(n2, device=_lib.device())
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    best = None
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 20
        best = ms if best is None else min(best, ms)
    cgbs = n2 * 8 / (best * 1e-3) / 1e9
    res["copy_fp32"] = {"avg_us": round(best * 1e3, 1), "achieved_GBs": round(cgbs, 1),
                        "frac": round(cgbs / HBM_PEAK_GBS, 4),
                        "what": "torch copy_ of the 512^3 FP32 vector (read + write)"}
    log(f"512^3 fp32 copy: {best * 1e3:.1f} us, {cgbs:.0f} GB/s")
    del x, y
    torch.cuda.empty_cache()
    return {"grid": "512^3 nodes, HBM-resident (x+y > 256 MB)", **res}


def _timed(torch, fn, reps=3):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps, out


def fd_evaluation(torch, _lib, fs, grid, c, n):
    """Values + gradients at 1e7 uniform random points (each point gathers 8
    corner coefficients from scattered cache lines), and at the 216^3 points
    of a linspace grid in the 'ij' order the reference evaluates on."""
    npts = 10_000_000
    g = torch.Generator(device="cuda").manual_seed(7)
    pts = torch.rand((npts, 3), dtype=torch.float64, device="cuda", generator=g)
    dt, _ = _timed(torch, lambda: fs.evaluate(grid, c, pts, True, True))
    bytes_alg = npts * (24 + 8 + 24) + 8 * n
    res = {"points": npts, "ms": round(dt * 1e3, 2), "Mpts_per_s": round(npts / dt / 1e6, 1),
           "GBs": round(bytes_alg / dt / 1e9, 1),
           "frac": round(bytes_alg / dt / 1e9 / HBM_PEAK_GBS, 4),
           "what": "values + gradients, uniform random points"}
    t = torch.linspace(0.0, 1.0, 216, dtype=torch.float64, device="cuda")
    xx, yy, zz = torch.meshgrid(t, t, t, indexing="ij")
    grid_pts = torch.stack([xx.reshape(-1), yy.reshape(-1), zz.reshape(-1)], 1).contiguous()
    del xx, yy, zz
    m = grid_pts.shape[0]
    dt2, _ = _timed(torch, lambda: fs.evaluate(grid, c, grid_pts, True, True))
    b2 = m * (24 + 8 + 24) + 8 * n
    res["ordered"] = {"points": m, "ms": round(dt2 * 1e3, 2),
                      "Mpts_per_s": round(m / dt2 / 1e6, 1), "GBs": round(b2 / dt2 / 1e9, 1),
                      "frac": round(b2 / dt2 / 1e9 / HBM_PEAK_GBS, 4),
                      "what": "values + gradients, 216^3 linspace points in 'ij' order"}
    log(f"FD evaluation: {res}")
    return res


def ensemble_c4(args, torch, _lib, fs, synthetic):
    """256 realisations, 32^3..96^3 nodes, FP32 + FP64 refinement."""
    dev = _lib.device()
    ens = synthetic.ensemble(args.ensemble, 1004)
    grids, blocks, ns = [], [], []
    for nnode, v, m in ens:
        c = nnode - 1
        g = _lib.make_grid((0, 0, 0), (1.0 / c,) * 3, (nnode,) * 3)
        grids.append(g)
        ns.append(nnode ** 3)
    reg = fs.reg_struct(W, (32, 32, 32), 1e-8)
    bs = fs.BatchSystem(grids, reg)
    vd = [(torch.from_numpy(v).to(dev), torch.from_numpy(m).to(dev)) for _, v, m in ens]

    def step():
        bs.set_rows([fd_rows(fs, _lib, g, v, m, k) for g, (v, m), k in zip(grids, vd, ns)])
        return bs.solve(rtol=args.rtol, precision="fp32", inner_rtol=1e-3, check_every=128)

    t0 = time.perf_counter()
    c, infos = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # accuracy on 8 realisations: FP64 single-system solves at rtol 1e-13
    errs = []
    parts = bs.split(c)
    for s in np.linspace(0, len(ens) - 1, 8).astype(int):
        single = fs.FdSystem(grids[s], fs.reg_struct(W, (ens[s][0],) * 3, 1e-8))
        single.set_rows(fd_rows(fs, _lib, grids[s], vd[s][0], vd[s][1], ns[s]))
        r, _ = single.solve(rtol=1e-13, precision="fp64", check_every=256)
        errs.append(float(torch.linalg.vector_norm(parts[s] - r) / torch.linalg.vector_norm(r)))
        single.close()
    sb_ms = bs.time_stencil("fp32", iters=10)
    tot = sum(ns)
    bs.close()
    del bs
    torch.cuda.empty_cache()
    log(f"ensemble (batched Jacobi PCG): {dt:.2f} s")
    # the same ensemble by per-system multigrid, shape groups on several streams
    blocks = [fd_rows(fs, _lib, g, v, m, k) for g, (v, m), k in zip(grids, vd, ns)]
    es = fs.EnsembleSolver(grids, reg, workers=args.workers)
    t0 = time.perf_counter()
    es.prepare(blocks, args.rtol)
    t_prep = time.perf_counter() - t0
    log(f"ensemble multigrid set-up ({len(es.shapes)} shapes): {t_prep:.2f} s")
    t0 = time.perf_counter()
    cm, infm = es.solve(blocks, rtol=args.rtol)
    dt_mg = time.perf_counter() - t0
    log(f"ensemble multigrid solves: {dt_mg:.2f} s")
    es_shapes, es_timing = len(es.shapes), dict(es.last_timing)
    es.close()
    del es
    # the same ensemble by batched multigrid: one hierarchy per system, every
    # system advanced by one launch sequence per iteration
    bm = fs.BatchedMultigrid(grids, reg, workers=args.workers)
    t0 = time.perf_counter()
    bm.prepare(blocks)
    t_bprep = time.perf_counter() - t0
    log(f"batched multigrid set-up ({len(ens)} hierarchies): {t_bprep:.2f} s")
    t0 = time.perf_counter()
    cb, infb = bm.solve(blocks, rtol=args.rtol)
    dt_b = time.perf_counter() - t0
    bm_timing = dict(bm.last_timing)
    log(f"batched multigrid solve: {dt_b:.2f} s ({bm_timing})")
    errs_mg, errs_b = [], []
    for s_ in np.linspace(0, len(ens) - 1, 8).astype(int):
        single = fs.FdSystem(grids[s_], fs.reg_struct(W, (ens[s_][0],) * 3, 1e-8))
        single.set_rows(blocks[s_])
        r, _ = single.solve(rtol=1e-13, precision="fp64", check_every=256)
        rn = torch.linalg.vector_norm(r)
        errs_mg.append(float(torch.linalg.vector_norm(cm[s_] - r) / rn))
        errs_b.append(float(torch.linalg.vector_norm(cb[s_] - r) / rn))
        single.close()
    bm.close()
    del bm
    res = {"systems": len(ens), "total_dof": tot, "seconds": round(dt, 3),
           "solves_per_s": round(len(ens) / dt, 2),
           "converged": sum(i["converged"] for i in infos),
           "iterations_max": max(i["iterations"] for i in infos),
           "rel_err_vs_fp64_max_of_8": max(errs),
           "stencil_fp32_us": round(sb_ms * 1e3, 1),
           "stencil_fp32_GBs": round(tot * 8 / (sb_ms * 1e-3) / 1e9, 1),
           "stencil_frac": round(tot * 8 / (sb_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "multigrid": {"seconds": round(dt_mg, 3), "solves_per_s": round(len(ens) / dt_mg, 2),
                         "setup_seconds": round(t_prep, 3), "shapes": es_shapes,
                         "converged": sum(i["converged"] for i in infm),
                         "iterations_max": max(i["iterations"] for i in infm),
                         "rel_err_vs_fp64_max_of_8": max(errs_mg),
                         "thread_seconds": {k: round(v, 3) for k, v in es_timing.items()},
                         "what": "per-system multigrid, one handle per grid shape (hierarchy "
                                 "built in setup_seconds), shape groups on 12 host threads "
                                 "with a stream each"},
           "batched_multigrid": {
               "seconds": round(dt_b, 3), "solves_per_s": round(len(ens) / dt_b, 2),
               "vs_batched_jacobi": round(dt / dt_b, 2), "setup_seconds": round(t_bprep, 3),
               "converged": sum(i["converged"] for i in infb),
               "iterations_max": max(i["iterations"] for i in infb),
               "rel_err_vs_fp64_max_of_8": max(errs_b),
               "phase_seconds": {k: round(v, 3) for k, v in bm_timing.items()},
               "what": "one hierarchy per system (grid part built in setup_seconds); timed: "
                       "rows of every system, row-dependent hierarchy parts (12 host "
                       "threads), and the solve of all systems by one launch sequence per "
                       "iteration"}}
    log(f"ensemble: {res}")
    torch.cuda.empty_cache()
    return res


def p1_c3(torch, _lib, fs, synthetic):
    """P1 on 128^3 cells: 10k values (the 2k gradient constraints are a no-op
    in the reference), cgw 0.1, FP64 and mixed solves, 1e7-point evaluation."""
    dev = _lib.device()
    cells = 128
    g = _lib.make_grid((0, 0, 0), (1.0 / cells,) * 3, (cells + 1,) * 3)
    n = (cells + 1) ** 3
    rng = np.random.default_rng(1003)
    pv = rng.uniform(0, 1, (10000, 3))
    f, _ = synthetic.field2(pv)
    f = f + rng.normal(0, 0.01, 10000)
    vals = torch.from_numpy(np.column_stack([pv, f])).to(dev)
    sysm = fs.P1System(g, 0.1, 1e-8)

    def assemble():
        _, _, ins, _, _ = fs.p1_locate(g, vals[:, :3].contiguous(), False)
        sysm.set_rows(fs.p1_build_rows(g, _lib.ROWS_VALUE, vals[ins].contiguous(), 1.0, n))

    assemble()
    t0 = time.perf_counter()
    c_ref, iref = sysm.solve(rtol=1e-12, precision="fp64", check_every=256)
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t0
    res = {"nodes": n, "tets": 5 * cells ** 3, "fp64_ref_iterations": iref["iterations"],
           "fp64_ref_seconds": round(t_ref, 3)}
    for prec in ("fp64", "fp32"):
        kw = dict(rtol=1e-10, precision=prec, check_every=128)
        if prec == "fp32":
            kw["inner_rtol"] = 1e-3
        dt, (c, info) = _timed(torch, lambda: (assemble(), sysm.solve(**kw))[1], reps=2)
        err = float(torch.linalg.vector_norm(c - c_ref) / torch.linalg.vector_norm(c_ref))
        res[prec] = {"ms_per_solve": round(dt * 1e3, 1), "iterations": info["iterations"],
                     "rel_err_vs_fp64": err}
    for prec, sb in (("fp64", 8), ("fp32", 4)):
        ms = sysm.time_stencil(prec, iters=20)
        res[f"stencil_{prec}_GBs"] = round(n * 2 * sb / (ms * 1e-3) / 1e9, 1)
        res[f"stencil_{prec}_us"] = round(ms * 1e3, 1)
    # the same operator on 256^3 cells (257^3 nodes): x + y in FP64 exceed the
    # Infinity Cache, so this is the HBM-bound figure
    big = 256
    gb = _lib.make_grid((0, 0, 0), (1.0 / big,) * 3, (big + 1,) * 3)
    nb = (big + 1) ** 3
    sb2 = fs.P1System(gb, 0.1, 1e-8)
    for prec, sb in (("fp64", 8), ("fp32", 4)):
        ms = sb2.time_stencil(prec, iters=20)
        gbs = nb * 2 * sb / (ms * 1e-3) / 1e9
        res[f"stencil_257_{prec}"] = {"avg_us": round(ms * 1e3, 1), "GBs": round(gbs, 1),
                                      "frac": round(gbs / HBM_PEAK_GBS, 4)}
    sb2.close()
    del sb2
    torch.cuda.empty_cache()
    npts = 10_000_000
    gen = torch.Generator(device="cuda").manual_seed(8)
    pts = torch.rand((npts, 3), dtype=torch.float64, device="cuda", generator=gen)
    dt, _ = _timed(torch, lambda: fs.p1_evaluate(g, c_ref, pts, True, False))
    res["evaluation_Mpts_per_s"] = round(npts / dt / 1e6, 1)
    log(f"P1: {res}")
    sysm.close()
    return res


def adjoint_c5(torch, synthetic):
    """Forward solve + adjoint gradient w.r.t. 5k values, 6k weights, 3k normal
    components at 128^3 (multigrid-preconditioned solves), loss at 2k held-out
    points."""
    from loopstructural_amd.adjoint import FDProblem

    values, normals = synthetic.constraints(5000, 1000, 1005)
    rng = np.random.default_rng(1005)
    held = rng.uniform(0.02, 0.98, (2000, 3))
    f_held, _ = synthetic.field1(held)
    prob = FDProblem((0, 0, 0), (127,) * 3, (1 / 127,) * 3, values[:, :3], normals[:, :3],
                     regularisation=0.1, precision="mg", rtol=1e-10)
    dev = prob.dev
    target = torch.tensor(f_held, device=dev)

    def run():
        ps = [torch.tensor(values[:, 3], device=dev, requires_grad=True),
              torch.ones(5000, dtype=torch.float64, device=dev, requires_grad=True),
              torch.tensor(normals[:, 3:6], device=dev, requires_grad=True),
              torch.ones(1000, dtype=torch.float64, device=dev, requires_grad=True)]
        loss = torch.mean((prob.evaluate(prob.solve(*ps), held) - target) ** 2)
        loss.backward()
        return loss

    dt, loss = _timed(torch, run, reps=2)
    res = {"ms_forward_backward": round(dt * 1e3, 1), "loss": float(loss),
           "parameters": 5000 + 5000 + 3000 + 1000}
    log(f"adjoint: {res}")
    return res


def host_env():
    """CPU model, visible cores and the BLAS/OMP thread settings of this host."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    keys = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "BLIS_NUM_THREADS")
    return {"cpu_model": model, "cores_visible": len(os.sched_getaffinity(0)),
            "thread_env": {k: os.environ.get(k) for k in keys},
            "note": "scipy's CSR SpMV is single-threaded: CG runs on one core"}


def c1_phases():
    """The 46^3-node FD workload (200 values, 50 normals of
    f = z + 0.1 sin 2 pi x cos 2 pi y, regularisation 0.1, cg at rtol 1e-10):
    a full CPU solve by the oracle's restatement of the reference path, timed
    phase by phase under the reference's timing keys, beside the GPU facade's
    own timing dict for the same solve (second call, kernels loaded)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from scipy.sparse import linalg as spla

    from loopstructural_amd import FiniteDifferenceInterpolator, StructuredGrid, synthetic
    from oracle import fd as ofd

    values, normals = synthetic.constraints(200, 50, 1001)
    gpu = None
    for _ in range(2):
        it = FiniteDifferenceInterpolator(StructuredGrid(origin=(0, 0, 0), nsteps=(45, 45, 45),
                                                         step_vector=(1 / 45,) * 3))
        it.set_value_constraints(values)
        it.set_normal_constraints(normals)
        it.setup_interpolator(regularisation=0.1)
        t0 = time.perf_counter()
        it.solve_system("cg", solver_kwargs={"rtol": 1e-10, "atol": 0})
        wall = time.perf_counter() - t0
        gpu = {k: v for k, v in it.get_last_solve_timing().items()
               if isinstance(v, (int, float, bool))}
        gpu["wall_seconds"] = wall
    og = ofd.Grid((0, 0, 0), (45, 45, 45), (1 / 45,) * 3)
    t0 = time.perf_counter()
    A, b = ofd.assemble_fd(og, {"value": values, "normal": normals},
                           regularisation=0.1).build_matrix()
    t1 = time.perf_counter()
    A, b, S = ofd.preprocess(A, b)
    t2 = time.perf_counter()
    count = [0]

    def cb(_):
        count[0] += 1

    y, info = spla.cg(A.T @ A, A.T @ b, rtol=1e-10, atol=0.0, callback=cb)
    c = S @ y
    t3 = time.perf_counter()
    cpu = {"assembly_seconds": t1 - t0, "preprocess_seconds": t2 - t1,
           "solve_seconds": t3 - t2, "total_seconds": t3 - t0, "cg_iterations": count[0],
           "cg_info": int(info), "matrix_rows": int(A.shape[0]), "matrix_nnz": int(A.nnz)}
    rel = float(np.linalg.norm(it.c - c) / np.linalg.norm(c))
    log(f"C1 phases: cpu total {cpu['total_seconds']:.1f}s ({count[0]} it), gpu wall "
        f"{gpu['wall_seconds']:.3f}s, rel diff {rel:.1e}")
    return {"workload": "46^3 nodes, 200 values, 50 normals, cg rtol 1e-10", "cpu": cpu,
            "gpu": gpu, "rel_l2_gpu_vs_cpu": rel, "host": host_env(), "cpu_kind": "port"}


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
            "host_cpus_visible": len(os.sched_getaffinity(0)), "host": host_env()}


if __name__ == "__main__":
    main()
