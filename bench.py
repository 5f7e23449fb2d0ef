"""Benchmark of the self-gravity Poisson solve on one MI355X.

Headline workload (the metric's "mixed-prec+batched" configuration): the block-structured AMR
hierarchy of the reference's fmg_poisson path -- 16^3 base blocks of 8^3 cells (128^3) plus four
nested refined levels, each covering the central half-width of its parent (20 480 blocks,
9 437 184 leaf cells), singular isothermal sphere source with a one-cell core and 5 % seeded
lognormal perturbations, multipole (sphere) boundary -- solved to relative residual
max|f - L phi| / max|f| < 1e-10.  One step = one full solve from phi = 0, inputs resident in HBM.
`value` is ms per solve of the mixed-precision solver (FP32 inner fmg_lin, FP64 outer residual and
update, hipGraph-replayed outer iteration); the plain FP64 GPU solve is timed beside it and the
CPU oracle (FP64, 1 core) on the same input is the CPU baseline.  The residual is reported in the
survey's three norms (L-inf, the reference's h^2-weighted L-inf of fmg.F90:1017-1039, volume-
weighted L2).

The other configurations run as sub-objects: "uniform" (periodic 256^3 clumpy density, the
mg.F90 path), "batch" (48 uneven periodic problems), "adjoint" (gradient of a sampled-force loss on
a periodic 128^3 grid) and "c0" (the periodic 64^3 CPU-runnable case).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--only SECTION]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/Poisson solve to rel. residual 1e-10: mixed-prec+batched vs FP64 GPU; % HBM peak"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed AMR solves (the headline steps)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256, help="uniform-grid size of the 'uniform' section")
    ap.add_argument("--uniform-steps", type=int, default=10)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baselines")
    ap.add_argument("--smoother-reps", type=int, default=50)
    ap.add_argument("--no-uniform", action="store_true", help="skip the periodic 256^3 section")
    ap.add_argument("--no-batch", action="store_true", help="skip the 48-problem uneven batch")
    ap.add_argument("--no-extra", action="store_true", help="skip the adjoint, 64^3 and block-memory sections")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--section", choices=["amr", "batch", "uniform", "adjoint", "c0", "blockmem"], default=None,
                    help="run only one section and print its JSON (profiling runs)")
    ap.add_argument("--amr-steps", type=int, default=None, help="alias of --steps for the AMR section")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (reference reading for HBM counters)")
    return ap.parse_args()


def calibrate_copy(mib):
    """5 device-to-device copies of a float32 tensor of `mib` MiB: a kernel whose HBM traffic is
    known (mib MiB read + mib MiB written per launch), to scale the FETCH_SIZE / WRITE_SIZE
    counters of the solver's kernels."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_mib": mib, "launches": 5}), flush=True)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


AMR_WORKLOAD = ("sphere AMR: 128^3 base (16^3 blocks of 8^3), 4 nested levels each refining the central "
                "half-width of its parent (20 480 blocks, 9 437 184 leaf cells), singular isothermal sphere "
                "(1-cell core, 5% lognormal), multipole BC, rel. residual 1e-10")


def _profile_traffic(name, kernel):
    """Scaled HBM bytes per launch of `kernel` from a committed rocprofv3 counter summary
    (lines "traffic_bytes_per_launch[<kernel>]: <bytes>")."""
    prof = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(prof):
        return None
    key = f"traffic_bytes_per_launch[{kernel}]:"
    for line in open(prof):
        if line.startswith(key):
            return int(line[len(key):])
    return None


def bench_amr(a, dev):
    """The headline: the sphere hierarchy solved by the plain FP64 solver and by the mixed solver;
    K timed mixed solves bracketed by device synchronisation."""
    import numpy as np
    import torch
    from code_only_amd import amr_hierarchy as ah
    from code_only_amd import problems
    from code_only_amd.amr import SG_PREC_MIXED, AmrSolver

    steps = a.amr_steps if a.amr_steps is not None else a.steps
    hier = ah.nested(base=16, levels=5, bc=ah.SPHERE)
    src = problems.sis_source(hier, seed=1)
    solver = AmrSolver(hier)
    s_t = torch.tensor(src, device=dev)
    g_t = torch.zeros((hier.nblk, 3, 8, 8, 8), dtype=torch.float64, device=dev)
    p_t = torch.zeros_like(s_t)
    log(f"[bench] AMR sphere hierarchy: {hier.nblk} blocks, {hier.cells()} leaf cells, plain FP64")

    def run(prec, k, w):
        st = None
        for _ in range(max(1, w)):
            p_t.zero_()
            st = solver.solve(p_t, s_t, g_t, tol=a.tol, maxit=30, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            p_t.zero_()
            st = solver.solve(p_t, s_t, g_t, tol=a.tol, maxit=30, precision=prec)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / k, st

    # the plain FP64 solve with its base-grid mg_cycle in one workgroup (as the reference's code
    # shape maps most directly), then spread over the chip (the default; same values)
    solver.set_base_solver(False)
    ms64_wg, _ = run(1, 1, 1)
    solver.set_base_solver(True)
    # and with the mixed solve's launch structure in FP64 (fused refresh, merged VMG steps: the
    # same values), so that mixed / this is the part of the speed-up precision buys
    solver.set_fp64_refresh(True)
    ms64_fz, st_fz = run(1, max(1, min(steps, 3)), 1)
    pfz = p_t.clone()
    solver.set_fp64_refresh(False)
    ms64, st = run(1, max(1, min(steps, 3)), 1)
    log(f"[bench] AMR fp64: {ms64:.2f} ms/solve ({ms64_wg:.2f} with the one-workgroup base solve, "
        f"{ms64_fz:.2f} with the fused refresh), {st['iterations']} it, rel {st['rel_linf']:.3e}")
    p64, g64 = p_t.clone(), g_t.clone()
    fz_same = bool(torch.equal(pfz, p64)) and st_fz["iterations"] == st["iterations"]
    msmx, stm = run(SG_PREC_MIXED, max(1, steps), a.warmup)
    err_p = ((p_t - p64).norm() / p64.norm()).item()
    err_g = ((g_t - g64).norm() / g64.norm()).item()
    log(f"[bench] AMR mixed: {msmx:.2f} ms/solve, {stm['iterations']} it, rel {stm['rel_linf']:.3e}, "
        f"psi err vs fp64 {err_p:.2e}")
    info = solver.info()
    # dominant kernel: the FMG-level-0 relaxation of the mixed solve (compact form, one colour per
    # launch), HIP events on the library's stream over whole sweeps (both colours); algorithmic
    # bytes of a launch = read the other colour + read f + write the colour = 1.5 x 4 B per leaf
    # cell (FP32)
    ms_sweep = solver.time_smoother(SG_PREC_MIXED, a.smoother_reps)
    tiles = solver.tile_info()
    nt = tiles["relax_fp32"]["threads_per_block"]
    launch_bytes = 6 * hier.cells()
    ms_launch = ms_sweep / 2
    achieved = launch_bytes / (ms_launch * 1e-3) / 1e9
    kname = f"sg::amr::k_blk8<true, float, {nt}, true>"
    traffic = _profile_traffic("amr_c2_hbm_traffic.md", kname)
    log(f"[bench] AMR smoother {ms_launch * 1e3:.1f} us per colour launch -> {achieved:.0f} GB/s algorithmic")
    ms64_sweep = solver.time_smoother(1, max(5, a.smoother_reps // 5))
    out = {
        "workload": AMR_WORKLOAD, "blocks": hier.nblk, "leaf_cells": hier.cells(),
        "mixed_gpu_ms": round(msmx, 3), "fp64_gpu_ms": round(ms64, 3), "speedup_mixed_vs_fp64": round(ms64 / msmx, 3),
        "fp64_gpu_ms_one_workgroup_base_solve": round(ms64_wg, 3),
        "fp64_gpu_ms_fused_refresh": round(ms64_fz, 3), "fp64_fused_refresh_bitwise_equal": fz_same,
        "speedup_mixed_vs_fp64_fused_refresh": round(ms64_fz / msmx, 3),
        "speedup_note": "against the reference's staged FP64 sequence; the FP64 solve with the mixed solve's "
                        "launch structure (same values) isolates the rest: FP32 inner fmg_lin plus its exact "
                        "transform base solve in place of the reference's base mg_cycle",
        "iterations": {"mixed": stm["iterations"], "fp64": st["iterations"]},
        "rel_residual": {"mixed_linf": stm["rel_linf"], "mixed_h2_linf": stm["rel_h2"], "mixed_l2": stm["rel_l2"],
                         "fp64_linf": st["rel_linf"], "fp64_h2_linf": st["rel_h2"], "fp64_l2": st["rel_l2"]},
        "history": {"mixed": stm["hist"], "fp64": st["hist"]},
        "mixed_vs_fp64_gpu": {"psi_rel_l2": err_p, "g_rel_l2": err_g},
        "kernel_nodes_per_outer_iteration": info["graph_nodes"],
        "fused": {"fmg_levels": info["fused_levels"], "vmg_gather_levels": info["vmg_fill_levels"]},
        "autotune": tiles,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kname + " (FMG level 0 relaxation, one colour per launch)",
                     "algorithmic_bytes_per_launch": launch_bytes, "ms_per_launch": ms_launch,
                     "traffic_vs_algorithmic": round(traffic / launch_bytes, 3) if traffic else None,
                     "hbm_rate_gbs": round(traffic / (ms_launch * 1e-3) / 1e9, 1) if traffic else None,
                     "fp64_ms_per_sweep": ms64_sweep},
    }
    cpu = None
    if not a.no_cpu:
        import oracle
        log("[bench] AMR CPU oracle (FP64, 1 core) on the same hierarchy")
        o = oracle.Amr(hier)
        t0 = time.perf_counter()
        po, go_, so = o.solve(src, tol=a.tol, maxit=30)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        pg, gg = p64.cpu().numpy(), g64.cpu().numpy()
        pm, gm = p_t.cpu().numpy(), g_t.cpu().numpy()
        out["fp64_gpu_vs_oracle"] = {"psi_rel_l2": float(np.linalg.norm(pg - po) / np.linalg.norm(po)),
                                     "g_rel_l2": float(np.linalg.norm(gg - go_) / np.linalg.norm(go_)),
                                     "bitwise": bool(np.array_equal(pg, po) and np.array_equal(gg, go_))}
        out["mixed_gpu_vs_oracle"] = {"psi_rel_l2": float(np.linalg.norm(pm - po) / np.linalg.norm(po)),
                                      "g_rel_l2": float(np.linalg.norm(gm - go_) / np.linalg.norm(go_))}
        cpu = {"value": round(cpu_ms, 1), "unit": "ms/solve", "cores": 1, "kind": "port",
               "sample": f"one full solve of the same sphere hierarchy to rel. residual {a.tol:g} "
                         f"({so['iterations']} outer iterations) by the FP64 C oracle on 1 core of {cpu_model()}"}
        log(f"[bench] AMR cpu oracle {cpu_ms:.0f} ms; parity {out['fp64_gpu_vs_oracle']}")
        o.close()
    solver.close()
    return out, cpu


def bench_batch(a, dev):
    """48 independent periodic problems, sizes drawn uniformly from 32, 64, 96, 128, 192 (seeded),
    each a clumpy density with 1..16 clumps, all solved together to rel. residual 1e-10: the
    mixed-precision batch (fused FP32 inner solve, graph-replayed) against the plain FP64 batch."""
    import numpy as np
    import torch
    from code_only_amd import problems
    from code_only_amd.batch import SG_PREC_FP64, SG_PREC_MIXED, BatchSolver

    ns, rhs = problems.batch_problems(48, seed=0, device=dev)
    cells = sum(n ** 3 for n in ns)
    log(f"[bench] batch: 48 problems, sizes {sorted(set(ns))}, {cells} cells")
    solver = BatchSolver(ns)

    def run(prec):
        psis = [torch.zeros_like(r) for r in rhs]
        st = solver.solve(psis, rhs, tol=a.tol, maxit=30, precision=prec)  # warm-up (autotune, plan, graph)
        times = []
        for _ in range(max(1, a.batch_steps)):
            for p in psis:
                p.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = solver.solve(psis, rhs, tol=a.tol, maxit=30, precision=prec)
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(times)), st, psis

    ms64, st64, p64 = run(SG_PREC_FP64)
    log(f"[bench] batch fp64: {ms64:.2f} ms")
    msmx, stmx, pmx = run(SG_PREC_MIXED)
    log(f"[bench] batch mixed: {msmx:.2f} ms")
    worst = 0.0
    for q in range(len(ns)):
        x, y = pmx[q] - pmx[q].mean(), p64[q] - p64[q].mean()
        worst = max(worst, (torch.linalg.norm(x - y) / torch.linalg.norm(y)).item())
    out = {"workload": "48 periodic problems, n uniform in {32, 64, 96, 128, 192} (seed 0), 1-16 clumps each",
           "sizes": ns, "cells": cells,
           "fp64_gpu_ms": round(ms64, 3), "mixed_gpu_ms": round(msmx, 3),
           "speedup_mixed_vs_fp64": round(ms64 / msmx, 3),
           "iterations": {"mixed": sorted(set(s["iterations"] for s in stmx)),
                          "fp64": sorted(set(s["iterations"] for s in st64))},
           "max_rel_linf": {"mixed": max(s["rel_linf"] for s in stmx), "fp64": max(s["rel_linf"] for s in st64)},
           "all_converged": all(s["status"] == 0 for s in stmx + st64),
           "mixed_vs_fp64_gpu_worst_psi_rel_l2": worst,
           "tiles": {str(v): solver.tile_info(v) for v in range(4)}}
    for v in out["tiles"].values():
        v.pop("ms", None)
    log(f"[bench] batch: mixed vs fp64 worst rel L2 {worst:.2e}, speedup {ms64 / msmx:.2f}")
    if not a.no_cpu:
        import oracle
        log("[bench] batch CPU oracle (FP64, 1 core) on all 48 problems")
        host = [r.cpu().numpy() for r in rhs]
        t0 = time.perf_counter()
        sols = [oracle.mg_cycle(f, 1.0 / n, tol=a.tol, maxit=30) for f, n in zip(host, ns)]
        cpu_ms = (time.perf_counter() - t0) * 1e3
        bit = all(np.array_equal(p64[q].cpu().numpy(), sols[q][0]) for q in range(len(ns)))
        wm = 0.0
        for q in range(len(ns)):
            po = sols[q][0] - sols[q][0].mean()
            pm = pmx[q].cpu().numpy()
            pm = pm - pm.mean()
            wm = max(wm, float(np.linalg.norm(pm - po) / np.linalg.norm(po)))
        out["cpu_oracle_ms"] = round(cpu_ms, 1)
        out["fp64_gpu_bitwise_vs_oracle"] = bit
        out["mixed_vs_oracle_worst_psi_rel_l2"] = wm
        log(f"[bench] batch cpu oracle {cpu_ms:.0f} ms; fp64 bitwise {bit}; mixed vs oracle worst {wm:.2e}")
    solver.close()
    return out


def bench_uniform(a, dev):
    """Periodic 256^3 clumpy density (the mg.F90 path): plain FP64 against the mixed solver, the
    fine-level sweep's roofline, the autotuner's tile timings, the CPU oracle."""
    import numpy as np
    import torch

    from code_only_amd import problems
    from code_only_amd.uniform import SG_PREC_FP64, SG_PREC_MIXED, UniformSolver
    from code_only_amd import _lib

    n = a.n
    h = 1.0 / n
    log(f"[bench] generating clumpy rhs n={n}")
    f = problems.clumpy_rhs(n, seed=0)
    ft = torch.tensor(f, device=dev)
    solver = UniformSolver(n, h, precisions=SG_PREC_FP64 | SG_PREC_MIXED)

    def solve(prec, steps, warm):
        psi = torch.zeros_like(ft)
        st = None
        for _ in range(warm):
            psi.zero_()
            st = solver.cycle(psi, ft, tol=a.tol, maxit=60, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            psi.zero_()
            st = solver.cycle(psi, ft, tol=a.tol, maxit=60, precision=prec)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / max(steps, 1)
        return ms, st, psi

    log("[bench] plain FP64 GPU solve")
    ms64, st64, psi64 = solve(SG_PREC_FP64, max(1, a.uniform_steps // 2), 1)
    log(f"[bench] fp64: {ms64:.3f} ms/solve, {st64['iterations']} it, rel {st64['rel_linf']:.3e}")
    log("[bench] mixed-precision GPU solve")
    msmx, stmx, psimx = solve(SG_PREC_MIXED, a.uniform_steps, 2)
    log(f"[bench] mixed: {msmx:.3f} ms/solve, {stmx['iterations']} it, rel {stmx['rel_linf']:.3e}")

    # dominant kernel: the fine-level smoother of the mixed solve, timed with HIP events on the
    # library's stream.  Algorithmic bytes of one full sweep = read u + read f + write u = 3 * 4 B
    # per fine cell (FP32).
    import ctypes
    ms_sweep = ctypes.c_double()
    _lib.check(solver._lib.sg_uniform_time_smoother(solver._ctx, SG_PREC_MIXED, a.smoother_reps,
                                                     ctypes.byref(ms_sweep)))
    sweep_bytes = 3 * 4 * n ** 3
    achieved = sweep_bytes / (ms_sweep.value * 1e-3) / 1e9
    log(f"[bench] smoother sweep {ms_sweep.value * 1e3:.1f} us -> {achieved:.0f} GB/s algorithmic")
    # The same sweep with a cold last-level cache: the 256^3 FP32 working set of one sweep (u, f and
    # the new u: 192 MiB) fits in MI355X's 256 MiB Infinity Cache, so back-to-back sweeps (and the
    # solve's own) partly hit it.  Before each timed sweep a 1 GiB fill evicts it; each sweep is
    # timed alone with HIP events, so the figure includes one launch's start-up.
    flush = torch.empty(1 << 28, dtype=torch.float32, device=dev)
    cold = []
    ms1 = ctypes.c_double()
    for r in range(a.smoother_reps // 5 + 1):
        flush.fill_(float(r))
        torch.cuda.synchronize()
        _lib.check(solver._lib.sg_uniform_time_smoother(solver._ctx, SG_PREC_MIXED, 1, ctypes.byref(ms1)))
        cold.append(ms1.value)
    del flush
    ms_cold = sorted(cold)[len(cold) // 2]
    achieved_cold = sweep_bytes / (ms_cold * 1e-3) / 1e9
    log(f"[bench] smoother sweep, cache cold: {ms_cold * 1e3:.1f} us -> {achieved_cold:.0f} GB/s algorithmic")

    tiles = {}
    dims = [(32, 16, 16, 1024), (16, 8, 8, 1024), (16, 8, 8, 256), (64, 8, 8, 512), (8, 8, 8, 512)]
    # streaming shapes: (SX, SY, ZC, threads)
    sdims = [(64, 16, 32, 512), (64, 16, 16, 512), (64, 32, 32, 512), (32, 16, 32, 256), (128, 16, 32, 512),
             (32, 16, 16, 256), (32, 8, 16, 256), (32, 32, 32, 512), (16, 16, 16, 256)]
    names = ["x".join(map(str, d[:3])) + f"/{d[3]}" for d in dims] + \
            ["stream " + "x".join(map(str, d[:2])) + f" z{d[2]}/{d[3]}" for d in sdims]
    lev, m = 0, n
    while m > 16:
        cfg, ns, fs = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = (ctypes.c_float * 32)()
        _lib.check(solver._lib.sg_uniform_tile_info(solver._ctx, lev, ctypes.byref(cfg), ctypes.byref(ns),
                                                    ctypes.byref(fs), ms, 32))
        cand = {}
        for q in range(4 * len(dims) + len(sdims)):
            if q < 4 * len(dims):
                key = names[q // 4] + f"/{1 + (q // 2) % 2}sw" + ("/fLDS" if q % 2 else "")
            else:
                key = names[len(dims) + q - 4 * len(dims)]
            if ms[q] > 0:
                cand[key] = round(ms[q] * 1e3, 1)
        tiles[str(m)] = {"shape": names[cfg.value], "sweeps_per_launch": ns.value, "source_in_lds": bool(fs.value),
                         "us_per_vcycle_visit": cand}
        lev += 1
        m //= 2
    log(f"[bench] autotuned tiles {json.dumps(tiles)}")

    # HBM bytes per launch of the roofline kernel (the autotuned shape at this size), from the
    # committed rocprofv3 counter summary: lines "traffic_bytes_per_launch[<shape>]: <bytes>"
    traffic = None
    prof = os.path.join(ROOT, "profiles", "uniform256_hbm_traffic.md")
    shape = tiles.get(str(n), {}).get("shape")
    if shape and tiles[str(n)]["source_in_lds"]:
        shape += "/fLDS"
    if n == 256 and shape and os.path.exists(prof):
        key = f"traffic_bytes_per_launch[{shape}]:"
        for line in open(prof):
            if line.startswith(key):
                traffic = int(line[len(key):])

    p64 = psi64 - psi64.mean()
    pmx = psimx - psimx.mean()
    err_mx_vs_64 = ((pmx - p64).norm() / p64.norm()).item()

    cpu = None
    err_vs_oracle = None
    if not a.no_cpu:
        import oracle
        log("[bench] CPU oracle (FP64, 1 core) on the same 256^3 input")
        t0 = time.perf_counter()
        psi_o, st_o = oracle.mg_cycle(f, h, tol=a.tol, maxit=60)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        po = psi_o - psi_o.mean()
        po_t = torch.tensor(po, device=dev)
        err_vs_oracle = ((pmx - po_t).norm() / po_t.norm()).item()
        cpu = {"value": round(cpu_ms, 1), "unit": "ms/solve", "cores": 1, "kind": "port",
               "sample": f"one full {n}^3 solve to rel. residual {a.tol:g} ({st_o['iterations']} outer "
                         f"iterations) by the FP64 C oracle on 1 core of {cpu_model()}"}
        log(f"[bench] cpu oracle {cpu_ms:.0f} ms, mixed vs oracle rel L2 {err_vs_oracle:.2e}")

    out = {
        "workload": f"periodic {n}^3 clumpy-density Poisson solve to rel. residual {a.tol:g}",
        "n": n, "clumps": 8, "sigma_cells": 4, "seed": 0,
        "mixed_gpu_ms": round(msmx, 4), "fp64_gpu_ms": round(ms64, 4),
        "speedup_mixed_vs_fp64": round(ms64 / msmx, 3),
        "iterations": {"mixed": stmx["iterations"], "fp64": st64["iterations"]},
        "rel_residual": {"mixed_linf": stmx["rel_linf"], "mixed_l2": stmx["rel_l2"], "fp64_linf": st64["rel_linf"]},
        "phi_err_vs_fp64_gpu": err_mx_vs_64,
        "phi_err_vs_cpu_oracle": err_vs_oracle,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "fine-level FP32 smoother sweep (both colours), " + (shape or "?"),
                     "algorithmic_bytes_per_launch": sweep_bytes,
                     "ms_per_launch": ms_sweep.value,
                     "cache_cold": {"achieved": round(achieved_cold, 1), "frac": round(achieved_cold / HBM_PEAK_GBS, 4),
                                    "ms_per_launch": ms_cold,
                                    "note": "Infinity Cache evicted by a 1 GiB fill before each sweep, median"}},
        "cpu_oracle": cpu,
        "autotune": tiles,
    }
    solver.close()
    return out


def bench_adjoint(a, dev):
    """Gradient of L = sum_p w_p . grad_h phi(c_p) (4096 seeded cells, w ~ N(0,1)^3) with respect to
    rho on a periodic 128^3 grid: forward solve, then the backward pass (the hand-written adjoint
    solve and the transpose of the sampled gradient), both in mixed precision."""
    import numpy as np
    import torch
    from code_only_amd import problems
    from code_only_amd.adjoint import sampled_gradient_loss
    from code_only_amd.uniform import SG_PREC_FP64, SG_PREC_MIXED, UniformSolver

    n = 128
    rng = np.random.default_rng(4)
    rho0 = problems.clumpy_density(n, seed=4, nclump=8)
    cells = torch.tensor(rng.integers(0, n ** 3, size=4096).astype(np.int32), device=dev)
    w = torch.tensor(rng.standard_normal((4096, 3)), device=dev)
    s = UniformSolver(n, precisions=SG_PREC_FP64 | SG_PREC_MIXED)

    def one(prec):
        rho = torch.tensor(rho0, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = sampled_gradient_loss(rho, s, cells, w, tol=a.tol, precision=prec)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return (t1 - t0) * 1e3, (t2 - t1) * 1e3, rho.grad

    one(SG_PREC_MIXED)
    fw, bw, gm = [], [], None
    for _ in range(3):
        f, b, gm = one(SG_PREC_MIXED)
        fw.append(f)
        bw.append(b)
    f64, b64, g64 = one(SG_PREC_FP64)
    err = ((gm - g64).norm() / g64.norm()).item()
    s.close()
    out = {"workload": "periodic 128^3, L = sum over 4096 seeded cells of w . grad phi, dL/drho by the "
                       "hand-written adjoint", "mixed_forward_ms": round(float(np.median(fw)), 3),
           "mixed_adjoint_ms": round(float(np.median(bw)), 3), "fp64_forward_ms": round(f64, 3),
           "fp64_adjoint_ms": round(b64, 3), "grad_mixed_vs_fp64_rel_l2": err}
    log(f"[bench] adjoint 128^3: forward {out['mixed_forward_ms']} ms, adjoint {out['mixed_adjoint_ms']} ms (mixed)")
    out["amr"] = bench_amr_adjoint(a, dev)
    return out


def bench_amr_adjoint(a, dev):
    """The sphere-boundary AMR adjoint where it converges (one refinement level): 16^3 base blocks plus
    one nested level, d(sum w psi^2)/d src through AmrPoissonSolve; the first backward pass includes
    the host assembly of the transposed operator, the timed ones do not."""
    import numpy as np
    import torch
    from code_only_amd import amr_hierarchy as ah
    from code_only_amd import problems
    from code_only_amd.adjoint import AmrPoissonSolve
    from code_only_amd.amr import SG_PREC_FP64, SG_PREC_MIXED, AmrSolver

    h = ah.nested(base=16, levels=2, bc=ah.SPHERE)
    g = AmrSolver(h)
    rng = np.random.default_rng(6)
    src0 = problems.sis_source(h, seed=6)
    w = torch.tensor(rng.standard_normal((h.nblk, 8, 8, 8)), device=dev)

    def one(prec):
        src = torch.tensor(src0, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        psi = AmrPoissonSolve.apply(src, g, a.tol, prec)
        loss = torch.sum(w * psi * psi)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        return (t1 - t0) * 1e3, (time.perf_counter() - t1) * 1e3, src.grad

    t_asm = time.perf_counter()
    one(SG_PREC_MIXED)
    first = (time.perf_counter() - t_asm) * 1e3
    fm, bm, gm = one(SG_PREC_MIXED)
    f64, b64, g64 = one(SG_PREC_FP64)
    out = {"workload": f"sphere AMR, 16^3 base blocks + 1 nested level ({h.nblk} blocks, {h.cells()} leaf cells), "
                       "d(sum w psi^2)/d src", "mixed_forward_ms": round(fm, 2), "mixed_adjoint_ms": round(bm, 2),
           "fp64_forward_ms": round(f64, 2), "fp64_adjoint_ms": round(b64, 2),
           "first_call_with_assembly_ms": round(first, 1),
           "grad_mixed_vs_fp64_rel_l2": ((gm - g64).norm() / g64.norm()).item()}
    g.close()
    log(f"[bench] AMR adjoint: forward {out['mixed_forward_ms']} ms, adjoint {out['mixed_adjoint_ms']} ms (mixed)")
    return out


def bench_blockmem(a, dev):
    """The C2 hierarchy with its hydro state in a device pool of per-grid arrays (ideal-gas HD layout,
    9 components, 2 ghost layers; own, predictor and past slots per grid, shuffled): the mixed solve
    staged from and written back to the pool against the same solve on dense arrays, and the
    gravity source step (source_g_all_level) over the pool."""
    import numpy as np
    import torch
    from code_only_amd import amr_hierarchy as ah
    from code_only_amd import problems
    from code_only_amd.amr import SG_PREC_MIXED, AmrSolver

    h = ah.nested(base=16, levels=5, bc=ah.SPHERE)
    lay = problems.BLOCKMEM_LAYOUTS["hd"]
    nb, W = h.nblk, 8 + 2 * lay["ng"]
    nslots = 3 * nb
    rng = np.random.default_rng(2)
    perm = rng.permutation(nslots).astype(np.int32)
    su, s1, s2 = perm[:nb], perm[nb:2 * nb], perm[2 * nb:]
    pool = torch.empty((nslots, lay["stride"]), dtype=torch.float64, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(2)
    pool.uniform_(0.5, 1.5, generator=gen)
    comp = pool.view(nslots, lay["ncomp"], W, W, W)
    for m in ("mvx", "mvy", "mvz"):
        comp[:, lay[m]].uniform_(-1.0, 1.0, generator=gen)
    src = torch.tensor(problems.sis_source(h, seed=1), device=dev)
    su_t = torch.tensor(su.astype(np.int64), device=dev)
    ins = (slice(lay["ng"], lay["ng"] + 8),) * 3
    comp[(su_t, lay["mrho"]) + ins] = src
    s = AmrSolver(h)
    s.set_blockmem(lay, nslots, su, s1, s2)

    def staged():
        comp[(su_t, lay["mpsi"]) + ins] = 0.0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = s.solve_blockmem(pool, 1.0, source=0, tol=a.tol, maxit=30, precision=SG_PREC_MIXED)
        return (time.perf_counter() - t0) * 1e3, st

    def dense():
        psi = torch.zeros_like(src)
        g = torch.empty((nb, 3, 8, 8, 8), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = s.solve(psi, src, g, tol=a.tol, maxit=30, precision=SG_PREC_MIXED)
        return (time.perf_counter() - t0) * 1e3, st

    staged()
    dense()
    ts, td = [], []
    for _ in range(3):
        t, st_s = staged()
        ts.append(t)
        t, st_d = dense()
        td.append(t)
    dt = [2e-3 / 2 ** l for l in range(h.nlev)]
    s.source_g(pool, dt)
    reps = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        s.source_g(pool, dt)
    src_ms = (time.perf_counter() - t0) * 1e3 / reps
    # algorithmic bytes of the source step per grid: all W^3 cells of 9 components read and
    # written, the interior's density and velocity of the predictor stage read
    per_grid = 2 * 9 * W ** 3 * 8 + 4 * 512 * 8
    gbs = per_grid * nb / (src_ms * 1e-3) / 1e9

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) * 1e3 / reps
    # the hydro steps around the solve, every level of the hierarchy (the reference calls them
    # level by level in the time step)
    hydro = {"boundary_grid_ms": [round(timed(lambda l=l: s.hydro_boundary(pool, l, 0)), 4) for l in range(h.nlev)],
             "fg2cg_ms": [round(timed(lambda l=l: s.fg2cg(pool, l)), 4) for l in range(1, h.nlev)]}
    fb = torch.randn((nb, 2, 6, 9, 64), dtype=torch.float64, device=dev, generator=gen)
    hydro["fluxcorrection_ms"] = [round(timed(lambda l=l: s.fluxcorrection(pool, fb, l, dt)), 4)
                                  for l in range(h.nlev - 1)]
    del fb
    s.close()
    out = {"workload": "C2 sphere hierarchy (20 480 grids) in a device pool of ideal-gas HD arrays "
                       "(9 components, 12^3 with 2 ghost layers, 3 shuffled slots per grid, 7.6 GB)",
           "mixed_blockmem_ms": round(float(np.median(ts)), 3), "mixed_dense_ms": round(float(np.median(td)), 3),
           "staging_overhead_ms": round(float(np.median(ts) - np.median(td)), 3),
           "iterations": [st_s["iterations"], st_d["iterations"]],
           "rel_linf": [st_s["rel_linf"], st_d["rel_linf"]],
           "source_g_ms": round(src_ms, 4), "source_g_bytes": per_grid * nb,
           "source_g_gbs": round(gbs, 1), "source_g_frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
           "hydro_steps_per_level": hydro}
    log(f"[bench] blockmem: {out}")
    return out


def bench_c0(a, dev):
    """The periodic 64^3 case of the mg.F90 path: CPU oracle (FP64, 1 core), plain FP64 GPU and mixed."""
    import numpy as np
    import torch
    from code_only_amd import problems
    from code_only_amd.uniform import SG_PREC_FP64, SG_PREC_MIXED, UniformSolver

    n = 64
    f = problems.clumpy_rhs(n, seed=0)
    ft = torch.tensor(f, device=dev)
    s = UniformSolver(n, 1.0 / n, precisions=SG_PREC_FP64 | SG_PREC_MIXED)
    res = {}
    for name, prec in (("fp64_gpu_ms", SG_PREC_FP64), ("mixed_gpu_ms", SG_PREC_MIXED)):
        psi = torch.zeros_like(ft)
        st = s.cycle(psi, ft, tol=a.tol, maxit=60, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            psi.zero_()
            st = s.cycle(psi, ft, tol=a.tol, maxit=60, precision=prec)
        torch.cuda.synchronize()
        res[name] = round((time.perf_counter() - t0) * 1e3 / 5, 4)
        res[name.replace("_ms", "_rel_linf")] = st["rel_linf"]
    s.close()
    res["workload"] = "periodic 64^3 clumpy-density Poisson solve to rel. residual 1e-10"
    if not a.no_cpu:
        import oracle
        t0 = time.perf_counter()
        _, so = oracle.mg_cycle(f, 1.0 / n, tol=a.tol, maxit=60)
        res["cpu_oracle_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
        res["cpu_oracle_iterations"] = so["iterations"]
    log(f"[bench] 64^3: {res}")
    return res


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build's solve runs on a single MI355X "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device("cuda:0")
    if a.section:
        if a.section == "amr":
            res, cpu = bench_amr(a, dev)
            res["cpu_baseline"] = cpu
        else:
            res = {"batch": bench_batch, "uniform": bench_uniform, "adjoint": bench_adjoint, "c0": bench_c0,
                   "blockmem": bench_blockmem}[a.section](a, dev)
        print(json.dumps({"section": a.section, **res}), flush=True)
        return
    amr, cpu = bench_amr(a, dev)
    steps = a.amr_steps if a.amr_steps is not None else a.steps
    out = {
        "metric": METRIC,
        "value": amr["mixed_gpu_ms"],
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": amr["mixed_gpu_ms"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner + f64 outer",
        "data": "synthetic (seeded singular isothermal sphere with lognormal perturbations)",
        "config": {"workload": AMR_WORKLOAD, "blocks": amr["blocks"], "leaf_cells": amr["leaf_cells"],
                   "base_blocks": 4096, "levels": 5, "boundary": "sphere multipole", "tol": a.tol},
        "fp64_gpu_ms": amr["fp64_gpu_ms"],
        "speedup_vs_fp64_gpu": amr["speedup_mixed_vs_fp64"],
        "fp64_gpu_ms_fused_refresh": amr["fp64_gpu_ms_fused_refresh"],
        "speedup_vs_fp64_gpu_fused_refresh": amr["speedup_mixed_vs_fp64_fused_refresh"],
        "iterations": amr["iterations"],
        "rel_residual": amr["rel_residual"],
        "roofline": amr["roofline"],
        "cpu_baseline": cpu,
        "amr": {k: v for k, v in amr.items() if k not in ("roofline",)},
    }
    if not a.no_uniform:
        out["uniform"] = bench_uniform(a, dev)
    if not a.no_batch:
        out["batch"] = bench_batch(a, dev)
    if not a.no_extra:
        out["adjoint"] = bench_adjoint(a, dev)
        out["c0"] = bench_c0(a, dev)
        out["blockmem"] = bench_blockmem(a, dev)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
