#!/usr/bin/env python3
"""Time-to-map of the GLS map-making solve on one MI355X.

A step is one complete drop-in solve of the workload from the caller's host
arrays (MLmap's arrays: global pixel indices, weights, signal, noise, invtt),
exactly as MLmap hands them over: upload to HBM, pointing build (MatInit),
Toeplitz plan, block-Jacobi preconditioner with the degenerate-pixel cut, PCG
to a relative residual of 1e-8, the map back in host memory and the
reference's side effects written into the caller's pix / weights / signal.
Default workload: 64 detectors x 2^20 samples on a 1024^2 grid, lambda = 8192,
2^17-sample Toeplitz blocks, FP64 (one rank's share).  Every step gets fresh
copies of the input arrays (the solve rewrites pix, weights and signal in
place); the copies are made outside the timed region.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]

Prints one JSON line.  Beside `value` it reports the same solve with the
inputs already resident in HBM (`device_resident`).  The roofline entry is for
the dominant kernel (the Toeplitz FFT tiles of N^-1): algorithmic bytes per
launch over its average launch duration, measured with HIP events on the
library's stream during the timed drop-in solves.  Further sections: mixed
precision (both refresh rules), the reference's own offset operator, the
two-level preconditioners, the CPU oracle on 1 and on all host cores, C1,
the noise-model builder, template map-making (C6), batched Monte Carlo
realizations (C4) and the request loop under Poisson arrivals (C5).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def algorithmic_bytes(m, n_map, w_t, w_p, unit_w0=True):
    """Compulsory HBM bytes per launch: pixel index + (2 or 3) weights + the TOD
    sample per sample for P and P^T, plus the map; one read and one write of the
    TOD for N^-1.  Map-sized vector kernels are not priced: their operands
    (n_map doubles each) stay resident in the 256 MB last-level cache."""
    nw = 2 if unit_w0 else 3
    bp = m * (4 + nw * w_p + w_t) + n_map * 8
    return {"P": bp, "Pt": bp, "tiles": 2 * m * w_t}


FP64_VECTOR_PEAK_TFLOPS = 78.6


def algorithmic_flops(block_sizes, lam):
    """FLOPs of one N^-1 launch as priced for the reference's tiles: per tile
    5 bs log2(bs) + 6 (bs/2 + 1) with bs = 2^ceil(log2 4 lambda) (4 lambda when
    that exceeds the block), bs - 2(lambda - 1) outputs per tile (toeplitz.c:200-206)."""
    import math

    f = 0.0
    for n in np.asarray(block_sizes, dtype=np.int64):
        if lam <= 1 or n < lam - 1:
            continue
        bs = 1 << math.ceil(math.log2(4 * lam))
        if bs > n:
            bs = 4 * lam
        eff = bs - 2 * (lam - 1)
        f += math.ceil(n / eff) * (5.0 * bs * math.log2(bs) + 6.0 * (bs // 2 + 1))
    return f


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ndet", type=int, default=64, help="detectors (64 = the full workload)")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--maxiter", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--cpu-ndet", type=int, default=8)
    ap.add_argument("--offset", choices=["symmetric", "reference", "none"], default="symmetric",
                    help="per-block offset handling of N^-1 for the timed solves")
    ap.add_argument("--no-reference-probe", action="store_true")
    ap.add_argument("--c4-real", type=int, default=32, help="batched Monte Carlo realizations (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=50, help="requests per serving rate (0: skip)")
    ap.add_argument("--serve-rates", default="1,2,5,10,20,50", help="Poisson arrival rates (req/s)")
    ap.add_argument("--serve-pool", type=int, default=16, help="distinct seeded requests cycled by the load")
    ap.add_argument("--no-noise-model", action="store_true", help="skip the invtt-from-noise section")
    ap.add_argument("--no-templates", action="store_true", help="skip the template map-making (MTmap) section")
    ap.add_argument("--twolvl-zn", type=int, default=8,
                    help="deflation space size of the two-level preconditioner section (0: skip)")
    ap.add_argument("--only-templates", action="store_true",
                    help="run only the template map-making section (C6) and print its JSON object")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="copy 1 GiB device to device once before the solves (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs one rank's share on one device "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    import torch

    from midapack_amd import lib as G
    from midapack_amd import workloads as W

    if args.only_templates:
        G.init(0)
        print(json.dumps({"templates_c6": templates_section(G, W, args)}), flush=True)
        return

    t = time.time()
    wl = W.config_c2(ndet=args.ndet)
    log(f"generated {wl.name}: m={wl.m} in {time.time() - t:.1f}s")
    dev = torch.device("cuda:0")
    G.init(0)
    d_pix = torch.from_numpy(wl.pix).to(dev)
    d_w = torch.from_numpy(wl.weights).to(dev)
    d_sig0 = torch.from_numpy(wl.signal).to(dev)
    d_sig = d_sig0.clone()
    d_noise = torch.from_numpy(wl.noise).to(dev)
    d_inv = torch.from_numpy(wl.invtt).to(dev)
    torch.cuda.synchronize()

    if args.copy_calibration:
        a = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    offset = {"symmetric": G.MGPU_OFFSET_SYMMETRIC, "reference": G.MGPU_OFFSET_REFERENCE,
              "none": G.MGPU_OFFSET_NONE}[args.offset]

    def solve(precision, profile=True, offset_mode=offset, maxiter=args.maxiter, refresh=10, precond=0, zn=0):
        d_sig.copy_(d_sig0)  # the cut zeroes signal samples in place (precond.c:605,610)
        torch.cuda.synchronize()
        pb = G.Problem.from_device(wl.m, d_pix.data_ptr(), d_w.data_ptr(), wl.block_sizes, wl.lam,
                                   d_inv.data_ptr(),
                                   opts=G.options(precision=precision, profile=profile, offset_mode=offset_mode,
                                                  refresh_every=refresh, precond=precond, z_2lvl=zn))
        x, hist, rep = pb.solve_device(d_sig.data_ptr(), d_noise.data_ptr(), tol=args.tol, maxiter=maxiter)
        kt = pb.kernel_times()
        pb.close()
        return x, hist, rep, kt

    # ---- the headline: drop-in solves from the caller's host arrays ---------
    hp, hw, hs = np.empty_like(wl.pix), np.empty_like(wl.weights), np.empty_like(wl.signal)

    def from_host(precision, profile=False, offset_mode=offset, maxiter=args.maxiter, refresh=10, tol=args.tol):
        """One drop-in solve on fresh copies of the inputs (made before the clock
        starts); returns (result, seconds)."""
        np.copyto(hp, wl.pix)
        np.copyto(hw, wl.weights)
        np.copyto(hs, wl.signal)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g = G.pcg_gls(hp, hw, hs, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=tol, maxiter=maxiter,
                      opts=G.options(precision=precision, profile=profile, offset_mode=offset_mode,
                                     refresh_every=refresh), inplace=True)
        torch.cuda.synchronize()
        return g, time.perf_counter() - t0

    for i in range(args.warmup):
        g, dt = from_host(G.MGPU_FP64)
        log(f"warmup {i}: {dt:.3f} s, {g.report['iterations']} iterations, converged={g.report['converged']}")
    G.dropin_kernel_times(reset=True)
    times, reps = [], []
    g = None
    for i in range(args.steps):
        g, dt = from_host(G.MGPU_FP64, profile=True)
        times.append(dt)
        reps.append(g.report)
    x64 = g.x
    kt_tot = G.dropin_kernel_times(reset=True)
    per_map = float(np.mean(times))
    log(f"FP64 from host arrays: {per_map:.3f} s per map (min {min(times):.3f}), iterations "
        f"{reps[-1]['iterations']}, converged {reps[-1]['converged']}")

    n_map = reps[-1]["n_map"]
    nbytes = algorithmic_bytes(wl.m, n_map, 8, 8)
    kernels = {}
    for k in G.KERNEL_CLASSES:
        ms, cnt = kt_tot[k]
        if cnt == 0:
            continue
        avg = ms / cnt
        ent = {"avg_ms": round(avg, 5), "launches": cnt, "total_ms_per_map": round(ms / args.steps, 3)}
        if nbytes.get(k):
            gbs = nbytes[k] / (avg * 1e-3) / 1e9
            ent.update(bytes=int(nbytes[k]), GBps=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4))
        kernels[k] = ent
    dominant = max((k for k in ("P", "tiles", "Pt") if k in kernels), key=lambda k: kernels[k]["total_ms_per_map"])
    dk = kernels[dominant]
    roofline = {"kernel": dominant, "bound": "hbm", "achieved": dk["GBps"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": dk["frac"], "traffic": measured_traffic(dominant, args)}
    # N^-1 sits at the FP64 vector ridge: its FLOP rate beside the bytes
    flops = algorithmic_flops(wl.block_sizes, wl.lam)
    roofline_flops = None
    if "tiles" in kernels and flops > 0:
        tf = flops / (kernels["tiles"]["avg_ms"] * 1e-3) / 1e12
        roofline_flops = {"kernel": "tiles", "bound": "fp64 vector", "achieved": round(tf, 2),
                          "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                          "frac": round(tf / FP64_VECTOR_PEAK_TFLOPS, 4), "flops_per_launch": flops}

    # ---- the same solve with the inputs resident in HBM ---------------------
    solve(G.MGPU_FP64, profile=False)
    ndev = max(1, min(args.steps, 3))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(ndev):
        _, _, rep_d, _ = solve(G.MGPU_FP64, profile=False)
    torch.cuda.synchronize()
    per_map_dev = (time.perf_counter() - t0) / ndev
    log(f"FP64 inputs resident in HBM: {per_map_dev:.3f} s per map")

    out = {
        "metric": "time-to-map at rel. residual 1e-8 (s)",
        "value": round(per_map, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(per_map * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded raster scan, 1/f noise, analytic invtt)",
        "config": {"workload": wl.name, "samples": wl.m, "pixels_seen": int(reps[-1]["n_map"] // 3 + reps[-1]["n_cut"]),
                   "lambda": wl.lam, "blocks": int(wl.block_sizes.size), "tol": args.tol,
                   "inputs": "host arrays (pageable numpy), side effects written back; per-step copies untimed",
                   "n_inv": {"symmetric": "C T C (symmetric per-block offset projection)",
                             "reference": "T C (reference stbmmProd mean removal)",
                             "none": "T"}[args.offset]},
        "iterations": reps[-1]["iterations"],
        "converged": bool(reps[-1]["converged"]),
        "n_cut": reps[-1]["n_cut"],
        "step_s": [round(t, 4) for t in times],
        "setup_ms": round(float(np.mean([r["t_setup_ms"] for r in reps])), 2),
        "solve_ms": round(float(np.mean([r["t_solve_ms"] for r in reps])), 2),
        "device_resident": {"time_to_map_s": round(per_map_dev, 4), "steps": ndev,
                            "iterations": int(rep_d.iterations),
                            "what": "same solve with pix, weights, signal, noise, invtt already in HBM"},
        "kernels": kernels,
        "roofline": roofline,
        "roofline_flops": roofline_flops,
    }

    if not args.no_mixed:
        try:
            # the converged map (FP64 to 1e-11): at tol 1e-8 the FP64 map itself
            # differs from it, which bounds what "error vs FP64" can resolve
            gt, _ = from_host(G.MGPU_FP64, tol=1e-11, maxiter=4 * args.maxiter)
            xt = gt.x
            err64 = float(np.max(np.abs(x64 - xt)) / np.max(np.abs(xt)))
            log(f"FP64 map at tol {args.tol:g} vs the converged map (tol 1e-11, {gt.report['iterations']} "
                f"iterations): {err64:.2e}")

            def mixed_entry(refresh):
                nm = max(1, min(args.steps, 5))
                from_host(G.MGPU_MIXED, refresh=refresh)
                tms = []
                for _ in range(nm):
                    gm, dt = from_host(G.MGPU_MIXED, refresh=refresh)
                    tms.append(dt)
                tm = float(np.mean(tms))
                solve(G.MGPU_MIXED, profile=False, refresh=refresh)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(ndev):
                    solve(G.MGPU_MIXED, profile=False, refresh=refresh)
                torch.cuda.synchronize()
                tmd = (time.perf_counter() - t0) / ndev
                err = float(np.max(np.abs(gm.x - x64)) / np.max(np.abs(x64)))
                errt = float(np.max(np.abs(gm.x - xt)) / np.max(np.abs(xt)))
                tag = "adaptive" if refresh == G.MGPU_REFRESH_ADAPTIVE else f"every {refresh}"
                log(f"mixed (refresh {tag}): {tm:.3f} s per map from host, {tmd:.3f} s resident, "
                    f"{gm.report['iterations']} iterations, {gm.report['refreshes']} refreshes, map error vs FP64 "
                    f"{err:.2e}, vs the converged map {errt:.2e}")
                return {"refresh": tag, "time_to_map_s": round(tm, 4), "speedup_vs_fp64": round(per_map / tm, 3),
                        "device_resident_s": round(tmd, 4),
                        "device_resident_speedup_vs_fp64": round(per_map_dev / tmd, 3),
                        "iterations": gm.report["iterations"], "converged": bool(gm.report["converged"]),
                        "refreshes": gm.report["refreshes"], "true_res_rel": gm.report["true_res_rel"],
                        "map_rel_err_vs_fp64": err, "map_rel_err_vs_converged": errt}

            out["mixed"] = mixed_entry(10)
            out["mixed"]["fp64_map_rel_err_vs_converged"] = err64
            out["mixed"]["adaptive"] = mixed_entry(G.MGPU_REFRESH_ADAPTIVE)
        except Exception as e:  # reported, not hidden
            out["mixed"] = {"error": str(e)}

    if not args.no_reference_probe and args.offset != "reference":
        # the reference's one-sided mean removal (T C) on the same workload: its
        # residual floor, then a timed solve stopped where the residual is lowest
        _, hr, rr, _ = solve(G.MGPU_FP64, profile=False, offset_mode=G.MGPU_OFFSET_REFERENCE, maxiter=800)
        kmin = int(np.nanargmin(hr))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, hr2, rr2, _ = solve(G.MGPU_FP64, profile=False, offset_mode=G.MGPU_OFFSET_REFERENCE,
                               maxiter=kmin + 1)
        torch.cuda.synchronize()
        tref = time.perf_counter() - t0
        out["reference_offset_probe"] = {"iterations": rr.iterations, "converged": bool(rr.converged),
                                         "min_rel_residual": float(np.nanmin(hr)),
                                         "at_iteration": kmin,
                                         "final_rel_residual": float(hr[-1]),
                                         "time_to_min_residual_s": round(tref, 4),
                                         "what": "T C (stbmmProd's offsets), inputs in HBM, stopped at the "
                                                 "iteration of its lowest residual"}
        log(f"reference offsets: converged={rr.converged}, min residual {np.nanmin(hr):.2e} at {kmin}, "
            f"{tref:.3f} s to reach it")

    if args.twolvl_zn > 0:
        # MLmap's precond 1 / 2 (two-level: block Jacobi plus a deflation space of
        # Zn vectors, rebuilt by every PCG call) on the same workload, inputs in HBM
        tl = {"zn": args.twolvl_zn, "block_jacobi_resident_s": round(per_map_dev, 4),
              "block_jacobi_iterations": int(rep_d.iterations)}
        for pc, name in ((1, "a_priori"), (2, "a_posteriori")):
            try:
                solve(G.MGPU_FP64, profile=False, precond=pc, zn=args.twolvl_zn)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                _, _, r2, _ = solve(G.MGPU_FP64, profile=False, precond=pc, zn=args.twolvl_zn)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                tl[name] = {"time_to_map_s": round(dt, 4), "iterations": int(r2.iterations),
                            "converged": bool(r2.converged), "build_ms": round(r2.t_2lvl_ms, 2),
                            "solve_ms": round(r2.t_solve_ms, 2)}
                log(f"two-level precond {pc} (Zn {args.twolvl_zn}): {dt:.3f} s resident, {r2.iterations} iterations "
                    f"(block Jacobi {rep_d.iterations}), deflation space built in {r2.t_2lvl_ms:.1f} ms")
            except Exception as e:  # reported, not hidden
                tl[name] = {"error": str(e)}
        out["twolvl_c2"] = tl

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, args, reps[-1]["iterations"])
        # all host cores the process may use (the GPU host shares its CPUs: at most 16)
        ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        out["cpu_baseline_all_cores"] = cpu_baseline(wl, args, reps[-1]["iterations"], threads=min(16, ncpu))
        out["c1_full_solve"] = c1_section(G, W, args)
    if not args.no_noise_model:
        out["noise_model_c2"] = noise_section(G, wl, args)
    del d_pix, d_w, d_sig0, d_sig, d_noise, d_inv, wl, hp, hw, hs
    torch.cuda.empty_cache()
    if not args.no_templates:
        out["templates_c6"] = templates_section(G, W, args)
    if args.c4_real > 0:
        out["batch_c4"] = c4_section(G, W, torch, args)
    if args.serve_requests > 0:
        out["serve_c5"] = serve_section(G, W, args)
    print(json.dumps(out), flush=True)


def templates_section(G, W, args):
    """C6: MTmap's template marginalisation at the C2 scale, from host arrays
    (problem, template tables, kernels and their pseudo-inverses, block-Jacobi
    preconditioner, PCG_GLS_templates), FP64; per-kernel-class device times of
    the solve."""
    wl = W.config_c6()
    log(f"C6: generated {wl.name}")

    def run(profile):
        t0 = time.perf_counter()
        pb = G.Problem(wl.pix, wl.weights, wl.block_sizes, 1, wl.invtt, opts=G.options(profile=profile))
        pb.set_templates(wl.meta)
        t_setup = time.perf_counter() - t0
        sig, _, _ = pb.precond_build(wl.signal)
        x, hist, rep = pb.pcg(sig, wl.noise, tol=args.tol, maxiter=args.maxiter)
        total = time.perf_counter() - t0
        kt = pb.kernel_times() if profile else None
        ntot, nt, ranks = pb.templates_info()
        pb.close()
        return total, t_setup, rep, kt, ntot, nt, ranks

    run(False)  # warm-up (library and solver initialisation)
    total, t_setup, rep, _, ntot, nt, ranks = run(False)
    _, _, rep_p, kt, _, _, _ = run(True)
    it = max(rep.iterations, 1)
    tm_ms = kt["templates"][0]
    log(f"C6: {total:.3f} s from host arrays (templates setup {t_setup:.3f} s), {rep.iterations} iterations, "
        f"templates {tm_ms / max(kt['templates'][1], 1):.3f} ms per application")
    return {"workload": wl.name, "time_to_map_from_host_s": round(total, 4), "setup_s": round(t_setup, 4),
            "pcg_device_ms": round(rep.t_solve_ms, 2), "iterations": int(rep.iterations),
            "converged": bool(rep.converged), "templates_total": int(ntot), "templates_per_block": int(nt[0]),
            "kernel_rank_min": int(ranks.min()), "kernel_rank_max": int(ranks.max()),
            "templates_ms_per_application": round(tm_ms / max(kt["templates"][1], 1), 4),
            "kernel_ms_per_map": {k: round(v[0], 2) for k, v in kt.items() if v[1] > 0}}


def noise_section(G, wl, args):
    """invtt of every C2 block built from the noise timestream on the device
    (noise2invtt, utils.py:500-601), from host arrays, against the oracle
    (scipy periodogram + curve_fit, 1 core) on the first 4 blocks."""
    from oracle import noise as N

    nb = int(wl.block_sizes.size)
    G.noise_invtt(wl.noise[: int(wl.block_sizes[0])], wl.block_sizes[:1], wl.lam, 10.0)  # warm-up
    t0 = time.perf_counter()
    inv, par = G.noise_invtt(wl.noise, wl.block_sizes, wl.lam, 10.0)
    tg = time.perf_counter() - t0
    k = 4
    n_k = int(np.sum(wl.block_sizes[:k]))
    t0 = time.perf_counter()
    ref, rpar = N.compute_invtt(wl.noise[:n_k], wl.block_sizes[:k], wl.lam, 10.0)
    tc = (time.perf_counter() - t0) * nb / k
    err = float(np.max(np.abs(inv[: k * wl.lam] - ref)) / np.max(np.abs(ref)))
    log(f"noise model: GPU {tg:.3f}s for {nb} blocks, CPU oracle {tc:.1f}s (extrapolated), row error {err:.1e}")
    return {"blocks": nb, "block_len": int(wl.block_sizes[0]), "lambda": wl.lam, "gpu_s_from_host": round(tg, 4),
            "cpu_oracle_s": round(tc, 2), "cpu_sample": f"first {k} blocks timed, scaled to {nb}",
            "max_row_rel_err_vs_oracle": err}


def c1_section(G, W, args):
    """C1 (the reference's CPU-runnable configuration) solved whole on both
    sides: the GPU drop-in from host arrays and the CPU oracle (1 core), same
    offset mode and tolerance -- no extrapolation."""
    from oracle import oracle as O

    wl = W.config_c1()
    mode = {"symmetric": O.OFFSET_SYMMETRIC, "reference": O.OFFSET_REFERENCE, "none": O.OFFSET_NONE}[args.offset]
    gmode = {"symmetric": G.MGPU_OFFSET_SYMMETRIC, "reference": G.MGPU_OFFSET_REFERENCE,
             "none": G.MGPU_OFFSET_NONE}[args.offset]
    o = G.options(offset_mode=gmode)
    G.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=args.tol,
              maxiter=args.maxiter, opts=o)  # warm-up
    t0 = time.perf_counter()
    g = G.pcg_gls(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=args.tol,
                  maxiter=args.maxiter, opts=o)
    tg = time.perf_counter() - t0
    t0 = time.perf_counter()
    with O.offset_mode(mode):
        r = O.gls_solve(wl.pix, wl.weights, wl.signal, wl.noise, wl.block_sizes, wl.lam, wl.invtt, tol=args.tol,
                        maxiter=args.maxiter)
    tc = time.perf_counter() - t0
    err = float(np.max(np.abs(g.x - r.x)) / np.max(np.abs(r.x)))
    log(f"C1: GPU {tg:.3f}s ({g.report['iterations']} it), CPU oracle {tc:.2f}s ({r.iterations} it), map diff {err:.1e}")
    return {"workload": wl.name, "gpu_s": round(tg, 4), "cpu_oracle_s": round(tc, 3), "cpu_cores": 1,
            "gpu_iterations": g.report["iterations"], "cpu_iterations": int(r.iterations),
            "map_rel_diff": err, "speedup": round(tc / tg, 1)}


def c4_section(G, W, torch, args):
    """Batched Monte Carlo (C4): independent realizations solved as segmented
    batches (one per lambda and tile length), inputs resident in HBM.  Compared
    with the plain FP64 path solving the realizations one after the other."""
    t = time.time()
    reals = W.config_c4(nreal=args.c4_real, progress=lambda name: log(f"C4: generated {name}"))
    total = int(sum(r.m for r in reals))
    log(f"C4: generated {len(reals)} realizations, {total} samples in {time.time() - t:.1f}s")
    dev = torch.device("cuda:0")
    items = [dict(pix=torch.from_numpy(r.pix).to(dev), weights=torch.from_numpy(r.weights).to(dev),
                  signal=torch.from_numpy(r.signal).to(dev), noise=torch.from_numpy(r.noise).to(dev),
                  invtt=torch.from_numpy(r.invtt).to(dev), block_sizes=r.block_sizes, lam=r.lam) for r in reals]
    grid = reals[0].nx * reals[0].ny
    del reals
    torch.cuda.synchronize()
    sym = G.MGPU_OFFSET_SYMMETRIC
    res = {}

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    o64 = G.options(precision=G.MGPU_FP64, offset_mode=sym, graph=1)
    omx = G.options(precision=G.MGPU_MIXED, offset_mode=sym, graph=1)
    def one_by_one():
        out = []
        for i, it in enumerate(items):
            out.append(G.pcg_batch([it], grid, tol=args.tol, maxiter=args.maxiter, opts=o64)[0])
            log(f"C4: FP64 alone r{i}: {out[-1].report['iterations']} iterations")
        return out

    solo, t_solo = timed(one_by_one)
    log(f"C4: FP64 one realization at a time {t_solo:.2f}s")
    b64, t64 = timed(lambda: G.pcg_batch(items, grid, tol=args.tol, maxiter=args.maxiter, opts=o64))
    log(f"C4: FP64 batched {t64:.2f}s")
    bmx, tmx = timed(lambda: G.pcg_batch(items, grid, tol=args.tol, maxiter=args.maxiter, opts=omx))
    log(f"C4: mixed batched {tmx:.2f}s")
    oad = G.options(precision=G.MGPU_MIXED, offset_mode=sym, graph=1, refresh_every=G.MGPU_REFRESH_ADAPTIVE)
    bad, tad = timed(lambda: G.pcg_batch(items, grid, tol=args.tol, maxiter=args.maxiter, opts=oad))
    log(f"C4: mixed batched, adaptive refreshes {tad:.2f}s")
    # the converged maps (FP64 to 1e-11)
    cv = G.pcg_batch(items, grid, tol=1e-11, maxiter=4 * args.maxiter, opts=o64)

    def worst(a_list, b_list):
        return max(float(np.max(np.abs(a.x - b.x)) / np.max(np.abs(b.x))) for a, b in zip(a_list, b_list))

    same = all(np.array_equal(a.x, b.x) for a, b in zip(solo, b64))
    err = max(float(np.max(np.abs(a.x - b.x)) / np.max(np.abs(a.x))) for a, b in zip(b64, bmx))
    err_ad = max(float(np.max(np.abs(a.x - b.x)) / np.max(np.abs(a.x))) for a, b in zip(b64, bad))
    log(f"C4: worst map error vs the converged maps: FP64 {worst(b64, cv):.2e}, mixed {worst(bmx, cv):.2e}, "
        f"mixed adaptive {worst(bad, cv):.2e}")
    it64 = [r.report["iterations"] for r in b64]
    itmx = [r.report["iterations"] for r in bmx]
    res.update(realizations=len(items), samples=total,
               lambdas={str(l): int(sum(1 for it in items if it["lam"] == l)) for l in (1024, 4096, 8192)},
               fp64_sequential_s=round(t_solo, 3), fp64_batched_s=round(t64, 3), mixed_batched_s=round(tmx, 3),
               maps_per_s_mixed_batched=round(len(items) / tmx, 3),
               speedup_mixed_batched_vs_fp64_sequential=round(t_solo / tmx, 3),
               speedup_fp64_batched_vs_sequential=round(t_solo / t64, 3),
               batched_fp64_bit_identical_to_solo=bool(same), max_map_rel_err_mixed_vs_fp64=err,
               mixed_adaptive_batched_s=round(tad, 3),
               speedup_mixed_adaptive_batched_vs_fp64_sequential=round(t_solo / tad, 3),
               max_map_rel_err_mixed_adaptive_vs_fp64=err_ad,
               max_map_rel_err_vs_converged={"fp64": worst(b64, cv), "mixed": worst(bmx, cv),
                                             "mixed_adaptive": worst(bad, cv)},
               all_converged=bool(all(r.report["converged"] for r in b64 + bmx + bad)),
               iterations_fp64=[min(it64), int(np.median(it64)), max(it64)],
               iterations_mixed=[min(itmx), int(np.median(itmx)), max(itmx)])
    del items
    torch.cuda.empty_cache()
    return res


def serve_section(G, W, args):
    """Serving (C5): Poisson arrivals of seeded requests (4-128 det x 2^16-2^20
    samples, 256^2-1024^2 grids, lambda 1024-8192) into the request loop,
    mixed precision; latency from submission to the map on the host."""
    import threading

    t = time.time()
    pool = W.c5_pool(5, args.serve_pool, progress=lambda name: log(f"serve: generated {name}"))
    reqs = [dict(pix=w.pix, weights=w.weights, signal=w.signal, noise=w.noise, block_sizes=w.block_sizes,
                 lam=w.lam, invtt=w.invtt) for w in pool]
    grids = [w.nx * w.ny for w in pool]
    log(f"serve: generated {len(pool)} requests ({sum(w.m for w in pool)} samples) in {time.time() - t:.1f}s")
    del pool
    sym = G.MGPU_OFFSET_SYMMETRIC
    # FP64 reference maps of the pool, one solo solve each
    ref = []
    for r, g in zip(reqs, grids):
        ref.append(G.pcg_batch([r], g, tol=args.tol, maxiter=args.maxiter,
                               opts=G.options(precision=G.MGPU_FP64, offset_mode=sym))[0])
        log(f"serve: FP64 reference map {len(ref)}: {ref[-1].report['iterations']} iterations")
    omx = G.options(precision=G.MGPU_MIXED, offset_mode=sym)
    out = {"requests_per_rate": args.serve_requests, "pool": len(reqs), "precision": "mixed",
           "slots": 1, "max_batch": 8, "rates": []}
    rng = np.random.default_rng(55)
    for rate in [float(v) for v in args.serve_rates.split(",")]:
        n = args.serve_requests
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
        tickets = [None] * n
        with G.Server(hbm_budget_gb=0, max_slots=1, max_batch=8, opts=omx) as srv:
            t0 = time.perf_counter()

            def feed():
                for i in range(n):
                    dt = arrivals[i] - (time.perf_counter() - t0)
                    if dt > 0:
                        time.sleep(dt)
                    j = i % len(reqs)
                    tickets[i] = (j, srv.submit(reqs[j], grids[j], tol=args.tol, maxiter=args.maxiter))

            th = threading.Thread(target=feed)
            th.start()
            lat, qs, errs = [], [], []
            for i in range(n):
                while tickets[i] is None:
                    time.sleep(0.001)
                j, tk = tickets[i]
                r, l_ms, q_ms = srv.wait(tk)
                lat.append(l_ms)
                qs.append(q_ms)
                a = ref[j].x
                errs.append(float(np.max(np.abs(r.x - a)) / np.max(np.abs(a))))
                if not r.report["converged"]:
                    errs[-1] = float("inf")
                if i % 5 == 4:
                    log(f"serve rate {rate}: {i + 1} requests answered")
            th.join()
            wall = time.perf_counter() - t0
            st = srv.stats()
        ent = {"rate": rate, "requests": n, "p50_ms": round(float(np.percentile(lat, 50)), 1),
               "p99_ms": round(float(np.percentile(lat, 99)), 1), "mean_queue_ms": round(float(np.mean(qs)), 1),
               "maps_per_s": round(n / wall, 3), "max_map_rel_err_vs_fp64": max(errs),
               "batched_groups": st["groups"], "peak_concurrent_solves": st["peak_running"],
               "peak_hbm_gb": round(st["peak_hbm_gb"], 2), "requeued": st["requeued"]}
        log(f"serve rate {rate}: p50 {ent['p50_ms']} ms, p99 {ent['p99_ms']} ms, {ent['maps_per_s']} maps/s")
        out["rates"].append(ent)
    return out


def measured_traffic(kernel_class, args):
    """HBM bytes per launch of a kernel class on this workload, from the
    calibrated FETCH_SIZE / WRITE_SIZE counter runs summarised in
    profiles/traffic_c2.json (None when absent or for another workload)."""
    fn = os.path.join(ROOT, "profiles", "traffic_c2.json")
    if args.ndet != 64 or not os.path.exists(fn):
        return None
    with open(fn) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel_class)
    return None if k is None else int(k["fetch_bytes"] + k["write_bytes"])


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, args, gpu_iters, threads=1):
    """The CPU oracle (a C restatement of the reference path) on the first
    cpu_ndet detectors of the same workload: setup + 1 iteration and setup + 5
    iterations are timed, and the time-to-map of the full workload is
    extrapolated linearly in samples at the GPU's iteration count.  threads = 1
    is the reference's default (one rank, no OpenMP); threads > 1 times the
    OpenMP build of the same restatement."""
    from oracle import oracle as O

    omp = threads > 1
    if omp:
        threads = O.set_threads(threads)
    nd = min(args.cpu_ndet, args.ndet)
    n = nd * (wl.m // args.ndet)
    nblk = int(np.searchsorted(np.cumsum(wl.block_sizes), n) + 1)
    sl = dict(pix=wl.pix[:3 * n], weights=wl.weights[:3 * n], signal=wl.signal[:n], noise=wl.noise[:n],
              block_sizes=wl.block_sizes[:nblk], invtt=wl.invtt[:nblk * wl.lam])
    times = {}
    mode = {"symmetric": O.OFFSET_SYMMETRIC, "reference": O.OFFSET_REFERENCE, "none": O.OFFSET_NONE}[args.offset]
    k_lo, k_hi = 1, 5
    for k in ((0, k_lo, k_hi) if omp else (k_lo, k_hi)):  # the OpenMP build gets an untimed first run
        t0 = time.perf_counter()
        with O.offset_mode(mode):
            O.gls_solve(sl["pix"], sl["weights"], sl["signal"], sl["noise"], sl["block_sizes"], wl.lam,
                        sl["invtt"], tol=1e-30, maxiter=max(k, 1) + 2, stop_after=max(k, 1), omp=omp)
        times[k] = time.perf_counter() - t0
    per_iter = (times[k_hi] - times[k_lo]) / (k_hi - k_lo)
    setup = times[k_lo] - k_lo * per_iter
    scale = args.ndet / nd
    est = (setup + gpu_iters * per_iter) * scale
    log(f"cpu oracle ({threads} thread(s)) slice {nd}/{args.ndet} det: setup {setup:.2f}s, {per_iter:.2f}s/iteration")
    out = {"value": round(est, 2), "unit": "s", "cores": threads, "kind": "port",
           "sample": f"first {nd} of {args.ndet} detectors ({n} samples): setup {setup:.2f} s + "
                     f"{per_iter:.3f} s/iteration measured, scaled x{scale:g} to the full workload at the "
                     f"GPU's {gpu_iters} iterations" + (" (OpenMP build)" if omp else ""),
           "cpu": cpu_model()}
    if not omp:
        cal = fft_calibration()
        out["fft_calibration"] = cal
        out["sample"] += (f"; the oracle's FFT takes {cal['ratio_c2']:.2f}x the time of scipy.fft (pocketfft, "
                          f"1 thread) at the C2 tile size")
    return out


def fft_calibration():
    """The oracle's complex FFT (what its real tile transforms run on, at half
    the real tile length) against scipy.fft's on one thread, at the tile sizes
    of the reference's rule for C1 (bs 4096) and C2 (bs 32768): the factor by
    which the oracle's FFT may overstate the CPU baseline's N^-1 time."""
    import scipy.fft as sfft
    from oracle import oracle as O

    res = {}
    for tag, n in (("c1", 2048), ("c2", 16384)):
        z = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
        reps = max(20, (1 << 22) // n)

        def best(fn):
            fn()
            t = []
            for _ in range(3):
                t0 = time.perf_counter()
                for _ in range(reps):
                    fn()
                t.append((time.perf_counter() - t0) / reps)
            return min(t)

        to = best(lambda: O.fft(z))
        ts = best(lambda: sfft.fft(z, workers=1))
        err = float(np.max(np.abs(O.fft(z) - sfft.fft(z))) / np.max(np.abs(sfft.fft(z))))
        res[f"oracle_us_{tag}"] = round(to * 1e6, 2)
        res[f"scipy_us_{tag}"] = round(ts * 1e6, 2)
        res[f"ratio_{tag}"] = round(to / ts, 3)
        res[f"max_rel_diff_{tag}"] = err
    log(f"FFT calibration: oracle/scipy = {res['ratio_c1']} (n=2048), {res['ratio_c2']} (n=16384)")
    return res


if __name__ == "__main__":
    main()
    # every helper thread has been joined; release the library's device and
    # pinned host memory, then leave at once (no lingering interpreter teardown)
    from midapack_amd import lib as _G

    if _G._LIB is not None:
        _G.finalize()
    sys.stdout.flush()
    sys.stderr.flush()
    # under a profiler (its library preloaded) the normal exit runs its finalisation
    if not any(k.startswith("ROCPROFILER") for k in os.environ) and "rocprofiler" not in os.environ.get("LD_PRELOAD", ""):
        os._exit(0)
