#!/usr/bin/env python3
"""Benchmark of the direct 2-D kernel-density sum on one MI355X.

Workload (the headline configuration): 1,000,000 reference points from a
seeded 3-component 2-D Gaussian mixture (weights uniform(0.5, 1.5),
normalised), 1,000,000 query points uniform on [-4, 4]^2, GaussianKernel
h = 0.05, FP32 pair terms with block-compensated FP64 accumulation.  The
references are in kd-tree order, the order KernelDensity hands its data
points to the device.  One step = one full 1M x 1M evaluation
(bk_eval_range_device) with points and output resident in HBM.  The
Epanechnikov product kernel (hx = 0.08, hy = 0.12) of the same
configuration is timed after it and reported beside it.

Also reported (object "cv_grid"): the batched cross-validation grid, four
seeded mixtures of 50k, 200k, 500k and 1M points x the 8 x 8 log grid
hx, hy in 0.01 * 50^(k/7) (GaussianProductKernel2d), LCV and convolution
LSCV for all 256 (dataset, bandwidth pair) cases in one device pipeline
(bk_cv_grid_run, graph replay).  Its unit is the (ordered pair, bandwidth
pair) evaluation: sum n^2 x 64 = 8.27e13 per step.

Prints one JSON line (see the keys below).  --gpus above 1 is refused: this
build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

N_REF = 1_000_000
N_QUERY = 1_000_000
H_GAUSS = 0.05
H_EPAN = (0.08, 0.12)
FLOP_PER_PAIR_GAUSS = 9     # dx, dy, dx^2, fma dy^2, scale, exp, fma (exp counted as 1)
FLOP_PER_PAIR_EPANPROD = 12
PEAK_FP32_TFLOPS = 157.3    # MI355X FP32 peak (vector and matrix)
# HBM bytes per launch of the Gaussian kernel at this workload, measured with
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (one run each) and scaled by a
# 256 MiB copy kernel's readings (fetch x2.00, write x1.00):
# 360,996 KiB fetched + 257,812 KiB written.
GAUSS_HBM_BYTES_PER_LAUNCH = (363_188 + 257_812) * 1024
# FP32 VALU issue peak of MI355X: 157.3 TFLOP/s / 2 FLOP per FMA = 256 CUs x
# 4 SIMDs x 32 lanes per cycle x 2.4 GHz (lane-instructions per second).
# Unpacked v_fma_f32 reaches it (a microbenchmark sustained 84%); v_exp_f32
# issues at a quarter of that rate, so an exp counts as 4 issue slots.
PEAK_VALU_LANE_INSTR = 256 * 4 * 32 * 2.4e9
EXP_ISSUE_SLOTS = 4
CV_SIZES = (50_000, 200_000, 500_000, 1_000_000)
CV_SEED = 10040
CV_GRID = 0.01 * 50.0 ** (np.arange(8) / 7.0)
CV_GAUSS_TILINGS = {0: (8, 8), 1: (4, 8), 2: (4, 4)}  # Gaussian tiling variants (x-, y-bandwidths per thread)
SUR_TRAIN, SUR_SERVE = 4_000_000, 10_000_000
SUR_STEPS, SUR_BATCH = 16_000, 16_384
MLP_FLOP_PER_QUERY = 66_304     # 2 * (2*128 + 128*128 + 128*128 + 128)
PEAK_BF16_TFLOPS = 2500.0       # MI355X dense BF16 MFMA
PEAK_FP64_TFLOPS = 78.6         # MI355X FP64 (vector and matrix)
G16_N, G16_D, G16_H = 200_000, 16, 0.5
G16_FLOP_PER_PAIR = 3 * G16_D + 4   # 16 differences, 16 squares+adds, scale, exp, fma (exp as 1)


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU work per baseline leg")
    ap.add_argument("--sur-steps", type=int, default=16_000, help="training steps of the surrogate")
    ap.add_argument("--kernels", default="gauss,epanprod,cvgrid,surrogate,gauss16,adaptive,dualtree,fp64",
                    help="comma list: gauss, epanprod, cvgrid, surrogate, gauss16, adaptive, dualtree, fp64")
    ap.add_argument("--no-spatial-order", action="store_true",
                    help="skip the timing of the same evaluations in device-sorted query order")
    ap.add_argument("--gauss-form", choices=["auto", "direct"], default="auto",
                    help="Gaussian exponent form (auto: tile-centred on compact tiles)")
    ap.add_argument("--no-cv-tune", action="store_true",
                    help="skip the CV grid autotuner (runs use the cached or default schedule)")
    return ap.parse_args()


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload():
    from bbrcit_kde_amd import datasets as ds

    pts, w = ds.gaussian_mixture(N_REF, seed=ds.SEED_MIXTURE_1M)
    perm = ds.kd_order(pts, 64)
    pts, w = pts[perm], w[perm]
    mass = w / w.sum()
    q = ds.uniform_box(N_QUERY, -4.0, 4.0, ds.SEED_MIXTURE_1M + 1)
    return pts, mass, q


def time_kernel(kde, desc, steps: int, warmup: int):
    import torch

    out = torch.empty(N_QUERY, dtype=torch.float64, device="cuda")
    # a stream of our own (default priority): the library launches on it and
    # the events are recorded on it
    stream = torch.cuda.Stream()
    kde.set_stream(stream.cuda_stream)
    for _ in range(warmup):
        kde.eval_device(desc, 0, N_REF - 1, 0, N_QUERY - 1, out.data_ptr(), True)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        starts[k].record(stream)
        kde.eval_device(desc, 0, N_REF - 1, 0, N_QUERY - 1, out.data_ptr(), True)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    kde.set_stream(0)
    return out.cpu().numpy(), ms, wall


def count_terms(bk, kde, desc, timed_vals):
    """One more launch, outside the timed region, of the counting variant of
    the same kernel: the pair terms its wavefronts evaluate.  Wavefronts pass
    over a (sub-)tile when each of its terms is below 2^-56 of its mass
    (Gaussian) or exactly zero (Epanechnikov) for all their 64 queries; those
    terms are not evaluated (not counted)."""
    import torch

    out = torch.empty(N_QUERY, dtype=torch.float64, device="cuda")
    kde.set_option(bk.BK_OPT_COUNT_TERMS, 1)
    kde.eval_device(desc, 0, N_REF - 1, 0, N_QUERY - 1, out.data_ptr(), True)
    torch.cuda.synchronize()
    lane_terms, query_terms = kde.term_count()
    kde.set_option(bk.BK_OPT_COUNT_TERMS, 0)
    same = bool(np.array_equal(out.cpu().numpy(), timed_vals))
    return {"lane_terms": lane_terms, "query_terms": query_terms, "results_identical_to_timed": same}


def time_host_boundary(bk, pts, mass, q, desc):
    """The headline evaluation through the host-buffer boundary: a new
    context, upload of the FP64 host arrays (rounded to FP32 on the host),
    feature preparation, evaluation and the copy of the results back into a
    host array, timed by the host clock around the whole sequence."""
    t0 = time.perf_counter()
    with bk.DirectKde(pts, q, mass) as kde:
        t1 = time.perf_counter()
        kde.eval(desc, 0, N_REF - 1, 0, N_QUERY - 1)
        t2 = time.perf_counter()
    pairs = float(N_REF) * float(N_QUERY)
    return {"value": pairs / (t2 - t0), "unit": "pair-evals/s", "s_total": t2 - t0, "s_upload": t1 - t0,
            "s_eval_to_host": t2 - t1,
            "note": "host arrays in, host results out (context creation and upload included); not the "
                    "headline value, which starts with the inputs resident in HBM"}


def cpu_baseline(pts, mass, q, gpu_vals, seconds: float):
    """Time the reference's CPU direct_eval (or the FP64 restatement) on a
    bounded sample of the same workload; check the GPU values there."""
    from bbrcit_kde_amd import datasets as ds
    from oracle.oracle import REF_LIB, Oracle, Reference, default_threads

    P, M = ds.round_f32(pts), ds.round_f32(mass)
    result = {}
    # parity sample: 256 queries spread over the whole range
    idx = np.linspace(0, N_QUERY - 1, 256).astype(np.int64)
    orc = Oracle()
    threads = default_threads()
    t0 = time.perf_counter()
    want = orc.direct(2, H_GAUSS, H_GAUSS, P, ds.round_f32(q[idx]), mass=M, normalize=True, threads=threads)
    dt = time.perf_counter() - t0
    # floor: 1e-10 of the density's upper bound C * sum(m) (U <= 1)
    floor = 1e-10 * (1.0 / (2.0 * np.pi * H_GAUSS ** 2)) * float(M.sum())
    above = np.abs(want) >= floor
    rel = np.abs(gpu_vals[idx][above] - want[above]) / np.abs(want[above])
    result["parity"] = {"oracle": "FP64 restatement", "queries": int(len(idx)), "max_rel_err": float(rel.max()),
                        "floor": floor, "frac_below_floor": float(1.0 - above.mean()), "tolerance": 1e-5}
    # below the floor: relative errors of the sampled queries down to 1e-20 of the bound
    deep = (np.abs(want) < floor) & (np.abs(want) >= 1e-10 * floor)
    if deep.any():
        drel = np.abs(gpu_vals[idx][deep] - want[deep]) / np.abs(want[deep])
        result["parity"]["below_floor"] = {
            "range": "1e-20 to 1e-10 of C*sum(m)", "queries": int(deep.sum()), "max_rel_err": float(drel.max()),
            "median_rel_err": float(np.median(drel)),
            "queries_zero": int(np.sum(np.abs(want) < 1e-10 * floor)),
            "note": "FP32 exponents: the relative error grows with the exponent's size; these values are below "
                    "1e-10 of the density's bound"}
    port_rate = len(idx) * N_REF / dt
    # port on all host cores, sized to about `seconds`
    nq = int(min(N_QUERY, max(256, port_rate * seconds / N_REF)))
    t0 = time.perf_counter()
    orc.direct(2, H_GAUSS, H_GAUSS, P, ds.round_f32(q[:nq]), mass=M, normalize=True, threads=threads)
    dt = time.perf_counter() - t0
    result["port"] = {"value": nq * N_REF / dt, "unit": "pair-evals/s", "cores": threads, "kind": "port",
                      "sample": f"{nq} queries x {N_REF} refs (FP64 restatement, {threads} threads)"}
    if REF_LIB.exists():
        ref = Reference()
        # the reference's own KernelDensity::direct_eval (FP64 Gaussian), its
        # KernelDensity (kd tree of the 1M references) built outside the timed
        # region; queries split over threads as test/test_kde_cppthread.cc does
        probe, dt_small = ref.direct_eval_timed(2, 0, H_GAUSS, H_GAUSS, pts, mass, q[:32], threads=1,
                                                leaf_nmax=1024)
        rate1 = 32 * N_REF / max(dt_small, 1e-9)
        for label, th, secs in (("reference_1thread", 1, 0.5 * seconds), ("reference", threads, seconds)):
            nq_ref = int(min(50000, max(64, rate1 * th * secs / N_REF)))
            vals, dt = ref.direct_eval_timed(2, 0, H_GAUSS, H_GAUSS, pts, mass, q[:nq_ref], threads=th,
                                             leaf_nmax=1024)
            result[label] = {"value": nq_ref * N_REF / dt, "unit": "pair-evals/s", "cores": th, "kind": "reference",
                             "seconds": dt,
                             "sample": f"{nq_ref} queries x {N_REF} refs, KernelDensity<2,GaussianKernel<2,double>>"
                                       f"::direct_eval per query on {th} thread(s); its kd-tree build is outside "
                                       "the timed region"}
    return result


def cv_datasets():
    from bbrcit_kde_amd import datasets as ds

    sets = []
    for k, n in enumerate(CV_SIZES):
        p, w = ds.gaussian_mixture(n, seed=CV_SEED + k)
        sets.append((p, w / w.sum()))
    return sets


def cv_instr_per_ref(A: int, B: int) -> tuple[int, int]:
    """(VALU instructions, of which v_exp_f32) per reference and lane of the
    Gaussian grid kernel (non-adaptive): 2 differences, 2 squares, A (fma +
    exp) for the x-factors, B (mul + exp) for the y-factors, A*B fma; matches
    the kernel's gfx950 ISA."""
    return 4 + 2 * A + 2 * B + A * B, A + B


def cv_schedule(bk, cv) -> dict:
    """The pair kernel's schedule of a plan's last run and where it came from."""
    src = cv.get_option(bk.BK_CVG_TUNE_SOURCE)
    return {"tiling_variant": cv.get_option(bk.BK_CVG_VARIANT), "unroll": cv.get_option(bk.BK_CVG_UNROLL),
            "workgroups": cv.get_option(bk.BK_CVG_WORKGROUPS),
            "chosen_by": {0: "forced", 1: "default", 2: "autotuner cache"}[src]}


def tune_table(timings):
    """bk_cv_grid_tune's candidates as JSON: pair-kernel ms per (tiling, unroll index)."""
    if timings is None:
        return None
    return {"pair_kernel_ms": [{"tiling_variant": t, "unroll_index": u, "ms": ms} for t, u, ms in timings],
            "note": "every schedule timed on one eager whole run of the grid; the fastest is recorded in the "
                    "autotuner cache and used by the runs below"}


def bench_cv_grid(bk, steps: int, warmup: int, cpu: bool, seconds: float, tune: bool = True):
    from bbrcit_kde_amd import datasets as ds

    sets = cv_datasets()
    units = float(sum(n * n for n in CV_SIZES)) * len(CV_GRID) ** 2
    res = {}
    with bk.CvGrid(bk.BK_GAUSSIAN_PRODUCT_2D, [p for p, _ in sets], [w for _, w in sets]) as cv:
        # the autotuner: every schedule timed on a whole run, the fastest recorded
        tuned = cv.tune(CV_GRID, CV_GRID) if tune else None
        # eager pass: the persistent pair kernel timed alone with HIP events on
        # the plan's stream
        cv.set_option(bk.BK_CVG_GRAPH, 0)
        lcv, lscv = cv.run(CV_GRID, CV_GRID)
        main_ms, _, rescued = cv.last_timing()
        variant = cv.get_option(bk.BK_CVG_VARIANT)
        schedule = cv_schedule(bk, cv)
        cv.set_option(bk.BK_CVG_GRAPH, 1)
        for _ in range(warmup):
            cv.run(CV_GRID, CV_GRID)
        ms = []
        for _ in range(steps):
            lcv2, lscv2 = cv.run(CV_GRID, CV_GRID)
            ms.append(cv.last_timing()[1])
        assert np.array_equal(lcv, lcv2) and np.array_equal(lscv, lscv2), "graph replay differs from eager run"
        # units the pair kernel evaluates (one instrumented eager run, untimed)
        cv.set_option(bk.BK_CVG_COUNT_UNITS, 1)
        gcount = cv.run(CV_GRID, CV_GRID)
        g_units = cv.unit_count()
        cv.set_option(bk.BK_CVG_COUNT_UNITS, 0)
        assert np.array_equal(gcount[1], lscv), "counting run differs"
    avg = float(np.mean(ms))
    A, B = CV_GAUSS_TILINGS[variant]
    n_instr, n_exp = cv_instr_per_ref(A, B)
    instr = n_instr / (A * B)
    slots = (n_instr + (EXP_ISSUE_SLOTS - 1) * n_exp) / (A * B)
    rate_main = g_units / (main_ms * 1e-3)  # evaluated units per second of the pair kernel
    res.update({
        "workload": "4 seeded 2-D mixtures (50k, 200k, 500k, 1M points, seeds 10040-10043) x 8x8 log grid "
                    "hx, hy = 0.01*50^(k/7); GaussianProductKernel2d; LCV + convolution LSCV per case",
        "unit": "pair-bw-evals/s", "value": units / (avg * 1e-3), "ms_per_step": avg, "steps": steps,
        "value_kind": "effective: every (ordered pair, bandwidth pair) unit counted",
        "units_per_step": units, "evaluated_units": g_units, "evaluated_frac": g_units / units,
        "evaluated_units_per_s": g_units / (avg * 1e-3),
        "main_kernel_ms": main_ms, "rescued_entries": int(rescued),
        "tiling": f"{A}x{B}", "schedule": schedule, "autotune": tune_table(tuned), "graph": True,
        "roofline": {"bound": "valu", "pipe": "FP32 VALU + v_exp_f32 issue",
                     "achieved": rate_main * instr, "peak": PEAK_VALU_LANE_INSTR, "unit": "VALU lane-instr/s",
                     "frac": rate_main * instr / PEAK_VALU_LANE_INSTR, "instr_per_pair_bw": instr,
                     "exp_weighted_frac": rate_main * slots / PEAK_VALU_LANE_INSTR,
                     "exp_weighted_note": f"each v_exp_f32 counted as {EXP_ISSUE_SLOTS} issue slots "
                                          "(quarter-rate transcendental)",
                     "note": f"over the evaluated units; separable product kernel: {A + B} exp per reference "
                             f"serve {A * B} bandwidth pairs, so the unfactored 9 FLOP/pair accounting exceeds "
                             "the FP32 peak; a wavefront passes over each 64-reference chunk whose products "
                             "are all below 2^-152 for its queries (no bit of any sum)",
                     "unfactored_tflops": rate_main * 9 / 1e12},
        "best_lcv": [list(map(int, np.unravel_index(np.argmax(np.where(np.isfinite(lcv[d]), lcv[d], -np.inf)),
                                                    lcv[d].shape))) for d in range(len(CV_SIZES))],
        "best_lscv": [list(map(int, np.unravel_index(np.argmin(lscv[d]), lscv[d].shape)))
                      for d in range(len(CV_SIZES))],
    })
    # the same plan with EpanechnikovProductKernel2d, the grid_search tool's
    # kernel (FP64 arithmetic; stages out of every bandwidth's support are
    # skipped, their terms being exactly zero)
    with bk.CvGrid(bk.BK_EPANECHNIKOV_PRODUCT_2D, [p for p, _ in sets], [w for _, w in sets]) as cv:
        e_tuned = cv.tune(CV_GRID, CV_GRID) if tune else None
        cv.set_option(bk.BK_CVG_GRAPH, 0)
        elcv, elscv = cv.run(CV_GRID, CV_GRID)
        e_main_ms = cv.last_timing()[0]
        e_schedule = cv_schedule(bk, cv)
        cv.set_option(bk.BK_CVG_GRAPH, 1)
        cv.run(CV_GRID, CV_GRID)
        e_ms = []
        for _ in range(max(1, min(steps, 2))):
            cv.run(CV_GRID, CV_GRID)
            e_ms.append(cv.last_timing()[1])
        e_variant = cv.get_option(bk.BK_CVG_VARIANT)
        # units the pair kernel evaluates (one instrumented eager run, untimed)
        cv.set_option(bk.BK_CVG_COUNT_UNITS, 1)
        ecount = cv.run(CV_GRID, CV_GRID)
        e_units = cv.unit_count()
        cv.set_option(bk.BK_CVG_COUNT_UNITS, 0)
        assert np.array_equal(ecount[1], elscv), "counting run differs"
    e_avg = float(np.mean(e_ms))
    res["epanechnikov_product"] = {
        "ms_per_step": e_avg, "value": units / (e_avg * 1e-3), "unit": "pair-bw-evals/s", "main_kernel_ms": e_main_ms,
        "value_kind": "effective: every (ordered pair, bandwidth pair) unit counted",
        "evaluated_units": e_units, "evaluated_frac": e_units / units,
        "evaluated_units_per_s": e_units / (e_avg * 1e-3),
        "tiling_variant": int(e_variant), "schedule": e_schedule, "autotune": tune_table(e_tuned),
        "note": "a wavefront passes over each 64-reference chunk out of the support of every bandwidth of its "
                "sub-grid for all its queries (exactly zero terms); evaluated_units counts the units of the "
                "chunks it runs",
        "best_lscv": [list(map(int, np.unravel_index(np.argmin(elscv[d]), elscv[d].shape)))
                      for d in range(len(CV_SIZES))]}
    # independent check: the single-bandwidth C-ABI path on the 50k dataset
    p0, w0 = sets[0]
    checks = []
    with bk.DirectKde(p0, None, w0) as kde:
        for ix, iy in [(2, 2), (4, 5), (6, 1)]:
            desc = bk.KernelDesc(bk.BK_GAUSSIAN_PRODUCT_2D, 2, CV_GRID[ix], CV_GRID[iy])
            checks.append(max(abs(lcv[0, ix, iy] - kde.likelihood_cv(desc, w0)),
                              abs(lscv[0, ix, iy] - kde.lsq_conv_cv(desc, w0)) / abs(lscv[0, ix, iy])))
    res["parity"] = {"vs": "bk_likelihood_cv / bk_lsq_conv_cv (single-bandwidth kernels), 50k dataset, 3 cases",
                     "max_err": float(max(checks))}
    if cpu:
        from oracle.oracle import REF_LIB, Oracle, Reference, default_threads

        # FP64 oracle on a 20k subset at one bandwidth pair (the grid on that subset)
        n_s = 20_000
        ps, ws = p0[:n_s], w0[:n_s] / w0[:n_s].sum()
        h = CV_GRID[3]
        with bk.CvGrid(bk.BK_GAUSSIAN_PRODUCT_2D, [ps], [ws]) as cv:
            gl, gs = cv.run([h], [h])
        orc = Oracle()
        M = ds.round_f32(ws)
        P = ds.round_f32(ps)
        want_l = orc.likelihood_cv(bk.BK_GAUSSIAN_PRODUCT_2D, h, h, P, ws, mass=M, threads=default_threads())
        want_s = orc.lsq_conv_cv(bk.BK_GAUSSIAN_PRODUCT_2D, h, h, P, ws, mass=M, threads=default_threads())
        res["parity"]["oracle"] = {"sample": f"{n_s} points of the 50k dataset, hx=hy={h:.4f}",
                                   "lcv_abs_err": float(abs(gl[0, 0, 0] - want_l)),
                                   "lscv_rel_err": float(abs(gs[0, 0, 0] - want_s) / abs(want_s))}
        if REF_LIB.exists():
            ref = Reference()
            n_r = 20_000
            pr, wr = p0[:n_r], w0[:n_r] / w0[:n_r].sum()
            hb = CV_GRID[2]
            t0 = time.perf_counter()
            ref.likelihood_cv(bk.BK_GAUSSIAN_PRODUCT_2D, 0, hb, hb, pr, wr, rel=1e-6, abs_err=1e-8, leaf_nmax=1024)
            ref.lsq_conv_cv(bk.BK_GAUSSIAN_PRODUCT_2D, 0, hb, hb, pr, wr, rel=1e-6, abs_err=1e-8, leaf_nmax=1024)
            dt = time.perf_counter() - t0
            res["cpu_baseline"] = {
                "value": n_r * n_r / dt, "unit": "pair-bw-evals/s", "cores": 1, "kind": "reference",
                "sample": f"KernelDensity<2,GaussianProductKernel2d<double>>::likelihood_cross_validate + "
                          f"lsq_convolution_cross_validate (rel 1e-6, abs 1e-8, leaf 1024, dual tree) on "
                          f"{n_r} points at hx=hy={hb:.4f}: one bandwidth pair of the grid, 1 thread"}
    return res


def bench_surrogate(bk, pts, mass, direct_qps: float, steps: int, warmup: int, train_steps: int = 16_000):
    import torch

    from bbrcit_kde_amd import datasets as ds

    desc = bk.KernelDesc(bk.BK_GAUSSIAN, 2, H_GAUSS)
    floor = 1e-10 * (1.0 / (2.0 * np.pi * H_GAUSS ** 2)) * float(ds.round_f32(mass).sum())
    qt = ds.uniform_box(SUR_TRAIN, -4.0, 4.0, ds.SEED_MIXTURE_1M + 3)
    t0 = time.perf_counter()
    with bk.DirectKde(pts, qt, mass) as lab:
        ft = lab.eval(desc, 0, len(pts) - 1, 0, len(qt) - 1)
    t_label = time.perf_counter() - t0
    qs = ds.uniform_box(SUR_SERVE, -4.0, 4.0, ds.SEED_MIXTURE_1M + 4)
    qx = torch.tensor(qs[:, 0], dtype=torch.float32, device="cuda")
    qy = torch.tensor(qs[:, 1], dtype=torch.float32, device="cuda")
    out = torch.empty(SUR_SERVE, dtype=torch.float64, device="cuda")
    raw = torch.empty(SUR_SERVE, dtype=torch.float64, device="cuda")
    res = {}
    with bk.DirectKde(pts, None, mass) as kde, bk.Surrogate(seed=1) as sur:
        sur.set_training_data(qt, ft, -4.0, 4.0, 1e-3 * floor)
        loss, train_ms = sur.train(train_steps, SUR_BATCH, 3e-3, 1e-5, seed=11)
        sigma = sur.params()[4]
        # raw inference, timed with events on the default stream around the call
        sur.infer_device(SUR_SERVE, qx.data_ptr(), qy.data_ptr(), raw.data_ptr())
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        sur.infer_device(SUR_SERVE, qx.data_ptr(), qy.data_ptr(), raw.data_ptr())
        ev1.record()
        torch.cuda.synchronize()
        infer_ms = ev0.elapsed_time(ev1)
        for _ in range(warmup):
            sur.serve_device(kde, desc, SUR_SERVE, qx.data_ptr(), qy.data_ptr(), out.data_ptr(), floor=floor,
                             sample_fraction=0.01, rel_tol=1e-3, cells=64, seed=7)
        stats = []
        for _ in range(steps):
            stats.append(sur.serve_device(kde, desc, SUR_SERVE, qx.data_ptr(), qy.data_ptr(), out.data_ptr(),
                                          floor=floor, sample_fraction=0.01, rel_tol=1e-3, cells=64, seed=7))
        torch.cuda.synchronize()
        # one more serve, outside the timed ones, with the exact kernel counting its pair terms
        kde.set_option(bk.BK_OPT_COUNT_TERMS, 1)
        sur.serve_device(kde, desc, SUR_SERVE, qx.data_ptr(), qy.data_ptr(), out.data_ptr(), floor=floor,
                         sample_fraction=0.01, rel_tol=1e-3, cells=64, seed=7)
        torch.cuda.synchronize()
        served_terms = kde.term_count()[1]
        kde.set_option(bk.BK_OPT_COUNT_TERMS, 0)
        served = out.cpu().numpy()
        rawv = raw.cpu().numpy()
    # the exact sum of every served query with the same device query sort and
    # tile culling the surrogate's check and fallback use: what the network
    # gains over the exact method without it
    exact = torch.empty(SUR_SERVE, dtype=torch.float64, device="cuda")
    with bk.DirectKde(pts, qs, mass) as ex:
        ex.set_option(bk.BK_OPT_QUERY_ORDER, bk.BK_QUERY_ORDER_SPATIAL)
        stream = torch.cuda.Stream()
        ex.set_stream(stream.cuda_stream)
        ex.eval_device(desc, 0, len(pts) - 1, 0, SUR_SERVE - 1, exact.data_ptr(), True)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        ex.eval_device(desc, 0, len(pts) - 1, 0, SUR_SERVE - 1, exact.data_ptr(), True)
        ev1.record(stream)
        torch.cuda.synchronize()
        exact_sorted_ms = ev0.elapsed_time(ev1)
        ex.set_option(bk.BK_OPT_COUNT_TERMS, 1)
        ex.eval_device(desc, 0, len(pts) - 1, 0, SUR_SERVE - 1, exact.data_ptr(), True)
        torch.cuda.synchronize()
        exact_terms = ex.term_count()[1]
        ex.set_option(bk.BK_OPT_COUNT_TERMS, 0)
        ex.set_stream(0)
    del exact
    ms = float(np.mean([st["ms_total"] for st in stats]))
    st = stats[-1]
    # independent audit: every 1000th query, exact direct sum
    idx = np.arange(0, SUR_SERVE, 1000)
    with bk.DirectKde(pts, qs[idx], mass) as ex:
        fe = ex.eval(desc, 0, len(pts) - 1, 0, len(idx) - 1)
    audit = np.abs(served[idx] - fe) / np.maximum(fe, floor)
    raw_err = np.abs(rawv[idx] - fe) / np.maximum(fe, floor)
    qps = SUR_SERVE / (ms * 1e-3)
    raw_qps = SUR_SERVE / (infer_ms * 1e-3)
    res.update({
        "workload": "MLP 2-128-128-128-1 (SiLU, BF16 weights, FP32 accumulation) trained on 4M (query, density) "
                    "pairs from the exact 1M-reference sum (h=0.05); serve 10M uniform queries on [-4,4]^2, 1% "
                    "exact sample, exact fallback of 64x64 cells with sampled rel. error > 1e-3",
        "unit": "queries/s", "value": qps, "ms_per_step": ms, "steps": steps,
        "vs_direct": qps / direct_qps, "direct_queries_per_s": direct_qps,
        "exact_sorted_queries_per_s": SUR_SERVE / (exact_sorted_ms * 1e-3), "exact_sorted_ms": exact_sorted_ms,
        "vs_exact_sorted": qps / (SUR_SERVE / (exact_sorted_ms * 1e-3)),
        "network_gain": {"vs_exact_sorted": qps / (SUR_SERVE / (exact_sorted_ms * 1e-3)),
                         "queries_served_by_network": int(SUR_SERVE - st["n_fallback"] - st["n_samples"]),
                         "exact_pair_terms_serving": served_terms, "exact_pair_terms_without_network": exact_terms,
                         "exact_pair_terms_saved_frac": 1.0 - served_terms / exact_terms,
                         "note": "the method without the network is the exact sum of all served queries in the "
                                 "same Z-curve order with the same tile culling; pair terms counted by the "
                                 "instrumented kernel (tests/test_gpu_surrogate_gain.py asserts both gains)"},
        "note": "the check and the fallback evaluate their queries in device-sorted Z-curve order, where each "
                "workgroup skips the reference tiles beyond 2^-100 of the kernel; vs_direct compares with the "
                "plain direct sum of random-order queries (the headline), vs_exact_sorted with the exact sum of "
                "all served queries in the same sorted order",
        "raw_inference_queries_per_s": raw_qps, "raw_vs_direct": raw_qps / direct_qps,
        "label_s": t_label, "train_ms": train_ms, "train_batch": SUR_BATCH,
        "final_loss": loss, "log_rmse": float(np.sqrt(loss) * sigma), "train_steps_run": train_steps,
        "cells_failed": st["cells_failed"], "cells": st["cells"], "fallback_queries": st["n_fallback"],
        "samples": st["n_samples"], "ms_infer": st["ms_infer"], "ms_check": st["ms_check"],
        "ms_fallback": st["ms_fallback"],
        "audit": {"queries": int(len(idx)), "served_frac_within_1e-3": float(np.mean(audit <= 1e-3)),
                  "served_max_err": float(audit.max()), "surrogate_alone_frac_within_1e-3":
                  float(np.mean(raw_err <= 1e-3)), "floor": floor},
        "roofline": {"bound": "mfma", "kernel": "mlp_infer_kernel (BF16 MFMA 16x16x32, hi+lo activations)",
                     "achieved": raw_qps * MLP_FLOP_PER_QUERY / 1e12, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": raw_qps * MLP_FLOP_PER_QUERY / 1e12 / PEAK_BF16_TFLOPS,
                     "flop_per_query": MLP_FLOP_PER_QUERY,
                     "note": "issued MFMA work is 2x the algorithmic FLOP (hi and lo activation halves)"},
    })
    return res


def bench_gauss16(bk, steps: int, warmup: int, cpu: bool, seconds: float):
    import torch

    from bbrcit_kde_amd import datasets as ds
    from oracle.oracle import Oracle, default_threads

    R = ds.standard_normal(G16_N, G16_D, 1004)
    Q = ds.standard_normal(G16_N, G16_D, 1005)
    m = np.full(G16_N, 1.0 / G16_N)
    desc = bk.KernelDesc(bk.BK_GAUSSIAN, G16_D, G16_H)
    out = torch.empty(G16_N, dtype=torch.float64, device="cuda")
    res = {"workload": "d=16 GaussianKernel h=0.5, 200,000 x 200,000 standard-normal points (seeds 1004/1005)",
           "unit": "pair-evals/s"}
    pairs = float(G16_N) * G16_N
    vals = {}
    with bk.DirectKde(R, Q, m) as kde:
        stream = torch.cuda.Stream()
        kde.set_stream(stream.cuda_stream)
        for name, form in (("gram_fp64_mfma", bk.BK_HIGH_DIM_AUTO), ("difference_fp64_valu", bk.BK_HIGH_DIM_DIRECT)):
            kde.set_option(bk.BK_OPT_HIGH_DIM_FORM, form)
            for _ in range(warmup):
                kde.eval_device(desc, 0, G16_N - 1, 0, G16_N - 1, out.data_ptr(), True)
            torch.cuda.synchronize()
            ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
            for e0, e1 in ev:
                e0.record(stream)
                kde.eval_device(desc, 0, G16_N - 1, 0, G16_N - 1, out.data_ptr(), True)
                e1.record(stream)
            torch.cuda.synchronize()
            ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev]))
            vals[name] = out.cpu().numpy()
            res[name] = {"ms_per_step": ms, "pair_evals_per_s": pairs / (ms * 1e-3)}
        kde.set_stream(0)
    g = res["gram_fp64_mfma"]
    res["value"] = g["pair_evals_per_s"]
    res["ms_per_step"] = g["ms_per_step"]
    res["speedup_vs_difference_form"] = g["pair_evals_per_s"] / res["difference_fp64_valu"]["pair_evals_per_s"]
    res["roofline"] = {"bound": "mfma", "pipe": "FP64 MFMA (dot products + C_r), then FP64/FP32 VALU (terms); on gfx950 the two do not overlap",
                       "achieved": g["pair_evals_per_s"] * 2 * G16_D / 1e12, "peak": PEAK_FP64_TFLOPS,
                       "unit": "TFLOP/s", "frac": g["pair_evals_per_s"] * 2 * G16_D / 1e12 / PEAK_FP64_TFLOPS,
                       "flop_per_pair_mfma": 2 * G16_D, "flop_per_pair_algorithmic": G16_FLOP_PER_PAIR,
                       "algorithmic_tflops": g["pair_evals_per_s"] * G16_FLOP_PER_PAIR / 1e12}
    # parity on 128 sampled queries against the FP64 oracle
    idx = np.linspace(0, G16_N - 1, 128).astype(np.int64)
    orc = Oracle()
    t0 = time.perf_counter()
    want = orc.direct(bk.BK_GAUSSIAN, G16_H, G16_H, ds.round_f32(R), ds.round_f32(Q[idx]), mass=ds.round_f32(m),
                      normalize=True, threads=default_threads())
    dt = time.perf_counter() - t0
    norm = (2 * np.pi) ** (-G16_D / 2) * G16_H ** (-G16_D)
    floor = 1e-10 * norm * float(ds.round_f32(m).sum())
    above = want >= floor
    res["parity"] = {"queries": int(len(idx)), "floor": floor, "frac_below_floor": float(1 - above.mean()),
                     "tolerance": 1e-5}
    for name, v in vals.items():
        rel = np.abs(v[idx][above] - want[above]) / want[above]
        res["parity"][f"max_rel_err_{name}"] = float(rel.max())
    if cpu:
        threads = default_threads()
        rate = len(idx) * G16_N / dt
        nq = int(min(G16_N, max(128, rate * seconds / G16_N)))
        t0 = time.perf_counter()
        orc.direct(bk.BK_GAUSSIAN, G16_H, G16_H, ds.round_f32(R), ds.round_f32(Q[:nq]), mass=ds.round_f32(m),
                   normalize=True, threads=threads)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": nq * G16_N / dt, "unit": "pair-evals/s", "cores": threads, "kind": "port",
                               "sample": f"{nq} queries x {G16_N} refs, FP64 restatement, {threads} threads (the "
                                         "reference evaluates d=16 on the CPU only)"}
    return res


DT_SIZES = (1 << 19, 1 << 22)   # test/test_kde11.cc sizes 2 << 18 and 2 << 21
DT_H, DT_REL, DT_ABS = 0.1, 1e-6, 1e-10
DT_CPU_N = 1 << 18


def bench_dual_tree(bk, steps: int, warmup: int, cpu: bool):
    """The reference's dual-tree benchmark (test/test_kde11.cc): N standard-
    normal 2-D points evaluated at themselves, EpanechnikovKernel<2,float>
    h = 0.1, rel 1e-6, abs 1e-10; here one bounded device evaluation over the
    points in kd order (bk_eval_bounded_device)."""
    import torch

    from bbrcit_kde_amd import datasets as ds

    desc = bk.KernelDesc(bk.BK_EPANECHNIKOV, 2, float(np.float32(DT_H)))
    res = {"workload": "test_kde11: N standard-normal 2-D points evaluated at themselves, "
                       "EpanechnikovKernel<2,float> h=0.1, rel 1e-6, abs 1e-10, points in kd order, bounds on "
                       "the device", "unit": "points/s", "sizes": {}}
    for n in DT_SIZES:
        pts = ds.standard_normal(n, 2, 1011)
        pts = pts[ds.kd_order(pts, 64)]
        mass = np.full(n, 1.0 / n)
        lo = torch.empty(n, dtype=torch.float64, device="cuda")
        up = torch.empty(n, dtype=torch.float64, device="cuda")
        with bk.DirectKde(pts, None, mass) as kde:
            stream = torch.cuda.Stream()
            kde.set_stream(stream.cuda_stream)
            for _ in range(warmup):
                kde.eval_bounded_device(desc, 0, n - 1, DT_REL, DT_ABS, lo.data_ptr(), up.data_ptr())
            torch.cuda.synchronize()
            ms = []
            for _ in range(steps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                kde.eval_bounded_device(desc, 0, n - 1, DT_REL, DT_ABS, lo.data_ptr(), up.data_ptr())
                e1.record(stream)
                torch.cuda.synchronize()
                ms.append(e0.elapsed_time(e1))
            kde.set_stream(0)
            _, _, st = kde.eval_bounded(desc, 0, n - 1, DT_REL, DT_ABS)
        l, u = lo.cpu().numpy(), up.cpu().numpy()
        gap_ok = float(np.mean((u - l) <= np.maximum(DT_ABS, DT_REL * l) * (1 + 1e-5)))
        avg = float(np.mean(ms))
        # (query block of 1024, reference tile of 256) pairs summed exactly
        summed = float(st["evaluated"]) * 1024 * 256
        res["sizes"][str(n)] = {"ms": avg, "points_per_s": n / (avg * 1e-3), "frac_within_tolerance": gap_ok,
                                "pruning": st, "pairs_summed": summed, "pairs_all": float(n) * n,
                                "summed_pair_evals_per_s": summed / (avg * 1e-3)}
        del lo, up
    # end to end through the C++ API, as test_kde11.cc:87-116 calls it:
    # KernelDensity<2, EpanechnikovKernel<2,float>, double>(data, 4096) and
    # eval(queries, 1e-6, 1e-10, 4096): host kd-tree builds, uploads, the
    # device evaluation, results back in the query tree's order
    import ctypes

    drv = ROOT / "tests" / "cpp" / "libapi_driver.so"
    if drv.exists():
        lib = ctypes.CDLL(str(drv))
        dp = ctypes.POINTER(ctypes.c_double)
        lib.api_tree_eval.restype = ctypes.c_long
        lib.api_tree_eval.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp,
                                      dp, ctypes.c_size_t, ctypes.c_int, dp, ctypes.c_size_t, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, dp, dp, dp]
        n = DT_SIZES[-1]
        pts = np.ascontiguousarray(ds.standard_normal(n, 2, 1011))
        w = np.ones(n)
        qo, lo_h, up_h = np.empty((n, 2)), np.empty(n), np.empty(n)
        ptr = lambda a: a.ctypes.data_as(dp)  # noqa: E731
        t0 = time.perf_counter()
        k = lib.api_tree_eval(bk.BK_EPANECHNIKOV, 2, 1, DT_H, DT_H, ptr(pts), ptr(w), n, 4096, ptr(pts), n, DT_REL,
                              DT_ABS, 4096, ptr(qo), ptr(lo_h), ptr(up_h))
        dt = time.perf_counter() - t0
        if k == n:
            res["end_to_end_cpp_api"] = {"n": n, "ms": dt * 1e3, "points_per_s": n / dt,
                                         "vs_reference_published_gpu": 37719.9 / (dt * 1e3),
                                         "vs_reference_published_cpu": 489527.0 / (dt * 1e3),
                                         "host_threads": os.cpu_count()}
    big = res["sizes"][str(DT_SIZES[-1])]
    res["value"] = big["points_per_s"]
    res["ms_per_step"] = big["ms"]
    res["reference_published"] = {"n": DT_SIZES[-1], "gpu_ms": 37719.9, "cpu_ms": 489527.0,
                                  "source": "test/cached_outputs/test_kde11_{gpu,cpu}.csv:15 (unknown sm_35 GPU "
                                            "and CPU; includes the host tree build and transfers)"}
    if cpu:
        from oracle.oracle import REF_LIB, Reference

        if REF_LIB.exists():
            pts = ds.standard_normal(DT_CPU_N, 2, 1011)
            t0 = time.perf_counter()
            Reference().tree_eval(bk.BK_EPANECHNIKOV, 1, DT_H, DT_H, pts, np.ones(DT_CPU_N), pts, DT_REL, DT_ABS,
                                  leaf_nmax=32, q_leaf_nmax=32)
            dt = time.perf_counter() - t0
            res["cpu_baseline"] = {"value": DT_CPU_N / dt, "unit": "points/s", "cores": 1, "kind": "reference",
                                   "sample": f"N={DT_CPU_N}: KernelDensity<2,EpanechnikovKernel<2,float>,double>"
                                             "::eval(queries, 1e-6, 1e-10, 32) (its dual tree, including the query "
                                             "tree build), 1 thread",
                                   "seconds": dt}
    return res


def bench_adaptive(bk, pts, mass, q):
    from bbrcit_kde_amd import datasets as ds

    import torch

    desc = bk.KernelDesc(bk.BK_GAUSSIAN_PRODUCT_2D, 2, 0.05, 0.05)
    n, nq = len(pts), len(q)
    res = {"workload": "1M mixture points: pilot GaussianProductKernel2d (0.05, 0.05) with self term, "
                       "lambda_i = (f_i/g)^-0.5 with FP64 geometric mean g, adaptive evaluation at 1M uniform "
                       "queries on [-4,4]^2, leave-one-out LCV of the adapted density", "unit": "pair-evals/s"}
    out = torch.empty(nq, dtype=torch.float64, device="cuda")
    with bk.DirectKde(pts, None, mass) as kde:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g, lam = kde.adapt(desc, 0.5, mass)
        t1 = time.perf_counter()
        lcv = kde.likelihood_cv(desc, mass)
        t2 = time.perf_counter()
        a_mass = mass * lam ** -2.0
    with bk.DirectKde(pts, q, a_mass, lam) as ev:
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        ev.eval_device(desc, 0, n - 1, 0, nq - 1, out.data_ptr(), True)
        torch.cuda.synchronize()
        t4 = time.perf_counter()
    pairs = 2.0 * n * n + float(n) * nq
    total = (t1 - t0) + (t2 - t1) + (t4 - t3)
    # pair terms the kernels evaluate (instrumented replays of the same three
    # calls, outside the timed calls; the FP64 recomputation of underflowed
    # leave-one-out sums is not counted)
    with bk.DirectKde(pts, None, mass) as kde:
        kde.set_option(bk.BK_OPT_COUNT_TERMS, 1)
        kde.adapt(desc, 0.5, mass)
        _, q_pilot = kde.term_count()
        kde.likelihood_cv(desc, mass)
        _, q_loo = kde.term_count()
    with bk.DirectKde(pts, q, a_mass, lam) as ev:
        ev.set_option(bk.BK_OPT_COUNT_TERMS, 1)
        ev.eval_device(desc, 0, n - 1, 0, nq - 1, out.data_ptr(), True)
        torch.cuda.synchronize()
        _, q_eval = ev.term_count()
    evaluated = float(q_pilot + q_loo + q_eval)
    res.update({"value": pairs / total, "value_kind": "effective: every ordered pair of the three passes counted",
                "evaluated_pairs": evaluated, "evaluated_frac": evaluated / pairs,
                "evaluated_pair_evals_per_s": evaluated / total,
                "evaluated_by_pass": {"pilot": q_pilot, "leave_one_out": q_loo, "adaptive_eval": q_eval},
                "pairs": pairs, "s_adapt": t1 - t0, "s_lcv": t2 - t1,
                "s_adaptive_eval": t4 - t3, "g": g, "lcv": lcv,
                "lambda_range": [float(lam.min()), float(lam.max())],
                "timing": "host wall clock around each synchronous C-ABI call (includes the FP64 reductions and "
                          "the host mass update of the adaptive pass)",
                "note": "the pilot and leave-one-out self sums run over kd-ordered points, so their workgroups "
                        "skip the reference tiles beyond 2^-100 of the kernel; in every pass wavefronts pass over "
                        "tiles each of whose terms is below 2^-56 of its mass for all their queries; value counts "
                        "every ordered pair, evaluated_pairs the terms the kernels computed"})
    return res


# FP64 instructions per pair of the FP64-context Gaussian kernel (gfx950
# ISA of this build: differences, exponent, the degree-12 exp2, ldexp,
# accumulate) and the measured cost of one FP64 VALU instruction per wave
F64_GAUSS_INSTR_PER_PAIR = 24
F64_CYCLES_PER_INSTR = 4.4


def bench_fp64(bk, pts, mass, q, steps: int, warmup: int):
    """The 1M x 1M workload in an FP64 context (the C++ API's default
    HipDirectKde<D, double>): both kernels in the given query order and in
    the device Z-curve order, parity on 256 sampled queries at 1e-12."""
    from oracle.oracle import Oracle, default_threads

    res = {"workload": "the headline 1M x 1M workload in a BK_PREC_F64 context (FP64 points and arithmetic)",
           "unit": "pair-evals/s"}
    pairs = float(N_REF) * float(N_QUERY)
    idx = np.linspace(0, N_QUERY - 1, 256).astype(np.int64)
    orc = Oracle()
    with bk.DirectKde(pts, q, mass, precision=bk.BK_PREC_F64) as kde:
        for name, kind, h in (("gauss", bk.BK_GAUSSIAN, (H_GAUSS, H_GAUSS)),
                              ("epanprod", bk.BK_EPANECHNIKOV_PRODUCT_2D, H_EPAN)):
            desc = bk.KernelDesc(kind, 2, *h)
            vals, ms, _ = time_kernel(kde, desc, steps, warmup)
            avg = float(np.mean(ms))
            kde.set_option(bk.BK_OPT_QUERY_ORDER, bk.BK_QUERY_ORDER_SPATIAL)
            svals, sms, _ = time_kernel(kde, desc, steps, 1)
            kde.set_option(bk.BK_OPT_QUERY_ORDER, bk.BK_QUERY_ORDER_GIVEN)
            want = orc.direct(kind, h[0], h[1], pts, q[idx], mass=mass, normalize=True, threads=default_threads())
            floor = 1e-10 * orc.normalization(kind, 2, *h) * float(mass.sum())
            above = want >= floor
            rel = np.abs(vals[idx][above] - want[above]) / want[above]
            res[name] = {"ms_per_step": avg, "pair_evals_per_s": pairs / (avg * 1e-3),
                         "spatial_order_ms": float(np.mean(sms)),
                         "max_abs_diff_spatial_vs_given": float(np.max(np.abs(svals - vals))),
                         "parity": {"queries": int(len(idx)), "tolerance": 1e-12,
                                    "frac_below_floor": float(1 - above.mean()),
                                    "max_rel_err": float(rel.max()) if above.any() else 0.0}}
    g = res["gauss"]["pair_evals_per_s"]
    bound = 256 * 4 * 2.4e9 * 64 / (F64_GAUSS_INSTR_PER_PAIR * F64_CYCLES_PER_INSTR)
    res["value"] = g
    res["roofline"] = {"bound": "valu", "pipe": "FP64 VALU issue",
                       "achieved": g, "peak": bound, "unit": "pair-evals/s", "frac": g / bound,
                       "note": f"{F64_GAUSS_INSTR_PER_PAIR} FP64 instructions per Gaussian pair at "
                               f"{F64_CYCLES_PER_INSTR} cycles per wave instruction (measured), 256 CUs x 4 SIMDs "
                               "x 2.4 GHz"}
    return res


def headline_roofline(head):
    """Roofline of the headline kernel.  `achieved` and `frac` count the pair
    terms the kernel evaluates (from its instrumented variant); frac_nominal
    counts all 1e12 pairs of the workload."""
    pairs = float(N_REF) * float(N_QUERY)
    sec = head["avg_ms"] * 1e-3
    t = head.get("terms") or {}
    executed = float(t.get("query_terms", pairs))
    ach = executed * FLOP_PER_PAIR_GAUSS / sec / 1e12
    return {"bound": "valu", "pipe": "FP32 VALU with v_exp_f32 (quarter-rate transcendental); compute-bound",
            "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": ach / PEAK_FP32_TFLOPS,
            "frac_executed": ach / PEAK_FP32_TFLOPS,
            "frac_nominal": pairs * FLOP_PER_PAIR_GAUSS / sec / 1e12 / PEAK_FP32_TFLOPS,
            "evaluated_pairs": executed, "nominal_pairs": pairs, "evaluated_frac_of_nominal": executed / pairs,
            "executed_pair_evals_per_s": executed / sec,
            "lane_terms": t.get("lane_terms"), "count_results_identical_to_timed": t.get("results_identical_to_timed"),
            "flop_per_pair": FLOP_PER_PAIR_GAUSS,
            "note": "achieved/frac: evaluated pair terms x 9 FLOP / kernel time; a wavefront passes over a "
                    "256-reference tile or 64-reference sub-tile each of whose terms is below 2^-56 of its mass "
                    "for all its 64 queries",
            "traffic": GAUSS_HBM_BYTES_PER_LAUNCH,
            "traffic_note": "bytes/launch, rocprofv3 FETCH_SIZE+WRITE_SIZE scaled by a copy-kernel calibration",
            "hbm_gb_s": GAUSS_HBM_BYTES_PER_LAUNCH / sec / 1e9}


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1", "n_gpus": args.gpus}))
        return 2
    import torch

    import bbrcit_kde_amd as bk

    log(f"host {platform.node()} cpus={os.cpu_count()}; generating 1M x 1M workload")
    pts, mass, q = make_workload()
    kinds = [k.strip() for k in args.kernels.split(",") if k.strip()]
    if not any(k in kinds for k in ("gauss", "epanprod")):
        kinds.insert(0, "gauss")
    results = {}
    with bk.DirectKde(pts, q, mass) as kde:
        if args.gauss_form == "direct":
            kde.set_option(bk.BK_OPT_GAUSS_FORM, bk.BK_GAUSS_FORM_DIRECT)
        for k in kinds:
            if k in ("cvgrid", "surrogate", "gauss16", "adaptive", "dualtree", "fp64"):
                continue
            if k == "gauss":
                desc, flop = bk.KernelDesc(bk.BK_GAUSSIAN, 2, H_GAUSS), FLOP_PER_PAIR_GAUSS
            elif k == "epanprod":
                desc, flop = bk.KernelDesc(bk.BK_EPANECHNIKOV_PRODUCT_2D, 2, *H_EPAN), FLOP_PER_PAIR_EPANPROD
            else:
                raise SystemExit(f"unknown kernel {k}")
            log(f"timing {k}: warmup {args.warmup}, steps {args.steps}")
            vals, ms, wall = time_kernel(kde, desc, args.steps, args.warmup)
            avg_ms = float(np.mean(ms))
            pairs = float(N_REF) * float(N_QUERY)
            results[k] = {"vals": vals, "ms": ms, "avg_ms": avg_ms, "pair_evals_per_s": pairs / (avg_ms * 1e-3),
                          "tflops": pairs * flop / (avg_ms * 1e-3) / 1e12, "wall_s": wall}
            log(f"{k}: {avg_ms:.3f} ms/step, {pairs / (avg_ms * 1e-3):.4e} pair-evals/s")
            results[k]["terms"] = count_terms(bk, kde, desc, vals)
            log(f"{k}: evaluated {results[k]['terms']['query_terms']:.4e} of {pairs:.4e} pair terms")
            if k == "gauss":
                results[k]["pcie_inclusive"] = time_host_boundary(bk, pts, mass, q, desc)
            if args.no_spatial_order:
                continue
            # the same evaluation with the queries sorted on the device first
            kde.set_option(bk.BK_OPT_QUERY_ORDER, bk.BK_QUERY_ORDER_SPATIAL)
            svals, sms, _ = time_kernel(kde, desc, min(args.steps, 3), 1)
            kde.set_option(bk.BK_OPT_QUERY_ORDER, bk.BK_QUERY_ORDER_GIVEN)
            results[k]["spatial_order"] = {
                "ms_per_step": float(np.mean(sms)), "speedup": avg_ms / float(np.mean(sms)),
                "max_abs_diff_vs_given": float(np.max(np.abs(svals - vals))),
                "note": "BK_QUERY_ORDER_SPATIAL: queries sorted along a Z curve on the device (sort, gather, "
                        "scatter inside the timed region); workgroups skip reference tiles beyond 2^-100 of "
                        "the kernel, so this is not a pair-evals/s figure"}
            log(f"{k} spatial order: {np.mean(sms):.3f} ms/step")

    cv_res = None
    if "cvgrid" in kinds:
        log(f"timing cvgrid: warmup {args.warmup}, steps {min(args.steps, 2)}")
        cv_res = bench_cv_grid(bk, min(args.steps, 2), args.warmup, not args.no_cpu_baseline, args.cpu_seconds,
                               tune=not args.no_cv_tune)
        log(f"cvgrid: {cv_res['ms_per_step']:.1f} ms/step, {cv_res['value']:.4e} pair-bw-evals/s")

    sur_res = None
    if "surrogate" in kinds and "gauss" in results:
        log("surrogate: labels, training, serving")
        sur_res = bench_surrogate(bk, pts, mass, results["gauss"]["pair_evals_per_s"] / N_REF, min(args.steps, 3),
                                  args.warmup, args.sur_steps)
        log(f"surrogate: {sur_res['ms_per_step']:.1f} ms per 10M, {sur_res['value']:.4e} q/s "
            f"({sur_res['vs_direct']:.2f}x direct)")

    g16_res = ad_res = None
    if "gauss16" in kinds:
        log("timing gauss16 (d=16, 200k x 200k)")
        g16_res = bench_gauss16(bk, min(args.steps, 3), args.warmup, not args.no_cpu_baseline, args.cpu_seconds)
        log(f"gauss16: {g16_res['ms_per_step']:.2f} ms/step, {g16_res['value']:.4e} pair-evals/s")
    if "adaptive" in kinds:
        log("adaptive pass + LOO LCV + adaptive evaluation (1M)")
        ad_res = bench_adaptive(bk, pts, mass, q)
        log(f"adaptive: {ad_res['value']:.4e} pair-evals/s")

    f64_res = None
    if "fp64" in kinds:
        log("FP64 context (1M x 1M, both kernels, given and spatial order)")
        f64_res = bench_fp64(bk, pts, mass, q, min(args.steps, 2), args.warmup)
        log(f"fp64: gauss {f64_res['gauss']['ms_per_step']:.1f} ms, epanprod "
            f"{f64_res['epanprod']['ms_per_step']:.1f} ms per step")

    dt_res = None
    if "dualtree" in kinds:
        log("dual tree (test_kde11 workload, bounded device evaluation)")
        dt_res = bench_dual_tree(bk, min(args.steps, 3), args.warmup, not args.no_cpu_baseline)
        log(f"dualtree: {dt_res['ms_per_step']:.2f} ms at N={DT_SIZES[-1]}, {dt_res['value']:.4e} points/s")

    head = results.get("gauss") or next(iter(results.values()))
    line = {
        "metric": "pair-evals/s 1M\u00d71M 2D FP32 vs FP64 oracle; surrogate queries/s at \u22641e-3 rel",
        "value": head["pair_evals_per_s"],
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["avg_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 3-component 2-D Gaussian mixture refs, uniform queries)",
        "config": {"workload": "direct 2-D KDE, 1,000,000 mixture refs (kd order) x 1,000,000 uniform queries "
                               "on [-4,4]^2, GaussianKernel h=0.05", "n_ref": N_REF, "n_query": N_QUERY,
                   "seed": 1002, "accumulation": "FP32 per 256-ref stage, FP64 across stages"},
        "roofline": headline_roofline(head),
    }
    if "epanprod" in results:
        e = results["epanprod"]
        pairs = float(N_REF) * float(N_QUERY)
        ev = float(e["terms"]["query_terms"])
        sec = e["avg_ms"] * 1e-3
        ach = ev * FLOP_PER_PAIR_EPANPROD / sec / 1e12
        line["epanechnikov_product"] = {
            "pair_evals_per_s": e["pair_evals_per_s"], "ms_per_step": e["avg_ms"],
            "spatial_order": e.get("spatial_order"), "arithmetic": "f64 on f32 points",
            "evaluated_pairs": ev, "nominal_pairs": pairs, "evaluated_frac": ev / pairs,
            "evaluated_pair_evals_per_s": ev / sec, "terms": e["terms"],
            "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                         "frac": ach / PEAK_FP64_TFLOPS, "frac_executed": ach / PEAK_FP64_TFLOPS,
                         "frac_nominal": e["tflops"] / PEAK_FP64_TFLOPS},
            "flop_per_pair": FLOP_PER_PAIR_EPANPROD, "hx": H_EPAN[0], "hy": H_EPAN[1],
            "note": "pair_evals_per_s counts all 1e12 pairs; a wavefront passes over every 64-reference "
                    "sub-tile whose terms are exactly zero for all its 64 queries (sums bit-identical), so "
                    "achieved/frac count the evaluated terms x 12 FLOP (FP64) against the FP64 peak"}
    line["spatial_order"] = head.get("spatial_order")
    line["pcie_inclusive"] = head.get("pcie_inclusive")
    if not args.no_cpu_baseline:
        log("CPU baseline leg")
        cb = cpu_baseline(pts, mass, q, head["vals"], args.cpu_seconds)
        main_cb = cb.get("reference") or cb["port"]
        line["cpu_baseline"] = main_cb
        if "reference" in cb:
            line["cpu_baseline_port"] = cb["port"]
            line["cpu_baseline_1thread"] = cb["reference_1thread"]
        line["parity"] = cb["parity"]
    else:
        line["cpu_baseline"] = None
    if cv_res is not None:
        line["cv_grid"] = cv_res
    if sur_res is not None:
        line["surrogate"] = sur_res
    if g16_res is not None:
        line["gauss16"] = g16_res
    if f64_res is not None:
        line["fp64_context"] = f64_res
    if ad_res is not None:
        line["adaptive"] = ad_res
    if dt_res is not None:
        line["dual_tree"] = dt_res
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
