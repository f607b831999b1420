"""Benchmark of the MI355X-native Chameleon Cholesky path.

One step = one FP64 CHAMELEON_dposv_Tile_Async (dpotrf + two tiled
triangular solves) on device-resident descriptors, N = 32768, NRHS = 64,
A = dplgsy(bump = N, seed 51), B = dplrnt(seed 5673) -- the single-GPU FP64
configuration of the reference's headline metric.  Inputs are regenerated
on the device before every step (untimed); each step is bracketed by a
device synchronisation on both sides.

Reported (one JSON line):
  value        dposv TFLOP/s = (flops_dpotrf(N) + flops_dpotrs(N, NRHS)) / t
  roofline     trailing SYRK+GEMM update kernel (the dominant kernel): its
               algorithmic flops / its device time, timed with HIP events on
               the stream it runs on, against the FP64 MFMA peak
  cpu_baseline the oracle (C restatement of Chameleon's OpenMP-task CPU path
               with OpenBLAS tile kernels) on the host's physical cores
               (capped by OMP_NUM_THREADS), on the benchmarked workload
               itself (one run, bounded to ~30 s of CPU work), with the
               reference's CPU configuration (N=4096, NB=320, NRHS=16) and
               scipy's LAPACK dposv beside it; CPU model and BLAS build
               recorded
  latency_ms   p50 of the per-step dposv times (p99 from 100 samples on,
               else the maximum, labelled)
  solve_latency_ms  p50 / p99 of dpotrs with the factor resident in HBM
               (NRHS = 1 and 64, 100 calls each), the per-request cost of a
               repeated solve
  lapack_layout the same dposv through the LAPACK-layout drop-in
               CHAMELEON_dposv on column-major host arrays, transfers
               included (end to end, the rate a LAPACK caller sees)
  mixed        for SPD matrices of condition number kappa (1e2, 1e4, 1e6;
               A = Q diag(kappa^(-i/(N-1))) Q^T, Q from the QR factorisation
               of a seeded Gaussian matrix, generated on the device), NRHS 1
               and 64: the FP64 dposv time and backward error next to the
               mixed-precision solve with FP64 iterative refinement (FP32,
               split-BF16 and BF16 trailing products), stopped by LAPACK's
               test and again at the FP64 solve's backward error (within
               2x) -- time, speed-up over FP64, refinement steps, FP64
               backward error ||b - A x||_inf / (||A||_inf ||x||_inf N eps)
               and forward error against the FP64 solution
  batch        20 000 independent SPD systems, n = floor(exp(U(ln 16, ln 1024))),
               nrhs uniform in 1..8, A_i = dplgsy(bump = n_i, seed + i),
               B_i = dplrnt(seed + i), generated on the device: time of the
               batched factor + solve and its flop rate, FP64 and FP32
               products + FP64 refinement, whole batch and per size bucket,
               with the tuner's workgroup shapes
  serve        Poisson arrivals (2000 requests/s, 10 s by default), n
               log-uniform in [64, 8192], nrhs in {1, 4, 16}, 30 % of requests
               reusing an earlier matrix, 200 GB factor budget, continuous
               batching: sustained request rate and flop rate, p50 / p99
               latency from scheduled arrival to solution, with FP64 factors
               and with split-BF16 factors + FP64 refinement of every solve
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
# HBM bytes per trailing-update launch measured with the FETCH_SIZE / WRITE_SIZE
# counters (scaled by a plain copy's readings) for the workload they were taken on
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "trailing_update_pmc_dposv_n32768.txt")
TRAFFIC_WORKLOAD = (32768, 512)


def measured_traffic(n, nb):
    if (n, nb) != TRAFFIC_WORKLOAD or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None
METRIC = "dposv TFLOP/s @n=32k + p99 solve latency, FP64 vs mixed-prec+IR, equal bwd error"


def flops_dpotrf(n):
    n = float(n)
    return n * (n * n / 3.0 + n / 2.0 + 1.0 / 6.0)


def flops_dpotrs(n, nrhs):
    return 2.0 * float(n) * float(n) * float(nrhs)


def gpu_dposv(ch, n, nrhs, nb, steps, warmup, stats=True):
    L = ch.lib
    L.CHAMAMD_Kernel_Stats.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong),
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    ch.set_param(ch.TILE_SIZE, nb)
    times = []
    with ch.Desc.alloc(n, n, nb) as dA, ch.Desc.alloc(n, nrhs, nb) as dB:
        for it in range(warmup + steps):
            seq = ch.Sequence()
            ch.check(L.CHAMELEON_dplgsy_Tile_Async(float(n), 123, dA.ptr, 51, seq.ptr, seq.req))
            ch.check(L.CHAMELEON_dplrnt_Tile_Async(dB.ptr, 5673, seq.ptr, seq.req))
            ch.check(seq.wait())
            seq.destroy()
            timed = it >= warmup
            if timed and it == warmup and stats:
                L.CHAMELEON_Enable(6)       # CHAMELEON_GENERATE_STATS
                L.CHAMAMD_Kernel_Stats_Reset()
            seq = ch.Sequence()
            t0 = time.perf_counter()
            ch.check(L.CHAMELEON_dposv_Tile_Async(ch.LOWER, dA.ptr, dB.ptr, seq.ptr, seq.req))
            info = seq.wait()
            t1 = time.perf_counter()
            seq.destroy()
            if info != 0:
                raise RuntimeError(f"dposv info={info}")
            if timed:
                times.append(t1 - t0)
            print(f"[bench] step {it} ({'timed' if timed else 'warmup'}) n={n} nb={nb}: {1e3 * (t1 - t0):.1f} ms",
                  file=sys.stderr, flush=True)
    kstats = {}
    if stats:
        names = {0: "update", 1: "potrf", 2: "trsm", 3: "solve", 4: "panel", 5: "colupd"}
        for cls, name in names.items():
            la, ms, fl = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
            L.CHAMAMD_Kernel_Stats(cls, ctypes.byref(la), ctypes.byref(ms), ctypes.byref(fl))
            kstats[name] = {"launches": la.value, "ms": ms.value, "flops": fl.value}
        L.CHAMELEON_Disable(6)
    return times, kstats


def latency_summary(ms):
    """p50 and p99 of per-call times; with fewer than 100 samples the 99th
    percentile is not resolved, so the maximum is reported instead and
    labelled as such."""
    ms = np.asarray(ms, dtype=np.float64)
    out = {"p50": round(float(np.percentile(ms, 50)), 3), "samples": int(ms.size)}
    if ms.size >= 100:
        out["p99"] = round(float(np.percentile(ms, 99)), 3)
    else:
        out["max"] = round(float(ms.max()), 3)
    return out


def solve_latency(ch, n, nb, reps):
    """dpotrs latency with the factor already resident on the device."""
    L = ch.lib
    ch.set_param(ch.TILE_SIZE, nb)
    out = {}
    with ch.Desc.alloc(n, n, nb) as dA:
        seq = ch.Sequence()
        ch.check(L.CHAMELEON_dplgsy_Tile_Async(float(n), 123, dA.ptr, 51, seq.ptr, seq.req))
        ch.check(L.CHAMELEON_dpotrf_Tile_Async(ch.LOWER, dA.ptr, seq.ptr, seq.req))
        ch.check(seq.wait())
        seq.destroy()
        for nrhs in (1, 64):
            lat = []
            with ch.Desc.alloc(n, nrhs, nb) as dB:
                for it in range(reps + 2):
                    seq = ch.Sequence()
                    ch.check(L.CHAMELEON_dplrnt_Tile_Async(dB.ptr, 5673 + it, seq.ptr, seq.req))
                    ch.check(seq.wait())
                    seq.destroy()
                    seq = ch.Sequence()
                    t0 = time.perf_counter()
                    ch.check(L.CHAMELEON_dpotrs_Tile_Async(ch.LOWER, dA.ptr, dB.ptr, seq.ptr, seq.req))
                    ch.check(seq.wait())
                    t1 = time.perf_counter()
                    seq.destroy()
                    if it >= 2:
                        lat.append((t1 - t0) * 1e3)
            out[f"nrhs{nrhs}"] = latency_summary(lat)
            print(f"[bench] dpotrs n={n} nrhs={nrhs}: {out[f'nrhs{nrhs}']}", file=sys.stderr, flush=True)
    return out


def mixed_leg(ch, n, nb, kappas, nrhs_list=(1, 64), precs=("fp32", "bf16x3", "bf16"), seed=1234):
    """FP64 dposv against mixed precision + FP64 iterative refinement on
    A = Q diag(lambda) Q^T, Q from the QR factorisation of a seeded Gaussian
    N x N matrix, lambda_i = kappa^(-i/(N-1)) (CHAMAMD_dplgsy_cond_Tile,
    ChamDistNormal; generated on the device, untimed), B = dplrnt(seed
    5673).  Every solve starts from the same A and B (device copies).
    Per precision two timings: refinement stopped by LAPACK dsposv's test,
    and refinement driven until the backward error is within 2x of the FP64
    dposv's on the same A and B ("matched").  Reported per result: time,
    speed-up over FP64, ITER, FP64 backward error
    ||b - A x||_inf / (||A||_inf ||x||_inf N eps) and forward error
    max_j ||x_j - x64_j||_inf / ||x64_j||_inf against the FP64 solution."""
    L = ch.lib
    ch.set_param(ch.TILE_SIZE, nb)
    codes = {"fp32": 0, "bf16": 1, "bf16x3": 2, "fp8": 3}
    res = {"matrix": "A = Q diag(kappa^(-i/(N-1))) Q^T, Q = qr(seeded Gaussian N x N), B = dplrnt(seed 5673)",
           "N": n, "NB": nb}

    def seqrun(fn):
        seq = ch.Sequence()
        rc = fn(seq)
        info = seq.wait()
        seq.destroy()
        ch.check(rc)
        return info

    def copy(src, dst):
        seqrun(lambda q: L.CHAMELEON_dlacpy_Tile_Async(123, src.ptr, dst.ptr, q.ptr, q.req))

    with ch.Desc.alloc(n, n, nb) as dA, ch.Desc.alloc(n, n, nb) as dA0:
        t0 = time.perf_counter()
        ch.check(L.CHAMAMD_dplgsy_cond_Tile(float(kappas[0]), 203, dA0.ptr, seed))
        res["generator_first_s"] = round(time.perf_counter() - t0, 2)
        for kappa in kappas:
            t0 = time.perf_counter()
            ch.check(L.CHAMAMD_dplgsy_cond_Tile(float(kappa), 203, dA0.ptr, seed))
            gen_s = time.perf_counter() - t0
            row = {"generator_s": round(gen_s, 2)}
            for nrhs in nrhs_list:
                flops = flops_dpotrf(n) + flops_dpotrs(n, nrhs)
                with ch.Desc.alloc(n, nrhs, nb) as dB, ch.Desc.alloc(n, nrhs, nb) as dX, \
                        ch.Desc.alloc(n, nrhs, nb) as dX64:
                    seqrun(lambda q: L.CHAMELEON_dplrnt_Tile_Async(dB.ptr, 5673, q.ptr, q.req))
                    # FP64 dposv: X64 = B, factor and solve in place
                    copy(dA0, dA)
                    copy(dB, dX64)
                    t0 = time.perf_counter()
                    info = seqrun(lambda q: L.CHAMELEON_dposv_Tile_Async(ch.LOWER, dA.ptr, dX64.ptr, q.ptr, q.req))
                    t64 = time.perf_counter() - t0
                    be64 = ctypes.c_double(-1.0)
                    ch.check(L.CHAMAMD_Backward_Error(ch.LOWER, dA0.ptr, dX64.ptr, dB.ptr, ctypes.byref(be64)))
                    x64 = dX64.to_array(n, nrhs)
                    x64n = np.abs(x64).max(axis=0)
                    cell = {"fp64": {"ms": round(t64 * 1e3, 2), "tflops": round(flops / t64 / 1e12, 3),
                                     "berr": be64.value, "info": info}}
                    print(f"[bench] kappa={kappa:g} nrhs={nrhs} fp64 dposv {t64 * 1e3:.1f} ms berr {be64.value:.3g}",
                          file=sys.stderr, flush=True)
                    copy(dA0, dA)
                    for pname in precs:
                        out = {}
                        for mode, target in (("lapack", 0.0), ("matched", 2.0 * be64.value)):
                            ch.check(L.CHAMAMD_Set_Refinement_Target(target))
                            it, berr = ctypes.c_int(-99), ctypes.c_double(-2.0)
                            t0 = time.perf_counter()
                            info = seqrun(lambda q: L.CHAMAMD_dposv_mixed_Tile_Async(
                                ch.LOWER, dA.ptr, dB.ptr, dX.ptr, codes[pname], ctypes.byref(it),
                                ctypes.byref(berr), q.ptr, q.req))
                            tm = time.perf_counter() - t0
                            if it.value < 0:       # the FP64 fallback factored A in place
                                copy(dA0, dA)
                            ch.check(L.CHAMAMD_Backward_Error(ch.LOWER, dA0.ptr, dX.ptr, dB.ptr, ctypes.byref(berr)))
                            x = dX.to_array(n, nrhs)
                            fwd = float(np.max(np.abs(x - x64).max(axis=0) / x64n))
                            out[mode] = {"ms": round(tm * 1e3, 2), "speedup_vs_fp64": round(t64 / tm, 3),
                                         "iter": it.value, "berr": berr.value, "fwd_err_vs_fp64": fwd,
                                         "info": info, "fp64_equiv_tflops": round(flops / tm / 1e12, 3)}
                            print(f"[bench] kappa={kappa:g} nrhs={nrhs} {pname}+IR ({mode}) {tm * 1e3:.1f} ms "
                                  f"iter {it.value} berr {berr.value:.3g} fwd {fwd:.3g}", file=sys.stderr, flush=True)
                        ch.check(L.CHAMAMD_Set_Refinement_Target(0.0))
                        cell[pname] = out
                row[f"nrhs{nrhs}"] = cell
            res[f"{kappa:g}"] = row
    # the target of the headline metric: time-to-solution at matched
    # backward error, kappa <= 1e4, best precision
    best = {}
    for kappa in kappas:
        if kappa > 1e4:
            continue
        for nrhs in nrhs_list:
            cell = res[f"{kappa:g}"][f"nrhs{nrhs}"]
            sp = [(cell[p]["matched"]["speedup_vs_fp64"], p) for p in precs
                  if cell[p]["matched"]["iter"] >= 0]
            if sp:
                best[f"kappa{kappa:g}_nrhs{nrhs}"] = {"speedup": max(sp)[0], "prec": max(sp)[1]}
    res["matched_error_best"] = best
    return res


def lapack_leg(ch, n, nrhs, nb, reps=2):
    """The LAPACK-layout drop-in on host arrays: CHAMELEON_dposv(Lower, N,
    NRHS, A, LDA, B, LDB) with A and B column-major in (pageable) host
    memory, as a LAPACK caller hands them over -- the transfers of A's
    lower triangle and of B to the device, the factorisation and solve,
    and the copies back of the factor and the solution, end to end.  A and
    B are generated on the device and copied to the host untimed
    (CHAMELEON_dplgsy / dplrnt, LAPACK layout).  Best of `reps`."""
    L = ch.lib
    ch.set_param(ch.TILE_SIZE, nb)
    A = np.zeros((n, n), order="F")
    B = np.zeros((n, nrhs), order="F")
    times = []
    for _ in range(reps):
        ch.check(L.CHAMELEON_dplgsy(float(n), ch.LOWER, n, ctypes.c_void_p(A.ctypes.data), n, 51))
        ch.check(L.CHAMELEON_dplrnt(n, nrhs, ctypes.c_void_p(B.ctypes.data), n, 5673))
        t0 = time.perf_counter()
        info = L.CHAMELEON_dposv(ch.LOWER, n, nrhs, ctypes.c_void_p(A.ctypes.data), n,
                                 ctypes.c_void_p(B.ctypes.data), n)
        times.append(time.perf_counter() - t0)
        if info != 0:
            raise RuntimeError(f"CHAMELEON_dposv info={info}")
        print(f"[bench] lapack-layout dposv n={n}: {times[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    best = min(times)
    flops = flops_dpotrf(n) + flops_dpotrs(n, nrhs)
    del A, B
    return {"call": "CHAMELEON_dposv(Lower, N, NRHS, A, LDA=N, B, LDB=N), host column-major arrays",
            "N": n, "NRHS": nrhs, "NB": nb, "ms": round(best * 1e3, 2), "tflops": round(flops / best / 1e12, 3),
            "ms_all": [round(t * 1e3, 2) for t in times]}


BUCKET_TOPS = (64, 128, 256, 512, 1024)


def _batch_run(ch, n, nrhs, seed, reps, mixed):
    """Best time of `reps` batched solves (FP64 dposv, or FP32-product
    factorisation + FP64 refinement) after one untimed run; inputs
    regenerated on the device before every run (untimed)."""
    L = ch.lib
    count = len(n)
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    b = ctypes.c_void_p()
    ch.check(L.CHAMAMD_Batch_Create(ctypes.byref(b), count, ip(n), ip(nrhs)))
    times, iters = [], None
    try:
        for it in range(reps + 1):
            ch.check(L.CHAMAMD_dplgsy_batch(b, seed))
            ch.check(L.CHAMAMD_dplrnt_batch(b, seed))
            info = np.zeros(count, dtype=np.int32)
            itv = np.zeros(count, dtype=np.int32)
            t0 = time.perf_counter()
            if mixed:
                rc = L.CHAMAMD_dposv_mixed_batch(b, 0, ip(itv), ip(info))
            else:
                rc = L.CHAMAMD_dposv_batch(b, ip(info))
            t1 = time.perf_counter()
            if rc != 0 or np.count_nonzero(info):
                raise RuntimeError(f"batch solve rc={rc}, {np.count_nonzero(info)} failing systems")
            if it > 0:
                times.append(t1 - t0)
            iters = itv
    finally:
        L.CHAMAMD_Batch_Destroy(ctypes.byref(b))
    return min(times), times, iters


def batch_leg(ch, count=20000, reps=3, seed=2024):
    """Uneven batch (every system factored and solved by one workgroup,
    the size buckets' workgroup shapes from CHAMAMD_Batch_Tune), in FP64
    and with FP32-product factorisations + FP64 refinement, for the whole
    batch and for the systems of each size bucket on their own."""
    L = ch.lib
    rng = np.random.default_rng(11)
    n = np.floor(np.exp(rng.uniform(np.log(16), np.log(1024), count))).astype(np.int32)
    nrhs = rng.integers(1, 9, count).astype(np.int32)

    def flops_of(nn_, rr_):
        nn_ = nn_.astype(np.float64)
        return float(np.sum(nn_ * (nn_ * nn_ / 3.0 + nn_ / 2.0 + 1.0 / 6.0) + 2.0 * nn_ * nn_ * rr_))

    flops = flops_of(n, nrhs)
    # the tuner on this batch (untimed), its table reported
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    b = ctypes.c_void_p()
    ch.check(L.CHAMAMD_Batch_Create(ctypes.byref(b), count, ip(n), ip(nrhs)))
    table = (ctypes.c_int * 6)()
    t0 = time.perf_counter()
    ch.check(L.CHAMAMD_Batch_Tune(b, table))
    tune_s = time.perf_counter() - t0
    L.CHAMAMD_Batch_Destroy(ctypes.byref(b))
    out = {"systems": count, "n_range": [int(n.min()), int(n.max())], "flops": flops,
           "tuner": {"waves_per_system": dict(zip([f"<={t}" for t in BUCKET_TOPS] + [">1024"], list(table))),
                     "tune_s": round(tune_s, 2)}}
    for mode, mixed in (("fp64", False), ("fp32_ir", True)):
        best, times, iters = _batch_run(ch, n, nrhs, seed, reps, mixed)
        out[mode] = {"ms": round(best * 1e3, 2), "tflops": round(flops / best / 1e12, 3),
                     "ms_all": [round(t * 1e3, 2) for t in times]}
        if mixed:
            out[mode]["ir_steps_max"] = int(iters.max())
            out[mode]["ir_steps_mean"] = round(float(iters.mean()), 3)
        print(f"[bench] batch {count} systems {mode}: {best * 1e3:.1f} ms", file=sys.stderr, flush=True)
    per = {}
    lo = 0
    for top in BUCKET_TOPS:
        sel = (((n + 15) // 16) * 16 > lo) & (((n + 15) // 16) * 16 <= top)
        lo = top
        if not sel.any():
            continue
        nb_, rb_ = np.ascontiguousarray(n[sel]), np.ascontiguousarray(nrhs[sel])
        fl = flops_of(nb_, rb_)
        row = {"systems": int(sel.sum())}
        for mode, mixed in (("fp64", False), ("fp32_ir", True)):
            best, _, _ = _batch_run(ch, nb_, rb_, seed, 2, mixed)
            row[mode] = {"ms": round(best * 1e3, 3), "tflops": round(fl / best / 1e12, 3)}
        per[f"<={top}"] = row
        print(f"[bench] batch bucket <={top}: {row}", file=sys.stderr, flush=True)
    out["per_bucket"] = per
    out["ms"] = out["fp64"]["ms"]
    out["tflops"] = out["fp64"]["tflops"]
    return out


def serve_leg(ch, seconds, rate=2000.0, budget=200 << 30, seed=7, drain=30.0, prec=-1, nb=256):
    """Open-loop Poisson arrivals against the solve server (continuous
    batching: CHAMAMD_Server_Progress never blocks, work is admitted onto
    free lanes while earlier work runs).  n log-uniform in [64, 8192],
    nrhs in {1, 4, 16}, 30 % of requests reuse an earlier matrix (new
    right-hand sides only), A = dplgsy(bump = n, seed) generated on the
    device, 200 GB factor budget.  prec = -1: FP64 factors; else the
    product precision of the large systems' factors (refined per request in
    FP64).  Latency is counted from each request's scheduled arrival time
    (queueing included).  Arrivals stop after `seconds`; serving stops
    `drain` seconds later and requests still unserved then are reported
    (their latency is unknown, so the percentiles are given only when fewer
    than 1 % are missing)."""
    L = ch.lib
    ch.set_param(ch.TILE_SIZE, nb)
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 1.5) + 16))
    arrivals = arrivals[arrivals < seconds]
    reqs, seen = [], []
    for i in range(len(arrivals)):
        if seen and rng.random() < 0.3:
            mid, n = seen[rng.integers(len(seen))]
        else:
            n = int(np.floor(np.exp(rng.uniform(np.log(64), np.log(8192)))))
            mid = i
            seen.append((mid, n))
        reqs.append((float(arrivals[i]), mid, n, int(rng.choice([1, 4, 16]))))
    # right-hand sides: copies of one random block (synthetic values)
    base = np.asfortranarray(rng.uniform(-0.5, 0.5, (8192, 16)))
    srv = ch.Server(budget)
    srv.set_precision(prec)
    lat, live, bad = [], {}, 0
    try:
        t0 = time.perf_counter()
        nxt = 0
        deadline = seconds + drain
        while (nxt < len(reqs) or live) and time.perf_counter() - t0 < deadline:
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt][0] <= now:
                ta, mid, n, r = reqs[nxt]
                B = base[:n, :r].copy(order="F")
                live[srv.submit(mid, n, B, seed=1000 + mid)] = ta
                nxt += 1
            srv.progress()
            done = srv.finished()
            if done:
                tdone = time.perf_counter() - t0
                for t in done:
                    ok, st = srv.test(t)
                    if st != 0:
                        bad += 1
                    lat.append(tdone - live.pop(t))
            elif not live and nxt < len(reqs):
                time.sleep(max(0.0, reqs[nxt][0] - (time.perf_counter() - t0)))
        wall = time.perf_counter() - t0
        st = srv.stats()
    finally:
        srv.destroy()
    if bad:
        raise RuntimeError(f"server: {bad} requests failed")
    lat = np.array(lat) * 1e3
    served = len(lat)
    unserved = len(reqs) - served
    full = unserved <= 0.01 * len(reqs)
    out = {"precision": "fp64" if prec < 0 else {0: "fp32", 1: "bf16", 2: "bf16x3", 3: "fp8"}[prec] + "+IR (n > 1024)",
           "requests": len(reqs), "served": served, "unserved": unserved, "offered_rate": rate,
           "window_s": seconds, "wall_s": round(wall, 3), "served_rate": round(served / wall, 1),
           "tflops": round(st["flops"] / wall / 1e12, 3),
           "p50_ms": round(float(np.percentile(lat, 50)), 2) if served and full else None,
           "p99_ms": round(float(np.percentile(lat, 99)), 2) if served and full else None,
           "p50_ms_served": round(float(np.percentile(lat, 50)), 2) if served else None,
           "p99_ms_served": round(float(np.percentile(lat, 99)), 2) if served else None,
           "factored": st["factored"], "reused": st["reused"], "resident": st["resident"],
           "max_lanes_busy": st["max_jobs"], "tile_size": nb}
    if prec >= 0:
        out.update({"mixed_solves": st["mixed_solves"], "ir_steps": st["ir_steps"], "fallbacks": st["fallbacks"],
                    "max_berr": st["max_berr"]})
    print(f"[bench] serve: {out}", file=sys.stderr, flush=True)
    return out


def copy_probe(ch, n, nb, reps=3):
    """Plain device copy of an n x n FP64 tile matrix (dlacpy over the tile
    mirrors): n*n*8 bytes read and written per launch.  Run under the
    FETCH_SIZE / WRITE_SIZE counters it gives the scale factor between the
    counters and real bytes."""
    L = ch.lib
    ch.set_param(ch.TILE_SIZE, nb)
    with ch.Desc.alloc(n, n, nb) as dA, ch.Desc.alloc(n, n, nb) as dB:
        seq = ch.Sequence()
        ch.check(L.CHAMELEON_dplgsy_Tile_Async(float(n), 123, dA.ptr, 51, seq.ptr, seq.req))
        ch.check(L.CHAMELEON_dplgsy_Tile_Async(float(n), 123, dB.ptr, 52, seq.ptr, seq.req))
        ch.check(seq.wait())
        seq.destroy()
        for _ in range(reps):
            seq = ch.Sequence()
            ch.check(L.CHAMELEON_dlacpy_Tile_Async(123, dA.ptr, dB.ptr, seq.ptr, seq.req))
            ch.check(seq.wait())
            seq.destroy()
    print(json.dumps({"copy_probe": {"n": n, "nb": nb, "launches": reps, "bytes_read_per_launch": n * n * 8,
                                     "bytes_written_per_launch": n * n * 8}}), flush=True)


def solve_probe(ch, n, nb, nrhs, reps=3):
    """A few dpotrs calls on a resident factor (N x N, NB, NRHS): a short
    program for counter passes over the strip solves.  Algorithmic traffic
    per strip-solve launch: the factor's lower triangle, 8 N (N+1) / 2 B."""
    L = ch.lib
    ch.set_param(ch.TILE_SIZE, nb)
    times = []
    with ch.Desc.alloc(n, n, nb) as dA, ch.Desc.alloc(n, nrhs, nb) as dB:
        seq = ch.Sequence()
        ch.check(L.CHAMELEON_dplgsy_Tile_Async(float(n), 122, dA.ptr, 51, seq.ptr, seq.req))
        ch.check(L.CHAMELEON_dpotrf_Tile_Async(122, dA.ptr, seq.ptr, seq.req))
        ch.check(seq.wait())
        seq.destroy()
        for _ in range(reps):
            seq = ch.Sequence()
            ch.check(L.CHAMELEON_dplrnt_Tile_Async(dB.ptr, 5673, seq.ptr, seq.req))
            ch.check(seq.wait())
            t0 = time.perf_counter()
            ch.check(L.CHAMELEON_dpotrs_Tile_Async(122, dA.ptr, dB.ptr, seq.ptr, seq.req))
            ch.check(seq.wait())
            times.append(round((time.perf_counter() - t0) * 1e3, 3))
            seq.destroy()
    print(json.dumps({"solve_probe": {"n": n, "nb": nb, "nrhs": nrhs, "launches_per_direction": reps,
                                      "factor_bytes": 8 * n * (n + 1) // 2, "ms": times}}), flush=True)


def host_cpu_info():
    """Cores this process may run on (os.sched_getaffinity), physical cores
    among them (SMT siblings counted once), the OMP_NUM_THREADS cap of the
    host's share, and the CPU model."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/thread_siblings_list") as f:
                cores.add(f.read().strip())
        except OSError:
            cores.add(str(c))
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    cap = os.environ.get("OMP_NUM_THREADS")
    cap = int(cap) if cap and cap.isdigit() and int(cap) > 0 else None
    return {"affinity_cpus": len(cpus), "physical_cores": len(cores), "omp_num_threads": cap, "model": model}


def cpu_baseline(n, nrhs, nb, budget_s=30.0):
    """Oracle dposv (Chameleon's OpenMP-task algorithm, OpenBLAS tile
    kernels, BLAS single-threaded inside tasks) on the host's physical
    cores (capped by OMP_NUM_THREADS, the host's CPU share).  First the
    reference's CPU configuration (N=4096, NB=320, NRHS=16, best of 3),
    whose rate predicts the time of the benchmarked workload; then the
    workload itself (N, NRHS, NB), one run, when the prediction is at most
    `budget_s` seconds, else the largest N (a multiple of NB) predicted to
    fit -- a bounded sample of it.  Beside it, scipy's LAPACK dposv
    (multithreaded OpenBLAS) on the same sample, the vendor-LAPACK
    reference.  Generation is untimed."""
    import oracle
    info = host_cpu_info()
    threads = info["physical_cores"]
    if info["omp_num_threads"]:
        threads = min(threads, info["omp_num_threads"])
    path = oracle.use_blas(True)
    if path is None:
        return None
    blas = oracle.blas_config() or os.path.basename(path)

    def run(n_, nb_, nrhs_):
        A = oracle.dplgsy(float(n_), oracle.LOWER, n_, 51, nb=nb_)
        B = oracle.dplrnt(n_, nrhs_, 5673, nb=nb_)
        t0 = time.perf_counter()
        info_ = oracle.dposv(oracle.LOWER, A, n_, B, nrhs_, nb_, nthreads=threads)
        t1 = time.perf_counter()
        assert info_ == 0
        return t1 - t0

    try:
        n0, nb0, r0 = 4096, 320, 16
        small = min(run(n0, nb0, r0) for _ in range(3))
        rate0 = (flops_dpotrf(n0) + flops_dpotrs(n0, r0)) / small
        ns = n
        while ns > nb and (flops_dpotrf(ns) + flops_dpotrs(ns, nrhs)) / rate0 > budget_s:
            ns -= nb
        print(f"[bench] cpu baseline: dposv N={ns} NB={nb} NRHS={nrhs} on {threads} threads", file=sys.stderr,
              flush=True)
        dt = run(ns, nb, nrhs)
    finally:
        oracle.use_blas(False)
    vendor = None
    try:
        import scipy.linalg.lapack as sl
        from threadpoolctl import threadpool_limits
        A = oracle.dplgsy(float(ns), oracle.LOWER, ns, 51)
        B = oracle.dplrnt(ns, nrhs, 5673)
        with threadpool_limits(limits=threads, user_api="blas"):
            t0 = time.perf_counter()
            _, _, vinfo = sl.dposv(A, B, lower=1, overwrite_a=1, overwrite_b=1)
            vt = time.perf_counter() - t0
        if vinfo == 0:
            vendor = {"value": round((flops_dpotrf(ns) + flops_dpotrs(ns, nrhs)) / vt / 1e12, 4),
                      "unit": "TFLOP/s", "ms": round(vt * 1e3, 1),
                      "sample": f"scipy.linalg.lapack.dposv N={ns} NRHS={nrhs}, {threads} BLAS threads"}
    except Exception as e:          # the vendor timing is informative only
        print(f"[bench] scipy dposv timing skipped: {e}", file=sys.stderr, flush=True)
    tf = (flops_dpotrf(ns) + flops_dpotrs(ns, nrhs)) / dt / 1e12
    what = "the benchmarked workload" if ns == n else f"a bounded sample of the workload (N={n} predicted over {budget_s:.0f} s)"
    return {"value": round(tf, 4), "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"{what}: dposv N={ns} NB={nb} NRHS={nrhs}, one run, {dt:.1f} s; OpenMP tasks + {blas} "
                      f"tile kernels, generation untimed",
            "host": info, "blas": blas,
            "reference_cpu_config": {"value": round(rate0 / 1e12, 4), "unit": "TFLOP/s",
                                     "sample": f"dposv N={n0} NB={nb0} NRHS={r0} (Chameleon's CPU configuration), "
                                               f"best of 3, {small * 1e3:.0f} ms"},
            "vendor_lapack": vendor}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--nb", type=int, default=0, help="tile size; 0 = pick the fastest of 256/384/512")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-latency", action="store_true")
    ap.add_argument("--no-lapack", action="store_true", help="skip the LAPACK-layout (host arrays) dposv leg")
    ap.add_argument("--kappas", default="1e2,1e4,1e6", help="condition numbers of the mixed-precision leg")
    ap.add_argument("--mixed-nb", type=int, default=0, help="tile size of the mixed-precision leg (0: the FP64 one)")
    ap.add_argument("--mixed-precs", default="fp32,bf16x3,bf16,fp8", help="products of the mixed-precision leg")
    ap.add_argument("--mixed-nrhs", default="1,64", help="right-hand-side counts of the mixed-precision leg")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve", type=float, default=10.0, metavar="SECONDS",
                    help="arrival window of the serving legs (0: skip)")
    ap.add_argument("--no-serve-mixed", dest="serve_mixed", action="store_false",
                    help="skip the serving leg with mixed-precision factors")
    ap.add_argument("--solve-probe", type=int, default=0, metavar="NRHS",
                    help="only run 3 dpotrs calls (N, NB 512, NRHS) on a resident factor and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="N",
                    help="only run an N x N device copy (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build drives exactly one GPU (asked for {args.gpus}); "
              "multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)

    import chameleon_amd as ch
    ch.init()
    if args.copy_probe:
        copy_probe(ch, args.copy_probe, 512)
        return
    if args.solve_probe:
        solve_probe(ch, args.n, 512, args.solve_probe)
        return
    n, nrhs = args.n, args.nrhs
    nb = args.nb
    tune = {}
    if nb == 0:
        # one untimed dposv first (code objects, workspaces), then the best
        # of two runs per candidate tile size
        gpu_dposv(ch, n, nrhs, 384, 1, 0, stats=False)
        for cand in (256, 384, 512):
            t, _ = gpu_dposv(ch, n, nrhs, cand, 2, 0, stats=False)
            tune[cand] = min(t)
        nb = min(tune, key=tune.get)
    times, kst = gpu_dposv(ch, n, nrhs, nb, args.steps, args.warmup)
    total = float(np.sum(times))
    flops = flops_dpotrf(n) + flops_dpotrs(n, nrhs)
    value = flops * len(times) / total / 1e12
    up = kst.get("update", {})
    achieved = up["flops"] / (up["ms"] * 1e-3) / 1e12 if up.get("ms") else None
    roofline = {"bound": "mfma", "kernel": "trailing_update_kernel (DSYRK+DGEMM)",
                "achieved": round(achieved, 3) if achieved else None, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 4) if achieved else None,
                "traffic": measured_traffic(n, nb), "traffic_unit": "bytes/launch",
                "launches": up.get("launches"),
                "avg_launch_ms": round(up["ms"] / up["launches"], 4) if up.get("launches") else None}
    slat = None if args.no_latency else solve_latency(ch, n, nb, 100)
    lapack = None if args.no_lapack else lapack_leg(ch, n, nrhs, nb)
    mixed = None if args.no_mixed else mixed_leg(ch, n, args.mixed_nb or nb,
                                                 tuple(float(x) for x in args.kappas.split(",")),
                                                 tuple(int(x) for x in args.mixed_nrhs.split(",")),
                                                 tuple(args.mixed_precs.split(",")))
    # the conditioned generator keeps its Q factor for the next call: free
    # it before the batch and serving legs, which size their pools to HBM
    ch.check(ch.lib.CHAMAMD_Cond_Cache_Release())
    batch = None if args.no_batch else batch_leg(ch)
    serve = None
    if args.serve > 0:
        serve = {"fp64": serve_leg(ch, args.serve)}
        if args.serve_mixed:
            serve["mixed"] = serve_leg(ch, args.serve, prec=2)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, nrhs, nb)
    lat = np.array(times) * 1e3
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": len(times),
        "warmup": args.warmup,
        "ms_per_step": round(total / len(times) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: A = dplgsy(bump=N, seed 51), B = dplrnt(seed 5673), generated on the device",
        "config": {"workload": f"FP64 dposv (dpotrf + dpotrs) N={n} NRHS={nrhs} NB={nb} uplo=Lower, "
                               "device-resident tile descriptors", "N": n, "NRHS": nrhs, "NB": nb,
                   "nb_tuning_ms": {str(k): round(v * 1e3, 1) for k, v in tune.items()}},
        "latency_ms": latency_summary(lat),
        "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                        "tflops": round(v["flops"] / (v["ms"] * 1e-3) / 1e12, 3) if v["ms"] else None}
                    for k, v in kst.items()},
        "solve_latency_ms": slat,
        "lapack_layout": lapack,
        "mixed": mixed,
        "batch": batch,
        "serve": serve,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ch.finalize()


if __name__ == "__main__":
    main()
