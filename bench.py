#!/usr/bin/env python3
"""Benchmark: time to a 1e-10 FP64 true residual at n = 16384 (synthetic
conditioned SPD, A = H Lambda H^T, Lambda log-spaced 1..1e6, block 256),
FP8-update mixed-precision Cholesky-PCG vs the FP32-update factor on the same
GPU, plus the SYRK roofline and the CPU restatement timed on a bounded sample.

One "step" is one full solve from A resident in HBM: copy A into the factor
workspace, factor (E4M3/FP32 per-tile switching), PCG to ||b - Ax|| / ||b||
<= 1e-10.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEED_A = 20240531
SEED_B = 7
HBM_PEAK_GBS = 8000.0
FP8_PEAK_TFLOPS = 5000.0
FP32_MFMA_PEAK_TFLOPS = 157.3


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nb", type=int, default=256)
    ap.add_argument("--eps", type=float, default=0.125)
    ap.add_argument("--agg", type=int, default=None,
                    help="panels per aggregated trailing update (library default 2)")
    ap.add_argument("--ladder", default="e4m3,fp32",
                    help="per-tile precision ladder of the mixed run, lowest first")
    ap.add_argument("--syrk-reps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="time only the first CPU_STEPS block steps of the CPU factor and "
                         "extrapolate by flops (0: the whole factor and PCG)")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP64 forward/backward error report")
    ap.add_argument("--syrk-only", action="store_true",
                    help="only the trailing update at the first-step shape and a 1 GiB device "
                         "copy of known size (for hardware-counter runs), one JSON line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (the factor and PCG run on a single "
              "MI355X; multi-GPU sharding is a design note, not a built path).", file=sys.stderr)
        sys.exit(2)

    import ctypes as C
    import numpy as np
    import torch
    from cholesky_gpu_amd import _lib as L
    from cholesky_gpu_amd import solver as S

    n, nb = args.n, args.nb
    h = S.Handle(0)
    stream = torch.cuda.Stream()
    L.check(h.lib.mpc_set_stream(h.h, C.c_void_p(stream.cuda_stream)), "mpc_set_stream")

    if args.syrk_only:
        sy = syrk_roofline(h, stream, n, nb, args.syrk_reps, C0=None)
        sya = syrk_roofline(h, stream, n, nb, args.syrk_reps, C0=None,
                            k=(args.agg or 2) * nb)
        cp = copy_reference(stream, args.syrk_reps)
        syd = syrk_roofline(h, stream, n, nb, args.syrk_reps, C0=None, k=2048)
        print(json.dumps({"syrk": sy, "syrk_aggregated": sya, "copy": cp, "syrk_deep": syd}),
              flush=True)
        h.close()
        return

    log("generating A", n)
    A, lam = S.matgen_householder(h, n, cond=1e6, lam_max=1e6, seed=SEED_A)
    B = torch.from_numpy(np.random.default_rng(SEED_B).standard_normal((1, n))).cuda()
    names = {"e4m3": L.PREC_E4M3, "e5m2": L.PREC_E5M2, "mxe4m3": L.PREC_MXE4M3,
             "mxe5m2": L.PREC_MXE5M2, "bf16": L.PREC_BF16, "fp16": L.PREC_FP16, "fp32": L.PREC_FP32}
    ladder = tuple(names[x] for x in args.ladder.split(","))
    fo8 = S.factor_opts(nb=nb, update_prec=L.PREC_SWITCH, ladder=ladder, eps_prime=args.eps,
                        fallback=1, agg_panels=args.agg)
    fo32 = S.factor_opts(nb=nb, update_prec=L.PREC_FP32, ladder=(L.PREC_FP32,), fallback=0,
                         agg_panels=args.agg)
    po = S.pcg_opts(tol=1e-10, maxit=500)

    def solve(fo):
        X, fi, pi, rc = S.precond_cg(h, A, B, fo, po)
        if rc != 0:
            raise RuntimeError(f"solve failed rc={rc} {S.info_dict(fi)} {S.info_dict(pi)}")
        return X, fi, pi

    def timed(fo, k, w):
        for _ in range(w):
            solve(fo)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = None
        for _ in range(k):
            out = solve(fo)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / k, out

    log("fp8-update solves")
    ms8, (X8, fi8, pi8) = timed(fo8, args.steps, args.warmup)
    log("fp32-update solves")
    ms32, (X32, fi32, pi32) = timed(fo32, max(1, args.steps), max(1, args.warmup))

    # verification outside the timed region: FP64 true residual of both answers
    A64 = A.to(torch.float64)
    b = B[0]
    rr8 = float(torch.linalg.norm(b - A64 @ X8[0]) / torch.linalg.norm(b))
    rr32 = float(torch.linalg.norm(b - A64 @ X32[0]) / torch.linalg.norm(b))
    del A64
    log(f"fp8: {ms8:.2f} ms its={pi8.iters} rr={rr8:.2e}; fp32: {ms32:.2f} ms its={pi32.iters} rr={rr32:.2e}")

    # ---- accuracy (outside the timed region): forward error against an FP64
    # host Cholesky solve of the same FP32 matrix, backward error of each
    # factor (the same factor the solve used: factored again into a buffer of
    # its own with the same options), Lanczos estimate of cond(M^-1 A)
    acc = None
    if not args.no_accuracy:
        from cholesky_gpu_amd import accuracy as ACC
        acc = {}
        Xs, where = ACC.fp64_solution(A, B)
        for name, fo, X, pi in (("fp8", fo8, X8, pi8), ("fp32", fo32, X32, pi32)):
            Lf = A.clone()
            rcf, _, _, _ = S.factor(h, Lf, fo, want_map=False)
            acc[name] = {"fwd_err": ACC.forward_error(X, Xs),
                         "bwd_err": ACC.backward_error(A, Lf) if rcf == 0 else None,
                         "kappa_est": round(float(pi.kappa_est), 4)}
            del Lf
        acc["fwd_ref"] = f"FP64 Cholesky of the FP32 matrix ({where})"
        del Xs
        log("accuracy", acc)

    # ---- SYRK roofline at the first-step shape (m = n - nb, k = nb), E4M3 tiles,
    # and at the aggregated shape the factor issues (k = agg_panels * nb)
    sy = syrk_roofline(h, stream, n, nb, args.syrk_reps, C0=A)
    m, k, syrk_ms, syrk_gbs, syrk_tf = sy["m"], sy["k"], sy["ms"], sy["gbs"], sy["tflops"]
    agg = args.agg if args.agg is not None else fo8.agg_panels
    sya = syrk_roofline(h, stream, n, nb, args.syrk_reps, C0=A, k=agg * nb) if agg > 1 else None
    # the same update at K = 2048 (a depth the factor does not issue): where the
    # arithmetic intensity reaches the FP8 ridge and the matrix cores take over
    syd = syrk_roofline(h, stream, n, nb, max(2, args.syrk_reps // 4), C0=A, k=2048)
    # the streaming rate this device attains: a 1 GiB elementwise copy on the same stream
    cpy = copy_reference(stream, max(2, args.syrk_reps // 2))

    # ---- CPU restatement on a bounded sample (rank 0, N = 1)
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(n, nb, args.eps, pi8.iters, args.cpu_steps, fi8, h, A, B)
        except Exception as e:  # the GPU result stands without the CPU estimate
            log("cpu baseline failed:", repr(e))

    # PCG bytes per iteration: the upper triangle of A (SYMV) and the lower
    # triangle of L twice (two sweeps), FP32; the start (x0 = M^-1 b, r0,
    # z0) is one more product and two more preconditioner applications
    tri = n * (n + 1) / 2 * 4.0
    def pcg_rate(pi, ms):
        nbytes = (pi.iters * 3 + 5) * tri
        return round(nbytes / (ms * 1e-3) / 1e9, 1)
    pcg_gbs = {"fp8": pcg_rate(pi8, pi8.ms_total), "fp32": pcg_rate(pi32, pi32.ms_total)}

    tot_flops = max(1.0, sum(fi8.syrk_flops))
    flop_frac = {L.PREC_NAMES[p]: round(fi8.syrk_flops[p] / tot_flops, 4)
                 for p in range(len(L.PREC_NAMES)) if fi8.syrk_flops[p] > 0}
    out = {
        "metric": "time to 1e-10 FP64 true residual, n=16384, FP8/FP32 per-tile-switching "
                  "Cholesky-PCG (vs FP32-update factor PCG)",
        "value": round(ms8, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms8, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "e4m3/fp32 update, fp32 factor, fp64 residual",
        "data": "synthetic (seeded H Lambda H^T, Lambda logspace(0,6,n); b ~ N(0,1))",
        "config": {"workload": "synthetic conditioned SPD n=16384, cond 1e6, block 256, "
                               "E4M3 per-tile-scale trailing update with FP32 switching, "
                               "PCG to 1e-10",
                   "n": n, "nb": nb, "eps_prime": args.eps, "ladder": args.ladder},
        "fp32_factor_ms": round(ms32, 3),
        "speedup_vs_fp32_factor": round(ms32 / ms8, 3),
        "pcg_iters": {"fp8": pi8.iters, "fp32": pi32.iters},
        "relres": {"fp8": rr8, "fp32": rr32},
        "factor_ms": {"fp8": round(fi8.ms_total, 3), "fp32": round(fi32.ms_total, 3)},
        "pcg_ms": {"fp8": round(pi8.ms_total, 3), "fp32": round(pi32.ms_total, 3)},
        "syrk_flop_fraction": flop_frac,
        "factor_tflops": {"fp8": round(fi8.flops_total / (fi8.ms_total * 1e-3) / 1e12, 2),
                          "fp32": round(fi32.flops_total / (fi32.ms_total * 1e-3) / 1e12, 2)},
        "factor_frac_peak": {
            "fp8_vs_fp8_peak": round(fi8.flops_total / (fi8.ms_total * 1e-3) / 1e12
                                     / FP8_PEAK_TFLOPS, 4),
            "fp32_vs_fp32_mfma_peak": round(fi32.flops_total / (fi32.ms_total * 1e-3) / 1e12
                                            / FP32_MFMA_PEAK_TFLOPS, 4)},
        "pcg_gbs": pcg_gbs,
        "pcg_frac_hbm": {kk: round(v / HBM_PEAK_GBS, 4) for kk, v in pcg_gbs.items()},
        "accuracy": acc,
        "roofline": {"kernel": "syrk_kernel<E4M3> (one trailing update, m=%d k=%d)" % (m, k),
                     "bound": "hbm", "achieved": round(syrk_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(syrk_gbs / HBM_PEAK_GBS, 4),
                     "traffic": syrk_traffic(m, k),
                     "avg_ms": round(syrk_ms, 4), "tflops": round(syrk_tf, 1),
                     "mfma_frac_fp8": round(syrk_tf / FP8_PEAK_TFLOPS, 4),
                     "copy_gbs": round(cpy["gbs"], 1),
                     "frac_of_copy": round(syrk_gbs / cpy["gbs"], 4)},
        "roofline_aggregated": None if sya is None else {
            "kernel": "syrk_kernel<E4M3> (the factor's aggregated update, m=%d k=%d)"
                      % (sya["m"], sya["k"]),
            "bound": "hbm", "achieved": round(sya["gbs"], 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(sya["gbs"] / HBM_PEAK_GBS, 4),
            "traffic": syrk_traffic(sya["m"], sya["k"]), "avg_ms": round(sya["ms"], 4),
            "tflops": round(sya["tflops"], 1),
            "mfma_frac_fp8": round(sya["tflops"] / FP8_PEAK_TFLOPS, 4)},
        "syrk_deep": {
            "kernel": "syrk_kernel<E4M3> (m=%d k=%d, not issued by the factor)" % (syd["m"], syd["k"]),
            "avg_ms": round(syd["ms"], 4), "gbs": round(syd["gbs"], 1),
            "hbm_frac": round(syd["gbs"] / HBM_PEAK_GBS, 4), "tflops": round(syd["tflops"], 1),
            "mfma_frac_fp8": round(syd["tflops"] / FP8_PEAK_TFLOPS, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    h.close()


def syrk_traffic(m, k):
    """HBM bytes per launch of the trailing update at this shape, from the
    committed FETCH_SIZE / WRITE_SIZE counter runs (each scaled by what the
    same counter reads for a copy kernel of known size); None when no counter
    run of this shape is on file."""
    for name in ("syrk_e4m3_n16384_step0_counters.json", "syrk_e4m3_n16384_agg_counters.json"):
        try:
            with open(os.path.join(ROOT, "profiles", name)) as f:
                rec = json.load(f)
        except (OSError, ValueError):
            continue
        if rec.get("m") == m and rec.get("k") == k:
            return int(round(rec["traffic_bytes"]))
    return None


def syrk_roofline(h, stream, n, nb, reps, C0=None, k=None):
    """One trailing update at the first-step shape (m = n - nb, k = nb, or the
    aggregated k = a nb) with all panel tiles in E4M3, timed with HIP events
    on the stream the kernel runs on.  Algorithmic bytes: the lower triangle
    of C read and written in FP32 (8 B per element incl. the diagonal), the
    E4M3 panel and its scales."""
    import ctypes as C
    import numpy as np
    import torch
    from cholesky_gpu_amd import _lib as L
    from cholesky_gpu_amd import solver as S
    m, tr = n - nb, nb
    k = k or nb
    rng = np.random.default_rng(1)
    P = torch.from_numpy(rng.standard_normal((k, m)).astype(np.float32)).cuda()
    ntile = (m + tr - 1) // tr
    pay, sc, pt = S.quantize_panel(h, P, [L.PREC_E4M3] * ntile, tr)
    Cbuf = torch.empty(m, m, dtype=torch.float32, device="cuda")
    if C0 is not None:
        Cbuf.copy_(C0[:m, :m])
    else:
        Cbuf.normal_()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    # the masked entry point: the formats (all E4M3) are named here, so the
    # timed calls queue only the update kernel (mpc_syrk reads the tile
    # formats back to the host and synchronises on every call)
    def call():
        return h.lib.mpc_syrk_masked(h.h, C.c_void_p(Cbuf.data_ptr()), m, m, k,
                                     C.c_void_p(pay.data_ptr()), pay.stride(0),
                                     C.c_void_p(sc.data_ptr()), tr, C.c_void_p(pt.data_ptr()),
                                     1 << L.PREC_E4M3, 1)

    torch.cuda.synchronize()
    for _ in range(2):
        L.check(call(), "mpc_syrk_masked")
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(reps):
        call()
    ev1.record(stream)
    ev1.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    flops = float(m) * (m + 1) * k
    nbytes = float(m) * (m + 1) / 2 * 8 + m * k * 1.0 + m * (k // 32)
    del Cbuf, pay, sc, P
    out = {"m": m, "k": k, "ms": ms, "bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9,
           "tflops": flops / (ms * 1e-3) / 1e12}
    log(f"syrk m={m} k={k}: {ms:.3f} ms  {out['gbs']:.0f} GB/s  {out['tflops']:.1f} TFLOP/s")
    return out


def copy_reference(stream, reps, nbytes=1 << 30):
    """A plain copy kernel of known size (an elementwise kernel that reads and
    writes nbytes; a contiguous copy_ would go to the DMA engine, which the
    kernel counters do not see), the scale for the hardware byte counters."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").normal_()
    b = torch.empty_like(a)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        torch.mul(a, 1.0, out=b)
        ev0.record(stream)
        for _ in range(reps):
            torch.mul(a, 1.0, out=b)
        ev1.record(stream)
    ev1.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    log(f"copy {nbytes >> 20} MiB: {ms:.3f} ms  {2 * nbytes / (ms * 1e-3) / 1e9:.0f} GB/s")
    return {"bytes_each_way": nbytes, "ms": ms, "gbs": 2 * nbytes / (ms * 1e-3) / 1e9}


def cpu_baseline(n, nb, eps, gpu_iters, cpu_steps, fi8, h, A, B):
    """The oracle's restatement of the reference path (FP32 micro_cholesky
    diagonal blocks, TRSM, quantised FP32-accumulate SYRK with the same
    per-tile switching, FP64 PCG) timed on the host's cores: by default the
    whole factor of the same n = 16384 matrix and the PCG to the same 1e-10
    true residual on that CPU factor, plus the small RBF workload (2048
    points, block 128, FP16 update, PCG to 1e-8) in full.  cpu_steps > 0
    times only the first block steps and extrapolates by flops."""
    import numpy as np
    from oracle import oracle as O
    log("cpu baseline: generating A on the host")
    Ah, _ = O.householder_spd(n, 1e6, 1e6, SEED_A)
    nt = (n + nb - 1) // nb
    b = B.cpu().numpy()[0]
    if cpu_steps <= 0:
        t0 = time.perf_counter()
        Lh, rc, mp, ud, fi = O.factor_mixed(Ah, nb=nb, update_prec=-1, ladder=(O.E4M3, O.FP32),
                                            eps_prime=eps)
        t_factor = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"CPU factor failed (info {rc})")
        t1 = time.perf_counter()
        _, it_c, rr_c, _ = O.pcg(Ah, np.tril(Lh), b, tol=1e-10, maxit=500)
        t_pcg = time.perf_counter() - t1
        del Lh
        total_s = t_factor + t_pcg
        sample = (f"full: factor of the n={n} matrix ({t_factor:.1f} s) + PCG on that factor to "
                  f"a 1e-10 FP64 true residual ({it_c} iterations, {t_pcg:.1f} s, relres "
                  f"{rr_c:.1e})")
        log(f"cpu: factor {t_factor:.1f}s, pcg {t_pcg:.1f}s ({it_c} its)")
    else:
        steps = min(cpu_steps, nt)
        t0 = time.perf_counter()
        Lh, rc, mp, ud, fi = O.factor_mixed(Ah, nb=nb, update_prec=-1, ladder=(O.E4M3, O.FP32),
                                            eps_prime=eps, max_steps=steps)
        t_steps = time.perf_counter() - t0
        del Lh

        def step_flops(j):
            s = n - (j + 1) * nb
            return s * (s + 1.0) * nb + s * nb * nb + nb ** 3 / 3.0
        t_factor = t_steps * sum(step_flops(j) for j in range(nt)) / sum(
            step_flops(j) for j in range(steps))
        total_s = t_factor
        sample = (f"first {steps} of {nt} block steps of the factor ({t_steps:.1f} s), "
                  f"extrapolated by flops; PCG not timed")
    del Ah
    # the small RBF workload in full
    t2 = time.perf_counter()
    Xp = O.points_gaussian(2048, 16, 11)
    K = np.asarray(O.rbf_matrix(Xp, [4.0], 1.0, 1e-2), np.float32)
    b1 = np.random.default_rng(1).standard_normal(2048)
    L1, rc1, _, _, _ = O.factor_mixed(K, nb=128, update_prec=O.FP16, ladder=(O.FP16,),
                                      tile_rows=128)
    _, it1, rr1, _ = O.pcg(K, np.tril(L1), b1, tol=1e-8, maxit=200)
    t_c1 = time.perf_counter() - t2
    return {"value": round(total_s * 1e3, 1), "unit": "ms", "cores": O.lib().orc_num_threads(),
            "kind": "port", "sample": sample,
            "rbf_2048_fp16_ms": round(t_c1 * 1e3, 1),
            "rbf_2048_fp16_sample": f"full: 2048-point RBF matrix, block-128 FP16-update factor "
                                    f"and PCG to 1e-8 ({it1} iterations), generation included"}


if __name__ == "__main__":
    main()
