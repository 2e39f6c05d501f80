"""Benchmark: verified interval GEMM at n = 8192 on one MI355X.

One step = one verified interval product C = IA * IB (the pidblas::mulmm
result, vcp/pidblas.hpp:239-294) of the 8192 x 8192 wide-exponent workload
(mid = +-m 2^e, m U[1,2), e U{-20..20}, radius 2^-45 |mid|), inputs resident
in HBM, through the Ozaki-II int8-MFMA form.  The plain FP64 RU/RD form is
timed on the same inputs for the ratio.  Work per step: 8 n^3 FP64-equivalent
flops (the four real GEMMs of pidblas::mulmm).

Prints ONE JSON line (see the keys below).  --gpus > 1 is not supported: the
path runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

PEAK_INT8_TOPS = 5000.0      # MI355X dense int8 MFMA, TOP/s
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector / matrix


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--samples", type=int, default=512)
    ap.add_argument("--no-resid", action="store_true")
    ap.add_argument("--resid-reps", type=int, default=120)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this path runs on one GPU; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import vcp_amd
    from vcp_amd import workloads as W

    n = args.n
    log("generating the n=%d wide-exponent interval workload" % n)
    t0 = time.time()
    if n == 8192:
        IA, IB = W.config3(n)
    else:
        rng = np.random.default_rng(W.SEED0 + 3)
        IA, IB = W.wide_exponent_intervals(rng, n, n), W.wide_exponent_intervals(rng, n, n)
    log("generated in %.1fs" % (time.time() - t0))

    dev = torch.device("cuda:0")
    # AoS column-major pairs: the (cols, rows, 2) C-contiguous buffer
    dA = torch.from_numpy(np.ascontiguousarray(IA.transpose(1, 0, 2))).to(dev)
    dB = torch.from_numpy(np.ascontiguousarray(IB.transpose(1, 0, 2))).to(dev)
    dC = torch.empty((n, n, 2), dtype=torch.float64, device=dev)
    dP = torch.empty((n, n, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    ctx = vcp_amd.Context(0)
    beta, mods = vcp_amd.ozaki_config(n)
    nmod = len(mods)

    def step(form, out):
        ctx.igemm_dev(form, n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, out.data_ptr(), n)

    # ---- Ozaki form (the headline) --------------------------------------
    log("ozaki warmup x%d" % args.warmup)
    for _ in range(args.warmup):
        step(vcp_amd.FORM_OZAKI, dC)
    ctx.synchronize()
    torch.cuda.synchronize()
    ctx.set_param(vcp_amd.PARAM_STAGE_TIMING, 1)
    log("ozaki timed x%d" % args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(vcp_amd.FORM_OZAKI, dC)
    ctx.synchronize()
    torch.cuda.synchronize()
    t_oz = (time.perf_counter() - t0) / args.steps
    stages = ctx.stage_times()
    ctx.set_param(vcp_amd.PARAM_STAGE_TIMING, 0)
    log("ozaki %.3f ms/step, stages %s" % (t_oz * 1e3, stages))

    # ---- plain FP64 directed-rounding form --------------------------------
    log("plain warmup x1, timed x%d" % args.plain_steps)
    step(vcp_amd.FORM_PLAIN, dP)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        step(vcp_amd.FORM_PLAIN, dP)
    ctx.synchronize()
    t_pl = (time.perf_counter() - t0) / max(args.plain_steps, 1)
    log("plain %.3f ms/step" % (t_pl * 1e3))

    flops = 8.0 * n * n * n
    tf_oz = flops / t_oz / 1e12
    tf_pl = flops / t_pl / 1e12

    # ---- radius ratio (all entries, on the GPU) and sampled containment ----
    rad_o = dC[..., 1] - dC[..., 0]
    rad_p = dP[..., 1] - dP[..., 0]
    ratio = (rad_o / torch.clamp(rad_p, min=torch.finfo(torch.float64).tiny))
    ratio_max = float(ratio.max())
    ratio_med = float(ratio.median())
    del rad_o, rad_p, ratio
    containment = None
    try:
        import oracle_ffi as O

        C = dC.cpu().numpy().transpose(1, 0, 2)
        P = dP.cpu().numpy().transpose(1, 0, 2)
        rng = np.random.default_rng(12345)
        ii = rng.integers(0, n, args.samples)
        jj = rng.integers(0, n, args.samples)
        ok_o, _, _ = O.check_containment(IA, IB, C, ii, jj)
        ok_p, _, _ = O.check_containment(IA, IB, P, ii, jj)
        containment = {"samples": int(args.samples), "ozaki": float(ok_o.mean()), "plain": float(ok_p.mean())}
        log("containment %s" % containment)
    except Exception as e:  # the checker is optional for the measurement
        log("containment check skipped: %r" % (e,))

    # ---- roofline of the dominant kernel (int8 MFMA GEMM stage) ------------
    t_gemm = stages.get("gemm")
    roof = None
    if t_gemm:
        # one launch: nmod residue products + 6 radius-digit plane products
        # (hi*hi, hi*lo + lo*hi for each of the two radius terms), 2 n^3 each
        ops = 2.0 * n * n * n * (nmod + 6)
        ach = ops / (t_gemm * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 2), "peak": PEAK_INT8_TOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_INT8_TOPS, 4), "traffic": None,
                "kernel": "k_oz_gemm (int8 MFMA: %d residue + 6 radius-digit plane products)" % nmod,
                "ops_per_launch": ops, "ms_per_launch": round(t_gemm, 4)}

    # ---- CPU baseline: the oracle's pidblas::mulmm restatement -------------
    cpu = None
    if not args.no_cpu:
        try:
            import oracle_ffi as O

            threads = O.lib().orc_get_threads()
            rows = 8
            t0 = time.perf_counter()
            O.pidblas_mulmm(np.ascontiguousarray(IA[:rows].transpose(1, 0, 2)).transpose(1, 0, 2), IB)
            t_small = time.perf_counter() - t0
            rows = int(max(8, min(n, rows * args.cpu_seconds / max(t_small, 1e-3))))
            sub = np.ascontiguousarray(IA[:rows].transpose(1, 0, 2)).transpose(1, 0, 2)
            log("cpu baseline: %d rows x %d x %d on %d threads" % (rows, n, n, threads))
            t0 = time.perf_counter()
            O.pidblas_mulmm(sub, IB)
            t_cpu = time.perf_counter() - t0
            cpu = {"value": round(8.0 * rows * n * n / t_cpu / 1e12, 5), "unit": "TFLOP/s", "cores": int(threads),
                   "kind": "port",
                   "sample": "pidblas::mulmm restated (oracle, sequential-k FMA chains under fesetround, OpenMP) "
                             "on the first %d rows of the n=%d workload (%d x %d x %d), %.1f s" % (rows, n, rows, n, n, t_cpu)}
        except Exception as e:
            log("cpu baseline skipped: %r" % (e,))

    resid = None if args.no_resid else bench_resid(ctx, args.resid_reps)

    line = {
        "metric": "verified interval GEMM FP64-equiv TFLOP/s n=8192: Ozaki-int8 vs FP64-RU/RD",
        "value": round(tf_oz, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_oz * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": "interval GEMM n=%d, mid=+-m*2^e (m U[1,2), e U{-20..20}), rad=2^-45|mid|" % n,
                   "n": n, "form": "ozaki-int8", "moduli": nmod, "beta": beta},
        "plain_fp64_rurd": {"tflops": round(tf_pl, 3), "ms_per_step": round(t_pl * 1e3, 3),
                            "frac_fp64_peak": round(tf_pl / PEAK_FP64_TFLOPS, 4)},
        "speedup_vs_plain": round(t_pl / t_oz, 3),
        "radius_ratio_ozaki_over_plain": {"max": ratio_max, "median": ratio_med},
        "containment": containment,
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched_residual": resid,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def bench_resid(ctx, reps):
    """256 verified residuals R_i A_i - I of uneven size (n log-uniform in
    [64, 1024]) in one graph-captured persistent launch; latency per batch."""
    import numpy as np
    import torch

    import vcp_amd
    from vcp_amd import workloads as W

    log("generating 256 residual problems")
    probs = W.config5()
    ns = [R.shape[0] for R, _ in probs]
    dev = torch.device("cuda:0")
    dR = [torch.from_numpy(np.ascontiguousarray(R.T)).to(dev) for R, _ in probs]
    dA = [torch.from_numpy(np.ascontiguousarray(IA.transpose(1, 0, 2))).to(dev) for _, IA in probs]
    dC = [torch.empty((n, n, 2), dtype=torch.float64, device=dev) for n in ns]
    nub = torch.empty(len(ns), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    args = (ns, [t.data_ptr() for t in dR], ns, [t.data_ptr() for t in dA], ns, [t.data_ptr() for t in dC], ns,
            nub.data_ptr())
    vcp_amd.residual_batched_dev(ctx, *args)      # builds the descriptors and captures the graph
    ctx.synchronize()
    lat = []
    for _ in range(reps):
        t0 = time.perf_counter()
        vcp_amd.residual_batched_dev(ctx, *args)  # graph replay
        ctx.synchronize()
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat) * 1e3
    flops = 6.0 * sum(float(n) ** 3 for n in ns)
    p50 = float(np.percentile(lat, 50))
    nb = nub.cpu().numpy()
    out = {"problems": len(ns), "n_min": int(min(ns)), "n_max": int(max(ns)), "graph_launches": reps,
           "p50_ms": round(p50, 4), "p99_ms": round(float(np.percentile(lat, 99)), 4),
           "tflops_p50": round(flops / (p50 * 1e-3) / 1e12, 3), "all_normub_below_1": bool((nb < 1.0).all()),
           "max_normub": float(nb.max())}
    try:
        import oracle_ffi as O

        idx = sorted(range(len(ns)), key=lambda i: ns[i])[:6]
        ok = True
        for i in idx:
            Co, nbo = O.residual(*probs[i])
            ok = ok and np.array_equal(dC[i].cpu().numpy().transpose(1, 0, 2), Co) and nb[i] == nbo
        out["oracle_bitexact_smallest6"] = bool(ok)
    except Exception as e:
        log("residual oracle check skipped: %r" % (e,))
    log("batched residual %s" % out)
    return out


if __name__ == "__main__":
    main()
