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
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB elementwise copy (reads 1 GiB, writes 1 GiB): the calibration "
                         "kernel for FETCH_SIZE / WRITE_SIZE counter runs")
    ap.add_argument("--no-infsup", action="store_true")
    ap.add_argument("--no-tuned", action="store_true")
    ap.add_argument("--no-config1", action="store_true")
    ap.add_argument("--no-config2", action="store_true")
    ap.add_argument("--no-galerkin", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-dd", action="store_true")
    ap.add_argument("--dd-n", type=int, default=4096)
    ap.add_argument("--dd-steps", type=int, default=5)
    ap.add_argument("--no-power", action="store_true",
                    help="skip the sustained runs that read socket power and GFX clock with amd-smi")
    ap.add_argument("--power-seconds", type=float, default=5.0)
    ap.add_argument("--variant", type=int, default=0,
                    help="int8 GEMM tile variant of the headline Ozaki product (0: the library default); "
                         "variants are bit-identical, so only the time changes")
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

    if args.copy_probe:
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        log("copy probe: 1 GiB read + 1 GiB written by one elementwise kernel")
        del src, dst

    ctx = vcp_amd.Context(0)
    if args.variant:
        ctx.set_param(vcp_amd.PARAM_OZ_VARIANT, args.variant)
    names = vcp_amd.OZ_RADIUS_NAMES
    wtiny = torch.finfo(torch.float64).tiny

    def run(c, form, out):
        c.igemm_dev(form, n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, out.data_ptr(), n)

    def timed_ozaki(c, out, warmup, steps):
        for _ in range(warmup):
            run(c, vcp_amd.FORM_OZAKI, out)
        c.synchronize()
        torch.cuda.synchronize()
        c.set_param(vcp_amd.PARAM_STAGE_TIMING, 1)
        t0 = time.perf_counter()
        for _ in range(steps):
            run(c, vcp_amd.FORM_OZAKI, out)
        c.synchronize()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / steps
        st = c.stage_times()
        c.set_param(vcp_amd.PARAM_STAGE_TIMING, 0)
        return t, st

    flops = 8.0 * n * n * n

    # ---- plain FP64 directed-rounding form (the radius reference) ----------
    log("plain warmup x1, timed x%d" % args.plain_steps)
    run(ctx, vcp_amd.FORM_PLAIN, dP)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        run(ctx, vcp_amd.FORM_PLAIN, dP)
    ctx.synchronize()
    t_pl = (time.perf_counter() - t0) / max(args.plain_steps, 1)
    tf_pl = flops / t_pl / 1e12
    log("plain %.3f ms/step" % (t_pl * 1e3))
    wp = torch.clamp(dP[..., 1] - dP[..., 0], min=wtiny)

    def ratio_of(out):
        r = (out[..., 1] - out[..., 0]) / wp
        return {"max": float(r.max()), "median": float(r.median())}

    # ---- Ozaki form, untuned defaults ---------------------------------------
    plan_d = ctx.ozaki_plan(n, n, n)
    log("ozaki default (%d moduli, %s radius digits): warmup x1, timed x%d"
        % (plan_d["moduli"], names[plan_d["radius"]], args.steps))
    t_def, st_def = timed_ozaki(ctx, dC, 1, args.steps)
    default = {"moduli": plan_d["moduli"], "radius_scheme": names[plan_d["radius"]],
               "tile_variant": plan_d["variant"], "plane_products": plan_d["products"],
               "ms_per_step": round(t_def * 1e3, 3), "tflops": round(flops / t_def / 1e12, 3),
               "speedup_vs_plain": round(t_pl / t_def, 3), "stages_ms": {k: round(x, 4) for k, x in st_def.items()},
               "radius_ratio_over_plain": ratio_of(dC)}
    log("ozaki default %s" % default)

    # ---- Ozaki form, autotuned: the headline --------------------------------
    # The tuner (before warmup, outside the timed region) picks the cheapest
    # moduli count and radius-digit scheme whose radii stay within 2x of the
    # plain form's on a row sample, then the fastest int8 tile variant for
    # that plane count.  Without tuning (--no-tuned) the headline is the
    # default configuration.
    hctx, tune = ctx, None
    dH = dC
    if not args.no_tuned:
        hctx = vcp_amd.Context(0)
        t0 = time.perf_counter()
        nm_t, rad_t, ratio_s = hctx.tune_igemm_moduli_dev(n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, 2.0, 0)
        v_t, _ = hctx.tune_igemm_tile(n, n, n)
        tune = {"moduli": nm_t, "radius_scheme": names[rad_t], "sampled_ratio": ratio_s,
                "sample_rows": max(1024, n // 8), "tile_variant": v_t, "target_ratio": 2.0,
                "tuning_s": round(time.perf_counter() - t0, 3)}
        log("tuned %s" % tune)
        dH = torch.empty_like(dC)
    plan = hctx.ozaki_plan(n, n, n)
    nmod, scheme = plan["moduli"], names[plan["radius"]]
    beta, _ = vcp_amd.ozaki_config(n, nmod)
    log("ozaki headline (%d moduli, %s radius digits, tile variant %d): warmup x%d, timed x%d"
        % (nmod, scheme, plan["variant"], args.warmup, args.steps))
    t_oz, stages = timed_ozaki(hctx, dH, args.warmup, args.steps)
    tf_oz = flops / t_oz / 1e12
    ratio_h = ratio_of(dH)
    log("ozaki %.3f ms/step, stages %s, radius ratio %s" % (t_oz * 1e3, stages, ratio_h))

    # ---- inf-sup form (pidblas_fma::mulmm), one timed product --------------
    infsup = None
    if not args.no_infsup:
        dI = torch.empty_like(dC)
        run(ctx, vcp_amd.FORM_INFSUP, dI)
        ctx.synchronize()
        t0 = time.perf_counter()
        run(ctx, vcp_amd.FORM_INFSUP, dI)
        ctx.synchronize()
        t_is = time.perf_counter() - t0
        # per (i, j, k): two chains of 4 FMAs and 3 maxima, or of 2 FMAs and
        # 1 maximum in k-slices whose operands are all strictly sign-definite
        # (every slice of this workload); counted as 2 ops per instruction
        definite = bool(((IA[..., 0] > 0) | (IA[..., 1] < 0)).all() and ((IB[..., 0] > 0) | (IB[..., 1] < 0)).all())
        per = 6 if definite else 14
        ops = 2.0 * per * n * n * n
        infsup = {"ms_per_step": round(t_is * 1e3, 3), "tflops_8n3": round(flops / t_is / 1e12, 3),
                  "two_fma_slices": definite, "fp64_instr_per_ijk": per, "fp64_ops_per_step": ops,
                  "frac_fp64_peak": round(ops / t_is / 1e12 / PEAK_FP64_TFLOPS, 4),
                  "radius_ratio_over_plain": ratio_of(dI)}
        log("infsup %s" % infsup)
        del dI

    # ---- sampled containment of the headline and default results -----------
    containment = None
    try:
        if args.samples <= 0:
            raise RuntimeError("no samples requested")
        import oracle_ffi as O

        rng = np.random.default_rng(12345)
        ii = rng.integers(0, n, args.samples)
        jj = rng.integers(0, n, args.samples)
        containment = {"samples": int(args.samples)}
        for key, X in (("ozaki", dH), ("ozaki_default", dC), ("plain", dP)):
            ok, _, _ = O.check_containment(IA, IB, X.cpu().numpy().transpose(1, 0, 2), ii, jj)
            containment[key] = float(ok.mean())
        log("containment %s" % containment)
    except Exception as e:  # the checker is optional for the measurement
        log("containment check skipped: %r" % (e,))

    # ---- roofline of the dominant kernel (int8 MFMA GEMM stage) ------------
    t_gemm = stages.get("gemm")
    roof = None
    if t_gemm:
        # one launch: nmod residue products + the radius-digit plane products
        # of the scheme (hi*hi, hi*lo + lo*hi and, full, lo*lo of each of the
        # two radius terms; half: hi*H, lo*H; ceil: H*H), 2 n^3 int8 ops each
        nprod = plan["products"]
        ops = 2.0 * n * n * n * nprod
        ach = ops / (t_gemm * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 2), "peak": PEAK_INT8_TOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_INT8_TOPS, 4), "traffic": measured_traffic(n, nmod, nprod, plan["variant"]),
                "kernel": "k_oz_gemm (int8 MFMA: %d residue + %d radius-digit plane products, %s radius scheme, "
                          "tile variant %d)" % (nmod, nprod - nmod, scheme, plan["variant"]),
                "ops_per_launch": ops, "ms_per_launch": round(t_gemm, 4)}

    # ---- power and clock under sustained load -------------------------------
    # The published peaks assume the 2400 MHz maximum GFX clock; under a
    # sustained product the board holds its socket power near the cap and
    # the clock drops.  Both readings (amd-smi) are taken while the headline
    # product, then the plain one, runs back to back; the GEMM stage's
    # fraction is also given against the int8 peak at the clock it ran at.
    power = None
    if not args.no_power:
        power = {"ozaki_headline": sustained_power(lambda: run(hctx, vcp_amd.FORM_OZAKI, dH), hctx.synchronize,
                                                   args.power_seconds),
                 "plain_fp64": sustained_power(lambda: run(ctx, vcp_amd.FORM_PLAIN, dP), ctx.synchronize,
                                               args.power_seconds)}
        power.update(power_limits())
        oz_p = power["ozaki_headline"]
        if roof and oz_p and oz_p.get("gfx_mhz") and power.get("max_gfx_mhz"):
            peak_clk = PEAK_INT8_TOPS * oz_p["gfx_mhz"] / power["max_gfx_mhz"]
            roof["gfx_mhz_sustained"] = oz_p["gfx_mhz"]
            roof["peak_at_sustained_clock"] = round(peak_clk, 1)
            roof["frac_at_sustained_clock"] = round(roof["achieved"] / peak_clk, 4)
        log("power %s" % power)

    # ---- CPU baseline: pidblas::mulmm on the reference's own vblas rmatmul
    # (oracle/_ref, built from the reference sources) when present, else the
    # oracle's restatement; a row sample of the same workload -----------------
    cpu = None
    if not args.no_cpu:
        try:
            import oracle_ffi as O

            threads = O.lib().orc_get_threads()
            use_ref = O.ref_lib() is not None
            product = O.ref_pidblas_mulmm if use_ref else O.pidblas_mulmm
            rows = 512 if use_ref else 8   # B's split is a fixed cost: calibrate on enough rows
            t0 = time.perf_counter()
            product(np.ascontiguousarray(IA[:rows].transpose(1, 0, 2)).transpose(1, 0, 2), IB)
            t_small = time.perf_counter() - t0
            rows = int(max(8, min(n, rows * args.cpu_seconds / max(t_small, 1e-3))))
            sub = np.ascontiguousarray(IA[:rows].transpose(1, 0, 2)).transpose(1, 0, 2)
            log("cpu baseline (%s): %d rows x %d x %d on %d threads" % ("reference" if use_ref else "port", rows, n,
                                                                          n, threads))
            t0 = time.perf_counter()
            Cref = product(sub, IB)
            t_cpu = time.perf_counter() - t0
            what = ("pidblas::mulmm with the reference's vblas rmatmul (AVX-512 / AVX2 micro-kernels, OpenMP) for its "
                    "four products" if use_ref else
                    "pidblas::mulmm restated (oracle, sequential-k FMA chains under fesetround, OpenMP)")
            cpu = {"value": round(8.0 * rows * n * n / t_cpu / 1e12, 5), "unit": "TFLOP/s", "cores": int(threads),
                   "kind": "reference" if use_ref else "port",
                   "sample": "%s on the first %d rows of the n=%d workload (%d x %d x %d), %.1f s, B's midpoint-radius "
                             "split included" % (what, rows, n, rows, n, n, t_cpu)}
            # the plain GPU form gives the same bits on those rows
            Cp_rows = dP[:, :rows, :].cpu().numpy().transpose(1, 0, 2)
            cpu["plain_gpu_rows_bitexact"] = bool(np.array_equal(Cp_rows, Cref))
        except Exception as e:
            log("cpu baseline skipped: %r" % (e,))
    if hctx is not ctx:
        hctx.close()

    resid = None if args.no_resid else bench_resid(ctx, args.resid_reps)
    dd = None if args.no_dd else bench_dd(ctx, args.dd_n, args.dd_steps, not args.no_cpu)
    cfg1 = None if args.no_config1 else bench_config1(ctx, with_cpu=not args.no_cpu)
    cfg2 = None if args.no_config2 else bench_config2(ctx)
    galerkin = None if args.no_galerkin else bench_galerkin(ctx)
    batched = None if args.no_batched else bench_igemm_batched(ctx)

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
                   "n": n, "form": "ozaki-int8", "moduli": nmod, "beta": beta, "radius_scheme": scheme,
                   "tile_variant": plan["variant"], "plane_products": plan["products"],
                   "configuration": "autotuned (radius ratio target 2)" if tune else "untuned defaults"},
        "plain_fp64_rurd": {"tflops": round(tf_pl, 3), "ms_per_step": round(t_pl * 1e3, 3),
                            "frac_fp64_peak": round(tf_pl / PEAK_FP64_TFLOPS, 4)},
        "speedup_vs_plain": round(t_pl / t_oz, 3),
        "radius_ratio_ozaki_over_plain": ratio_h,
        "tuning": tune,
        "untuned_default": default,
        "infsup_fp64": infsup,
        "containment": containment,
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched_residual": resid,
        "dd_gemm": dd,
        "interval_n1024": cfg1,
        "plain_n4096": cfg2,
        "galerkin_uhphiphi": galerkin,
        "igemm_batched": batched,
        "power_clock": power,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def _smi(args, tool="amd-smi"):
    import shutil
    import subprocess

    exe = shutil.which(tool)
    if not exe:
        return ""
    try:
        return subprocess.run([exe] + args, capture_output=True, text=True, timeout=20).stdout
    except Exception:
        return ""


def parse_power_clock(text):
    """(socket power in W, mean GFX clock over the XCDs in MHz) from the text
    of `amd-smi metric --power --clock`, or None when either is missing."""
    import re

    pw = re.search(r"SOCKET_POWER:\s*([0-9.]+)", text)
    clks = [int(x) for x in re.findall(r"GFX_\d+:\s*\n\s*CLK:\s*(\d+)", text)]
    if not pw or not clks:
        return None
    return float(pw.group(1)), sum(clks) / len(clks)


def sustained_power(step, sync, seconds, period=0.25):
    """Socket power (W) and mean GFX clock over the XCDs (MHz), read with
    amd-smi every `period` s while step() runs back to back for `seconds`:
    medians of the samples after the first second (the clocks settle).  None
    when amd-smi is absent or reads nothing."""
    import statistics
    import threading

    samples = []
    done = threading.Event()

    def poll():
        while not done.is_set():
            pc = parse_power_clock(_smi(["metric", "--power", "--clock"]))
            if pc:
                samples.append((time.perf_counter(),) + pc)
            done.wait(period)

    th = threading.Thread(target=poll, daemon=True)
    t0 = time.perf_counter()
    th.start()
    calls = 0
    while time.perf_counter() - t0 < seconds:
        step()
        calls += 1
        if calls % 4 == 0:
            sync()
    sync()
    done.set()
    th.join()
    use = [s for s in samples if s[0] - t0 > 1.0]
    if not use:
        return None
    return {"socket_power_w": statistics.median(s[1] for s in use),
            "gfx_mhz": round(statistics.median(s[2] for s in use)), "samples": len(use), "calls": calls,
            "seconds": round(time.perf_counter() - t0, 2)}


def power_limits():
    """The maximum GFX clock (MHz, amd-smi) and the maximum graphics package
    power (W, rocm-smi), as the tools report them (empty when they do not)."""
    import re

    out = {}
    m = re.search(r"MAX_CLK:\s*(\d+)", _smi(["metric", "--clock"]))
    if m:
        out["max_gfx_mhz"] = int(m.group(1))
    m = re.search(r"Max Graphics Package Power \(W\)\"?:\s*\"?([0-9.]+)", _smi(["-M", "--json"], tool="rocm-smi"))
    if m:
        out["power_cap_w"] = float(m.group(1))
    return out


def measured_traffic(n, nmod, nprod, variant):
    """HBM bytes per launch of the int8 GEMM stage, from the committed
    FETCH_SIZE / WRITE_SIZE counter runs (profiles/ozaki_gemm_traffic.json),
    when one was taken on this workload (size, moduli, plane products and
    tile variant); otherwise None."""
    path = os.path.join(ROOT, "profiles", "ozaki_gemm_traffic.json")
    try:
        with open(path) as f:
            recs = json.load(f).get("records", [])
    except (OSError, ValueError, AttributeError):
        return None
    for rec in recs:
        if (rec.get("n") == n and rec.get("moduli") == nmod and rec.get("plane_products") == nprod
                and rec.get("variant") == variant):
            return rec.get("bytes_per_launch")
    return None


def padded_fraction(ms, ns, ks, shapes):
    """padded fraction of the multiply-adds of a batched launch's tiles:
    rows in 64- or 32-row blocks by the shape of each problem's size class
    (shapes: ceil(log2 rows) -> mode; missing = 64 x 64), columns in 64"""
    import math

    tile = useful = 0.0
    for m, n, k in zip(ms, ns, ks):
        mode = shapes.get(max(0, math.ceil(math.log2(m))) if m > 0 else 0, 1)
        rows = 0
        while rows < m:
            rows += 32 if mode == 3 or (mode == 2 and m - rows <= 32) else 64
        tile += float(rows) * 64 * math.ceil(n / 64) * k
        useful += float(m) * n * k
    return 1.0 - useful / tile if tile else 0.0


def tuned_shapes(ctx, kind):
    """the tile shape chosen per size class (from the saved tuning table)"""
    import tempfile

    import vcp_amd

    out = {}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tune.txt")
        ctx.tune_save(path)
        with open(path) as f:
            for ln in f:
                fs = ln.split()
                if len(fs) == 4 and fs[0] == "tiles" and fs[1] == kind:
                    out["n<=%d" % (1 << int(fs[2]))] = vcp_amd.BATCH_TILES_NAMES[int(fs[3])]
    return out


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
    def replays():
        vcp_amd.residual_batched_dev(ctx, *args)      # builds the descriptors and captures the graph
        ctx.synchronize()
        lat = []
        for _ in range(reps):
            t0 = time.perf_counter()
            vcp_amd.residual_batched_dev(ctx, *args)  # graph replay
            ctx.synchronize()
            lat.append(time.perf_counter() - t0)
        return np.array(lat) * 1e3

    flops = 6.0 * sum(float(n) ** 3 for n in ns)
    lat = replays()   # untuned: 64 x 64 tiles, three workgroups per CU
    untuned = {"p50_ms": round(float(np.percentile(lat, 50)), 4),
               "frac_fp64_peak": round(flops / (float(np.percentile(lat, 50)) * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
               "padded_fraction": round(padded_fraction(ns, ns, ns, {}), 4)}
    t0 = time.perf_counter()
    per_cu, _, padded = vcp_amd.tune_residual_dev(ctx, *args)
    tuning = {"seconds": round(time.perf_counter() - t0, 3), "workgroups_per_cu": per_cu,
              "tile_shapes": tuned_shapes(ctx, "residual"), "padded_fraction": round(padded, 4)}
    log("batched residual untuned %s, tuning %s" % (untuned, tuning))
    lat = replays()
    p50 = float(np.percentile(lat, 50))
    nb = nub.cpu().numpy()
    # bytes the batched residual must move: R (8 n^2) and IA (16 n^2) read, C (16 n^2)
    # written by the tiles, C read again by the row sums (16 n^2)
    hbm = 56.0 * sum(float(n) ** 2 for n in ns)
    out = {"problems": len(ns), "n_min": int(min(ns)), "n_max": int(max(ns)), "graph_launches": reps,
           "p50_ms": round(p50, 4), "p99_ms": round(float(np.percentile(lat, 99)), 4),
           "tflops_p50": round(flops / (p50 * 1e-3) / 1e12, 3),
           "frac_fp64_peak": round(flops / (p50 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
           "hbm_algorithmic_bytes": hbm, "hbm_algorithmic_GBps_p50": round(hbm / (p50 * 1e-3) / 1e9, 1),
           "frac_hbm_peak": round(hbm / (p50 * 1e-3) / 8e12, 4),
           "all_normub_below_1": bool((nb < 1.0).all()), "max_normub": float(nb.max()),
           "configuration": "autotuned tile shapes per size class and persistent grid", "tuning": tuning,
           "untuned_64x64": untuned}
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
    # the whole verified solve on the same problems (imats::linearsolve after
    # its LAPACK step): one right-hand side each, radii 2^-50 |b|
    rng = np.random.default_rng(W.SEED0 + 51)
    Ibs = []
    for n in ns:
        b = rng.uniform(-1.0, 1.0, (n, 1))
        Ibs.append(W.interval_from_midrad(b, np.ldexp(np.abs(b), -50)))
    dB = [torch.from_numpy(np.ascontiguousarray(b.transpose(1, 0, 2))).to(dev) for b in Ibs]
    dX = [torch.empty((1, n, 2), dtype=torch.float64, device=dev) for n in ns]
    torch.cuda.synchronize()
    one = [1] * len(ns)
    largs = (ns, one, [t.data_ptr() for t in dA], ns, [t.data_ptr() for t in dB], ns, [t.data_ptr() for t in dR], ns,
             [t.data_ptr() for t in dX], ns)
    gs = ctx.lss_batched_dev(*largs)
    lat = []
    for _ in range(max(10, reps // 4)):
        t0 = time.perf_counter()
        ctx.lss_batched_dev(*largs)   # returns after its stream synchronise
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat) * 1e3
    p50 = float(np.percentile(lat, 50))
    lss = {"problems": len(ns), "nrhs": 1, "calls": len(lat), "p50_ms": round(p50, 4),
           "p99_ms": round(float(np.percentile(lat, 99)), 4), "tflops_p50": round(flops / (p50 * 1e-3) / 1e12, 3),
           "all_verified": bool(max(gs) < 1.0)}
    try:
        import oracle_ffi as O

        ok = True
        for i in sorted(range(len(ns)), key=lambda i: ns[i])[:3]:
            Xo, go, vok = O.lss(probs[i][1], Ibs[i], probs[i][0])
            ok = ok and vok and go == gs[i] and np.array_equal(dX[i].cpu().numpy().transpose(1, 0, 2), Xo)
        lss["oracle_bitexact_smallest3"] = bool(ok)
    except Exception as e:
        log("lss oracle check skipped: %r" % (e,))
    # the same solves with every R_p = inv(mid(IA_p)) formed on the device
    # first (imats.hpp:298-301; rocSOLVER getrf + getrs, one batched call per
    # order rounded up to a multiple of 64), so the whole imats::linearsolve
    # runs on the GPU
    largs_inv = largs[:6] + (None, None) + largs[8:]
    gi = ctx.lss_batched_dev(*largs_inv)
    lat = []
    for _ in range(max(3, reps // 16)):
        t0 = time.perf_counter()
        ctx.lss_batched_dev(*largs_inv)
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat) * 1e3
    lss["with_device_inverse"] = {"calls": len(lat), "orders": len(set(ns)),
                                  "p50_ms": round(float(np.percentile(lat, 50)), 4),
                                  "p99_ms": round(float(np.percentile(lat, 99)), 4),
                                  "all_verified": bool(max(gi) < 1.0)}
    log("batched verified solve %s" % lss)
    out["lss_batched"] = lss
    return out


def bench_igemm_batched(ctx, count=256, reps=20):
    """count interval products n_p x n_p x n_p (n_p log-uniform in [64, 1024],
    the batched-residual size distribution; midpoints U[-1,1], radii
    2^-40 |mid|) in one graph-captured persistent launch, against the same
    products launched one by one.  FP64 flops: 8 n^3 per product (four FMA
    chains per output)."""
    import numpy as np
    import torch

    from vcp_amd import workloads as W

    ns = W.config5_sizes(count)
    rng = np.random.default_rng(W.SEED0 + 61)
    dev = torch.device("cuda:0")
    log("generating %d batched interval products" % count)
    dA = [torch.from_numpy(np.ascontiguousarray(W.uniform_intervals(rng, n, n, -40).transpose(1, 0, 2))).to(dev)
          for n in ns]
    dB = [torch.from_numpy(np.ascontiguousarray(W.uniform_intervals(rng, n, n, -40).transpose(1, 0, 2))).to(dev)
          for n in ns]
    dC = [torch.empty((n, n, 2), dtype=torch.float64, device=dev) for n in ns]
    torch.cuda.synchronize()
    args = (ns, ns, ns, [t.data_ptr() for t in dA], ns, [t.data_ptr() for t in dB], ns, [t.data_ptr() for t in dC], ns)
    flops = 8.0 * sum(float(n) ** 3 for n in ns)

    def replays():
        ctx.igemm_batched_dev(*args)   # builds the tile list and captures the graph
        ctx.synchronize()
        lat = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ctx.igemm_batched_dev(*args)
            ctx.synchronize()
            lat.append(time.perf_counter() - t0)
        return float(np.percentile(np.array(lat) * 1e3, 50))

    p50u = replays()   # untuned: 64 x 64 tiles
    untuned = {"p50_ms": round(p50u, 4), "frac_fp64_peak": round(flops / (p50u * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
               "padded_fraction": round(padded_fraction(ns, ns, ns, {}), 4)}
    t0 = time.perf_counter()
    _, padded = ctx.tune_igemm_batched_dev(*args)
    tuning = {"seconds": round(time.perf_counter() - t0, 3), "tile_shapes": tuned_shapes(ctx, "igemm"),
              "padded_fraction": round(padded, 4)}
    log("batched interval products untuned %s, tuning %s" % (untuned, tuning))
    p50 = replays()
    # every product against its single plain product, on the smallest and largest
    import vcp_amd

    check = sorted(range(len(ns)), key=lambda i: ns[i])
    same = True
    for i in check[:4] + check[-2:]:
        C1 = torch.empty_like(dC[i])
        ctx.igemm_dev(vcp_amd.FORM_PLAIN, ns[i], ns[i], ns[i], dA[i].data_ptr(), ns[i], dB[i].data_ptr(), ns[i],
                      C1.data_ptr(), ns[i])
        ctx.synchronize()
        same = same and bool(torch.equal(C1, dC[i]))
    # the same products one launch each
    for i, n in enumerate(ns):
        ctx.igemm_dev(vcp_amd.FORM_PLAIN, n, n, n, dA[i].data_ptr(), n, dB[i].data_ptr(), n, dC[i].data_ptr(), n)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        for i, n in enumerate(ns):
            ctx.igemm_dev(vcp_amd.FORM_PLAIN, n, n, n, dA[i].data_ptr(), n, dB[i].data_ptr(), n, dC[i].data_ptr(), n)
    ctx.synchronize()
    t_single = (time.perf_counter() - t0) / 3 * 1e3
    out = {"problems": count, "n_min": int(min(ns)), "n_max": int(max(ns)), "graph_launches": reps,
           "p50_ms": round(p50, 4), "tflops_p50": round(flops / (p50 * 1e-3) / 1e12, 3),
           "frac_fp64_peak": round(flops / (p50 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
           "configuration": "autotuned tile shapes per size class", "tuning": tuning, "untuned_64x64": untuned,
           "single_product_bitexact_6": same, "one_launch_per_product_ms": round(t_single, 4)}
    log("batched interval products %s" % out)
    del dA, dB, dC
    return out


def bench_galerkin(ctx, p=40, q=80, steps=5):
    """The Galerkin matrix producer (Legendre_Bases_Generator::uhphiphi) on a
    2D Legendre basis: m = p^2 basis functions, q^2 Gauss points, interval
    arithmetic.  Work: m (m + 1) / 2 q^2 interval multiply-adds.  Each is 8
    FP64 products, 6 maxima and 2 additions in the general form, and 4
    products, 2 maxima and 2 additions where the operands of a stage of 16
    points all lie strictly on one side of zero (every stage of this basis
    and solution, which have no zero at the Gauss points)."""
    import numpy as np
    import torch

    import vcp_amd
    from vcp_amd import workloads as W

    phi, w, u = W.galerkin(p, q)
    m, npts = phi.shape[:2]
    dev = torch.device("cuda:0")
    dP = torch.from_numpy(np.ascontiguousarray(phi.transpose(1, 0, 2))).to(dev)
    dW = torch.from_numpy(w).to(dev)
    dU = torch.from_numpy(u).to(dev)
    dA = torch.empty((m, m, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    L = vcp_amd.lib()
    args = (ctx._h, m, npts, dP.data_ptr(), m, dW.data_ptr(), dU.data_ptr(), dA.data_ptr(), m)
    assert L.vcphip_uhphiphi_dev(*args) == 0
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        L.vcphip_uhphiphi_dev(*args)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    ima = m * (m + 1) / 2 * npts
    # every product and sum is separately rounded (no FMA): the bound is the
    # FP64 vector instruction rate, half the FMA flop peak, at the executed
    # instructions per multiply-add (8 when every operand is sign-definite)
    definite = bool(((phi[..., 0] > 0) | (phi[..., 1] < 0)).all() and ((u[:, 0] > 0) | (u[:, 1] < 0)).all()
                    and (w[:, 0] > 0).all())
    per = 8 if definite else 16
    instr = per * ima / (ms * 1e-3) / 1e12
    out = {"basis": "2D Legendre, p=%d (m=%d), %dx%d Gauss points" % (p, m, q, q), "ms": round(ms, 3),
           "interval_madds_per_s": round(ima / (ms * 1e-3), 1), "two_product_stages": definite,
           "fp64_instr_per_madd": per, "fp64_instr_per_s_e12": round(instr, 3),
           "frac_fp64_issue_peak": round(instr / (PEAK_FP64_TFLOPS / 2), 4)}
    try:
        import oracle_ffi as O

        idx = np.linspace(0, m - 1, 8).astype(int)
        A = dA.cpu().numpy().transpose(1, 0, 2)
        sub = O.pairs(len(idx), npts)
        sub[...] = phi[idx]
        t1 = time.perf_counter()
        ref = O.uhphiphi(sub, w, u)
        out["oracle_bitexact_8rows"] = bool(np.array_equal(A[np.ix_(idx, idx)], ref))
        out["cpu_oracle_8rows_s"] = round(time.perf_counter() - t1, 3)
    except Exception as e:
        log("galerkin oracle check skipped: %r" % (e,))
    log("galerkin %s" % out)
    return out


def bench_config1(ctx, n=1024, steps=100, with_cpu=True):
    """The interval product at n=1024 (mid U[-1,1], radius 2^-40 |mid|): the
    configuration of the reference's timing demo of this path
    (test_matrix/test_interval_matrix_mul.cpp:46-72).  Plain and Ozaki forms
    on the GPU (8 n^3 FP64-equivalent flops each), 4096 sampled entries
    checked for exact containment, radii against the reference's, and the
    reference's own pidblas::mulmm (vblas rmatmul, OpenMP) on the host cores
    for the whole product, median of three."""
    import numpy as np
    import torch

    import vcp_amd
    from vcp_amd import workloads as W

    IA, IB = W.config1(n)
    dev = torch.device("cuda:0")
    dA = torch.from_numpy(np.ascontiguousarray(IA.transpose(1, 0, 2))).to(dev)
    dB = torch.from_numpy(np.ascontiguousarray(IB.transpose(1, 0, 2))).to(dev)
    dP = torch.empty((n, n, 2), dtype=torch.float64, device=dev)
    dZ = torch.empty((n, n, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    flops = 8.0 * n ** 3

    def timed(form, out):
        def run():
            ctx.igemm_dev(form, n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, out.data_ptr(), n)
        for _ in range(10):   # a sub-millisecond product: let the clocks settle first
            run()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            run()
        ctx.synchronize()
        return (time.perf_counter() - t0) / steps

    t_p = timed(vcp_amd.FORM_PLAIN, dP)
    t_z = timed(vcp_amd.FORM_OZAKI, dZ)
    plan = ctx.ozaki_plan(n, n, n)
    out = {"workload": "interval GEMM n=%d, mid U[-1,1], rad=2^-40|mid|" % n, "n": n,
           "plain": {"ms": round(t_p * 1e3, 4), "tflops": round(flops / t_p / 1e12, 3),
                     "frac_fp64_peak": round(flops / t_p / 1e12 / PEAK_FP64_TFLOPS, 4)},
           "ozaki": {"ms": round(t_z * 1e3, 4), "tflops": round(flops / t_z / 1e12, 3), "moduli": plan["moduli"],
                     "radius_scheme": {vcp_amd.OZ_RADIUS_FULL: "full", vcp_amd.OZ_RADIUS_FOLDED: "folded"}[plan["radius"]]}}
    P = dP.cpu().numpy().transpose(1, 0, 2)
    Z = dZ.cpu().numpy().transpose(1, 0, 2)
    try:
        import oracle_ffi as O

        rng = np.random.default_rng(W.SEED0 + 1)
        ii = rng.integers(0, n, 4096)
        jj = rng.integers(0, n, 4096)
        ok_p, _, _ = O.check_containment(IA, IB, P, ii, jj)
        ok_z, _, _ = O.check_containment(IA, IB, Z, ii, jj)
        out["containment_4096"] = {"plain": float(ok_p.mean()), "ozaki": float(ok_z.mean())}
        if with_cpu:
            use_ref = O.ref_lib() is not None
            product = O.ref_pidblas_mulmm if use_ref else O.pidblas_mulmm
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                ref = product(IA, IB)
                ts.append(time.perf_counter() - t0)
            t_c = sorted(ts)[1]
            w_ref = np.maximum(ref[..., 1] - ref[..., 0], np.finfo(float).tiny)
            rz = (Z[..., 1] - Z[..., 0]) / w_ref
            out["cpu_reference"] = {
                "ms": round(t_c * 1e3, 2), "tflops": round(flops / t_c / 1e12, 4),
                "cores": int(O.lib().orc_get_threads()), "kind": "reference" if use_ref else "port",
                "what": "pidblas::mulmm with the reference's vblas rmatmul for its four products" if use_ref else
                        "pidblas::mulmm restated (oracle)",
                "plain_gpu_bitexact": bool(np.array_equal(P, ref))}
            out["radius_ratio_ozaki_over_reference"] = {"max": float(rz.max()), "median": float(np.median(rz))}
            out["gpu_plain_speedup_vs_cpu"] = round(t_c / t_p, 2)
            out["gpu_ozaki_speedup_vs_cpu"] = round(t_c / t_z, 2)
    except Exception as e:
        log("config1 checks skipped: %r" % (e,))
    log("config1 n=%d %s" % (n, out))
    return out


def bench_config2(ctx, n=4096, steps=3):
    """The plain FP64 directed-rounding forms at n=4096: point RU/RD product
    (rmatmul +1/-1 in one pass, 4 n^3 flops) on A, B U[-1,1], and the
    interval product (pidblas::mulmm, 8 n^3) with radii 2^-40 |mid|.  Checks
    C_down <= exact <= C_up on sampled entries."""
    import numpy as np
    import torch

    import vcp_amd
    from vcp_amd import workloads as W

    A, B = W.config2(n)
    rng = np.random.default_rng(W.SEED0 + 21)
    IA, IB = W.uniform_intervals(rng, n, n, -40), W.uniform_intervals(rng, n, n, -40)
    dev = torch.device("cuda:0")
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).to(dev)
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)
    dU = torch.empty((n, n), dtype=torch.float64, device=dev)
    dD = torch.empty((n, n), dtype=torch.float64, device=dev)
    dIA = torch.from_numpy(np.ascontiguousarray(IA.transpose(1, 0, 2))).to(dev)
    dIB = torch.from_numpy(np.ascontiguousarray(IB.transpose(1, 0, 2))).to(dev)
    dC = torch.empty((n, n, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def timed(fn):
        fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t0) / steps

    t_pp = timed(lambda: ctx.gemm_rurd_dev(n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, dU.data_ptr(),
                                           dD.data_ptr(), n))
    t_ii = timed(lambda: ctx.igemm_dev(vcp_amd.FORM_PLAIN, n, n, n, dIA.data_ptr(), n, dIB.data_ptr(), n,
                                       dC.data_ptr(), n))
    C_ii = dC.cpu().numpy().transpose(1, 0, 2)   # the other forms below reuse dC
    out = {"n": n,
           "rurd_point": {"ms": round(t_pp * 1e3, 3), "tflops_4n3": round(4.0 * n ** 3 / t_pp / 1e12, 3),
                          "frac_fp64_peak": round(4.0 * n ** 3 / t_pp / 1e12 / PEAK_FP64_TFLOPS, 4)},
           "interval_plain": {"ms": round(t_ii * 1e3, 3), "tflops_8n3": round(8.0 * n ** 3 / t_ii / 1e12, 3),
                              "frac_fp64_peak": round(8.0 * n ** 3 / t_ii / 1e12 / PEAK_FP64_TFLOPS, 4)}}
    # the other pidblas forms on the same operands; flops count the FMA
    # chains each output carries (2 per chain and k): mul_im_m / mul_m_im 3
    # chains (6 n^3), mulltmm 5 chains on the upper half (5 n^3), vmulmm 2
    # chains on the upper half (2 n^3)
    L = vcp_amd.lib()

    def call(rc, name):
        if rc:
            raise RuntimeError("%s: %d" % (name, rc))

    forms = [
        ("mul_im_m", 6.0, lambda: call(L.vcphip_igemm_ip_dev(ctx._h, n, n, n, dIA.data_ptr(), n, dA.data_ptr(), n,
                                                             dC.data_ptr(), n), "vcphip_igemm_ip_dev")),
        ("mul_m_im", 6.0, lambda: ctx.igemm_pi_dev(n, n, n, dA.data_ptr(), n, dIB.data_ptr(), n, dC.data_ptr(), n)),
        ("mulltmm", 5.0, lambda: call(L.vcphip_igemm_tn_dev(ctx._h, n, n, dIA.data_ptr(), n, dC.data_ptr(), n),
                                      "vcphip_igemm_tn_dev")),
        ("vmulmm", 2.0, lambda: call(L.vcphip_vgemm_tn_dev(ctx._h, n, n, dA.data_ptr(), n, dC.data_ptr(), n),
                                     "vcphip_vgemm_tn_dev")),
    ]
    for name, f, fn in forms:
        t = timed(fn)
        out[name] = {"ms": round(t * 1e3, 3), "chain_flops_per_n3": f,
                     "tflops": round(f * n ** 3 / t / 1e12, 3),
                     "frac_fp64_peak": round(f * n ** 3 / t / 1e12 / PEAK_FP64_TFLOPS, 4)}
    try:
        import oracle_ffi as O

        U = dU.cpu().numpy().T
        D = dD.cpu().numpy().T
        ii = rng.integers(0, n, 4096)
        jj = rng.integers(0, n, 4096)
        ok = 0
        for i, j in zip(ii, jj):
            lo, hi, _ = O.exact_dot(A[i], B[:, j])
            ok += int(D[i, j] <= lo and hi <= U[i, j])
        out["rurd_sampled_enclosure"] = {"samples": 4096, "ok": ok}
        okc, _, _ = O.check_containment(IA, IB, C_ii, ii, jj)
        out["interval_sampled_containment"] = {"samples": 4096, "ok": int(okc.sum())}
    except Exception as e:
        log("config2 check skipped: %r" % (e,))
    log("plain n=4096 %s" % out)
    return out


def bench_dd(ctx, n, steps, with_cpu):
    """Double-double product C = A * B (pddblas::mulmm, vcp/pddblas.hpp:318)
    at n x n x n: hi U[-1,1], lo = hi 2^-53 U[-1,1].  Reports time, the error
    against the exact product on 4096 sampled entries in dd ulps (units of
    2^-105 |C_ij|), and the error of the pddblas restatement beside the
    GPU's on the same entries of its first rows."""
    import numpy as np
    import torch

    import vcp_amd
    from vcp_amd import workloads as W

    log("generating the n=%d dd workload" % n)
    if n == 4096:
        A, B = W.config4(n)
    else:
        rng = np.random.default_rng(W.SEED0 + 4)
        A, B = W.dd_matrix(rng, n, n), W.dd_matrix(rng, n, n)
    dev = torch.device("cuda:0")
    dA = torch.from_numpy(np.ascontiguousarray(A.transpose(1, 0, 2))).to(dev)
    dB = torch.from_numpy(np.ascontiguousarray(B.transpose(1, 0, 2))).to(dev)
    dC = torch.empty((n, n, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def step():
        ctx.ddgemm_dev(n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, dC.data_ptr(), n)

    # the int8 GEMM tile variant for this shape (autotuner; bit-identical choices)
    t0 = time.perf_counter()
    variant, _ = ctx.tune_ddgemm_tile(n, n, n)
    tune_s = time.perf_counter() - t0
    for _ in range(2):
        step()
    ctx.synchronize()
    ctx.set_param(vcp_amd.PARAM_STAGE_TIMING, 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.synchronize()
    t = (time.perf_counter() - t0) / steps
    stages = ctx.stage_times()
    ctx.set_param(vcp_amd.PARAM_STAGE_TIMING, 0)
    nmod = ctx.get_param(vcp_amd.PARAM_DD_NMOD) or None
    out = {"n": n, "ms_per_step": round(t * 1e3, 3), "dd_tflops": round(2.0 * n ** 3 / t / 1e12, 3),
           "stages_ms": {k: round(v, 4) for k, v in stages.items()},
           "moduli": nmod if nmod else "default: fewest giving beta >= 110",
           "tile_variant": variant, "configuration": "autotuned tile variant (%.2f s)" % tune_s}
    log("dd gemm %s" % out)
    try:
        import oracle_ffi as O

        C = dC.cpu().numpy().transpose(1, 0, 2)
        # 4096 sampled entries against the exact product (superaccumulator),
        # in dd ulps: units of 2^-105 |c_ij|
        rng = np.random.default_rng(777)
        ii = rng.integers(0, n, 4096)
        jj = rng.integers(0, n, 4096)
        err_gpu = []
        for i, j in zip(ii, jj):
            h, _, e = O.exact_dd_dot(A[i], B[:, j], C[i, j, 0], C[i, j, 1])
            err_gpu.append(e / (abs(h) * 2.0 ** -105))
        out["err_vs_exact_dd_ulps"] = {"samples": len(err_gpu), "unit": "2^-105 |c_ij|",
                                       "max": float(np.max(err_gpu)), "median": float(np.median(err_gpu))}
        if with_cpu:
            # the pddblas restatement (Ozaki scheme I, dd accumulation) on the
            # first rows, compared with the GPU on the same entries
            rows = 8
            t0 = time.perf_counter()
            Cr, _, _ = O.pddblas_mulmm(np.ascontiguousarray(A[:rows].transpose(1, 0, 2)).transpose(1, 0, 2), B)
            cpu_t = time.perf_counter() - t0
            jc = rng.integers(0, n, 64)
            eg, er = [], []
            for i in range(rows):
                for j in jc:
                    h, _, e = O.exact_dd_dot(A[i], B[:, j], C[i, j, 0], C[i, j, 1])
                    _, _, e2 = O.exact_dd_dot(A[i], B[:, j], Cr[i, j, 0], Cr[i, j, 1])
                    eg.append(e / (abs(h) * 2.0 ** -105))
                    er.append(e2 / (abs(h) * 2.0 ** -105))
            out["same_entries_gpu_vs_pddblas_restated_dd_ulps"] = {
                "samples": len(eg), "gpu": {"max": float(np.max(eg)), "median": float(np.median(eg))},
                "pddblas_restated": {"max": float(np.max(er)), "median": float(np.median(er))}}
            out["cpu_pddblas_restated"] = {"dd_tflops": round(2.0 * rows * n * n / cpu_t / 1e12, 6),
                                           "threads": int(O.lib().orc_get_threads()),
                                           "sample": "first %d rows (%d x %d x %d), %.1f s" % (rows, rows, n, n, cpu_t)}
        log("dd accuracy %s" % {k: v for k, v in out.items() if "err" in k or "cpu" in k})
    except Exception as e:
        log("dd accuracy check skipped: %r" % (e,))
    # pddblas::inv (A X = I by pddblas::linearsolve's refinement) on the same
    # A, every step on the device; max |I - A X| by the dd residual product
    eye = torch.zeros((n, n, 2), dtype=torch.float64, device=dev)
    eye[..., 0] = torch.eye(n, dtype=torch.float64, device=dev)
    dX = torch.empty((n, n, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx.dd_solve_dev(n, n, dA.data_ptr(), n, eye.data_ptr(), n, dX.data_ptr(), n)
    ts = []
    for _ in range(2):
        t0 = time.perf_counter()
        rounds = ctx.dd_solve_dev(n, n, dA.data_ptr(), n, eye.data_ptr(), n, dX.data_ptr(), n)
        ts.append(time.perf_counter() - t0)
    L = vcp_amd.lib()
    rc = L.vcphip_dd_residual_dev(ctx._h, n, n, n, dA.data_ptr(), n, dX.data_ptr(), n, eye.data_ptr(), n,
                                  dC.data_ptr(), n)
    ctx.synchronize()
    res = float(dC[..., 0].abs().max().item()) if rc == 0 else None
    out["dd_inv"] = {"n": n, "ms": round(min(ts) * 1e3, 3), "rounds": int(rounds), "max_abs_I_minus_AX": res}
    log("dd inverse %s" % out["dd_inv"])
    del dA, dB, dX
    out.update(bench_dd_factor(ctx, n, dC, eye))
    del dC, eye
    return out


def _dd_sym(n, seed, dev):
    """symmetric positive definite dd matrix on the device, (columns, rows, 2)
    layout: hi = G G^T / n + I, lo = hi * 2^-54 * S with S symmetric in [-1, 1]"""
    import torch

    g = torch.Generator(device=dev).manual_seed(seed)
    G = torch.randn((n, n), dtype=torch.float64, device=dev, generator=g)
    H = G @ G.T / n
    H = (H + H.T) * 0.5 + torch.eye(n, dtype=torch.float64, device=dev)
    S = torch.rand((n, n), dtype=torch.float64, device=dev, generator=g) * 2.0 - 1.0
    S = (S + S.T) * 0.5
    return torch.stack([H, H * S * 2.0 ** -54], dim=2).contiguous()


def _dd_err(P, Q):
    """max |P - Q| of two dd matrices held as (.., .., 2) tensors"""
    return float(((P[..., 0] - Q[..., 0]) + (P[..., 1] - Q[..., 1])).abs().max().item())


def bench_dd_factor(ctx, n, dC, eye):
    """pddblas::Cholesky (vcp/pddblas.hpp:488-563) at n and pddblas::eigsym
    (Ogita-Aishima, :219-314, 566-606) at n/2, every step on the device.
    Reports the time of one call (best of two after a warm-up) and
    max |U^T U - A| / max |X^T X - I| from the dd product A^T B."""
    import torch

    import vcp_amd

    L = vcp_amd.lib()
    dev = torch.device("cuda:0")
    out = {}
    A = _dd_sym(n, 11, dev)
    U = torch.empty_like(A)
    torch.cuda.synchronize()
    ts = []
    for i in range(3):
        t0 = time.perf_counter()
        ctx.dd_cholesky_dev(n, A.data_ptr(), n, U.data_ptr(), n)
        ctx.synchronize()
        if i:
            ts.append(time.perf_counter() - t0)
    rc = L.vcphip_ddgemm_ta_dev(ctx._h, n, n, n, U.data_ptr(), n, U.data_ptr(), n, dC.data_ptr(), n)
    ctx.synchronize()
    out["dd_cholesky"] = {"n": n, "ms": round(min(ts) * 1e3, 3),
                          "max_abs_UtU_minus_A": _dd_err(dC, A) if rc == 0 else None,
                          "max_abs_A": float(A[..., 0].abs().max().item())}
    log("dd Cholesky %s" % out["dd_cholesky"])
    del A, U
    ne = n // 2
    A = _dd_sym(ne, 12, dev)
    X = torch.empty_like(A)
    lam = torch.empty((ne, 2), dtype=torch.float64, device=dev)
    C = torch.empty_like(A)
    torch.cuda.synchronize()
    ts = []
    for i in range(3):
        t0 = time.perf_counter()
        rc = L.vcphip_dd_eigsym_dev(ctx._h, ne, A.data_ptr(), ne, lam.data_ptr(), X.data_ptr(), ne, 4)
        ctx.synchronize()
        if rc:
            raise RuntimeError("vcphip_dd_eigsym_dev: %d" % rc)
        if i:
            ts.append(time.perf_counter() - t0)
    rc = L.vcphip_ddgemm_ta_dev(ctx._h, ne, ne, ne, X.data_ptr(), ne, X.data_ptr(), ne, C.data_ptr(), ne)
    ctx.synchronize()
    out["dd_eigsym"] = {"n": ne, "ms": round(min(ts) * 1e3, 3), "itep": 4,
                        "max_abs_XtX_minus_I": _dd_err(C, eye[:ne, :ne]) if rc == 0 else None}
    log("dd eigsym %s" % out["dd_eigsym"])
    return out


if __name__ == "__main__":
    main()
