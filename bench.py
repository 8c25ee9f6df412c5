"""Benchmark: T170L60 spherical-harmonic transform pairs/s on one MI355X.

Workload (one "step" = one dynamics-step transform set at T170, I=512, J=256, KM=60):
  4 prognostic fields (vorticity, divergence, temperature, salinity) spectral -> grid,
  each with its longitude and latitude derivatives (xy_w, xy_GradLon_w, xy_GradLat_w):
      12 inverse units per layer;
  6 nonlinear tendency fields grid -> spectral (w_xy): 6 forward units per layer.
  18 units = 9 transform pairs per layer, 540 pairs per step.  A unit is one single-field,
  single-layer transform in one direction; pairs/s = units / 2 / s.
The value and both gradients of the 4 fields come from one fused inverse call
(dsht_inv_vgrad_d: one Legendre pass per column chunk for all three); the same step with
one call per kind (xy_w, xy_GradLon_w, xy_GradLat_w separately) is timed beside it and
reported per form as `separate_calls`.
Inputs are seeded red-spectrum fields (variance ~ (n+1)^-3) with depth-dependent zonal-mean
T/S backgrounds, already resident in HBM when the timed region starts.  The precision form
reported as `value` is the fastest Legendre form whose output stays within relative L2
1e-6 of the FP64 path on every output group (the FP64 form itself if none does).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--prec auto|fp64|bf16x3|...]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dennou-ogcm_amd"))

METRIC = "T170L60 transform pairs/s, mixed vs FP64 at rel err≤1e-6; p99 step latency"
NMAX, IM, JM, KM = 170, 512, 256, 60
RADIUS = 6.371e6
DEPTH = 5000.0
N_PROG, N_TEND = 4, 6
INV_KINDS = ("VAL", "GRADLON", "GRADLAT")
UNITS_PER_STEP = KM * (N_PROG * len(INV_KINDS) + N_TEND)
PAIRS_PER_STEP = UNITS_PER_STEP / 2
TOL_MIXED = 1e-6
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def synth_fields(torch, dev, L, KM=KM):
    """Spectral state (L, 4*KM) in the fast LC layout (seeded red spectrum, grid RMS per field,
    depth-dependent T/S backgrounds)."""
    g = torch.Generator(device=dev)
    g.manual_seed(0x0D0C2026)
    n = torch.floor(torch.sqrt(torch.arange(L, device=dev, dtype=torch.float64)))
    sd = torch.sqrt(0.5 * (n + 1.0) ** -3)
    spec = torch.randn((L, N_PROG * KM), generator=g, device=dev, dtype=torch.float64) * sd[:, None]
    # grid RMS targets: vor 1e-5 1/s, div 1e-6 1/s, T' 1 K, S' 0.05 psu
    rms = torch.tensor([1e-5, 1e-6, 1.0, 0.05], device=dev, dtype=torch.float64)
    m = torch.arange(L, device=dev) - (n * (n + 1)).long()
    wgt = torch.where(m == 0, 1.0, 2.0).to(torch.float64)
    for f in range(N_PROG):
        blk = spec[:, f * KM:(f + 1) * KM]
        blk[0] = 0.0
        cur = torch.sqrt((wgt[:, None] * blk ** 2).sum(0))
        blk *= rms[f] / cur
    z = -(torch.arange(KM, device=dev, dtype=torch.float64) + 0.5) * (DEPTH / KM)
    spec[0, 2 * KM:3 * KM] = 2.0 + 23.0 * torch.exp(z / 1000.0)  # T background, 25 -> 2 C
    spec[0, 3 * KM:4 * KM] = 34.5 + 1.0 * torch.exp(z / 1000.0)  # S background, 35 +/- 0.5
    return spec.contiguous()


def form_parts(name):
    """(Legendre precision form, FP32 grid I/O) of a bench form name such as 'bf16x6+f32grid'."""
    base, _, io = name.partition("+")
    return base, io == "f32grid"


def run_step(plan, dsht, bufs, stream, fused=True):
    """One step.  fused: the three inverse kinds of each field come from one
    dsht_inv_vgrad_d call (one contraction pass for value and both gradients); otherwise
    three dsht_inv_d calls, as the SpmlUtil function-per-kind API makes them.  A fifth
    element True in bufs: FP32 grids (the *_f32_d entry points)."""
    spec, g_out, tend, spec_t = bufs[:4]
    f32 = len(bufs) > 4 and bufs[4]
    if fused:
        (plan.inv_vgrad_f32_d if f32 else plan.inv_vgrad_d)(spec, *g_out, stream=stream)
    else:
        for k, kind in enumerate(INV_KINDS):
            (plan.inv_f32_d if f32 else plan.inv_d)(getattr(dsht, kind), spec, g_out[k], stream=stream)
    (plan.fwd_f32_d if f32 else plan.fwd_d)(dsht.VAL, tend, spec_t, stream=stream)


def rel_l2(a, b):
    d = float((a - b).double().norm())
    n = float(b.double().norm())
    return d / n if n > 0 else d


def time_steps(torch, plan, dsht, bufs, steps, warmup, min_warm_s=0.3, fused=True):
    """W warmup steps (continued until `min_warm_s` of GPU work so clocks have settled),
    then K timed steps back to back; per-step event pairs give the latency distribution."""
    stream = torch.cuda.current_stream().cuda_stream
    t0 = time.perf_counter()
    done = 0
    while done < warmup or time.perf_counter() - t0 < min_warm_s:
        run_step(plan, dsht, bufs, stream, fused)
        done += 1
        if done % 8 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        ev[i][0].record()
        run_step(plan, dsht, bufs, stream, fused)
        ev[i][1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([a.elapsed_time(b) for a, b in ev])
    return wall, lat


def latency_loop(torch, plan, dsht, bufs, nsteps):
    """Per-step latency distribution from `nsteps` back-to-back steps, each bracketed by its own
    event pair on the launch stream (p50/p99 over >= 1000 steps)."""
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(5):
        run_step(plan, dsht, bufs, stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(nsteps)]
    for i in range(nsteps):
        ev[i][0].record()
        run_step(plan, dsht, bufs, stream)
        ev[i][1].record()
    torch.cuda.synchronize()
    lat = np.array([a.elapsed_time(b) for a, b in ev])
    return {"steps": nsteps, "p50_step_ms": float(np.percentile(lat, 50)), "p99_step_ms": float(np.percentile(lat, 99)),
            "max_step_ms": float(lat.max()), "mean_step_ms": float(lat.mean())}


def kernel_profile(torch, plan, dsht, bufs, steps):
    stream = torch.cuda.current_stream().cuda_stream
    dsht.prof_reset()
    dsht.prof_enable(True)
    for _ in range(steps):
        run_step(plan, dsht, bufs, stream)
    torch.cuda.synchronize()
    dsht.prof_enable(False)
    out = {}
    for name in ("leg_inv", "leg_fwd", "fft_inv", "fft_fwd"):
        ms, cnt, by, fl = dsht.prof_query(name)
        if cnt:
            out[name] = {"ms_total": ms, "launches": cnt, "avg_us": 1e3 * ms / cnt,
                         "bytes_per_launch": by / cnt, "flops_per_launch": fl / cnt,
                         "GBps": by / (ms * 1e-3) / 1e9, "TFLOPs": fl / (ms * 1e-3) / 1e12}
    return out


ORACLE_BUILD = ["prebuilt oracle/libdsht_oracle.so"]


def prepare_cpu_oracle():
    """Compile the oracle with -O3 -march=native -fopenmp for this host and load that build
    (before anything else loads the oracle); keeps the prebuilt library if that fails."""
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    try:
        oracle.use_library(oracle.build_native(tempfile.mkdtemp(prefix="dsht_oracle_native_")))
        ORACLE_BUILD[0] = "gcc -O3 -march=native -fopenmp on this host"
    except Exception as e:  # noqa: BLE001 - keep the prebuilt library if the host compiler fails
        print(f"[bench] native oracle build failed ({e}); timing the prebuilt library", file=sys.stderr, flush=True)


def cpu_baseline(spec, tend, reps=21, warm=2):
    """The FP64 CPU oracle on the benchmarked step itself: the 4 prognostic fields x 60 layers
    to grid with both derivatives (VAL, GRADLON, GRADLAT of 240 columns) and the 6 x 60
    tendency grids back to spectral, on the same inputs as the GPU step.  The oracle is
    compiled here with -O3 -march=native -fopenmp (OpenMP over columns) for the host that
    runs it; the value is the median of `reps` repetitions after `warm` untimed ones.  A
    one-thread figure comes from the same step restricted to 4 layers."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    build = ORACLE_BUILD[0]
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    o = oracle.Oracle(NMAX, IM, JM, RADIUS)
    w = spec.T.cpu().numpy().copy()      # (4*KM, L)
    xy = tend.cpu().numpy()              # (6*KM, J, I)

    def step(wc, xc):
        for k in (oracle.VAL, oracle.GRADLON, oracle.GRADLAT):
            o.inv(wc, k)
        o.fwd(xc)

    for _ in range(warm):
        step(w, xy)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        step(w, xy)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    # one thread: the same step over 4 of the 60 layers (16 inverse and 24 forward columns)
    oracle.set_threads(1)
    lay = [f * KM + k for f in range(N_PROG) for k in range(4)]
    layt = [f * KM + k for f in range(N_TEND) for k in range(4)]
    t1 = []
    for _ in range(3):
        t0 = time.perf_counter()
        step(w[lay], xy[layt])
        t1.append(time.perf_counter() - t0)
    oracle.set_threads(threads)
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": PAIRS_PER_STEP / med, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"the benchmarked T170L60 step ({int(PAIRS_PER_STEP)} pairs: 240 columns x VAL/GradLon/"
                      f"GradLat inverse + 360 columns forward) on the GPU step's inputs, median of {reps} "
                      f"repetitions ({med:.3f} s each); FP64 oracle (oracle/dsht_oracle.c, {build}), "
                      f"OpenMP over columns",
            "ms_per_step": 1e3 * med, "one_thread_value": (PAIRS_PER_STEP * 4 / KM) / float(np.median(t1)),
            "cpu_model": model, "nproc": os.cpu_count()}


def oracle_check(torch, spec, g_out, NCHK=3):
    """FP64 GPU outputs against the CPU oracle on a few columns (parity at full size)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    o = oracle.Oracle(NMAX, IM, JM, RADIUS)
    cols = [0, KM + 7, 3 * KM - 1][:NCHK]
    w = spec[:, cols].T.cpu().numpy().copy()
    ref = o.inv(w)
    got = g_out[0][cols].cpu().numpy()
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


FIELDS = ("vorticity", "divergence", "temperature", "salinity")


def per_field_errors(outs, ref):
    """Relative L2 and relative max error per prognostic field over its three grid outputs
    (value, lon and lat derivative), against the FP64 form."""
    out = {}
    for f, name in enumerate(FIELDS):
        sl = slice(f * KM, (f + 1) * KM)
        d2 = n2 = 0.0
        dmax = rmax = 0.0
        for a, b in zip(outs[:3], ref[:3]):
            d = (a[sl] - b[sl]).double()
            d2 += float((d * d).sum())
            n2 += float((b[sl].double() ** 2).sum())
            dmax = max(dmax, float(d.abs().max()))
            rmax = max(rmax, float(b[sl].abs().max()))
        out[name] = {"rel_l2": (d2 / n2) ** 0.5 if n2 > 0 else d2 ** 0.5, "rel_max": dmax / rmax if rmax else dmax}
    return out


def energy_drift(torch, plan, dsht, form, spec, trips=100):
    """Relative change of spectral energy sum_l w_l^2 over `trips` inverse+forward round
    trips of the prognostic state in `form`."""
    stream = torch.cuda.current_stream().cuda_stream
    plan.set_precision({v: k for k, v in dsht.PREC_NAMES.items()}[form])
    w = spec.clone()
    w[0] = 0.0  # drop the depth backgrounds: drift of the perturbation energy
    g = torch.empty((w.shape[1], JM, IM), dtype=torch.float64, device=spec.device)
    e0 = float((w * w).sum())
    for _ in range(trips):
        plan.inv_d(dsht.VAL, w, g, stream=stream)
        plan.fwd_d(dsht.VAL, g, w, stream=stream)
    torch.cuda.synchronize()
    plan.set_precision(dsht.PREC_FP64)
    return (float((w * w).sum()) - e0) / e0


def traffic_table(path=os.path.join(ROOT, "profiles", "hbm_traffic_T170L60.md")):
    """Measured HBM bytes per launch of the step's calls (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-
    calibrated) from the committed profile summary: {(form, scope): bytes}."""
    out = {}
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 4 and f[0] == "traffic":
                out[(f[1], f[2])] = float(f[3])
    except OSError:
        pass
    return out


def roofline_probe(torch, plan, dsht, form, scope, bufs, reps=10):
    """Average duration and algorithmic bytes/flops of `scope` on the launches of the
    step's own call (the launch shape the committed counter profile measured)."""
    spec, g_out, tend, spec_t = bufs[:4]
    f32 = len(bufs) > 4 and bufs[4]
    stream = torch.cuda.current_stream().cuda_stream
    pid = {v: k for k, v in dsht.PREC_NAMES.items()}[form_parts(form)[0]]
    plan.set_precision(pid)
    inv = scope.endswith("inv")
    nc = spec.shape[1] if inv else tend.shape[0]

    def call():
        if inv:
            (plan.inv_vgrad_f32_d if f32 else plan.inv_vgrad_d)(spec, *g_out, stream=stream)
        else:
            (plan.fwd_f32_d if f32 else plan.fwd_d)(dsht.VAL, tend, spec_t, stream=stream)

    for _ in range(3):
        call()
    torch.cuda.synchronize()
    dsht.prof_reset()
    dsht.prof_enable(True)
    for _ in range(reps):
        call()
    torch.cuda.synchronize()
    dsht.prof_enable(False)
    ms, cnt, by, fl = dsht.prof_query(scope)
    plan.set_precision(dsht.PREC_FP64)
    t = ms * 1e-3 / cnt
    gbs, tfs = by / cnt / t / 1e9, fl / cnt / t / 1e12
    mpeak = FP64_PEAK_TFS if form == "fp64" else 2500.0
    if scope.startswith("leg") and tfs / mpeak > gbs / HBM_PEAK_GBS:
        roof = {"bound": "mfma", "achieved": tfs, "peak": mpeak, "unit": "TFLOP/s", "frac": tfs / mpeak}
    else:
        roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}
    roof["traffic"] = traffic_table().get((form, scope))
    # FFT launches per call of the fused inverse: three in FP64 (one per output grid), two in
    # the split forms (the field with its longitude derivative from one read, the latitude
    # derivative); launch-average bytes and times cover both kinds
    per_call = (3 if form == "fp64" else 2) if scope == "fft_inv" else 1
    roof.update({"kernel": scope, "columns_per_launch": nc * reps * per_call / cnt, "avg_launch_us": 1e6 * t,
                 "algorithmic_bytes_per_launch": by / cnt, "algorithmic_flops_per_launch": fl / cnt,
                 "traffic_source": "profiles/hbm_traffic_T170L60.md" if roof["traffic"] else None})
    return roof


def t85_line(torch, dsht, steps, warmup, N=85, I=256, J=128, km=60):
    """The same dynamics-step transform set at T85L60 (I=256, J=128) in FP64: throughput, the
    FP64 Legendre contractions' achieved rate against the 78.6 TFLOP/s MFMA peak (useful
    flops: 4 J T per pair in the even/odd form) and the FFTs' achieved HBM rate."""
    dev = torch.device("cuda:0")
    plan = dsht.Plan(N, I, J, kmax=km, radius=RADIUS)
    spec = synth_fields(torch, dev, plan.lm, km)
    g_out = [torch.empty((N_PROG * km, J, I), dtype=torch.float64, device=dev) for _ in INV_KINDS]
    stream = torch.cuda.current_stream().cuda_stream
    plan.inv_d(dsht.VAL, spec, g_out[0], stream=stream)
    v = g_out[0].view(N_PROG, km, J, I)
    tend = torch.cat([v[0] * v[2], v[1] * v[3], v[0] * v[3], v[1] * v[2], v[2] * v[2] * 1e-3, v[3] * v[3]],
                     0).contiguous()
    spec_t = torch.empty((plan.lm, N_TEND * km), dtype=torch.float64, device=dev)
    bufs = (spec, g_out, tend, spec_t)
    pairs = km * (N_PROG * len(INV_KINDS) + N_TEND) / 2
    wall, _ = time_steps(torch, plan, dsht, bufs, steps, warmup)
    lat = latency_loop(torch, plan, dsht, bufs, 1000)
    prof = kernel_profile(torch, plan, dsht, bufs, 5)
    leg = {k: {"avg_us": v["avg_us"], "TFLOPs": v["TFLOPs"], "mfma_fp64_frac": v["TFLOPs"] / FP64_PEAK_TFS,
               "GBps": v["GBps"]} for k, v in prof.items() if k.startswith("leg")}
    fft = {k: {"avg_us": v["avg_us"], "GBps": v["GBps"], "frac_hbm_peak": v["GBps"] / HBM_PEAK_GBS}
           for k, v in prof.items() if k.startswith("fft")}
    return {"workload": f"T{N}L{km} (I={I}, J={J}) dynamics-step transform set, FP64", "pairs_per_s": pairs * steps / wall,
            "ms_per_step": 1e3 * wall / steps, "p50_step_ms": lat["p50_step_ms"], "p99_step_ms": lat["p99_step_ms"],
            "legendre": leg, "fft": fft,
            "mfma_note": "mfma_fp64_frac = useful FP64 flops / time / 78.6 TFLOP/s; the counter-measured MfmaUtil "
                         "is in profiles/"}


# ---------------------------------------------------------------- batched form
BATCH_TRUNC = (21, 42, 85, 170)
BATCH_KM = (20, 30, 60)
# IM = 3(N+1) rounded up to a length with a register FFT (M = IM/2 = R1*R2), JM = IM/2 made even
BATCH_GRIDS = {21: (72, 36), 42: (144, 72), 85: (270, 136), 170: (540, 270)}


def batched_bench(torch, dsht, n_inst=64, steps=10, seq_steps=2, seed=0xB47C4, prec="bf16x3"):
    """64 independent instances of uneven truncation and layer count advanced by one
    hipGraph-captured batched step (tile/chunk knobs from the autotuner; truncation groups are
    parallel branches of the graph), against the same step enqueued eagerly (groups on
    concurrent streams) and against the instances run one after another, two calls each
    (one inverse over its 4 KM prognostic columns, one forward over its 6 KM tendency
    columns).  Error of the split form against the FP64 batched step (itself checked against
    the oracle in the test suite)."""
    from dsht import autotune

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    inst = [(int(rng.choice(BATCH_TRUNC)), int(rng.choice(BATCH_KM))) for _ in range(n_inst)]
    plans = {N: dsht.Plan(N, *BATCH_GRIDS[N], radius=RADIUS) for N in sorted({n for n, _ in inst})}
    b = dsht.Batch()
    ids = [b.add(plans[N], km) for N, km in inst]
    b.finalize()
    groups = {}
    for iid, (N, km) in zip(ids, inst):
        g, _, gc, nbytes = b.info(iid)
        groups[g] = (N, gc)
    views = {}
    for g, (N, gc) in groups.items():
        p = plans[N]
        spec, grid, tend, spect = b.group_views(g, p, gc)
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed + g)
        n = torch.floor(torch.sqrt(torch.arange(p.lm, device=dev, dtype=torch.float64)))
        spec.copy_(torch.randn(tuple(spec.shape), generator=gen, device=dev, dtype=torch.float64)
                   * torch.sqrt(0.5 * (n + 1.0) ** -3)[:, None])
        tend.copy_(torch.randn(tuple(tend.shape), generator=gen, device=dev, dtype=torch.float64))
        views[g] = (p, gc, spec, grid, tend, spect)
    units = sum(km for _, km in inst) * (N_PROG * len(INV_KINDS) + N_TEND)
    s = torch.cuda.Stream()
    ss = s.cuda_stream
    # FP64 reference step
    for p in plans.values():
        p.set_precision(dsht.PREC_FP64)
    b.step(ss)
    s.synchronize()
    ref = {g: (v[3].clone(), v[5].clone()) for g, v in views.items()}
    pid = {v: k for k, v in dsht.PREC_NAMES.items()}[prec]
    tuning = {}
    with torch.cuda.stream(s):
        for g, (p, gc, *_r) in views.items():
            p.set_precision(pid)
            t = autotune.tune(p, N_PROG * gc, use_cache=False)
            tuning[f"T{p.nmax}"] = {"chunk": t["chunk"], "cb": t["cb"], "leg": t["leg"], "ms": t["ms"],
                                    "candidates": len(t["candidates"]), "columns": N_PROG * gc}
    s.synchronize()
    b.step(ss)
    s.synchronize()
    err = {}
    for g, (p, gc, spec, grid, tend, spect) in views.items():
        err[f"T{p.nmax}"] = {"grid": rel_l2(grid, ref[g][0]), "spectral": rel_l2(spect, ref[g][1])}
    b.capture(ss)

    def timed(fn, k):
        for _ in range(2):
            fn()
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        s.synchronize()
        return (time.perf_counter() - t0) / k

    t_graph = timed(lambda: b.launch(ss), steps)
    t_eager = timed(lambda: b.step(ss), steps)
    t_seq = timed(lambda: b.step_sequential(ss), seq_steps)
    b.close()
    return {
        "workload": f"{n_inst} instances, truncation from {list(BATCH_TRUNC)}, KM from {list(BATCH_KM)}, "
                    "grids " + ", ".join(f"T{n}: {i}x{j}" for n, (i, j) in BATCH_GRIDS.items()) +
                    "; per instance and step the T170L60 transform set (12 inverse + 6 forward units/layer)",
        "precision": prec,
        "instances": [{"nmax": n, "km": km} for n, km in inst],
        "units_per_step": units,
        "graph_pairs_per_s": units / 2 / t_graph,
        "graph_ms_per_step": 1e3 * t_graph,
        "eager_batched_ms_per_step": 1e3 * t_eager,
        "sequential_ms_per_step": 1e3 * t_seq,
        "sequential_pairs_per_s": units / 2 / t_seq,
        "speedup_vs_sequential": t_seq / t_graph,
        "graph_vs_eager": t_eager / t_graph,
        "rel_l2_vs_fp64": err,
        "autotuned": tuning,
    }


# ---------------------------------------------------------------- serving form
def column_bench(torch, dsht, reps=21, nsets=3):
    """Vertical column diagnostics (HBEDiagnose_OMG, _HydPres, _UVBarot and the QUICK vertical
    advective flux of calc_ADVFlxZ_QUICK) on T170L60 device
    grids (60 layers of 512 x 256 points).  Calls rotate over `nsets` buffer sets (each
    call touches about 190 MB, so a set is out of the 256 MB last-level cache again when it
    comes back round).  Each kernel's average duration comes from events on the stream it
    is launched on; its algorithmic bytes are every input layer read once and every output
    level written once (8 bytes a value)."""
    plan = dsht.Plan(NMAX, IM, JM, kmax=KM, radius=6.371e6)
    g = torch.Generator(device="cuda").manual_seed(0xC01)
    shape = (KM, JM, IM)
    cdk = torch.full((KM,), 1.0 / KM, dtype=torch.float64, device="cuda")
    rfdk = torch.full((KM + 1,), float(KM), dtype=torch.float64, device="cuda")

    def rnd(scale=1.0, shift=0.0):
        return shift + scale * torch.randn(shape, dtype=torch.float64, device="cuda", generator=g)

    sets = []
    for _ in range(nsets):
        sets.append({"div": rnd(1e-6), "h": rnd(0.05, 1.0), "dens": rnd(), "u": rnd(), "v": rnd(),
                     "omg": torch.empty((KM + 1, JM, IM), dtype=torch.float64, device="cuda"),
                     "hp": torch.empty(shape, dtype=torch.float64, device="cuda"),
                     "flx": torch.empty((KM + 1, JM, IM), dtype=torch.float64, device="cuda"),
                     "kv": rnd(1e-4, 1e-3), "e3": rnd(200.0, 4000.0),
                     "qa": torch.empty(shape, dtype=torch.float64, device="cuda"),
                     "ub": torch.empty((JM, IM), dtype=torch.float64, device="cuda"),
                     "vb": torch.empty((JM, IM), dtype=torch.float64, device="cuda")})
    npts = IM * JM
    fdk = torch.full((KM + 1,), 1.0 / KM, dtype=torch.float64, device="cuda")
    fdk[0] = fdk[KM] = 0.5 / KM
    work = torch.empty(3 * KM * npts, dtype=torch.float64, device="cuda")
    calls = {
        "omega": (lambda b: plan.col_omega_d(b["div"], b["h"], cdk, out=b["omg"]), (2 * KM + KM + 1) * npts * 8),
        "hydpres": (lambda b: plan.col_hydpres_d(b["dens"], b["h"], cdk, 9.8, out=b["hp"]), 3 * KM * npts * 8),
        "uvbarot": (lambda b: plan.col_uvbarot_d(b["u"], b["v"], b["h"], cdk, b["ub"], b["vb"]),
                    (3 * KM + 2) * npts * 8),
        # the flux reads the tracer layers and omega at half levels 1..km-1 (no free surface)
        "quick": (lambda b: plan.col_quick_d(b["omg"], b["dens"], cdk, rfdk, out=b["flx"]),
                  (KM + (KM - 1) + (KM + 1)) * npts * 8),
        # minimum: q0, rhs, kv, e3 read and qa written once (the elimination's rows go
        # through `work` and qa in addition)
        "vdiff": (lambda b: plan.col_vdiff_d(b["u"], b["v"], b["kv"], b["e3"], cdk, fdk, 3600.0, 1.0, "D", "N",
                                             out=b["qa"], work=work), 5 * KM * npts * 8),
    }
    out = {"workload": f"T{NMAX}L{KM} grids ({IM} x {JM} points, {KM} layers), FP64; calls rotate over "
                       f"{nsets} buffer sets"}
    for name, (fn, nbytes) in calls.items():
        for k in range(nsets):
            fn(sets[k])
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in range(reps):
            fn(sets[k % nsets])
        e1.record()
        e1.synchronize()
        us = 1e3 * e0.elapsed_time(e1) / reps
        gbps = nbytes / (us * 1e-6) / 1e9
        out[name] = {"avg_us": us, "algorithmic_bytes": nbytes, "GBps": gbps, "frac_hbm_peak": gbps / 8000.0}
    return out


def lfam3_bench(torch, dsht, steps=20, N=NMAX, km=KM):
    """The device LF-AM3 dynamics step at T170L60 (I=512, J=256): one run, and four runs of
    the same truncation in one pool (each transform over all their layers in one call).
    One step = (42 KM + 9) forward + (28 KM + 9) inverse units per run."""
    from dsht import serve

    out = {"workload": f"LF-AM3 dynamics step (rigid lid, linear EOS), T{N}L{km} (I={IM}, J={JM}), FP64, "
                       f"dt {serve.DT[N]:g} s"}
    plan = dsht.Plan(N, IM, JM, kmax=km, radius=RADIUS)
    for nr in (1, 4):
        pool = serve.Pool(plan, km, nr, serve.DT[N], torch.device("cuda:0"))
        for r in range(nr):
            with torch.cuda.stream(pool.stream):
                serve.initial_state(pool, r, 100 + r)
            pool.lf.set_rotation(r, serve.EARTH_OMEGA * (0.5 + r))
            pool.runs.append(None)
        g = pool._graph(nr)
        with torch.cuda.stream(pool.stream):
            for _ in range(3):
                g.replay()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(pool.stream)
            for _ in range(steps):
                g.replay()
            e1.record(pool.stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / steps
        f, i = pool.lf.units(nr)
        ok = bool(torch.isfinite(pool.lf.state(dsht.LF_U, dsht.LF_N)[:nr * km]).all())
        out[f"runs_{nr}"] = {"ms_per_step": ms, "pairs_per_s": (f + i) / 2 / (ms * 1e-3), "pairs_per_step": (f + i) / 2,
                             "finite": ok}
        pool = None
    return out


def _redigm_fields(torch, dsht, plan, km, seed):
    """A stratified ocean on the plan's grid: potential temperature [K], salinity, thickness
    factor, layer depths, bottom depth (3-5 km), layer thicknesses and reciprocal half-level
    spacings of the Chebyshev-Lobatto sigma levels."""
    import numpy as np
    fk, _ = dsht.sig_grid(km, km)
    cdk = fk[:-1] - fk[1:]
    ck = np.empty(km + 2)
    ck[1:-1] = 0.5 * (fk[:-1] + fk[1:])
    ck[0], ck[-1] = fk[0] + 0.5 * cdk[0], fk[-1] - 0.5 * cdk[-1]
    rfdk = 1.0 / (ck[:-1] - ck[1:])
    J, I = plan.jm, plan.im
    lon = plan.x_Lon[None, :]
    mu = np.sin(plan.y_Lat)[:, None]
    topo = 3000.0 + 2000.0 * (0.5 + 0.5 * np.cos(2 * lon) * (1 - mu ** 2))
    z = topo[None] * ck[1:-1, None, None]
    rng = np.random.default_rng(seed)
    pt = 275.0 + 20.0 * np.exp(z / 800.0) + 6.0 * (mu ** 2) * np.cos(lon)[None] * np.exp(z / 1500.0) \
        + 0.02 * rng.standard_normal(z.shape)
    sa = 34.5 + 0.5 * np.exp(z / 1000.0) + 0.3 * mu * np.sin(3 * lon)[None] * np.exp(z / 2000.0)
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device="cuda")
    trc = t(np.concatenate([pt, sa]))  # both tracers in one array, as the model's xyza_TRC
    return {"pt": trc[:km], "sa": trc[km:], "h": t(np.broadcast_to(topo, z.shape)), "z": t(z), "topo": t(topo),
            "cdk": t(cdk), "rfdk": t(rfdk)}, (pt, sa, np.broadcast_to(topo, z.shape).copy(), z, topo, cdk, rfdk)


def redigm_bench(torch, dsht, reps=10, km=KM, cpu=True):
    """Redi/GM isopycnal mixing of both tracers (LPhys_RediGM_hspm_vfvm_AddMixingTerm) on
    T170L60 device grids, FP64 and BF16X6 transforms.  One call: the forward transform of the
    2 km tracer layers, their GradLon and GradLat (one fused inverse), the column kernel
    (slopes, tapering, Redi and skew fluxes), the AlphaOptr forward of the fluxes and the
    inverse of the divergence, and the accumulation into both right-hand sides.  Transform
    units per call: 2 km + 4 km forward (AlphaOptr takes two grids), 4 km + 2 km inverse.
    The column kernel's roofline: its algorithmic bytes (8 input and 6 output layer fields
    once, plus the bottom depth) over its average duration (library timers, HIP events on
    its stream).  The CPU leg times the oracle's restatement at T42L20 on this host."""
    out = {"workload": f"Redi/GM tendency of ptemp and salt, T{NMAX}L{km} (I={IM}, J={JM}), "
                       "DM95 x LDD95 tapering, kappa 800 m2/s",
           "units_per_call": {"forward": 6 * km, "inverse": 6 * km}}
    plan = dsht.Plan(NMAX, IM, JM, kmax=km, radius=RADIUS)
    d, _ = _redigm_fields(torch, dsht, plan, km, seed=0x4ED1)
    rt = torch.zeros((km, JM, IM), dtype=torch.float64, device="cuda")
    rs = torch.zeros_like(rt)
    work = torch.empty(dsht.load().dsht_redigm_work_size(plan.h, km), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    args = (d["pt"], d["sa"], d["h"], d["z"], d["topo"], d["cdk"], d["rfdk"], rt, rs)
    for form, pid in (("fp64", dsht.PREC_FP64), ("bf16x6", dsht.PREC_BF16X6)):
        plan.set_precision(pid)
        for _ in range(2):
            plan.redigm_d(*args, work=work, stream=stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            plan.redigm_d(*args, work=work, stream=stream)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        dsht.prof_reset()
        dsht.prof_enable(True)
        for _ in range(reps):
            plan.redigm_d(*args, work=work, stream=stream)
        torch.cuda.synchronize()
        dsht.prof_enable(False)
        kern = {}
        for name in ("redigm_col", "redigm_rhs", "leg_inv", "leg_fwd", "fft_inv", "fft_fwd"):
            kms, cnt, by, _ = dsht.prof_query(name)
            if cnt:
                us = 1e3 * kms / cnt
                kern[name] = {"avg_us": us, "launches_per_call": cnt / reps, "bytes_per_launch": by / cnt,
                              "GBps": by / cnt / (us * 1e-6) / 1e9}
        col = kern.get("redigm_col", {})
        out[form] = {"ms_per_call": ms, "pairs_per_s": 6 * km / (ms * 1e-3), "kernels": kern,
                     "finite": bool(torch.isfinite(rt).all() and torch.isfinite(rs).all()),
                     "roofline_col": {"bound": "hbm", "achieved": col.get("GBps"), "peak": HBM_PEAK_GBS,
                                      "unit": "GB/s", "frac": (col.get("GBps") or 0) / HBM_PEAK_GBS}}
    plan.set_precision(dsht.PREC_FP64)
    if cpu:
        out["cpu_baseline"] = redigm_cpu(torch, dsht)
    return out


def redigm_cpu(torch, dsht, N=42, I=128, J=64, km=20, reps=3):
    """The oracle's Redi/GM restatement (numpy column part, one thread, around the C oracle's
    transforms, OpenMP over columns) timed at T42L20 on this host, and the device call at the
    same size."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import redigm as RG
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    plan = dsht.Plan(N, I, J, kmax=km, radius=RADIUS)
    d, (pt, sa, h, z, topo, cdk, rfdk) = _redigm_fields(torch, dsht, plan, km, seed=0x4ED2)
    o = O.Oracle(N, I, J, RADIUS)
    cfg = RG.Config()
    RG.add_mixing_term(o, pt, sa, h, z, topo, cdk, rfdk, cfg, lat=plan.y_Lat)
    t0 = time.perf_counter()
    for _ in range(reps):
        ref = RG.add_mixing_term(o, pt, sa, h, z, topo, cdk, rfdk, cfg, lat=plan.y_Lat)
    cpu_s = (time.perf_counter() - t0) / reps
    rt = torch.zeros((km, J, I), dtype=torch.float64, device="cuda")
    rs = torch.zeros_like(rt)
    args = (d["pt"], d["sa"], d["h"], d["z"], d["topo"], d["cdk"], d["rfdk"], rt, rs)
    plan.redigm_d(*args)
    torch.cuda.synchronize()
    err = max(float(np.linalg.norm(g.cpu().numpy() - r) / np.linalg.norm(r)) for g, r in zip((rt, rs), ref))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        rt.zero_()
        rs.zero_()
        plan.redigm_d(*args)
    e1.record()
    e1.synchronize()
    return {"sample": f"T{N}L{km} (I={I}, J={J}) Redi/GM tendency, {reps} calls", "cpu_ms_per_call": 1e3 * cpu_s,
            "gpu_ms_per_call": e0.elapsed_time(e1) / 20, "cores": threads, "kind": "port",
            "rel_l2_gpu_vs_oracle": err}


def serving_bench(torch, dsht, n_req=32, rate=20.0, nsteps=200, seed=0x5E4E, factor=1.8):
    """Poisson arrivals of independent LF-AM3 runs (truncation, KM, rotation rate drawn per
    run), served four ways:
      batched            continuous batching, one pool step on the GPU at a time (pools take
                         turns), admission capped so a pool's step stays within `factor` x
                         its one-run step time;
      batched_concurrent the same cap, every pool stepping on its own stream as soon as its
                         previous step is done (pools share the GPU);
      batched_uncapped   one pool at a time without the latency cap;
      one_at_a_time      the runs in arrival order, each alone.
    Per-run p99 step latency is compared with the median step time of that (truncation, KM)
    run alone."""
    from dsht import serve

    def stream():
        return serve.poisson_stream(n_req, rate, seed=seed, nsteps=nsteps)

    kinds = sorted({(r.nmax, r.km) for r in stream()})
    out = {}
    for mode in ("batched", "batched_concurrent", "batched_uncapped", "one_at_a_time"):
        srv = serve.Server(latency_factor=factor if mode in ("batched", "batched_concurrent") else None,
                           serialize=mode != "batched_concurrent")
        # every mode builds (and, capped, calibrates) its pools and captures their step graphs
        # before the clock starts
        for n, km in kinds:
            srv.pool(n, km)
        torch.cuda.synchronize()
        reqs = stream()
        print(f"[bench] serving: {mode}", file=sys.stderr, flush=True)
        out[mode] = srv.run(reqs, batched=(mode != "one_at_a_time"))
        out[mode]["runs"] = reqs
        if mode == "batched":
            # throughput bound of batching on this mix: every run's one-run step time against
            # its marginal cost in a full pool (t(n) = a + b n, calibrated per pool)
            t1 = sum(srv.pools[(r.nmax, r.km)].t_ref for r in reqs)
            tb = sum(srv.pools[(r.nmax, r.km)].t_per_run for r in reqs)
            marginal = {f"T{n}L{km}": {"one_run_ms": pl.t_ref, "per_added_run_ms": pl.t_per_run}
                        for (n, km), pl in srv.pools.items()}
        srv = None
        torch.cuda.synchronize()
    # each (truncation, KM) alone: the latency reference
    solo = {}
    srv = serve.Server()
    for n, km in kinds:
        r = serve.Request(0, n, km, 1.0, 40, 0.0, 1)
        srv.run([r], batched=True)
        solo[(n, km)] = float(np.median(r.step_ms[5:]))
    srv = None
    res = {}
    for mode, v in out.items():
        reqs = v.pop("runs")
        ratio = [float(np.percentile(r.step_ms, 99)) / solo[(r.nmax, r.km)] for r in reqs]
        by_t = {}
        for r, q in zip(reqs, ratio):
            by_t.setdefault(f"T{r.nmax}", []).append(q)
        v["p99_over_solo_max"] = max(ratio)
        v["p99_over_solo_median"] = float(np.median(ratio))
        v["p99_over_solo_by_truncation"] = {k: max(q) for k, q in sorted(by_t.items())}
        res[mode] = v
    b, u = res["batched"], res["one_at_a_time"]
    return {
        "workload": f"{n_req} LF-AM3 runs, Poisson arrivals at {rate:g}/s, {nsteps} steps each, truncation from "
                    f"{list(serve.GRIDS)}, KM from [20, 30, 60], rotation 0.25-4x Earth; FP64; one step = "
                    f"(42 KM + 9) forward + (28 KM + 9) inverse units per run",
        "batched": b,
        "batched_concurrent": res["batched_concurrent"],
        "batched_uncapped": res["batched_uncapped"],
        "one_at_a_time": u,
        "throughput_ratio": b["pairs_per_s"] / u["pairs_per_s"],
        "throughput_ratio_concurrent": res["batched_concurrent"]["pairs_per_s"] / u["pairs_per_s"],
        "throughput_ratio_uncapped": res["batched_uncapped"]["pairs_per_s"] / u["pairs_per_s"],
        "throughput_ratio_bound": t1 / tb,
        "pool_step_model": marginal,
        "solo_step_ms": {f"T{n}L{km}": t for (n, km), t in solo.items()},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--prec", default="auto")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy 3 times (HBM counter calibration) and exit")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--latency-steps", type=int, default=1000,
                    help="steps of the separate per-step latency loop (p50/p99) of each form")
    ap.add_argument("--no-t85", action="store_true")
    ap.add_argument("--no-lfam3", action="store_true")
    ap.add_argument("--lfam3-only", action="store_true",
                    help="only run the LF-AM3 step section (T170L60, 1 and 4 runs) and print its JSON")
    ap.add_argument("--column-only", action="store_true",
                    help="only run the vertical column diagnostics section and print its JSON")
    ap.add_argument("--batch-only", action="store_true",
                    help="only run the batched-instances section and print its JSON")
    ap.add_argument("--serve-only", action="store_true",
                    help="only run the serving section and print its JSON")
    ap.add_argument("--redigm-only", action="store_true",
                    help="only run the Redi/GM mixing section and print its JSON")
    ap.add_argument("--no-redigm", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU only: the transform path has no "
                                   "multi-GPU form here; layers, fields and instances are independent columns, so "
                                   "it would shard by column blocks, one process per GPU", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        # three kernels of known traffic for the HBM counters: a 256 MiB device copy (reads and
        # writes 256 MiB), a sum over 256 MiB (reads only) and a 256 MiB fill (writes only)
        a = torch.ones(32 << 20, dtype=torch.float64, device=dev)  # 256 MiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        s = torch.empty((), dtype=torch.float64, device=dev)
        for _ in range(3):
            torch.sum(a, dim=0, out=s)
        torch.cuda.synchronize()
        for _ in range(3):
            b.fill_(2.0)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8,
                          "sum_bytes_read": a.numel() * 8, "fill_bytes_written": b.numel() * 8, "launches": 3,
                          "sum": float(s)}))
        return
    if args.column_only:
        import dsht
        print(json.dumps({"column_diagnostics": column_bench(torch, dsht)}))
        return
    if args.lfam3_only:
        import dsht
        print(json.dumps({"lfam3": lfam3_bench(torch, dsht, steps=args.steps)}))
        return
    if args.batch_only:
        import dsht
        print(json.dumps({"batched": batched_bench(torch, dsht)}))
        return
    if args.serve_only:
        import dsht
        print(json.dumps({"serving": serving_bench(torch, dsht)}))
        return
    if args.redigm_only:
        import dsht
        print(json.dumps({"redigm": redigm_bench(torch, dsht, cpu=not args.no_cpu)}))
        return

    import dsht

    if not args.no_cpu:
        prepare_cpu_oracle()
    print(f"[bench] T{NMAX} I={IM} J={JM} KM={KM}: building plan", file=sys.stderr, flush=True)
    plan = dsht.Plan(NMAX, IM, JM, kmax=KM, radius=RADIUS)
    L = plan.lm
    spec = synth_fields(torch, dev, L)
    g_out = [torch.empty((N_PROG * KM, JM, IM), dtype=torch.float64, device=dev) for _ in INV_KINDS]
    # tendency fields: nonlinear products of the prognostic grids (computed once, untimed)
    stream = torch.cuda.current_stream().cuda_stream
    plan.inv_d(dsht.VAL, spec, g_out[0], stream=stream)
    v = g_out[0].view(N_PROG, KM, JM, IM)
    tend = torch.cat([v[0] * v[2], v[1] * v[3], v[0] * v[3], v[1] * v[2], v[2] * v[2] * 1e-3, v[3] * v[3]], 0)
    tend = tend.contiguous()
    spec_t = torch.empty((L, N_TEND * KM), dtype=torch.float64, device=dev)
    bufs = (spec, g_out, tend, spec_t)
    torch.cuda.synchronize()

    forms = ["fp64"] if args.prec == "fp64" else (
        ["fp64", args.prec] if args.prec != "auto" else
        ["fp64", "bf16x3", "fp16x3", "bf16x6", "fp16", "bf16x6+f32grid", "fp16x3+f32grid"])
    # FP32 grid I/O: the same prognostic spectra, FP32 output grids, FP32 copies of the tendencies
    g_out32 = [torch.empty((N_PROG * KM, JM, IM), dtype=torch.float32, device=dev) for _ in INV_KINDS]
    bufs32 = (spec, g_out32, tend.float().contiguous(), spec_t, True)
    results = {}
    field_err = {}
    ref_out = None
    for form in forms:
        base, f32 = form_parts(form)
        cur = bufs32 if f32 else bufs
        gcur = cur[1]
        pid = {v: k for k, v in dsht.PREC_NAMES.items()}[base]
        try:
            plan.set_precision(pid)
            run_step(plan, dsht, cur, stream)
            torch.cuda.synchronize()
        except dsht.DshtError as e:
            print(f"[bench] form {form} unavailable: {e}", file=sys.stderr, flush=True)
            continue
        outs = [g.clone().double() for g in gcur] + [spec_t.clone()]
        if ref_out is None:
            ref_out = outs
            err = {"oracle_rel_l2_inverse": oracle_check(torch, spec, g_out)}
        else:
            err = {f"rel_l2_{nm}": rel_l2(a, b) for nm, a, b in
                   zip(list(INV_KINDS) + ["forward"], outs, ref_out)}
            field_err[form] = per_field_errors(outs, ref_out)
        print(f"[bench] form {form}: timing {args.steps} steps", file=sys.stderr, flush=True)
        wall, _ = time_steps(torch, plan, dsht, cur, args.steps, args.warmup)
        lat = latency_loop(torch, plan, dsht, cur, args.latency_steps)
        prof = kernel_profile(torch, plan, dsht, cur, max(2, min(args.steps, 5)))
        # the same step with one dsht_inv_d call per inverse kind (function-per-kind API)
        wall_s, _ = time_steps(torch, plan, dsht, cur, args.steps, args.warmup, fused=False)
        sep_err = max(rel_l2(a, b) for a, b in zip(gcur, outs[:3]))
        results[form] = {
            "pairs_per_s": PAIRS_PER_STEP * args.steps / wall,
            "ms_per_step": 1e3 * wall / args.steps,
            "p50_step_ms": lat["p50_step_ms"],
            "p99_step_ms": lat["p99_step_ms"],
            "latency": lat,
            "errors": err,
            "kernels": prof,
            "separate_calls": {"pairs_per_s": PAIRS_PER_STEP * args.steps / wall_s,
                               "ms_per_step": 1e3 * wall_s / args.steps,
                               "rel_l2_vs_fused": sep_err},
        }
        print(f"[bench] form {form}: {results[form]['pairs_per_s']:.0f} pairs/s "
              f"{results[form]['ms_per_step']:.3f} ms/step err {err}", file=sys.stderr, flush=True)
    plan.set_precision(dsht.PREC_FP64)

    ok = [f for f in results if f == "fp64" or max(results[f]["errors"].values()) <= TOL_MIXED]
    best = max(ok, key=lambda f: results[f]["pairs_per_s"])
    r = results[best]
    drift = {f: energy_drift(torch, plan, dsht, f, spec) for f in results if not form_parts(f)[1]}
    # roofline of the dominant kernel of the reported form, probed on the step's own calls
    kname = max(r["kernels"].items(), key=lambda kv: kv[1]["ms_total"])[0]
    roof = roofline_probe(torch, plan, dsht, best, kname, bufs32 if form_parts(best)[1] else bufs)
    t85 = None
    if not args.no_t85:
        print("[bench] T85L60 FP64 line", file=sys.stderr, flush=True)
        t85 = t85_line(torch, dsht, args.steps, args.warmup)
        print(f"[bench] T85: {t85['pairs_per_s']:.0f} pairs/s, {t85['ms_per_step']:.3f} ms/step", file=sys.stderr,
              flush=True)
    print("[bench] column diagnostics", file=sys.stderr, flush=True)
    column = column_bench(torch, dsht)
    print("[bench] column: " + ", ".join(f"{k} {column[k]['avg_us']:.1f} us {column[k]['GBps']:.0f} GB/s"
                                         for k in ("omega", "hydpres", "uvbarot", "quick", "vdiff")), file=sys.stderr, flush=True)
    batched = serving = lf = rg = None
    if not args.no_redigm:
        print("[bench] Redi/GM mixing", file=sys.stderr, flush=True)
        rg = redigm_bench(torch, dsht, cpu=not args.no_cpu)
        print("[bench] Redi/GM: " + ", ".join(f"{f} {rg[f]['ms_per_call']:.2f} ms/call" for f in ("fp64", "bf16x6")),
              file=sys.stderr, flush=True)
    if not args.no_lfam3:
        print("[bench] LF-AM3 step", file=sys.stderr, flush=True)
        lf = lfam3_bench(torch, dsht)
        print("[bench] LF-AM3: " + ", ".join(f"{k} {v['ms_per_step']:.2f} ms/step {v['pairs_per_s']:.0f} pairs/s"
                                             for k, v in lf.items() if k.startswith("runs")), file=sys.stderr,
              flush=True)
    if not args.no_batch:
        print("[bench] batched form (64 instances)", file=sys.stderr, flush=True)
        batched = batched_bench(torch, dsht)
        print(f"[bench] batched: graph {batched['graph_ms_per_step']:.2f} ms/step, sequential "
              f"{batched['sequential_ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    if not args.no_serve:
        print("[bench] serving form", file=sys.stderr, flush=True)
        serving = serving_bench(torch, dsht)
        print(f"[bench] serving: throughput ratio {serving['throughput_ratio']:.2f}, p99/solo "
              f"{serving['batched']['p99_over_solo_max']:.2f}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(spec, tend)

    line = {
        "metric": METRIC,
        "value": r["pairs_per_s"],
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best == "fp64" else best,
        "mixed_speedup_vs_fp64_fp64grids": (max(results[f]["pairs_per_s"] for f in ok if f != "fp64" and
                                                 not form_parts(f)[1]) / results["fp64"]["pairs_per_s"])
        if any(f != "fp64" and not form_parts(f)[1] for f in ok) and "fp64" in results else None,
        "data": "synthetic",
        "config": {"workload": "T170L60 dynamics-step transform set (4 fields to grid with lon/lat "
                               "derivatives, 6 tendencies to spectral); inverse kinds of a field in one "
                               "fused call (forms.*.separate_calls: one call per kind)",
                   "nmax": NMAX, "im": IM, "jm": JM, "km": KM, "units_per_step": UNITS_PER_STEP,
                   "pairs_per_step": PAIRS_PER_STEP, "parallelism": "single GPU"},
        "p50_step_ms": r["p50_step_ms"],
        "p99_step_ms": r["p99_step_ms"],
        "latency_steps": args.latency_steps,
        "forms": {f: {"pairs_per_s": v["pairs_per_s"], "ms_per_step": v["ms_per_step"],
                      "p50_step_ms": v["p50_step_ms"], "p99_step_ms": v["p99_step_ms"], "errors": v["errors"],
                      "separate_calls": v["separate_calls"],
                      "kernels": {kk: {"avg_us": kv["avg_us"], "GBps": kv["GBps"], "TFLOPs": kv["TFLOPs"]}
                                  for kk, kv in v["kernels"].items()}}
                  for f, v in results.items()},
        "per_field_errors_vs_fp64": field_err,
        "energy_drift_100_roundtrips": drift,
        "mixed_speedup_vs_fp64": (r["pairs_per_s"] / results["fp64"]["pairs_per_s"]) if "fp64" in results else None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": serving,
        "column_diagnostics": column,
        "t85_fp64": t85,
        "lfam3": lf,
        "redigm": rg,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
