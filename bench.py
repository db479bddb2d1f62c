"""Benchmark of the TiltRec SIRT path on one MI355X.

Workload (one step): SIRT x20, gamma 1.0, of a 61-tilt series of
1024 x 1024 float32 projections (-60..+60 degrees in 2 degree steps, the
ellipsoid-and-bead phantom scaled x4, noise and per-tilt shifts/rotations)
into a 1024 x 1024 x 256 volume, TiltRec-mpi numeric rules, plain FP32.
The stack is resident in HBM before timing starts; a step runs the whole
reconstruction (layout transposes, ray weights, 20 iterations, output).

Metric: voxel-projection updates/s = X * Y * T * P * iterations / seconds
per step, with p50/p99 step latency.  The roofline of the dominant kernel
is measured live with HIP events on its own stream; the CPU baseline times
the CPU restatement of TiltRec-mpi's loops on a bounded sample of rows.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SIRT voxel·projection updates/s at 1 GPU vs plain FP32; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q
    lo = int(k)
    hi = min(lo + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def time_form(trec, ctx, torch, geom, ang, stack_d, storage, iters, steps, warmup):
    """Times `steps` reconstructions; returns per-step milliseconds."""
    method = trec.make_method(trec.SIRT, iters, 1.0, storage=storage)
    vol = torch.empty((geom.ny, geom.thickness, geom.nx), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        ctx.reconstruct(geom, ang, stack_d, method, vol, stream=stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    ev[0].record(stream)
    for k in range(steps):
        ctx.reconstruct(geom, ang, stack_d, method, vol, stream=stream)
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    return [ev[k].elapsed_time(ev[k + 1]) for k in range(steps)], vol


def methods_form(trec, ctx, torch, geom, ang, stack_d, reps=3):
    """Milliseconds per call of the one-pass methods, of ADMM, and of SIRT x20
    with a 0.5 degree pitch, alone and with a 1 degree tilt-angle offset
    (the geometries a GeoParGen estimate and FindOffset feed back, which run
    the general path, bilinear in both detector axes; with both, the
    detector row also moves along x), and with a 5 degree pitch and a 4
    degree offset (the detector row moves by 3 rows over 512 voxels), all on
    the same series (stack resident, volume in layout Y)."""
    vol = torch.empty((geom.ny, geom.thickness, geom.nx), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    pitched = trec.make_geom(geom.nx, geom.ny, geom.nproj, geom.thickness, pitch=0.5)
    pitched_off = trec.make_geom(geom.nx, geom.ny, geom.nproj, geom.thickness, offset=1.0, pitch=0.5)
    pitched_off54 = trec.make_geom(geom.nx, geom.ny, geom.nproj, geom.thickness, offset=4.0, pitch=5.0)
    cases = [("BPT", "BPT", geom), ("WBP", "WBP", geom), ("FBP", "FBP", geom),
             ("ADMM,2,5,0.5,0.001", "ADMM,2,5,0.5,0.001", geom), ("SIRT,20,1.0 pitch 0.5", "SIRT,20,1.0", pitched),
             ("SIRT,20,1.0 pitch 0.5 offset 1.0", "SIRT,20,1.0", pitched_off),
             ("SIRT,20,1.0 pitch 5.0 offset 4.0", "SIRT,20,1.0", pitched_off54)]
    res = {}
    for label, text, g in cases:
        m = trec.make_method(text)
        ctx.reconstruct(g, ang, stack_d, m, vol, stream=stream)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
        ev[0].record(stream)
        for k in range(reps):
            ctx.reconstruct(g, ang, stack_d, m, vol, stream=stream)
            ev[k + 1].record(stream)
        torch.cuda.synchronize()
        ms = sum(ev[k].elapsed_time(ev[k + 1]) for k in range(reps)) / reps
        res[label] = {"ms_per_call": round(ms, 3)}
        if m.method == trec.ADMM:
            # per outer iteration the operator is applied to the start and
            # twice per CG step (the direction, then the true residual)
            res[label]["operator_applications"] = m.iterations * (2 * m.cg_iterations + 1)
        log(f"[bench] {label}: {ms:.2f} ms/call")
    return res


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_cores():
    """The cores this process may use, at most 16 (one GPU's share of a box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(ang, stack, T, iters):
    """The CPU restatement of TiltRec-mpi's SIRT (the oracle, FP32) on a
    bounded sample of rows of the same series: rows are independent at pitch
    0, so a row's time scales to the whole volume.  Three builds of the same
    C sources:
      O0      the reference's shipped flags (none: CMakeLists.txt:21-24),
              one process, one row;
      native  -O3 -march=native, built here on this host, one row;
      split   native, N processes on N disjoint rows at once (the ideal form
              of TiltRec-mpi's y split, main-mpi.cpp:710-740, without its
              margin bug).
    The headline is the split."""
    import subprocess
    import tempfile

    import numpy as np
    from oracle import orc
    subprocess.check_call(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "lib/liborc_native.so",
                           "lib/liborc_O0.so"])
    P, Y, X = stack.shape
    coef = orc.coeffs_mpi(ang)
    org = orc.origin(X, T, T / 2)
    per_row = float(X) * T * P * iters
    res = {"cpu_model": cpu_model()}
    for name in ("O0", "native"):
        L = orc.variant(name)
        row = stack[:, Y // 2:Y // 2 + 1, :].copy()
        t = time.perf_counter()
        orc.sirt_f32(coef, row, T, org, iters, 1.0, L=L)
        dt = time.perf_counter() - t
        res[name] = {"value": per_row / dt, "cores": 1, "seconds": round(dt, 2), "rows": 1}
        log(f"[bench] CPU {name}, one row: {per_row / dt:.4e} updates/s ({dt:.1f} s)")
    n = cpu_cores()
    with tempfile.TemporaryDirectory() as td:
        procs = []
        for k in range(n):
            y = (k * Y) // n
            f = os.path.join(td, f"rows{k}.npz")
            np.savez(f, coef=coef, rows=stack[:, y:y + 1, :].copy(), T=T, iters=iters)
            procs.append(subprocess.Popen([sys.executable, "-m", "oracle.cpu_worker", f, "native"], cwd=ROOT,
                                          stdout=subprocess.PIPE, text=True))
        secs = []
        for pr in procs:
            out, _ = pr.communicate()
            if pr.returncode != 0:
                raise RuntimeError("CPU baseline worker failed")
            secs.append(json.loads(out.strip().splitlines()[-1])["seconds"])
    wall = max(secs)
    res["split"] = {"value": n * per_row / wall, "cores": n, "seconds": round(wall, 2), "rows": n}
    log(f"[bench] CPU native, {n} processes on {n} rows: {n * per_row / wall:.4e} updates/s ({wall:.1f} s)")
    return res


def end_to_end(ang, stack, T, iters):
    """The TiltRec drop-in end to end: the CLI (installed as TiltRec-mpi, the
    bench's numeric rules) reads the MRC stack and tilt file, reconstructs
    and writes the MRC volume.  Wall clock of the whole process, and the
    CLI's own read / upload / reconstruct / write split."""
    import subprocess
    import tempfile

    import numpy as np
    from autom2_amd import mrc
    P, Y, X = stack.shape
    with tempfile.TemporaryDirectory() as td:
        fin, fang, fout = (os.path.join(td, n) for n in ("stack.mrc", "tilts.tlt", "vol.mrc"))
        mrc.write(fin, stack)
        np.savetxt(fang, ang, fmt="%.6f")
        cmd = [os.path.join(ROOT, "autom2_amd", "bin", "TiltRec-mpi"), "--input", fin, "--output", fout,
               "--tiltfile", fang, "--geometry", f"0,0,0,{T}", "--method", f"SIRT,{iters},1.0"]
        t = time.perf_counter()
        out = subprocess.run(cmd, check=True, capture_output=True, text=True).stdout
        wall = time.perf_counter() - t
    split = {}
    for line in out.splitlines():
        if line.startswith("read "):
            parts = [p.strip().split() for p in line.split(",")]
            split = {p[0]: float(p[1]) for p in parts if len(p) == 3}
    return {"seconds": round(wall, 3), "value": float(X) * Y * T * P * iters / wall,
            "unit": "voxel-projection updates/s", "cli_split_s": split,
            "workload": f"TiltRec-mpi CLI: MRC stack {P} x {Y} x {X} float32 in, SIRT x{iters}, "
                        f"MRC volume {X} x {T} x {Y} float32 out (local files)"}


def serving_mix(n, seed):
    """Requests drawn from the batching mix: size {512,1024,2048} px square,
    tilts {41,61,81} (3, 2, 1.5 degree steps), thickness 128..512, method
    WBP, SIRT x10 or SIRT x20.  Stacks are random (their content does not
    change the work); one stack per (size, tilts) is shared."""
    import numpy as np
    rng = np.random.default_rng(seed)
    steps = {41: 3.0, 61: 2.0, 81: 1.5}
    reqs, pool = [], {}
    for _ in range(n):
        X = int(rng.choice([512, 1024, 2048]))
        P = int(rng.choice([41, 61, 81]))
        T = int(rng.integers(128, 513))
        method = str(rng.choice(["WBP", "SIRT,10,1.0", "SIRT,20,1.0"]))
        if (X, P) not in pool:
            pool[(X, P)] = (rng.random((P, X, X), dtype=np.float32) * 50.0)
        ang = (np.arange(P) * steps[P] - (P - 1) * steps[P] / 2).astype(np.float32)
        reqs.append(dict(X=X, P=P, T=T, method=method, ang=ang, stack=pool[(X, P)]))
    return reqs


def out_pool(reqs, dtype, k=8):
    """Host output buffers for the served volumes: k arrays of the largest
    volume, request i writing into buffer i % k.  The bench checks no
    values, and holding every request's volume at once would take more host
    memory than a box gives one run."""
    import numpy as np
    big = max(r["X"] * r["T"] * r["X"] for r in reqs)
    return [np.empty(big, dtype) for _ in range(min(k, len(reqs)))]


def serve(trec, reqs, arrivals, storage, max_active, half_io=False):
    """Submits the requests to a server at their arrival times and waits for
    all: latencies (submission to volume in host memory) and the span from
    the first submission to the last completion.  half_io: the stacks are
    IEEE half (as MRC mode 12 files hold them) and the volumes come back as
    half."""
    import numpy as np
    pool = out_pool(reqs, np.float16 if half_io else np.float32)
    stacks = {}
    if half_io:
        for r in reqs:
            if id(r["stack"]) not in stacks:
                stacks[id(r["stack"])] = r["stack"].astype(np.float16)
    # WBP keeps a float stack: its ramp filter would raise the stack's half
    # rounding above 1e-3 of the volume
    half_in = lambda r: half_io and not r["method"].startswith("WBP")
    srv = trec.Server(0, budget=96 << 30, max_active=max_active)
    t0 = time.perf_counter()
    ids = []
    for k, (r, ta) in enumerate(zip(reqs, arrivals)):
        dt = ta - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        if k and k % 50 == 0:
            log(f"[bench] serving: {k} of {len(reqs)} submitted at {time.perf_counter() - t0:.1f} s")
        out = pool[k % len(pool)][:r["X"] * r["T"] * r["X"]]
        st = stacks[id(r["stack"])] if half_in(r) else r["stack"]
        ids.append(srv.submit(trec.make_geom(r["X"], r["X"], r["P"], r["T"]),
                              trec.make_method(r["method"], storage=storage), r["ang"], st, out))
    recs = [srv.wait(i) for i in ids]
    srv.close()
    lat = [(rec["done_s"] - rec["submit_s"]) * 1e3 for rec in recs]
    span = max(rec["done_s"] for rec in recs) - min(rec["submit_s"] for rec in recs)
    return lat, span


def serving_form(trec, torch, n, rate, seed=1004, active=4, saturated_only=False, poisson_only=False, arms_only=None):
    """The serving configuration: n requests of the batching mix, Poisson
    arrivals at `rate` per second and all at once (saturated).  Every arm
    runs through the same server and copy machinery (pinned double-buffered
    staging, threaded host copies):
      plain_f32       one request at a time (max_active 1), FP32 storage;
      batched_f16     continuous batching (up to 4 active), FP16 storage;
      batched_f16_half_io  the same with half volumes out (MRC mode 12) and,
                      for SIRT, half stacks in: half the bytes over PCIe;
      batched_f32     continuous batching, FP32 storage (saturated only).
    Up to `active` (4) requests at once: most requests of the mix fill the
    GPU alone, so more at once only dilute the L2.
    Latency is from submission to the volume in host memory.  saturated_only
    and poisson_only keep one load; arms_only (names) keeps those arms."""
    import numpy as np
    reqs = serving_mix(n, seed)
    gaps = np.random.default_rng(seed + 1).exponential(1.0 / rate, n)
    arrivals = list(np.cumsum(gaps) - gaps[0])
    out = {"requests": n, "arrival_rate_per_s": rate,
           "copies": "every arm through the server's pinned staging; latency includes host copies"}
    stat = lambda lat, span: {"series_per_s": round(n / span, 3), "p50_ms": round(percentile(lat, 0.5), 1),
                              "p99_ms": round(percentile(lat, 0.99), 1)}
    arms = [("plain_f32", trec.STORE_F32, 1, False), ("batched_f16", trec.STORE_F16, active, False),
            ("batched_f16_half_io", trec.STORE_F16, active, True)]
    out["max_active"] = active
    modes = ("saturated",) if saturated_only else (("poisson",) if poisson_only else ("poisson", "saturated"))
    for mode in modes:
        arr = arrivals if mode == "poisson" else [0.0] * n
        res = {}
        for name, sto, act, half in arms + ([("batched_f32", trec.STORE_F32, active, False)] if mode == "saturated"
                                            else []):
            if arms_only and name not in arms_only:
                continue
            log(f"[bench] serving ({mode}): {name}")
            lat, span = serve(trec, reqs, arr, sto, act, half)
            res[name] = stat(lat, span)
            res[name]["span_s"] = round(span, 3)
        base = res["plain_f32"]["span_s"] if "plain_f32" in res else None
        for name in res:
            if name != "plain_f32" and base:
                res[name]["throughput_vs_plain_f32"] = round(base / res[name]["span_s"], 3)
        out[mode] = res
        log(f"[bench] serving ({mode}): {res}")
    return out


def xf_stage(trec, ctx, torch, stack_d, xf, reps=5, cpu=True):
    """The .xf application of the raw stack (mrcstack, cubic): per-call and
    per-kernel times (HIP events on the launch stream), the warp kernel's
    HBM roofline (4 bytes read + 4 written per pixel), and the CPU oracle on
    a few sections."""
    P, H, W = stack_d.shape
    out = torch.empty_like(stack_d)
    stream = torch.cuda.current_stream()
    for _ in range(2):
        ctx.apply_xf(stack_d, xf, out, stream=stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ev[0].record(stream)
    for k in range(reps):
        ctx.apply_xf(stack_d, xf, out, stream=stream)
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    call_ms = sorted(ev[k].elapsed_time(ev[k + 1]) for k in range(reps))[reps // 2]
    ctx.profile(True)
    for _ in range(reps):
        ctx.apply_xf(stack_d, xf, out, stream=stream)
    torch.cuda.synchronize()
    warp_ms, nw = ctx.profile_read("xf_warp")
    mean_ms, nm = ctx.profile_read("xf_mean")
    ctx.profile(False)
    warp_ms /= max(nw, 1)
    mean_ms /= max(nm, 1)
    px = float(P) * H * W
    achieved = 8 * px / (warp_ms / 1e3) / 1e9
    res = {"workload": f"cubic .xf application of {P} sections of {W}x{H} float32 (the bench stack "
                       f"and its true alignment)",
           "value": px / (call_ms / 1e3), "unit": "pixels/s", "ms_per_call": round(call_ms, 4),
           "kernels_ms": {"xf_warp": round(warp_ms, 4), "xf_mean": round(mean_ms, 4)},
           "roofline": {"bound": "hbm", "kernel": "xf_warp", "achieved": round(achieved, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                        "bytes_per_launch": 8 * px,
                        "traffic": profiled_traffic("xf_warp<4", "xf1024_cubic_counters.md")
                        if (P, H, W) == (61, 1024, 1024) else None}}
    if cpu:
        from oracle import orc
        n = 4
        host = stack_d[:n].cpu().numpy()
        t0 = time.perf_counter()
        orc.apply_xf(host, xf[:n], mode=1)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": n * H * W / dt, "unit": "pixels/s", "cores": 1, "kind": "port",
                               "sample": f"{n} of {P} sections ({dt:.2f} s)"}
    del out
    return res


def geopar_stage(trec, ctx, torch, ang, stack_d, reps=5, cpu=True):
    """AuTom2's geometry workflow (CommandChainManager.qml:210-252) with
    everything resident in HBM: TiltRec-cuda's SIRT x10 at thickness 300 of
    the bench series (MRC (X, 300, Y), which is layout Y), then GeoParGen's
    x views of that volume.  Per-call and per-kernel times (HIP events on the
    launch stream), the exhaustive search for comparison, and the CPU oracle
    on a sample of views."""
    P, Y, X = stack_d.shape
    T = 300
    g = trec.make_geom(X, Y, P, T, rules=trec.RULES_CUDA)
    m = trec.make_method("SIRT,10,1.0")
    vol = torch.empty((Y, T, X), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    ctx.reconstruct(g, ang, stack_d, m, vol, stream=stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record(stream)
    ctx.reconstruct(g, ang, stack_d, m, vol, stream=stream)
    ev[1].record(stream)
    avg, nex = ctx.geopar(vol, trec.GP_VIEW_X, stream=stream)
    ev[2].record(stream)
    torch.cuda.synchronize()
    rec_ms, first_ms = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
    times = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.geopar(vol, trec.GP_VIEW_X, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    call_ms = sorted(times)[reps // 2]
    ctx.profile(True)
    for _ in range(reps):
        ctx.geopar(vol, trec.GP_VIEW_X, stream=stream)
    torch.cuda.synchronize()
    kern = {}
    for k in ("gp_profile", "gp_search", "gp_views"):
        tot, n = ctx.profile_read(k)
        kern[k] = tot / max(n, 1)
    ctx.profile(False)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    avg_x, _ = ctx.geopar(vol, trec.GP_VIEW_X, exhaustive=True, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    exh_ms = e0.elapsed_time(e1)
    nchips = len(range(0, Y - 50, 50))
    pairs = float(X) * nchips * (T - 30) * (T - 29) / 2
    # the profile pass reads the volume twice: min/max, then the chips' sections
    prof_bytes = 4.0 * X * T * Y + 4.0 * X * T * 50 * nchips
    achieved = prof_bytes / (kern["gp_profile"] / 1e3) / 1e9
    res = {"workload": f"GeoParGen x views of TiltRec-cuda's SIRT x10 volume of the bench series "
                       f"(MRC {X} x {T} x {Y}): {X} views, {nchips} chips of {T} rows each",
           "value": X / (call_ms / 1e3), "unit": "views/s", "ms_per_call": round(call_ms, 4),
           "first_call_ms": round(first_ms, 4), "reconstruction_ms": round(rec_ms, 3),
           "averages": [round(v, 5) for v in avg], "exhaustive_chips": nex,
           "exhaustive_search_ms_per_call": round(exh_ms, 3), "same_as_exhaustive": list(avg) == list(avg_x),
           "pairs_per_s": pairs / (call_ms / 1e3),
           "kernels_ms": {k: round(v, 4) for k, v in kern.items()},
           "roofline": {"bound": "hbm", "kernel": "gp_profile", "achieved": round(achieved, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                        "bytes_per_launch": prof_bytes, "traffic": None}}
    if (P, Y, X) == (61, 1024, 1024):
        t = [profiled_traffic(k, "geopar1024_kernel_trace.md") for k in ("gp_minmax_x", "gp_profiles_x")]
        if None not in t:
            res["roofline"]["traffic"] = sum(t)
    if cpu:
        from oracle import orc
        n = 2
        sub = vol[:, :, 500:500 + n].contiguous().cpu().numpy()
        t0 = time.perf_counter()
        orc.geopar(sub, 0)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": n / dt, "unit": "views/s", "cores": 1, "kind": "port",
                               "sample": f"{n} of {X} views ({dt:.2f} s); the reference evaluates each "
                                         "view's search twice, the oracle once"}
    del vol
    return res


def counters_run(iters, config="sirt1024"):
    """The run the counter profiles measure: `iters` SIRT iterations of the
    configuration's series through the library after autotuning, as the
    bench runs them (FP32, and FP16 storage too for sirt2048), then a 1 GiB
    device copy whose FETCH_SIZE / WRITE_SIZE calibrate the HBM byte
    counters."""
    import torch
    from autom2_amd import synth, tiltrec as trec
    c = synth.CONFIGS[config]
    ang, stack, _, _ = synth.config_series(config)
    s = torch.from_numpy(stack).cuda()
    del stack
    g = trec.make_geom(c["X"], c["Y"], c["P"], c["T"])
    ctx = trec.Context(0)
    vol = torch.empty((c["Y"], c["T"], c["X"]), device="cuda")
    for sto in ([trec.STORE_F32] if config == "sirt1024" else [trec.STORE_F32, trec.STORE_F16]):
        m = trec.make_method(trec.SIRT, iters, 1.0, storage=sto)
        ctx.autotune(g, m)  # the launch shapes the bench uses
        log(f"[bench] counters run: storage {sto}, plan {ctx.plan(g, ang, m)}")
        ctx.reconstruct(g, ang, s, m, vol)
        torch.cuda.synchronize()
    del vol, s
    torch.cuda.empty_cache()
    a = torch.empty(256 * 1024 * 1024, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log("[bench] counters run done")


def profiled_traffic(kernel, summary):
    """Read plus write bytes per dispatch of the kernel whose name starts with
    k_<kernel> in a committed counter summary, or None."""
    rd = profiled("k_" + kernel, "read GB", summary, prefix=True)
    wr = profiled("k_" + kernel, "write GB", summary, prefix=True)
    return None if rd is None or wr is None else (rd + wr) * 1e9


def profiled(kernel, column, summary="sirt1024_f32_counters.md", prefix=False):
    """A per-dispatch figure of the kernel named `kernel` (its name up to the
    argument list) from a committed counter summary under profiles/: the
    cell in `column` of the first table that has that column and a row for
    that kernel, or None."""
    path = os.path.join(ROOT, "profiles", summary)
    try:
        col = None
        for line in open(path):
            cells = [c.strip() for c in line.split("|")]
            if len(cells) < 4:
                col = None
                continue
            if column in cells:
                col = cells.index(column)
            elif col is not None:
                m = re.match(r"`([^`]*)`", cells[1])
                name = (m.group(1) if m else cells[1]).split("(")[0].strip()
                if name == kernel or (prefix and name.startswith(kernel)):
                    return float(cells[col].replace(",", ""))
    except (OSError, ValueError):
        pass
    return None


# MI355X: 256 CUs at 2.4 GHz, each issuing one 64-lane FP32 multiply-add per
# cycle without packed instructions (half the published 157.3 TFLOP/s, which
# counts packed FP32; the kernels here emit none).  Mixed instruction streams can
# issue somewhat faster, so the fraction is a rate, not a hard ceiling.
VALU_PEAK_WAVE_INSTS = 256 * 2.4e9


def kernel_roofline(kern, plan, X, Y, T, P, storage_bytes=4, summary=None):
    """The dominant kernel of a SIRT step, its algorithmic bytes per launch and
    HBM roofline fraction (event-timed average duration), with the HBM
    traffic and VALU instruction count of the same kernel from a committed
    counter summary when there is one.
    Algorithmic bytes: V = X Y T voxels, PA = P X Y pixels, s = storage bytes;
    forward 4V + 2 s PA when it writes the residual itself (one z chunk: reads
    p, writes r), else 4V + 4 nzc PA (writes its chunk partials); combine
    (4 nzc + 2 s) PA; back 8V + s PA."""
    V = float(X) * Y * T
    PA = float(P) * X * Y
    nzc = plan["zchunks"]
    sb = storage_bytes
    alg = {"forward": 4 * V + (2 * sb * PA if nzc == 1 else 4 * nzc * PA), "combine": (4 * nzc + 2 * sb) * PA,
           "back": 8 * V + sb * PA}
    dom = max(("forward", "back"), key=lambda k: kern[k]["ms_total"])
    st = {4: "float", 2: "_Float16"}[sb]
    hv = "true" if sb == 2 else "false"  # FP16 forwards read the half volume copy
    names = {"forward": (f"k_forward_g<{st}, true, {hv}, {plan['band_cells']}>" if nzc == 1
                         else f"k_forward_g<float, false, {hv}, 400>"),
             "back": f"k_back<{st}, 0, {plan['back_tilts']}, {plan['back_voxels_z']}>"}
    name = names[dom]
    ms = kern[dom]["ms_avg"]
    achieved = alg[dom] / (ms / 1e3) / 1e9
    out = {"bound": "hbm", "kernel": dom, "kernel_name": name, "achieved": round(achieved, 2),
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
           "bytes_per_launch": alg[dom], "ms_per_launch": round(ms, 4),
           "kernels_ms": {k: round(v["ms_avg"], 4) for k, v in kern.items()}}
    if summary:
        rd, wr = profiled(name, "read GB", summary), profiled(name, "write GB", summary)
        if rd is not None and wr is not None:
            out["traffic"] = (rd + wr) * 1e9
        valu = profiled(name, "SQ_INSTS_VALU", summary)
        if valu is not None:
            out["valu"] = {"insts_per_launch": valu,
                           "frac": round(valu / (ms / 1e3) / VALU_PEAK_WAVE_INSTS, 4),
                           "peak": "256 CUs x 1 wave64 FP32 multiply-add per cycle x 2.4 GHz (non-packed FP32 rate)"}
    return out


def live_kernels(trec, ctx, torch, geom, ang, stack_d, method, vol, runs=2):
    """Average per-kernel milliseconds over `runs` reconstructions, from HIP
    events recorded on the launch stream."""
    ctx.profile(True)
    for _ in range(runs):
        ctx.reconstruct(geom, ang, stack_d, method, vol)
    torch.cuda.synchronize()
    kern = {}
    for k in ("prep", "forward", "combine", "back", "finish"):
        tot, n = ctx.profile_read(k)
        kern[k] = {"ms_total": tot, "launches": n, "ms_avg": tot / n if n else 0.0}
    ctx.profile(False)
    return kern


def sirt2048_form(trec, torch, iters=20, steps=2):
    """SIRT x20 of 61 x 2048^2 -> 2048^2 x 512 (the precision
    configuration's series, scale x8 phantom), FP32 and FP16
    storage, each with its step time and its dominant kernel's roofline."""
    from autom2_amd import synth
    c = synth.CONFIGS["sirt2048"]
    X, Y, T, P = c["X"], c["Y"], c["T"], c["P"]
    log(f"[bench] generating sirt2048: {P} x {Y} x {X} -> {Y} x {T} x {X}")
    ang, stack, _, _ = synth.config_series("sirt2048")
    stack_d = torch.from_numpy(stack).cuda()
    del stack
    geom = trec.make_geom(X, Y, P, T)
    ctx = trec.Context(0)
    res = {"workload": f"SIRT x{iters} gamma 1.0, {P} tilts of {X}x{Y} float32 (-60..+60 deg) -> {X}x{Y}x{T}"}
    updates = float(X) * Y * T * P * iters
    for name, sto, sb in (("f32", trec.STORE_F32, 4), ("f16", trec.STORE_F16, 2)):
        m = trec.make_method(trec.SIRT, iters, 1.0, storage=sto)
        ctx.autotune(geom, m)
        plan = ctx.plan(geom, ang, m)
        ms, vol = time_form(trec, ctx, torch, geom, ang, stack_d, sto, iters, steps, 1)
        kern = live_kernels(trec, ctx, torch, geom, ang, stack_d, m, vol, runs=1)
        del vol
        msm = sum(ms) / len(ms)
        res[name] = {"ms_per_step": round(msm, 2), "value": updates / (msm / 1e3), "plan": plan,
                     "roofline": kernel_roofline(kern, plan, X, Y, T, P, sb,
                                                 "sirt2048_counters.md")}
        log(f"[bench] 2048 series {name}: {msm:.1f} ms/step, {updates / (msm / 1e3):.4e} updates/s, "
            f"forward {kern['forward']['ms_avg']:.2f} ms, back {kern['back']['ms_avg']:.2f} ms")
    res["f16_speedup_vs_f32"] = round(res["f32"]["ms_per_step"] / res["f16"]["ms_per_step"], 3)
    ctx.close()
    del stack_d
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="sirt1024")
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-2048", action="store_true", help="skip the 61 x 2048^2 -> 2048^2 x 512 form")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end CLI run with MRC files")
    ap.add_argument("--no-forms", action="store_true", help="skip the FP16/BF16 and serving forms")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-active", type=int, default=4, help="most active requests of the batched serving arms")
    ap.add_argument("--saturated-only", action="store_true", help="serving form: only the saturated load")
    ap.add_argument("--poisson-only", action="store_true", help="serving form: only the Poisson load")
    ap.add_argument("--serve-arms", default="", help="serving form: comma-separated arms to run (default all)")
    ap.add_argument("--serving-only", action="store_true",
                    help="only the serving form (its own JSON line), e.g. with --serve-requests 120")
    ap.add_argument("--counters", action="store_true", help="only the run the counter profiles measure")
    ap.add_argument("--no-xf", action="store_true", help="skip the .xf application stage")
    ap.add_argument("--no-geopar", action="store_true", help="skip the GeoParGen stage")
    ap.add_argument("--geopar-only", action="store_true",
                    help="only the GeoParGen stage (its own JSON line), as its profiles measure it")
    ap.add_argument("--xf-only", action="store_true",
                    help="only the .xf application stage (its own JSON line), as its profiles measure it")
    ap.add_argument("--methods-only", action="store_true",
                    help="only the other methods and the pitched SIRT (their own JSON line), as their profiles "
                         "measure them")
    args = ap.parse_args()
    if args.counters:
        counters_run(2, args.config)
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library runs each series on one device (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from autom2_amd import synth, tiltrec as trec

    c = synth.CONFIGS[args.config]
    X, Y, T, P = c["X"], c["Y"], c["T"], c["P"]
    log(f"[bench] generating {args.config}: {P} x {Y} x {X} -> {Y} x {T} x {X}")
    ang, stack, _xf, _ = synth.config_series(args.config)
    stack_d = torch.from_numpy(stack).cuda()
    geom = trec.make_geom(X, Y, P, T)
    ctx = trec.Context(0)
    if args.geopar_only:
        print(json.dumps({"geopar": geopar_stage(trec, ctx, torch, ang, stack_d, cpu=not args.no_cpu)}), flush=True)
        return
    if args.xf_only:
        print(json.dumps({"xf_apply": xf_stage(trec, ctx, torch, stack_d, _xf, cpu=not args.no_cpu)}), flush=True)
        return
    if args.serving_only:
        del stack_d
        torch.cuda.empty_cache()
        arms_only = [a for a in args.serve_arms.split(",") if a] or None
        print(json.dumps({"serving": serving_form(trec, torch, args.serve_requests, 2.0, active=args.serve_active,
                                                  saturated_only=args.saturated_only, poisson_only=args.poisson_only,
                                                  arms_only=arms_only)}), flush=True)
        return
    if args.methods_only:
        print(json.dumps({"methods": methods_form(trec, ctx, torch, geom, ang, stack_d)}), flush=True)
        return
    t0 = time.perf_counter()
    method = trec.make_method(trec.SIRT, args.iters, 1.0)
    ctx.autotune(geom, method)
    plan = ctx.plan(geom, ang, method)
    t_tune = time.perf_counter() - t0
    # the tuner again on a fresh context: the same choice?
    ctx2 = trec.Context(0)
    ctx2.autotune(geom, method)
    plan["autotune_repeat_same"] = ctx2.plan(geom, ang, method) == plan
    ctx2.close()
    ctx.autotune(geom, method)  # the process table now holds this context's choice again
    plan = dict(ctx.plan(geom, ang, method), autotune_repeat_same=plan["autotune_repeat_same"])
    log(f"[bench] autotune {t_tune:.2f} s; plan {plan}")

    updates = float(X) * Y * T * P * args.iters
    log(f"[bench] timing FP32 SIRT x{args.iters}: warmup {args.warmup}, steps {args.steps}")
    ms, vol = time_form(trec, ctx, torch, geom, ang, stack_d, trec.STORE_F32, args.iters, args.steps, args.warmup)
    ms_mean = sum(ms) / len(ms)
    value = updates / (ms_mean / 1e3)
    log(f"[bench] FP32: {ms_mean:.2f} ms/step, {value:.4e} updates/s")

    # live per-kernel durations (HIP events on the launch stream)
    kern = live_kernels(trec, ctx, torch, geom, ang, stack_d, method, vol)
    log(f"[bench] kernels: " + ", ".join(f"{k} {v['ms_avg']:.3f} ms" for k, v in kern.items()))
    roofline = kernel_roofline(kern, plan, X, Y, T, P, 4,
                               "sirt1024_f32_counters.md" if args.config == "sirt1024" else None)

    forms = {}
    if not args.no_forms:
        for name, sto in (("f16", trec.STORE_F16), ("bf16", trec.STORE_BF16)):
            fms, _ = time_form(trec, ctx, torch, geom, ang, stack_d, sto, args.iters, max(3, args.steps // 2), 1)
            fm = sum(fms) / len(fms)
            forms[name] = {"ms_per_step": round(fm, 3), "value": updates / (fm / 1e3),
                           "speedup_vs_f32": round(ms_mean / fm, 3)}
            log(f"[bench] {name}: {fm:.2f} ms/step")

    if not args.no_forms:
        forms["methods"] = methods_form(trec, ctx, torch, geom, ang, stack_d)
        vol = None
        torch.cuda.empty_cache()
        forms["serving"] = serving_form(trec, torch, args.serve_requests, 2.0)

    xf_res = None
    if not args.no_xf:
        xf_res = xf_stage(trec, ctx, torch, stack_d, _xf, cpu=not args.no_cpu)
        log(f"[bench] .xf application: {xf_res['ms_per_call']:.3f} ms/call, warp "
            f"{xf_res['kernels_ms']['xf_warp']:.3f} ms ({xf_res['roofline']['frac']:.3f} of HBM peak)")

    gp_res = None
    if not args.no_geopar:
        gp_res = geopar_stage(trec, ctx, torch, ang, stack_d, cpu=not args.no_cpu)
        log(f"[bench] GeoParGen: {gp_res['ms_per_call']:.3f} ms/call, search {gp_res['kernels_ms']['gp_search']:.3f} ms, "
            f"exhaustive {gp_res['exhaustive_search_ms_per_call']:.1f} ms")

    s2048 = None
    if not args.no_2048:
        vol = None
        torch.cuda.empty_cache()
        s2048 = sirt2048_form(trec, torch)

    e2e = None
    if not args.no_e2e:
        e2e = end_to_end(ang, stack, T, args.iters)
        log(f"[bench] end to end (CLI, MRC in and out): {e2e['seconds']:.2f} s, {e2e['value']:.4e} updates/s, "
            f"{e2e['cli_split_s']}")

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline: oracle SIRT x{args.iters}, O0, native and a {cpu_cores()}-process split")
        cb = cpu_baseline(ang, stack, T, args.iters)
        sp = cb["split"]
        cpu = {"value": sp["value"], "unit": "voxel-projection updates/s", "cores": sp["cores"], "kind": "port",
               "sample": f"{sp['rows']} of {Y} rows of the same SIRT x{args.iters} workload, one row per "
                         f"process ({sp['seconds']} s); rows are independent at pitch 0",
               "build": "-O3 -march=native", "cpu_model": cb["cpu_model"],
               "single_process": {"O0 (the reference's shipped flags)": cb["O0"], "-O3 -march=native": cb["native"]},
               "gpu_over_cpu": round(value / sp["value"], 1)}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-projection updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mean, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"SIRT x{args.iters} gamma 1.0, {P} tilts of {X}x{Y} float32 "
                               f"(-60..+60 deg) -> {X}x{Y}x{T} volume, TiltRec-mpi rules",
                   "global_batch": 1, "iterations": args.iters, "tilts": P, "detector": [X, Y],
                   "volume": [X, Y, T], "parallelism": "single GPU"},
        "latency_ms": {"p50": round(percentile(ms, 0.5), 3), "p99": round(percentile(ms, 0.99), 3)},
        "plan": plan,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "end_to_end": e2e,
        "sirt2048": s2048,
        "forms": forms,
        "xf_apply": xf_res,
        "geopar": gp_res,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
