"""Benchmark: GUPS of the cone-beam operator pair at 512^3 / 512 views on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--nphi 512]
                    [--no-cpu-baseline]

A step is one SIDDONVI sweep in FP32: a clean Siddon projection of the
volume followed by a VI backprojection of the sinogram, device-resident
(512^3 volume of 64 seeded random ellipsoids, 512 views over 360 deg,
512x512 detector, FOD 1000 mm, FDD 2000 mm, 1 mm voxels, 2 mm pixels).
GUPS = voxel-view updates / s = 2 N^3 Nphi / t_step.  The RS projector is
timed alongside.  Kernel times come from HIP events recorded on the stream
the kernels run on.  One JSON line is printed, with these further legs:
  loop       10 gradient iterations at 512^3: SIDDONVI and RSVI in FP32,
             SIDDONVI in its FP16 / BF16 mixed-precision forms
  mixed      FP16 / BF16 storage forms against FP32 (speed, error vs FP64
             and vs FP32 on the GPU, relative L2 and Linf)
  mixed_c2   the 720-view, 768x512-detector workload at 512^3 and 1024^3:
             FP32 checked against the oracle, then the 16-bit forms
  serving    mixed-size reconstruction stream, served vs sequential, p50/p99
  batched    uneven SIRT problems in shared launches vs one at a time
  cpu_baseline  the FP32 oracle on the host cores (views 0..3), and the
             reference C-path Siddon on one thread
The GPU is released before the process ends: every server is closed and
the device synchronised.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GUPS proj & backproj at 512³/512 views, %HBM roofline; mixed vs FP32 @L2≤1e-3; p99 lat"
HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--nphi", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-views", type=int, default=4, help="views in the CPU baseline sample")
    ap.add_argument("--serving-requests", type=int, default=200,
                    help="requests in the serving sample (0 skips the serving leg)")
    ap.add_argument("--serving-rates", default="10,50", help="Poisson arrival rates (req/s)")
    ap.add_argument("--serving-schedules", default="recurrent,direct",
                    help="schedules of the FP32-copy serving runs (the direct one on --serving-direct-requests)")
    ap.add_argument("--serving-direct-requests", type=int, default=64)
    ap.add_argument("--serving-bf16-rate", type=float, default=50.0, help="rate of the BF16-copy run (0: none)")
    ap.add_argument("--mixed-oracle-views", type=int, default=8,
                    help="views of the FP64 oracle check of the mixed forms at 512^3 / 512 views "
                         "(0 skips the mixed legs)")
    ap.add_argument("--c2-oracle-views", type=int, default=32,
                    help="views of the oracle checks of the 720-view workload at edges up to 512")
    ap.add_argument("--c2-oracle-views-large", type=int, default=8,
                    help="views of the oracle checks of the 720-view workload at edges above 512")
    ap.add_argument("--c2-sizes", default="512,1024",
                    help="edges of the mixed-precision workload (720 views, 768x512 detector); empty skips it")
    ap.add_argument("--batch-problems", type=int, default=64,
                    help="uneven SIRT problems in the batched leg (0 skips it)")
    ap.add_argument("--loop-iters", type=int, default=10,
                    help="gradient iterations of the 512^3 reconstruction leg (0: skip)")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="run one element-wise copy of the volume (known bytes) before the timed steps, "
                         "as a reference kernel for FETCH_SIZE / WRITE_SIZE counter scaling")
    return ap.parse_args()


def load_traffic(n, nphi):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (profiles/hbm_traffic_512.md), for the workload they were measured on."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_512.md")
    if not os.path.exists(path):
        return {}
    for line in open(path):
        if line.startswith("traffic_json:"):
            d = json.loads(line.split(":", 1)[1])
            if d.get("n") == n and d.get("nphi") == nphi:
                return d.get("kernels", {})
    return {}


def load_valu(n, nphi):
    """VALU issue fractions per kernel (profiles/valu_issue_512.md), for the
    workload they were measured on."""
    path = os.path.join(ROOT, "profiles", "valu_issue_512.md")
    if not os.path.exists(path):
        return {}
    for line in open(path):
        if line.startswith("valu_json:"):
            d = json.loads(line.split(":", 1)[1])
            if d.get("n") == n and d.get("nphi") == nphi:
                return d.get("kernels", {})
    return {}


def timed(fn, reps, stream):
    """Mean ms of `fn` over `reps` launches, HIP events on `stream`."""
    import torch
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    for i in range(reps):
        starts[i].record(stream)
        fn()
        ends[i].record(stream)
    torch.cuda.synchronize()
    ts = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return sum(ts) / len(ts), ts


def cpu_baseline(vol_host, n, nphi, views, threads):
    """FP32 oracle (test-infrastructure C restatement) on the host cores.
    value: VI backprojection and clean Siddon projection over the first
    `views` views of the same acquisition on `threads` threads.  Beside it:
    RS on the same views, Siddon single-threaded on one view (the reference's
    C path is single-threaded), the 256^3 Shepp-Logan workload in full
    (256 views; Siddon, VI and RS on `threads` threads) and the host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ffi import Oracle, scaled_geometry, set_threads, shepp_logan

    def sub(nn, k, total):
        g = scaled_geometry(nn, k)
        g.num_stop, g.den_stop = 2.0 * k, float(total)   # same angles as views 0..k-1
        return Oracle(g)

    def clock(fn):
        t0 = time.perf_counter()
        out = fn()
        return out, time.perf_counter() - t0

    set_threads(threads)
    o = sub(n, views, nphi)
    sino, t_sid = clock(lambda: o.siddon(vol_host))
    _, t_vi = clock(lambda: o.vi(sino))
    _, t_rs = clock(lambda: o.rs(vol_host))
    upd = float(n) ** 3 * views
    set_threads(1)
    o1 = sub(n, 1, nphi)
    _, t_sid1 = clock(lambda: o1.siddon(vol_host))
    set_threads(threads)
    # the 256^3 workload in full
    o0 = Oracle(scaled_geometry(256, 256))
    v0 = shepp_logan(256)
    s0, t0_sid = clock(lambda: o0.siddon(v0))
    _, t0_vi = clock(lambda: o0.vi(s0))
    _, t0_rs = clock(lambda: o0.rs(v0))
    upd0 = 256.0 ** 3 * 256
    # The reference's own C path (SiddonProjector_compute_C_mem_CPU.cpp:20-276, single-threaded as there), on
    # views 0..compat_views-1 of the same 256-view acquisition.
    compat_views = 32
    oc = sub(256, compat_views, 256)
    _, tc = clock(lambda: oc.siddon_compat(v0, threads=1))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": round(2 * upd / (t_sid + t_vi) / 1e9, 4),
        "unit": "GUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"clean Siddon projection + VI backprojection of the {n}^3 phantom over views 0..{views - 1} "
                  f"of {nphi} (FP32 oracle, OpenMP, {threads} threads): siddon {t_sid:.2f} s, vi {t_vi:.2f} s",
        "siddon_gups": round(upd / t_sid / 1e9, 4),
        "vi_gups": round(upd / t_vi / 1e9, 4),
        "rs_gups": round(upd / t_rs / 1e9, 4),
        "siddon_1thread_gups": round(float(n) ** 3 / t_sid1 / 1e9, 4),
        "siddon_compat_1thread": {"s": round(tc, 3), "gups": round(256.0 ** 3 * compat_views / tc / 1e9, 4),
                                  "sample": f"reference C-path Siddon (compat form of the oracle), one thread, "
                                            f"256^3 Shepp-Logan, views 0..{compat_views - 1} of 256"},
        "c0_256_full": {"siddon_s": round(t0_sid, 3), "vi_s": round(t0_vi, 3), "rs_s": round(t0_rs, 3),
                        "siddon_gups": round(upd0 / t0_sid / 1e9, 4), "vi_gups": round(upd0 / t0_vi / 1e9, 4),
                        "rs_gups": round(upd0 / t0_rs / 1e9, 4),
                        "workload": "256^3 Shepp-Logan, 256 views, 256x256 detector, FP32 oracle"},
        "host": {"cpu_model": model, "nproc": os.cpu_count(), "threads": threads},
    }


def serving_leg(ctx, n_req, rates, schedules, seed=42, bf16_rate=None, direct_requests=64):
    """Mixed-size reconstruction stream: the plain sequential FP32 loop
    (tomo_gd_run, one request after another) against the serving loop at
    Poisson arrival rates.  Requests: edge n in {128,256,384,512} with
    weights .4/.3/.2/.1, 180 or 360 views (p = .5), n x n detector, 20
    gradient iterations with the quadratic regulariser (lambda 0.01, every
    step optimal), pairs alternating RSVI / SIDDONVI, 16-ellipsoid phantom
    seeded by the request id.  Sinograms y = P(phantom) are made on the
    device before the clock starts (excluded from latency).

    The sequential form's service time depends only on the request's
    configuration (edge, views, pair): each of the stream's configurations is
    run twice (after a warm-up call) and the faster time stands for every
    request of that configuration; the sequential throughput is n_req / sum of service
    times and its latencies come from a FIFO queue at the same arrivals.  The
    sequential results of those runs are the references the served results
    of the same requests are compared with.  The serving loop runs FP32
    operator copies with each schedule of `schedules` at every rate, and
    BF16 copies at `bf16_rate`; the direct schedule (no recurrence) runs the
    first `direct_requests` requests only."""
    import numpy as np
    import torch
    from tomogpi_amd import lib
    rng = np.random.default_rng(seed)
    edges = rng.choice([128, 256, 384, 512], size=n_req, p=[0.4, 0.3, 0.2, 0.1])
    views = rng.choice([180, 360], size=n_req)
    reqs = []
    for i in range(n_req):
        n, nphi = int(edges[i]), int(views[i])
        desc = lib.AcquisitionDesc.scaled(n, nphi)
        G = lib.Geometry(ctx, desc)
        pair = lib.PAIR_RSVI if i % 2 == 0 else lib.PAIR_SIDDONVI
        truth = torch.empty(G.vol_shape, dtype=torch.float32, device="cuda")
        G.ellipsoids(truth, count=16, seed=1000 + i, amin=0.02 * n, amax=0.3 * n)
        y = torch.empty(G.sino_shape, dtype=torch.float32, device="cuda")
        G.project(truth, y, lib.PROJ_RS if pair == lib.PAIR_RSVI else lib.PROJ_SIDDON)
        del truth
        reqs.append(dict(n=n, nphi=nphi, desc=desc, G=G, pair=pair, y=y, key=(n, nphi, pair)))
    torch.cuda.synchronize()
    gd = dict(iters=20, kopt=21, lam=0.01)
    # Sequential plain form: the faster of two timed runs per configuration
    # (repeated runs now agree within 1%; the minimum keeps a one-off host
    # stall out of the baseline).
    service_of, refs = {}, {}
    for i, r in enumerate(reqs):
        if r["key"] in service_of:
            continue
        vol = torch.zeros(r["G"].vol_shape, dtype=torch.float32, device="cuda")
        r["G"].gd_run(r["y"], vol, pair=r["pair"], iters=1, kopt=2, lam=0.01)   # warm-up (allocations, code)
        times = []
        for _ in range(2):
            vol.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r["G"].gd_run(r["y"], vol, pair=r["pair"], **gd)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        service_of[r["key"]] = min(times)
        refs[i] = vol
    service = [service_of[r["key"]] for r in reqs]
    seq_total = sum(service)
    out = {"requests": n_req, "sizes": {str(k): int((edges == k).sum()) for k in (128, 256, 384, 512)},
           "configurations": len(service_of),
           "sequential_rps": round(n_req / seq_total, 4), "sequential_busy_s": round(seq_total, 3),
           "sequential_note": "tomo_gd_run per configuration (edge, views, pair), the faster of two timed runs after "
                              "a warm-up call; the stream's sequential time is the sum over its requests",
           "runs": []}
    runs = [(rate, lib.F32, sch, n_req) for sch in schedules for rate in rates
            if sch == lib.SCHEDULE_RECURRENT]
    if lib.SCHEDULE_DIRECT in schedules:
        runs.append((max(rates), lib.F32, lib.SCHEDULE_DIRECT, min(direct_requests, n_req)))
    if bf16_rate:
        runs.append((bf16_rate, lib.BF16, lib.SCHEDULE_RECURRENT, n_req))
    for rate, storage, sch, count in runs:
        arrivals = np.cumsum(np.random.default_rng(seed + int(rate)).exponential(1.0 / rate, size=count))
        t_free, lat_seq = 0.0, []
        for a_t, s_t in zip(arrivals, service[:count]):
            start = max(a_t, t_free)
            t_free = start + s_t
            lat_seq.append(t_free - a_t)
        srv = lib.Server(ctx, max_active=64, use_graph=True, schedule=sch)
        outs = [torch.zeros(r["G"].vol_shape, dtype=torch.float32, device="cuda") for r in reqs[:count]]
        torch.cuda.synchronize()
        ids = []
        t0 = time.perf_counter()
        for i, r in enumerate(reqs[:count]):
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(r["desc"], r["y"], outs[i], pair=r["pair"], storage=storage, **gd))
        for rid in ids:
            srv.wait(rid)
        st = srv.stats()
        srv.close()
        torch.cuda.synchronize()
        errs = [float(torch.linalg.vector_norm(outs[i] - v) / torch.linalg.vector_norm(v))
                for i, v in refs.items() if i < count]
        lat = np.array(lat_seq)
        seq_s = sum(service[:count])
        rec = {
            "rate_rps": rate, "copies": "bf16" if storage == lib.BF16 else "f32",
            "schedule": "recurrent" if sch == lib.SCHEDULE_RECURRENT else "direct", "requests": count,
            "serving_rps": round(st["throughput_rps"], 4),
            "speedup_vs_sequential": round(st["throughput_rps"] * seq_s / count, 3),
            "p50_s": round(st["latency_p50_s"], 3), "p99_s": round(st["latency_p99_s"], 3),
            "sequential_p50_s": round(float(np.percentile(lat, 50, method="inverted_cdf")), 3),
            "sequential_p99_s": round(float(np.percentile(lat, 99, method="inverted_cdf")), 3),
            "peak_active": st["peak_active"], "peak_gib": round(st["peak_bytes"] / 2**30, 2),
            "max_rel_l2_vs_sequential": float(f"{max(errs):.3e}") if errs else None,
            "compared": len(errs)}
        out["runs"].append(rec)
        print(f"[serving] {rec}", file=sys.stderr, flush=True)
        del outs
    for r in reqs:
        r["G"].close()
    return out


def mixed_leg(ctx, G, desc, vol, fp32_ms, stream, oracle_views, check_fp32=False):
    """Mixed-precision forms of one acquisition: volumes and sinograms stored
    in FP16 (scaled so max|x| = 2^14) or BF16, FP32 accumulation, FP64
    geometry.  Speed of each operator against its FP32 time, and the error of
    each 16-bit result against the FP64 oracle and against the FP32 result on
    the GPU, on views 0..oracle_views-1 (a CPU-bounded subset of the same
    acquisition): relative L2 and relative Linf (max |a - b| / max |b|).
    With check_fp32 the FP32 GPU results on the subset are also compared with
    the FP32 oracle (the forms the 16-bit speed-ups are measured against)."""
    import numpy as np
    import torch
    from tomogpi_amd import lib
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ffi import Geometry, Oracle
    n, nphi = G.info["nx"], G.info["nphi"]
    nu, nv = G.info["nu"], G.info["nv"]
    out = {}
    vmax = float(vol.abs().max())
    # FP64 references on the view subset (same angles: the stop angle keeps the step)
    gsub = Geometry(fod=desc.fod_mm, fdd=desc.fdd_mm, nphi=oracle_views, su=desc.u_detector_size_mm,
                    sv=desc.v_detector_size_mm, nu=nu, nv=nv, sx=desc.x_volume_size_mm, sy=desc.y_volume_size_mm,
                    sz=desc.z_volume_size_mm, nx=n, ny=n, nz=n, num_stop=2.0 * oracle_views, den_stop=float(nphi))
    o = Oracle(gsub)
    vol_host = vol.cpu().numpy()
    t0 = time.perf_counter()

    def oracle_step(name, fn):
        t = time.perf_counter()
        r = fn()
        print(f"[mixed] n={n}: oracle {name} on {oracle_views} views: {time.perf_counter() - t:.1f} s",
              file=sys.stderr, flush=True)
        return r

    ref = {"siddon": oracle_step("siddon f64", lambda: o.siddon(vol_host.astype(np.float64), f64=True)),
           "rs": oracle_step("rs f64", lambda: o.rs(vol_host.astype(np.float64), f64=True))}
    # FP32 sinogram subset, the backprojection input
    sino_sub = oracle_step("siddon f32", lambda: o.siddon(vol_host))
    ref["vi"] = oracle_step("vi f64", lambda: o.vi(sino_sub.astype(np.float64), f64=True))
    ref32 = {}
    if check_fp32:
        ref32 = {"siddon": sino_sub, "rs": oracle_step("rs f32", lambda: o.rs(vol_host)),
                 "vi": oracle_step("vi f32", lambda: o.vi(sino_sub))}
    del vol_host
    print(f"[mixed] n={n}: oracle on {oracle_views} views: {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    dsub = dataclasses.replace(desc, projection_nb=oracle_views, num_stop_angle=2.0 * oracle_views,
                               den_stop_angle=float(nphi))
    Gs = lib.Geometry(ctx, dsub)

    def errs(a, b):
        a = np.asarray(a, dtype=np.float64)
        b = np.asarray(b, dtype=np.float64)
        return {"rel_l2": float(f"{np.linalg.norm(a - b) / np.linalg.norm(b):.3e}"),
                "rel_linf": float(f"{np.abs(a - b).max() / np.abs(b).max():.3e}")}

    # FP32 results on the subset (the reference point of the 16-bit forms)
    sin32 = torch.from_numpy(sino_sub).cuda()
    got32 = {}
    for op in ("siddon", "rs"):
        t = torch.empty(Gs.sino_shape, dtype=torch.float32, device="cuda")
        Gs.project(vol, t, lib.PROJ_SIDDON if op == "siddon" else lib.PROJ_RS)
        got32[op] = t.cpu().numpy()
    t = torch.empty(Gs.vol_shape, dtype=torch.float32, device="cuda")
    Gs.backproject(sin32, t)
    got32["vi"] = t.cpu().numpy()
    del t
    if check_fp32:
        out["fp32_vs_fp32_oracle"] = {op: {"bit_exact": bool(np.array_equal(got32[op], ref32[op])),
                                           **errs(got32[op], ref32[op])} for op in ("siddon", "rs", "vi")}
        out["fp32_vs_fp64_oracle"] = {op: errs(got32[op], ref[op]) for op in ("siddon", "rs", "vi")}
        print(f"[mixed] n={n} FP32 vs oracle: {out['fp32_vs_fp32_oracle']}", file=sys.stderr, flush=True)

    for name, tdt, code in (("f16", torch.float16, lib.F16), ("bf16", torch.bfloat16, lib.BF16)):
        sv = 2.0 ** 14 / vmax if name == "f16" else 1.0
        v16 = (vol * sv).to(tdt)
        s16 = torch.empty(G.sino_shape, dtype=tdt, device="cuda")
        b16 = torch.empty(G.vol_shape, dtype=tdt, device="cuda")
        ss = 2.0 ** 14 / (vmax * 2.0 * n) if name == "f16" else 1.0     # ray sums <= max * chord
        sb = 2.0 ** 14 / (vmax * 2.0 * n * nphi * 2.0) if name == "f16" else 1.0
        ms = {}
        for op, fn in (("siddon", lambda: G.project(v16, s16, lib.PROJ_SIDDON, vol_scale=sv, sino_scale=ss,
                                                     stream=stream)),
                       ("vi", lambda: G.backproject(s16, b16, sino_scale=ss, vol_scale=sb, stream=stream)),
                       ("rs", lambda: G.project(v16, s16, lib.PROJ_RS, vol_scale=sv, sino_scale=ss, stream=stream))):
            fn()
            ms[op], _ = timed(fn, 3, stream)
        torch.cuda.synchronize()
        del b16, s16
        # errors on the view subset
        got = {}
        for op in ("siddon", "rs"):
            e_s = torch.empty(Gs.sino_shape, dtype=tdt, device="cuda")
            Gs.project(v16, e_s, lib.PROJ_SIDDON if op == "siddon" else lib.PROJ_RS, vol_scale=sv, sino_scale=ss)
            got[op] = (e_s.float() / ss).cpu().numpy()
        sin16 = (sin32 * ss).to(tdt)
        e_b = torch.empty(Gs.vol_shape, dtype=tdt, device="cuda")
        Gs.backproject(sin16, e_b, sino_scale=ss, vol_scale=sb)
        torch.cuda.synchronize()
        got["vi"] = (e_b.float() / sb).cpu().numpy()
        rec = {
            "siddon_ms": round(ms["siddon"], 3), "vi_ms": round(ms["vi"], 3), "rs_ms": round(ms["rs"], 3),
            "siddon_speedup": round(fp32_ms["siddon"] / ms["siddon"], 3),
            "vi_speedup": round(fp32_ms["vi"] / ms["vi"], 3), "rs_speedup": round(fp32_ms["rs"] / ms["rs"], 3),
            "siddonvi_pair_speedup": round((fp32_ms["siddon"] + fp32_ms["vi"]) / (ms["siddon"] + ms["vi"]), 3),
            "rsvi_pair_speedup": round((fp32_ms["rs"] + fp32_ms["vi"]) / (ms["rs"] + ms["vi"]), 3),
            "vs_fp64_oracle": {op: errs(got[op], ref[op]) for op in ("siddon", "rs", "vi")},
            "vs_fp32_gpu": {op: errs(got[op], got32[op]) for op in ("siddon", "rs", "vi")},
            "error_sample": f"views 0..{oracle_views - 1} of {nphi}, 16-bit inputs and outputs"}
        rec["siddon_rel_l2_vs_fp64"] = rec["vs_fp64_oracle"]["siddon"]["rel_l2"]
        rec["vi_rel_l2_vs_fp64"] = rec["vs_fp64_oracle"]["vi"]["rel_l2"]
        out[name] = rec
        print(f"[mixed] n={n} {name}: {rec}", file=sys.stderr, flush=True)
        del v16, e_s, e_b, sin16
        torch.cuda.empty_cache()
    Gs.close()
    return out


def c2_leg(ctx, sizes, stream, oracle_views, oracle_views_large):
    """The mixed-precision workload at each edge n: 720 views over 360 deg,
    768 x 512 detector with pixels of 2 n / 512 mm (full vertical coverage),
    FOD / FDD 500 / 1000 mm scaled by n / 256, 1 mm voxels; volume: 64
    ellipsoids, seed 7, semi-axes 5..120 voxels scaled by n / 512.  FP32
    times of the three operators, then the 16-bit forms against them."""
    import torch
    from tomogpi_amd import lib
    out = {}
    for n in sizes:
        desc = lib.AcquisitionDesc.scaled(n, 720, nu=768, nv=512, pixel_mm=2.0 * n / 512.0)
        G = lib.Geometry(ctx, desc)
        vol = torch.empty(G.vol_shape, dtype=torch.float32, device="cuda")
        G.ellipsoids(vol, count=64, seed=7, amin=5.0 * n / 512.0, amax=120.0 * n / 512.0)
        sino = torch.empty(G.sino_shape, dtype=torch.float32, device="cuda")
        back = torch.empty(G.vol_shape, dtype=torch.float32, device="cuda")
        fp32 = {}
        for op, fn in (("siddon", lambda: G.project(vol, sino, lib.PROJ_SIDDON, stream=stream)),
                       ("vi", lambda: G.backproject(sino, back, stream=stream)),
                       ("rs", lambda: G.project(vol, sino, lib.PROJ_RS, stream=stream))):
            fn()
            fp32[op], _ = timed(fn, 3, stream)
        upd = float(n) ** 3 * 720
        print(f"[c2] n={n} FP32: " + ", ".join(f"{k} {v:.2f} ms" for k, v in fp32.items()), file=sys.stderr,
              flush=True)
        del sino, back
        torch.cuda.empty_cache()
        res = mixed_leg(ctx, G, desc, vol, fp32, stream, oracle_views if n <= 512 else max(1, oracle_views_large),
                        check_fp32=True)
        res["fp32"] = {k: {"ms": round(v, 3), "gups": round(upd / v / 1e6, 2)} for k, v in fp32.items()}
        out[str(n)] = res
        del vol
        G.close()
        torch.cuda.empty_cache()
    return out


def batch_leg(ctx, n_prob, iters=20, seed=7):
    """Batched uneven SIRT reconstructions: problems with edge n uniform in
    {128,192,256,320,384}, views uniform in [90,360], n x n detector,
    16-ellipsoid phantoms seeded per problem, pairs alternating RSVI /
    SIDDONVI, `iters` SIRT iterations each.  One at a time (tomo_sirt_run,
    FP32) against all at once in the server's shared launches, with FP32
    operator copies (the volumes must agree bit for bit) and with BF16
    operator copies (error against the FP32 volumes reported)."""
    import numpy as np
    import torch
    from tomogpi_amd import lib
    rng = np.random.default_rng(seed)
    edges = rng.choice([128, 192, 256, 320, 384], size=n_prob)
    views = rng.integers(90, 361, size=n_prob)
    probs = []
    for i in range(n_prob):
        n, nphi = int(edges[i]), int(views[i])
        desc = lib.AcquisitionDesc.scaled(n, nphi)
        G = lib.Geometry(ctx, desc)
        pair = lib.PAIR_RSVI if i % 2 == 0 else lib.PAIR_SIDDONVI
        truth = torch.empty(G.vol_shape, dtype=torch.float32, device="cuda")
        G.ellipsoids(truth, count=16, seed=2000 + i, amin=0.02 * n, amax=0.3 * n)
        y = torch.empty(G.sino_shape, dtype=torch.float32, device="cuda")
        G.project(truth, y, lib.PROJ_RS if pair == lib.PAIR_RSVI else lib.PROJ_SIDDON)
        del truth
        probs.append(dict(desc=desc, G=G, pair=pair, y=y, n=n, nphi=nphi))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p in probs:
        p["ref"] = torch.empty(p["G"].vol_shape, dtype=torch.float32, device="cuda")
        p["G"].sirt_run(p["y"], p["ref"], pair=p["pair"], iters=iters)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    print(f"[batch] one at a time: {t_seq:.2f} s", file=sys.stderr, flush=True)

    def batched(storage):
        outs = [torch.zeros(p["G"].vol_shape, dtype=torch.float32, device="cuda") for p in probs]
        srv = lib.Server(ctx, max_active=min(64, n_prob), use_graph=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ids = [srv.submit(p["desc"], p["y"], o, pair=p["pair"], iters=iters, algorithm=lib.ALGO_SIRT,
                          storage=storage) for p, o in zip(probs, outs)]
        for rid in ids:
            srv.wait(rid)
        t = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        torch.cuda.synchronize()
        return outs, t, st

    outs, t_batch, st = batched(lib.F32)
    same = all(bool(torch.equal(o, p["ref"])) for o, p in zip(outs, probs))
    print(f"[batch] batched: {t_batch:.2f} s, identical={same}", file=sys.stderr, flush=True)
    outs16, t_bf16, _ = batched(lib.BF16)
    err16 = max(float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)) for a, b in zip(outs16, outs))
    print(f"[batch] batched, BF16 copies: {t_bf16:.2f} s, max rel-L2 vs FP32 {err16:.3e}", file=sys.stderr,
          flush=True)
    upd = sum(2.0 * iters * p["n"] ** 3 * p["nphi"] for p in probs)
    for p in probs:
        p["G"].close()
    return {"problems": n_prob, "iterations": iters, "algorithm": "SIRT", "storage": "f32",
            "one_at_a_time_s": round(t_seq, 3), "batched_s": round(t_batch, 3),
            "speedup": round(t_seq / t_batch, 3), "batched_gups": round(upd / t_batch / 1e9, 2),
            "peak_active": st["peak_active"], "rounds": st["rounds"], "identical_to_one_at_a_time": same,
            "bf16_copies": {"batched_s": round(t_bf16, 3), "speedup_vs_one_at_a_time_fp32": round(t_seq / t_bf16, 3),
                            "max_rel_l2_vs_fp32_batched": float(f"{err16:.3e}"),
                            "note": "BF16 operator copies (z-pair; A-pair for RS problems) rounded from the FP32 state; VI, sums and state FP32"}}


def loop_leg(G, vol, sino, iters, upd, stream):
    """The workload's reconstruction: gradient descent with the quadratic
    regulariser (lambda 0.01) from zero on y = P(phantom), every step optimal
    (kopt = iterations + 1), for the SIDDONVI and RSVI pairs in FP32, and
    the SIDDONVI pair in its FP16 and BF16 mixed-precision forms (each
    operator reads a 16-bit copy of its FP32 input, FP32 accumulation and
    state) with the per-iteration residual against the FP32 run's."""
    import torch
    from tomogpi_amd import lib
    out = {}
    for name, pair, proj in (("siddonvi", lib.PAIR_SIDDONVI, lib.PROJ_SIDDON), ("rsvi", lib.PAIR_RSVI, lib.PROJ_RS)):
        G.project(vol, sino, proj, stream=stream)
        y = sino.clone()
        runs = [("f32", lib.F32)] + ([("f16", lib.F16), ("bf16", lib.BF16)] if pair == lib.PAIR_SIDDONVI else [])
        base = None
        for dname, dt in runs:
            x = torch.zeros_like(vol)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            hist = G.gd_run(y, x, pair=pair, iters=iters, kopt=iters + 1, lam=0.01, stream=stream, storage=dt)
            torch.cuda.synchronize()
            t_loop = time.perf_counter() - t0
            rec = {"iterations": iters, "lambda": 0.01, "kopt": iters + 1, "s": round(t_loop, 3),
                   "ms_per_iteration": round(t_loop * 1e3 / iters, 2),
                   "gups": round(3.0 * upd * iters / t_loop / 1e9, 2),
                   "residual_sq_first": hist[0]["residual_sq"], "residual_sq_last": hist[-1]["residual_sq"]}
            if base is None:
                base = (hist, x, t_loop)
                rec["note"] = ("per iteration two projections (P x, P dJ) and one VI backprojection (3 N^3 Nphi "
                               "updates), plus the Laplacian terms, norms and step rule")
            else:
                r32 = [h["residual_sq"] for h in base[0]]
                r16 = [h["residual_sq"] for h in hist]
                rel = [abs(u - v) / v for u, v in zip(r16, r32)]
                rec["residual_rel_diff_vs_f32"] = [float(f"{v:.3e}") for v in rel]
                rec["max_residual_rel_diff_vs_f32"] = float(f"{max(rel):.3e}")
                rec["volume_rel_l2_vs_f32"] = float(
                    f"{float(torch.linalg.vector_norm(x - base[1]) / torch.linalg.vector_norm(base[1])):.3e}")
                rec["speedup_vs_f32"] = round(base[2] / t_loop, 3)
            out[name if dname == "f32" else f"{name}_{dname}"] = rec
            print(f"[loop] {name} {dname}: {rec}", file=sys.stderr, flush=True)
            del x
        del y
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": f"bench.py runs on exactly one GPU; --gpus {a.gpus} is not supported "
                                   "(no multi-GPU path in this build)"}))
        sys.exit(2)
    import torch
    from tomogpi_amd import lib

    n, nphi = a.n, a.nphi
    ctx = lib.Context(0)
    desc = lib.AcquisitionDesc.scaled(n, nphi)
    G = lib.Geometry(ctx, desc)
    stream = torch.cuda.current_stream()
    vol = torch.empty(G.vol_shape, dtype=torch.float32, device="cuda")
    G.ellipsoids(vol, count=64, seed=0, amin=5.0 * n / 512, amax=120.0 * n / 512)
    sino = torch.empty(G.sino_shape, dtype=torch.float32, device="cuda")
    back = torch.empty(G.vol_shape, dtype=torch.float32, device="cuda")
    vol_host = vol.cpu().numpy()
    n_seg = G.project_work(lib.PROJ_SIDDON)
    n_samp = G.project_work(lib.PROJ_RS)

    def siddon():
        G.project(vol, sino, lib.PROJ_SIDDON, stream=stream)

    def vi():
        G.backproject(sino, back, stream=stream)

    def rs():
        G.project(vol, sino, lib.PROJ_RS, stream=stream)

    def rs_staged():
        G.project(vol, sino, lib.PROJ_RS, lib.FLAG_STAGED, stream=stream)

    def siddon_staged():
        G.project(vol, sino, lib.PROJ_SIDDON, lib.FLAG_STAGED, stream=stream)

    if a.calibration_copy:
        copy = vol + 0.0  # an element-wise kernel: reads and writes the volume once
        torch.cuda.synchronize()
        del copy
    for _ in range(a.warmup):
        siddon(); vi()
    rs()
    torch.cuda.synchronize()

    # Timed region: K steps, synchronised on both sides; per-kernel events.
    sid_ev, vi_ev = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        s0, s1, s2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        s0.record(stream); siddon(); s1.record(stream); vi(); s2.record(stream)
        sid_ev.append((s0, s1)); vi_ev.append((s1, s2))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / a.steps
    sid_ms = sum(x.elapsed_time(y) for x, y in sid_ev) / a.steps
    vi_ms = sum(x.elapsed_time(y) for x, y in vi_ev) / a.steps
    rs_ms, _ = timed(rs, max(2, min(a.steps, 3)), stream)
    rs_staged()
    rs_staged_ms, _ = timed(rs_staged, max(2, min(a.steps, 3)), stream)
    siddon_staged()
    sid_staged_ms, _ = timed(siddon_staged, max(2, min(a.steps, 3)), stream)
    siddon()  # the timed form's sinogram again (VI input of later legs)
    print(f"[bench] siddon {sid_ms:.2f} ms, vi {vi_ms:.2f} ms, rs {rs_ms:.2f} ms (LDS-staged: siddon "
          f"{sid_staged_ms:.2f} ms, rs {rs_staged_ms:.2f} ms)", file=sys.stderr, flush=True)

    upd = float(n) ** 3 * nphi
    nsino = float(G.info["nu"]) * G.info["nv"] * nphi
    b_vi = upd * 4 * 4 + float(n) ** 3 * 4
    b_sid = n_seg * 4 + nsino * 4
    b_rs = n_samp * 8 * 4 + nsino * 4
    # Three fractions per operator:
    #   effective_gather_frac  algorithmic bytes (every chord, tap or sample a DRAM read) / t / 8 TB/s: the rate
    #                          the gathers are served at, most of them from L2 / the Infinity Cache; above 1
    #                          when the caches serve more than HBM could, so it is not an HBM occupancy
    #   physical_hbm_frac      FETCH_SIZE + WRITE_SIZE bytes per launch (calibrated) / t / 8 TB/s
    #   valu_issue_frac        VALU wave-instructions per launch x cycles per instruction / (4 SIMDs x CUs x
    #                          kernel cycles), from the counters and the issue-rate probe (the kernels' bound)
    measured = load_traffic(n, nphi)
    valu = load_valu(n, nphi)
    ops = {}
    for name, ms, alg, key in (("siddon", sid_ms, b_sid, "siddon_kernel"),
                               ("vi", vi_ms, b_vi, "vi_staged_kernel"),
                               ("rs", rs_ms, b_rs, "rs_kernel"),
                               ("siddon_staged", sid_staged_ms, b_sid, "siddon_staged_kernel"),
                               ("rs_staged", rs_staged_ms, b_rs, "rs_staged_kernel")):
        rec = {"ms": round(ms, 3), "gups": round(upd / ms / 1e6, 2), "alg_bytes": alg,
               "effective_gather_frac": round(alg / (ms * 1e-3) / HBM_PEAK, 4)}
        if key in measured:
            rec["hbm_traffic_bytes"] = measured[key]
            rec["physical_hbm_frac"] = round(measured[key] / (ms * 1e-3) / HBM_PEAK, 4)
        if key in valu:
            rec["valu_issue_frac"] = valu[key]["issue_frac"]
            rec["texture_data_busy"] = valu[key].get("td_busy")
        ops[name] = rec
    ops["siddon"]["n_seg"] = n_seg
    ops["rs"]["n_samp"] = n_samp
    ops["rs"]["form"] = "per-lane gathers from the z-pair copies (rs_kernel)"
    ops["rs_staged"]["form"] = ("taps staged in LDS per window of up to 8 sample planes, copied as 16-byte "
                                "chunks (TOMO_FLAG_STAGED, rs_staged_kernel)")
    ops["siddon"]["form"] = "per-lane gathers from the z-pair copies (siddon_kernel)"
    ops["siddon_staged"]["form"] = ("quads staged in LDS per window of up to 8 slabs, copied as 16-byte chunks "
                                    "(TOMO_FLAG_STAGED, siddon_staged_kernel)")
    ops["siddon_staged"]["n_seg"] = n_seg
    ops["rs_staged"]["n_samp"] = n_samp
    # compulsory traffic: the volume and the sinogram once each
    comp = float(n) ** 3 * 4 + nsino * 4
    for name, ms in (("siddon", sid_ms), ("vi", vi_ms), ("rs", rs_ms), ("siddon_staged", sid_staged_ms),
                     ("rs_staged", rs_staged_ms)):
        ops[name]["compulsory_frac"] = round(comp / (ms * 1e-3) / HBM_PEAK, 4)
    # The roofline line is the step's dominant kernel, the Siddon projector: its algorithmic bytes (every positive
    # chord counted as one 4-byte read, plus the sinogram write) over its HIP-event time.
    dom = "vi" if vi_ms >= sid_ms else "siddon"
    dom_bytes = b_vi if dom == "vi" else b_sid
    dom_ms = vi_ms if dom == "vi" else sid_ms
    dom_kernel = "vi_staged_kernel" if dom == "vi" else "siddon_kernel"
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom_kernel,
                "achieved": round(achieved, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(achieved * 1e9 / HBM_PEAK, 4), "traffic": measured.get(dom_kernel),
                "basis": "algorithmic bytes per launch (every positive chord one 4-byte read, plus the sinogram) / "
                         "launch time: an effective gather rate, mostly served by L2 and the Infinity Cache",
                "physical_frac": ops[dom].get("physical_hbm_frac"),
                "valu_issue_frac": ops[dom].get("valu_issue_frac")}

    mixed = None
    if a.mixed_oracle_views > 0:
        mixed = mixed_leg(ctx, G, desc, vol, {"siddon": sid_ms, "vi": vi_ms, "rs": rs_ms}, stream,
                          a.mixed_oracle_views)

    loop = None
    if a.loop_iters > 0:
        loop = loop_leg(G, vol, sino, a.loop_iters, upd, stream)

    serving = None
    del vol, sino, back
    torch.cuda.empty_cache()
    mixed_c2 = None
    if a.c2_sizes.strip() and a.mixed_oracle_views > 0:
        mixed_c2 = c2_leg(ctx, [int(x) for x in a.c2_sizes.split(",")], stream, a.c2_oracle_views,
                          a.c2_oracle_views_large)
    if a.serving_requests > 0:
        names = {"recurrent": lib.SCHEDULE_RECURRENT, "direct": lib.SCHEDULE_DIRECT}
        scheds = [names[x] for x in a.serving_schedules.split(",") if x]
        serving = serving_leg(ctx, a.serving_requests, [float(x) for x in a.serving_rates.split(",")], scheds,
                              bf16_rate=a.serving_bf16_rate or None, direct_requests=a.serving_direct_requests)

    batched = None
    if a.batch_problems > 0:
        batched = batch_leg(ctx, a.batch_problems)

    base = None
    if not a.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        base = cpu_baseline(vol_host, n, nphi, a.cpu_views, threads)

    out = {
        "metric": METRIC,
        "value": round(2 * upd / (ms_step * 1e-3) / 1e9, 3),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"SIDDONVI sweep (clean Siddon projection + VI backprojection), {n}^3 volume of "
                               f"64 random ellipsoids (seed 0), {nphi} views over 360 deg, {n}x{n} detector, "
                               f"FOD {desc.fod_mm:g} mm, FDD {desc.fdd_mm:g} mm, FP32 device-resident",
                   "n": n, "nphi": nphi, "detector": [G.info["nu"], G.info["nv"]]},
        "roofline": roofline,
        "cpu_baseline": base,
        "operators": ops,
        "loop": loop,
        "mixed": mixed,
        "mixed_c2": mixed_c2,
        "serving": serving,
        "batched": batched,
    }
    # release the device before printing the line: geometry, context, queued work
    G.close()
    torch.cuda.synchronize()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
