#!/usr/bin/env python3
"""Benchmark: precompute LUT sets/s (+ p50/p99 latency) at <= 1e-3 relative
error vs the FP64 oracle, on one MI355X.

Workload (one "step"): the whole precompute of one default Earth atmosphere
(Parameters::default(), src/precompute.rs:849-935: transmittance 256x64,
irradiance 64x16, scattering 32x128x32x8, order 4) through the plain
single-atmosphere HIP pipeline in precision mode, re-launched into the same
resident buffers (the reference's bench likewise re-submits one recorded
command buffer, benches/precompute.rs:138-148).  Inputs (the parameters) are
resident; no host transfer is inside the timed region.

Prints ONE JSON line.  Extra fields: latency percentiles of individual
build requests, FP32-mode and binary16-storage throughput, the reference's
own benchmark configuration (scattering 8x32x8x2, median of 100
submit-and-wait repetitions), sky Mpixel/s (64 frames of 1920x1080 rays,
camera altitude U[0,10] km, sun zenith U[0,108] deg), the roofline of the
dominant kernel (scattering density) and a CPU baseline (the FP32
restatement of the reference passes on the host cores).
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

# HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default);
# streams sharing a queue run in FIFO order, so a render could wait behind a
# whole precompute.  The serving run uses 6 streams: give each its own
# queue.  Must be set before the HIP runtime initialises.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

from fuzzyblue_amd import (  # noqa: E402
    Atmosphere,
    Builder,
    DeviceBuffer,
    Parameters,
    Renderer,
    Stream,
    abi,
    costmodel,
    synchronize,
)
from fuzzyblue_amd.lib import check, load  # noqa: E402
from fuzzyblue_amd.params import RESOLUTIONS  # noqa: E402

METRIC = "LUT-sets/s + p50/p99 latency at ≤1e-3 rel err vs FP64, 1×MI355X"
PROFILE_STEPS = 10
# rocprofv3 --pmc summary of the density kernel on this workload (FETCH_SIZE
# and WRITE_SIZE per dispatch, calibrated against a plain copy).
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "scattering_density_default_earth_hbm.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class Timer:
    def __init__(self):
        import ctypes as C

        self.C = C
        self.a = C.c_void_p()
        self.b = C.c_void_p()
        check(load().fb_event_create(C.byref(self.a)))
        check(load().fb_event_create(C.byref(self.b)))

    def start(self, stream):
        check(load().fb_event_record(self.a, stream.handle if stream else None))

    def stop(self, stream):
        check(load().fb_event_record(self.b, stream.handle if stream else None))

    def ms(self):
        v = self.C.c_float()
        check(load().fb_event_elapsed_ms(self.a, self.b, self.C.byref(v)))
        return v.value


def measure_throughput(b, stream, p, steps, warmup):
    pa = Atmosphere.build(b, stream, p)
    pa.wait()
    for _ in range(warmup):
        pa.relaunch(stream)
    synchronize()
    t = Timer()
    t.start(stream)
    for _ in range(steps):
        pa.relaunch(stream)
    t.stop(stream)
    synchronize()
    ms = t.ms()
    return pa, ms


def measure_latency(b, stream, p, n):
    """Submit -> complete wall time of individual build requests (allocation
    from the builder's cached pool included)."""
    lat = []
    for _ in range(n):
        t0 = time.perf_counter()
        pa = Atmosphere.build(b, stream, p)
        pa.wait()
        lat.append((time.perf_counter() - t0) * 1e3)
        pa.close()
    lat = np.array(lat)
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


def measure_reference_bench(b, stream, reps=100, warmup=10):
    """The reference's own benchmark configuration (benches/precompute.rs:
    123-147: default parameters with scattering 8x32x8x2, one recorded
    command buffer submitted and waited for per iteration): the captured
    graph relaunched and waited for per iteration, 10 warm-up + 100 timed
    repetitions, median wall time (host clock around relaunch -> complete)."""
    out = {"scattering_r_mu_mus_nu": [8, 32, 8, 2], "order": 4, "reps": reps}
    base = Parameters().replace(scattering_r_size=8, scattering_mu_size=32,
                                scattering_mu_s_size=8, scattering_nu_size=2)
    for name, flags in (("fp32", 0), ("precise", abi.FB_FLAG_PRECISE)):
        pa = Atmosphere.build(b, stream, base.replace(flags=flags))
        pa.wait()
        for _ in range(warmup):
            pa.relaunch(stream)
            stream.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            pa.relaunch(stream)
            stream.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        pa.close()
        med = float(np.median(ts))
        out[name] = {"median_ms": med, "lut_sets_s": 1e3 / med}
    return out


def density_traffic():
    """HBM bytes per density launch (average over the order-2 and the two
    order>=3 launches of one LUT set) from the committed rocprofv3 --pmc
    summary; None if it is absent."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split()[1])
    except OSError:
        pass
    return None


def max_err_vs_fixture(atm, floor=1e-6):
    path = os.path.join(ROOT, "tests", "golden", "earth_default_fp64.npz")
    if not os.path.exists(path):
        return None
    g = np.load(path)

    def rel(x, y):
        x = x.astype(np.float64)
        y = y.astype(np.float64)
        return float((np.abs(x - y) / np.maximum(np.abs(y), floor * np.abs(y).max())).max())

    T = atm.download("transmittance")[..., :3]
    E = atm.download("irradiance")[..., :3]
    S = atm.download("scattering").reshape(-1, 4)[g["S_idx"]]
    return {"transmittance": rel(T, g["T"]), "irradiance": rel(E, g["E"]),
            "scattering": rel(S[:, :3], g["S"][:, :3])}


def measure_render(b, stream, atm, frames=64, w=1920, h=1080, reps=3):
    rng = np.random.default_rng(2)
    alt = rng.uniform(0.0, 10.0, frames).astype(np.float32)
    zen = np.radians(rng.uniform(0.0, 108.0, frames)).astype(np.float32)
    d_alt = DeviceBuffer.from_array(alt)
    d_zen = DeviceBuffer.from_array(zen)
    px = w * h
    rad = DeviceBuffer(frames * px * 16)
    tr = DeviceBuffer(frames * px * 16)
    r = Renderer(b)
    r.frames(stream, atm, d_alt, d_zen, frames, px, 2, rad, tr)
    synchronize()
    t = Timer()
    t.start(stream)
    for _ in range(reps):
        r.frames(stream, atm, d_alt, d_zen, frames, px, 2, rad, tr)
    t.stop(stream)
    ms = t.ms() / reps
    for buf in (rad, tr, d_alt, d_zen):
        buf.close()
    return frames * px / (ms * 1e-3) / 1e6, ms


def measure_batch(b, stream, n=256, seed=4):
    """The seeded mixed set (workloads.mixed_batch: sizes half/default/double,
    orders 2..8) as one batch (one launch per pass, graph replay) against the
    plain pipeline (each atmosphere's own schedule), both timed with HIP
    events on the same stream."""
    from fuzzyblue_amd.workloads import mixed_batch

    ps = mixed_batch(n, seed, precise=True)
    # One autotune per LUT-size class of the mix (block sizes only; the
    # LUTs are bit-identical for every choice).
    for res in RESOLUTIONS.values():
        b.autotune(Parameters().replace(**res, flags=abi.FB_FLAG_PRECISE), stream)
    work = sum(costmodel.lut_set_work(p).roof_seconds() for p in ps)
    pend = Atmosphere.build_batch(b, stream, ps)
    for pa in pend:
        pa.wait()
    t = Timer()
    reps = 2
    t.start(stream)
    for _ in range(reps):
        pend[0].relaunch(stream)
    t.stop(stream)
    ms_batch = t.ms() / reps
    for pa in pend:
        pa.close()
    ms_plain = 0.0
    for p in ps:
        pa = Atmosphere.build(b, stream, p)
        pa.wait()
        t.start(stream)
        pa.relaunch(stream)
        t.stop(stream)
        ms_plain += t.ms()
        pa.close()
    return {
        "atmospheres": n,
        "seed": seed,
        "batched_lut_sets_s": n / (ms_batch * 1e-3),
        "plain_lut_sets_s": n / (ms_plain * 1e-3),
        "speedup_vs_plain": ms_plain / ms_batch,
        "batched_ms": ms_batch,
        "plain_ms": ms_plain,
        "roofline_ms": work * 1e3,
        "roofline_frac": work * 1e3 / ms_batch,
    }


def _rel(x, y, scale):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return np.abs(x - y) / np.maximum(np.abs(y), 1e-6 * scale)


def measure_precision(b, stream):
    """The precision workload (Earth, LUTs at twice the default resolution,
    order 8) in precision mode against its FP64-oracle fixture: max / mean
    relative error per LUT, and per scattering order, for texels whose view
    cosine is within 1e-3 of the horizon and for the rest."""
    path = os.path.join(ROOT, "tests", "golden", "earth_double_order8_fp64.npz")
    if not os.path.exists(path):
        return None
    g = np.load(path)
    band = g["S_band"]

    def split(e, m):
        return {"horizon_max": float(e[m].max()), "horizon_mean": float(e[m].mean()),
                "other_max": float(e[~m].max()), "other_mean": float(e[~m].mean())}

    def build(order):
        pa = Atmosphere.build(b, stream, Parameters().replace(
            **RESOLUTIONS["double"], order=order, flags=abi.FB_FLAG_PRECISE))
        pa.wait()
        return pa

    pa = build(8)
    a = pa.atmosphere()
    T = a.download("transmittance")[..., :3]
    E = a.download("irradiance")[..., :3]
    S = a.download("scattering").reshape(-1, 4)[g["S_idx"]]
    pa.close()
    out = {
        "workload": "Earth, transmittance 512x128, irradiance 128x32, scattering 64x256x64x16, "
                    "order 8, precision mode",
        "transmittance_max": float(_rel(T, g["T"], np.abs(g["T"]).max()).max()),
        "irradiance_max": float(_rel(E, g["E"], np.abs(g["E"]).max()).max()),
        "scattering": split(_rel(S[:, :3], g["S"][:, :3],
                                 np.abs(g["S"][:, :3]).max(axis=0)).max(axis=1), band),
        "per_order": [],
    }
    sel, oband = g["S_idx"][g["O_sel"]], band[g["O_sel"]]
    for order in range(1, 9):
        pa = build(order)
        which = "delta_rayleigh" if order == 1 else "delta_multiple_scattering"
        s_o = pa.download_scratch(which).reshape(-1, 4)[sel, :3]
        e_o = pa.download_scratch("delta_irradiance")[..., :3]
        pa.close()
        ref_s, ref_e = g["orders_S"][order - 1], g["orders_E"][order - 1]
        out["per_order"].append({
            "order": order,
            "scattering": split(_rel(s_o, ref_s, np.abs(ref_s).max(axis=0)).max(axis=1), oband),
            "irradiance_max": float(_rel(e_o, ref_e, np.abs(ref_e).max()).max())})
    return out


def measure_mixed_accuracy(b, stream):
    """Eight random atmospheres drawn like the mixed set (half resolution,
    orders 2-8) as one precision-mode batch against their FP64-oracle LUTs:
    the worst member's max relative error per LUT."""
    import json as _json

    path = os.path.join(ROOT, "tests", "golden", "mixed_half_fp64.npz")
    if not os.path.exists(path):
        return None
    g = np.load(path)
    ps = [Parameters.from_dict(d) for d in _json.loads(str(g["params"]))]
    pend = Atmosphere.build_batch(b, stream, ps)
    worst = {"transmittance": 0.0, "irradiance": 0.0, "scattering": 0.0}
    for i, pa in enumerate(pend):
        pa.wait()
        a = pa.atmosphere()
        for name, key, x in (("transmittance", "T", a.download("transmittance")[..., :3]),
                             ("irradiance", "E", a.download("irradiance")[..., :3]),
                             ("scattering", "S",
                              a.download("scattering").reshape(-1, 4)[g["S_idx"]])):
            y = g[f"{key}{i}"]
            scale = np.abs(y).reshape(-1, y.shape[-1]).max(axis=0)
            worst[name] = max(worst[name], float(_rel(x, y, scale).max()))
    for pa in pend:
        pa.close()
    return {"atmospheres": len(ps), "max_rel_err": worst}


def measure_served_mixed(b, n=256, seed=4, budget_gb=200.0):
    """The whole seeded mixed set submitted to the serving loop at once:
    LUT sets per second from first submission to the last completion."""
    from fuzzyblue_amd.serving import Server
    from fuzzyblue_amd.workloads import mixed_batch

    ps = mixed_batch(n, seed, precise=True)
    srv = Server(b, int(budget_gb * 1e9), max_inflight=64, n_streams=4)
    t0 = time.perf_counter()
    for p in ps:
        srv.submit_precompute(p)
    srv.drain()
    wall = time.perf_counter() - t0
    infos = srv.poll()
    lat = np.array([1e3 * (i.t_done - i.t_submit) for i in infos if i.status == abi.FB_OK])
    failed = sum(1 for i in infos if i.status != abi.FB_OK)
    srv.close()
    return {"atmospheres": n, "lut_sets_s": float(lat.size / wall), "wall_s": wall,
            "failed": int(failed), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99))}


def measure_serving(b, rates, duration, budget_gb=200.0):
    """Poisson arrivals (workloads.arrivals: 20% precompute drawn like the
    mixed set, 80% sky renders of 640x360 / 1920x1080 / 3840x2160 rays
    against a random resident LUT set) replayed in real time into the
    native serving loop with an HBM budget; p50/p99 per request type."""
    from fuzzyblue_amd.serving import Server, replay
    from fuzzyblue_amd.workloads import arrivals

    out = {}
    for rate in rates:
        reqs = arrivals(rate, duration, seed=5 + int(rate))
        srv = Server(b, int(budget_gb * 1e9), max_inflight=16, n_streams=4)
        res = replay(srv, reqs, seed=int(rate))
        srv.close()
        out[f"lambda_{rate:g}"] = res
        log(f"[bench] serving lambda={rate}: {res}")
    return {"duration_s": duration, "hbm_budget_gb": budget_gb, "runs": out}


def cpu_baseline(p):
    """FP32 restatement of the reference passes on the host cores, timed on a
    bounded sample: full transmittance / direct / indirect passes, and texel
    subsets of single scattering, density and multiple scattering scaled to
    the full LUT."""
    from oracle import Oracle

    o = Oracle("f32")
    nt, ni, ns = costmodel.texel_counts(p)
    rng = np.random.default_rng(0)
    t0 = time.perf_counter()
    T = o.pass_transmittance(p)
    t_T = time.perf_counter() - t0
    t0 = time.perf_counter()
    dE = o.pass_direct_irradiance(p, T)
    t_D = time.perf_counter() - t0
    n_ss = 65536
    idx = rng.choice(ns, n_ss, replace=False)
    t0 = time.perf_counter()
    o.single_texels(p, T, idx)
    t_S = (time.perf_counter() - t0) * ns / n_ss
    # Inputs for the later passes: the restatement's own single scattering.
    dR, dM, S = o.pass_single_scattering(p, T)
    zeros = np.zeros_like(dR)
    n_j = 8192
    jdx = rng.choice(ns, n_j, replace=False)
    t0 = time.perf_counter()
    o.density_texels(p, 2, T, dR, dM, zeros, dE, jdx)
    t_J2 = (time.perf_counter() - t0) * ns / n_j
    t0 = time.perf_counter()
    o.density_texels(p, 3, T, dR, dM, dR, dE, jdx)
    t_J3 = (time.perf_counter() - t0) * ns / n_j
    t0 = time.perf_counter()
    o.pass_indirect(p, 1, dR, dM, zeros, np.zeros_like(dE))
    t_I = time.perf_counter() - t0
    n_m = 65536
    mdx = rng.choice(ns, n_m, replace=False)
    t0 = time.perf_counter()
    o.multiple_texels(p, T, dR, mdx)
    t_M = (time.perf_counter() - t0) * ns / n_m
    n_orders = p.order - 1
    t_set = t_T + t_D + t_S + t_J2 + (n_orders - 1) * t_J3 + n_orders * (t_I + t_M)
    cores = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", cores))
    return {
        "value": 1.0 / t_set,
        "unit": "LUT-sets/s",
        "cores": min(cores, threads),
        "kind": "port",
        "sample": (
            "FP32 restatement (oracle/atmo_oracle.c, OpenMP) of the reference passes for the "
            "default Earth LUT set, order 4: full transmittance/direct/indirect passes; "
            f"{n_ss} single-scattering, {n_j} density (orders 2 and 3) and {n_m} "
            "multiple-scattering texels, scaled to the full 1,048,576-texel LUT"
        ),
        "seconds_per_set": t_set,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--latency-requests", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-render", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-reference-config", action="store_true",
                    help="skip the reference's own benchmark configuration (8x32x8x2)")
    ap.add_argument("--serving-rates", default="50",
                    help="comma-separated Poisson rates (req/s) for the serving run; "
                         "empty to skip")
    ap.add_argument("--serving-seconds", type=float, default=10.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 5 device-to-device copies of BYTES (the known-size "
                         "kernel that calibrates FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        src = DeviceBuffer(args.copy_calibration)
        dst = DeviceBuffer(args.copy_calibration)
        src.zero()
        for _ in range(5):
            check(load().fb_device_copy(dst.ptr, src.ptr, args.copy_calibration, None))
        synchronize()
        print(json.dumps({"copy_bytes": args.copy_calibration, "copies": 5}))
        return
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: the workload is one atmosphere's LUT set on one "
            "MI355X; runs on several GPUs are not supported (atmospheres would shard "
            "across GPUs with no exchange, one whole LUT set per GPU)", file=sys.stderr)
        sys.exit(2)

    b = Builder(0)
    stream = Stream()
    p = Parameters().replace(flags=abi.FB_FLAG_PRECISE)
    log("[bench] autotune")
    tuning = b.autotune(p, stream)
    tuning32 = b.autotune(p.replace(flags=0), stream)
    log(f"[bench] tuning: {tuning}")
    log("[bench] precision-mode throughput")
    pa, ms = measure_throughput(b, stream, p, args.steps, args.warmup)
    value = args.steps / (ms * 1e-3)
    ms_per_step = ms / args.steps
    log(f"[bench] {value:.1f} LUT-sets/s ({ms_per_step:.3f} ms/set)")

    # Per-launch profile of the same schedule (HIP events on the launch
    # stream around each launch), over PROFILE_STEPS consecutive steps run
    # right after the timed region.
    profs = [pa.profile(stream) for _ in range(PROFILE_STEPS)]
    prof = [(name, order, t / PROFILE_STEPS) for name, order, t in profs[0]]
    for other in profs[1:]:
        prof = [(name, order, t + u / PROFILE_STEPS)
                for (name, order, t), (_, _, u) in zip(prof, other)]
    passes = {}
    for name, order, t in prof:
        passes.setdefault(name, []).append(t)
    dens = [(order, t) for name, order, t in prof if name == "scattering_density"]
    dens_flop = sum(costmodel.pass_work(p, "scattering_density", o).flop for o, _ in dens)
    dens_ms = sum(t for _, t in dens)
    achieved = dens_flop / (dens_ms * 1e-3) / 1e12
    dens_bytes = sum(costmodel.pass_work(p, "scattering_density", o).bytes for o, _ in dens)

    err = max_err_vs_fixture(pa.atmosphere())
    log(f"[bench] max rel err vs FP64 oracle: {err}")

    log("[bench] latency")
    p50, p99 = measure_latency(b, stream, p, args.latency_requests)

    log("[bench] fp32-mode throughput")
    pa32, ms32 = measure_throughput(b, stream, p.replace(flags=0), args.steps, args.warmup)
    pa32.close()

    # The reference's storage formats (binary16 scattering and 3D scratch),
    # precision-mode arithmetic, default block shapes.
    log("[bench] binary16-storage throughput")
    p16 = p.replace(flags=abi.FB_FLAG_PRECISE | abi.FB_FLAG_F16_STORAGE)
    pa16, ms16 = measure_throughput(b, stream, p16, args.steps, args.warmup)
    f16_storage = {"lut_sets_s": args.steps / (ms16 * 1e-3),
                   "max_rel_err_vs_fp64": max_err_vs_fixture(pa16.atmosphere()),
                   # values below the binary16 range flush to 0 (floor 1e-6
                   # x max counts them at 1.0); with a floor of 1e-3 x max:
                   "max_rel_err_vs_fp64_floor_1e-3": max_err_vs_fixture(pa16.atmosphere(),
                                                                         1e-3),
                   "scattering_lut_bytes": pa16.atmosphere().lut("scattering").bytes}
    pa16.close()

    ref_bench = None
    if not args.no_reference_config:
        log("[bench] the reference's benchmark configuration")
        ref_bench = measure_reference_bench(b, stream)
        log(f"[bench] reference-bench sizes: {ref_bench}")

    sky = None
    if not args.no_render:
        log("[bench] render")
        mpx, rms = measure_render(b, stream, pa.atmosphere())
        sky = {"mpixel_s": mpx, "ms_per_64_frames": rms, "frames": 64, "width": 1920,
               "height": 1080}

    batch = None
    if not args.no_batch:
        log("[bench] mixed batch of 256 atmospheres")
        batch = measure_batch(b, stream)
        batch["served"] = measure_served_mixed(b)
        batch["served_speedup_vs_plain"] = (batch["served"]["lut_sets_s"] /
                                            batch["plain_lut_sets_s"])
        log(f"[bench] batch: {batch}")

    precision = mixed = None
    if not args.no_precision:
        log("[bench] precision workload and mixed-set accuracy")
        precision = measure_precision(b, stream)
        mixed = measure_mixed_accuracy(b, stream)

    serving = None
    if args.serving_rates:
        rates = [float(x) for x in args.serving_rates.split(",") if x]
        log(f"[bench] serving {rates} req/s for {args.serving_seconds} s each")
        serving = measure_serving(b, rates, args.serving_seconds)

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline(Parameters())

    work = costmodel.lut_set_work(p)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "LUT-sets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain single-atmosphere precompute of the default Earth atmosphere "
                        "(order 4; LUTs 256x64, 64x16, 32x128x32x8), precision mode (FP64 "
                        "geometry, FP32 radiance)",
            "order": p.order,
            "transmittance": list(p.transmittance_extent()),
            "irradiance": list(p.irradiance_extent()),
            "scattering_r_mu_mus_nu": [p.scattering_r_size, p.scattering_mu_size,
                                       p.scattering_mu_s_size, p.scattering_nu_size],
            "precision": "precise",
        },
        "latency_ms": {"p50": p50, "p99": p99, "requests": args.latency_requests},
        "tuning": {"precise": tuning, "fp32": tuning32},
        "max_rel_err_vs_fp64": err,
        "fp32_mode_lut_sets_s": args.steps / (ms32 * 1e-3),
        "f16_storage": f16_storage,
        "reference_bench_config": ref_bench,
        "sky": sky,
        "batch_mixed256": batch,
        "serving": serving,
        "precision_double_order8": precision,
        "mixed_half_vs_fp64": mixed,
        "roofline": {
            "bound": "valu",
            "kernel": "scattering_density",
            "achieved": achieved,
            "peak": costmodel.PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / costmodel.PEAK_FP32_TFLOPS,
            "traffic": density_traffic(),
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE x2.0 + WRITE_SIZE x1.0 per dispatch, "
                              "profiles/scattering_density_default_earth_hbm.txt",
            "algorithmic_bytes_per_launch": dens_bytes / max(len(dens), 1),
            "avg_launch_ms": dens_ms / max(len(dens), 1),
            "launches": len(dens),
            "profile_steps": PROFILE_STEPS,
        },
        "lut_set_roofline_ms": work.roof_seconds() * 1e3,
        "passes_ms": {k: float(np.sum(v)) for k, v in passes.items()},
        "cpu_baseline": cpu,
    }
    pa.close()
    stream.close()
    b.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
