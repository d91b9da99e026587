"""Python side of the serving loop (fb_server_* in include/fuzzyblue_hip.h).

``Server`` wraps the native request loop; ``replay()`` drives it with a
timed arrival stream (workloads.arrivals) in real time from one client
thread and reports per-type p50/p99 latency (submit -> completion observed
by the server) and throughput.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import Dict, List, Optional

import numpy as np

from . import abi
from .api import Builder
from .lib import check, load
from .params import Parameters
from .workloads import Request


class Server:
    def __init__(self, builder: Builder, hbm_budget_bytes: int, max_inflight: int = 8,
                 n_streams: int = 4, keep_outputs: bool = False, precompute_share_pct: int = 0,
                 admit_min_free: int = 0):
        cfg = abi.fb_server_config(int(hbm_budget_bytes), int(max_inflight), int(n_streams),
                                   1 if keep_outputs else 0, int(precompute_share_pct),
                                   int(admit_min_free))
        self.handle = C.c_void_p()
        check(load().fb_server_create(builder.handle, C.byref(cfg), C.byref(self.handle)))
        self.builder = builder

    def submit_precompute(self, p: Parameters):
        rid, lid = C.c_uint64(), C.c_uint64()
        c = p.to_c()
        check(load().fb_server_submit_precompute(self.handle, C.byref(c), C.byref(rid),
                                                 C.byref(lid)))
        return rid.value, lid.value

    def submit_precomputes(self, ps):
        """Submit many precomputes in one call (fb_server_submit_precomputes):
        [(request_id, lut_id)]."""
        n = len(ps)
        arr = (abi.fb_parameters * n)(*[p.to_c() for p in ps])
        rids, lids = (C.c_uint64 * n)(), (C.c_uint64 * n)()
        check(load().fb_server_submit_precomputes(self.handle, arr, n, rids, lids))
        return list(zip(rids, lids))

    def submit_render(self, lut_id: int, altitude_km: float, sun_zenith: float, width: int,
                      height: int, seed: int) -> int:
        rid = C.c_uint64()
        check(load().fb_server_submit_render(self.handle, lut_id, altitude_km, sun_zenith, width,
                                             height, seed, C.byref(rid)))
        return rid.value

    def poll(self, cap: int = 4096) -> List[abi.fb_request_info]:
        arr = (abi.fb_request_info * cap)()
        n = C.c_size_t()
        check(load().fb_server_poll(self.handle, arr, cap, C.byref(n)))
        return [arr[i] for i in range(n.value)]

    def drain(self):
        check(load().fb_server_drain(self.handle))

    def resident(self, cap: int = 4096) -> List[int]:
        arr = (C.c_uint64 * cap)()
        n = C.c_size_t()
        check(load().fb_server_resident(self.handle, arr, cap, C.byref(n)))
        return [arr[i] for i in range(n.value)]

    def download(self, lut_id: int, which: int, shape) -> np.ndarray:
        out = np.empty(shape, np.float32)
        check(load().fb_server_download(self.handle, lut_id, which,
                                        out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def render_result(self, request_id: int, width: int, height: int) -> np.ndarray:
        out = np.empty((height, width, 4), np.float32)
        check(load().fb_server_render_result(self.handle, request_id,
                                             out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def stats(self) -> abi.fb_server_stats:
        s = abi.fb_server_stats()
        check(load().fb_server_get_stats(self.handle, C.byref(s)))
        return s

    def close(self):
        if self.handle:
            load().fb_server_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def replay(server: Server, requests: List[Request], warm_luts: Optional[List[int]] = None,
           seed: int = 0) -> Dict:
    """Submit `requests` at their arrival times; renders target a uniformly
    random resident LUT set (or, while none is resident yet, the newest
    submitted one).  Returns latency percentiles per type and throughput."""
    rng = np.random.default_rng(seed)
    kinds: Dict[int, str] = {}
    pixels: Dict[int, int] = {}
    submitted_luts = list(warm_luts or [])
    infos = []
    t0 = time.perf_counter()
    for r in requests:
        delay = r.t - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        if r.kind == "precompute":
            rid, lid = server.submit_precompute(r.params)
            submitted_luts.append(lid)
            kinds[rid] = "precompute"
        else:
            res = server.resident()
            pool = res if res else submitted_luts[-1:]
            if not pool:
                continue
            lut = int(pool[int(rng.integers(0, len(pool)))])
            rid = server.submit_render(lut, r.altitude_km, r.sun_zenith, r.width, r.height,
                                       r.seed)
            kinds[rid] = "render"
            pixels[rid] = r.width * r.height
        infos.extend(server.poll())
    server.drain()
    infos.extend(server.poll())
    wall = time.perf_counter() - t0
    out: Dict = {"wall_s": wall, "requests": len(kinds)}
    for kind in ("precompute", "render"):
        lat = np.array([1e3 * (i.t_done - i.t_submit) for i in infos
                        if kinds.get(i.id) == kind and i.status == abi.FB_OK])
        failed = sum(1 for i in infos if kinds.get(i.id) == kind and i.status != abi.FB_OK)
        entry = {"completed": int(lat.size), "failed": int(failed)}
        if lat.size:
            ok = [i for i in infos if kinds.get(i.id) == kind and i.status == abi.FB_OK]
            wait = np.array([1e3 * (i.t_start - i.t_submit) for i in ok])
            run = np.array([1e3 * (i.t_done - i.t_start) for i in ok])
            entry.update(p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                         mean_ms=float(lat.mean()), per_s=float(lat.size / wall),
                         # split: queued (submit -> admitted) and on device
                         # (admitted -> completion observed)
                         queue_p99_ms=float(np.percentile(wait, 99)),
                         device_p99_ms=float(np.percentile(run, 99)))
            if kind == "render":
                gpu = np.array([i.gpu_ms for i in ok])
                gwait = np.array([i.gpu_wait_ms for i in ok])
                entry.update(gpu_p50_ms=float(np.percentile(gpu, 50)),
                             gpu_p99_ms=float(np.percentile(gpu, 99)),
                             gpu_wait_p99_ms=float(np.percentile(gwait, 99)))
        out[kind] = entry
    done_px = sum(pixels[i.id] for i in infos if i.id in pixels and i.status == abi.FB_OK)
    out["render_mpixel_s"] = done_px / wall / 1e6
    st = server.stats()
    out["evictions"] = int(st.evictions)
    out["hbm_peak_gb"] = st.hbm_peak / 1e9
    out["worker_step_max_ms"] = {"retire": float(st.step_max_ms[0]),
                                 "admit_renders": float(st.step_max_ms[1]),
                                 "admit_precomputes": float(st.step_max_ms[2])}
    out["worker_slow_steps"] = int(st.slow_steps)
    return out
