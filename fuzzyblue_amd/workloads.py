"""Seeded synthetic workloads.

mixed_batch(): many atmospheres of uneven size and order — bottom radius
U[2000, 10000] km, thickness U[20, 200] km, Rayleigh RGB log-U[1e-4, 1e-1]
/km with scale height U[2, 20] km, grey Mie log-U[1e-4, 1e-1] /km
(extinction = scattering / 0.9, the Earth ratio of src/precompute.rs:909-910)
with scale height U[0.5, 5] km and g U[0.6, 0.9], the Earth ozone tent
(src/precompute.rs:912-929) with probability 1/2 and extinction scaled by
U[0, 2], grey ground albedo U[0, 0.5], LUT resolution uniform over
{half, default, double}, scattering order uniform over {2..8}; solar
irradiance, sun radius and mu_s_min keep their defaults.

arrivals(): a Poisson request stream for the serving loop: 20% precompute
requests (drawn as mixed_batch), 80% sky renders of {640x360, 1920x1080,
3840x2160} rays against a random resident LUT set, with camera altitude
U[0, 10] km and sun zenith U[0, 108] deg.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .params import RESOLUTIONS, DensityProfile, DensityProfileLayer, Parameters

RES_NAMES = ("half", "default", "double")


def _log_uniform(rng, lo, hi, size=None):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size))


def random_atmosphere(rng, precise: bool = True, resolution: Optional[str] = None,
                      order: Optional[int] = None) -> Parameters:
    res = resolution or RES_NAMES[int(rng.integers(0, 3))]
    p = Parameters().replace(**RESOLUTIONS[res])
    p.order = int(order if order is not None else rng.integers(2, 9))
    p.bottom_radius = float(rng.uniform(2000.0, 10000.0))
    p.top_radius = p.bottom_radius + float(rng.uniform(20.0, 200.0))
    p.rayleigh_scattering = [float(x) for x in _log_uniform(rng, 1e-4, 1e-1, 3)]
    hr = float(rng.uniform(2.0, 20.0))
    p.rayleigh_density = DensityProfile(
        [DensityProfileLayer(), DensityProfileLayer(0.0, 1.0, -1.0 / hr, 0.0, 0.0)])
    m = float(_log_uniform(rng, 1e-4, 1e-1))
    p.mie_scattering = [m, m, m]
    p.mie_extinction = [m / 0.9] * 3
    hm = float(rng.uniform(0.5, 5.0))
    p.mie_density = DensityProfile(
        [DensityProfileLayer(), DensityProfileLayer(0.0, 1.0, -1.0 / hm, 0.0, 0.0)])
    p.mie_phase_function_g = float(rng.uniform(0.6, 0.9))
    if rng.uniform() < 0.5:
        k = float(rng.uniform(0.0, 2.0))
        p.absorption_extinction = [6.5e-4 * k, 1.881e-3 * k, 8.5e-5 * k]
    else:
        p.absorption_extinction = [0.0, 0.0, 0.0]
    a = float(rng.uniform(0.0, 0.5))
    p.ground_albedo = [a, a, a]
    p.flags = 1 if precise else 0
    return p


def mixed_batch(n: int = 256, seed: int = 4, precise: bool = True) -> List[Parameters]:
    rng = np.random.default_rng(seed)
    return [random_atmosphere(rng, precise) for _ in range(n)]


RENDER_SIZES = ((640, 360), (1920, 1080), (3840, 2160))


@dataclass
class Request:
    t: float                 # arrival time, seconds from the start
    kind: str                # "precompute" | "render"
    params: Optional[Parameters] = None
    width: int = 0
    height: int = 0
    altitude_km: float = 0.0
    sun_zenith: float = 0.0
    seed: int = 0


def arrivals(rate: float, duration: float, seed: int, precise: bool = True,
             precompute_fraction: float = 0.2) -> List[Request]:
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        if rng.uniform() < precompute_fraction:
            out.append(Request(t, "precompute", params=random_atmosphere(rng, precise)))
        else:
            w, h = RENDER_SIZES[int(rng.integers(0, 3))]
            out.append(Request(t, "render", width=w, height=h,
                               altitude_km=float(rng.uniform(0.0, 10.0)),
                               sun_zenith=float(math.radians(rng.uniform(0.0, 108.0))),
                               seed=int(rng.integers(0, 2 ** 31))))
    return out
