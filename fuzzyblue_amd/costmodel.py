"""Algorithmic work of the precompute passes (roofline accounting, batch
planning and autotuning).

Per-sample costs count FP32 FLOPs (an FMA = 2), transcendental operations
(exp, log, sqrt, rsqrt, rcp, sin, cos) and bytes, with loop-invariant terms
hoisted and multilinear interpolation counted at its separable minimum.
Samples per texel follow the reference's quadratures:
  transmittance      501   (transmittance.comp:14, trapezoid of 500 intervals)
  direct irradiance    1   (direct_irradiance.comp:10-27)
  single scattering   51   (single_scattering.comp:42)
  scattering density 512   (16 x 32 directions, scattering_density.comp:34)
  indirect irradiance 1024 (16 x 64 directions, indirect_irradiance.comp:20)
  multiple scattering 51   (multiple_scattering.comp:21)
"""
from __future__ import annotations

from dataclasses import dataclass

# (samples per texel, FLOP per sample, transcendental per sample)
PASS_COST = {
    "transmittance": (501, 35, 4),
    "direct_irradiance": (1, 60, 3),
    "single_scattering": (51, 170, 13),
    "scattering_density_2": (512, 118, 5),
    "scattering_density_n": (512, 90, 3),
    "indirect_irradiance_1": (1024, 53, 2),
    "indirect_irradiance_n": (1024, 25, 0),
    "multiple_scattering": (51, 254, 10),
}

TEXEL_BYTES = 16  # float4

# MI355X vendor peaks (dense, no sparsity).
PEAK_FP32_TFLOPS = 157.3
PEAK_TRANS_TOPS = 9.83
PEAK_HBM_GBS = 8000.0


@dataclass
class Work:
    flop: float = 0.0
    trans: float = 0.0
    bytes: float = 0.0

    def __add__(self, o: "Work") -> "Work":
        return Work(self.flop + o.flop, self.trans + o.trans, self.bytes + o.bytes)

    def roof_seconds(self) -> float:
        return max(
            self.flop / (PEAK_FP32_TFLOPS * 1e12),
            self.trans / (PEAK_TRANS_TOPS * 1e12),
            self.bytes / (PEAK_HBM_GBS * 1e9),
        )


def texel_counts(p):
    nt = p.transmittance_mu_size * p.transmittance_r_size
    ni = p.irradiance_mu_s_size * p.irradiance_r_size
    ns = p.scattering_nu_size * p.scattering_mu_s_size * p.scattering_mu_size * p.scattering_r_size
    return nt, ni, ns


def pass_work(p, name: str, order: int = 2) -> Work:
    """Work of one launch of pass `name` (fb_pass names) at scattering order
    `order` (the density launch of order n, the indirect launch of order n-1,
    the multiple launch of order n)."""
    nt, ni, ns = texel_counts(p)
    b = TEXEL_BYTES
    if name == "transmittance":
        s, f, x = PASS_COST["transmittance"]
        return Work(nt * s * f, nt * s * x, nt * b)
    if name == "direct_irradiance":
        s, f, x = PASS_COST["direct_irradiance"]
        return Work(ni * s * f, ni * s * x, ni * b + nt * b)
    if name == "single_scattering":
        s, f, x = PASS_COST["single_scattering"]
        return Work(ns * s * f, ns * s * x, 3 * ns * b + nt * b)
    if name == "clear_irradiance":
        return Work(0, 0, ni * b)
    if name == "geometry":
        # per-set tables: the density pass's taps and ground terms (33 texels
        # per r, NU + 2 per mu_s) and the 51 row samples of single (5
        # texels each) and multiple (4 texels each) scattering
        r, mus, nu = p.scattering_r_size, p.scattering_mu_s_size, p.scattering_nu_size
        rows = p.scattering_r_size * p.scattering_mu_size
        return Work(0, 0, (r * 33 + mus * (nu + 2) + rows * 51 * 9) * b)
    if name == "scattering_density":
        key = "scattering_density_2" if order == 2 else "scattering_density_n"
        s, f, x = PASS_COST[key]
        reads = (2 if order == 2 else 1) * ns * b + nt * b + ni * b
        return Work(ns * s * f, ns * s * x, reads + ns * b)
    if name == "indirect_irradiance":
        key = "indirect_irradiance_1" if order == 2 else "indirect_irradiance_n"
        s, f, x = PASS_COST[key]
        reads = (2 if order == 2 else 1) * ns * b
        return Work(ni * s * f, ni * s * x, reads + ni * b + 2 * ni * b)
    if name == "multiple_scattering":
        s, f, x = PASS_COST["multiple_scattering"]
        return Work(ns * s * f, ns * s * x, ns * b + nt * b + ns * b + 2 * ns * b)
    raise KeyError(name)


def lut_set_work(p) -> Work:
    """Whole precompute of one LUT set (src/precompute.rs:1726-1999)."""
    w = Work()
    for name in ("transmittance", "direct_irradiance", "single_scattering", "clear_irradiance"):
        w = w + pass_work(p, name)
    for order in range(2, p.order + 1):
        for name in ("scattering_density", "indirect_irradiance", "multiple_scattering"):
            w = w + pass_work(p, name, order)
    return w


def density_ground_rows(p) -> list:
    """Per r index z of the scattering LUT: how many of the 16 theta rows of
    the density quadrature (theta_l = (l + 1/2) pi/16) meet the ground
    (RayIntersectsGround, params.h:126-133) at the texel radius
    r_z = sqrt(rho_z^2 + bottom^2), rho_z = H z/(R-1) (scattering.h:62-75)."""
    import math

    b, t = float(p.bottom_radius), float(p.top_radius)
    H = math.sqrt(t * t - b * b)
    R = p.scattering_r_size
    out = []
    for z in range(R):
        rho = H * z / (R - 1)
        r = math.sqrt(rho * rho + b * b)
        n = 0
        for l in range(16):
            mu = math.cos((l + 0.5) * math.pi / 16)
            if mu < 0.0 and r * r * (mu * mu - 1.0) + b * b >= 0.0:
                n += 1
        out.append(n)
    return out


def density_lds_bytes(p, order: int) -> float:
    """LDS bytes the scattering-density kernel reads in its sample loop per
    launch: each of the 16 x 32 direction samples of a texel reads one
    (value, delta) profile entry of the incident radiance — 24 bytes (one
    b128 + one b64 read) of multiple scattering, or 48 bytes (three b128) of
    the single Rayleigh + Mie pair at order 2 — and, on theta rows that meet
    the ground, one 24-byte entry of the ground irradiance."""
    per = 48 if order == 2 else 24
    texels_per_r = p.scattering_nu_size * p.scattering_mu_s_size * p.scattering_mu_size
    total = 0.0
    for g in density_ground_rows(p):
        total += texels_per_r * 32 * (16 * per + g * 24)
    return total
