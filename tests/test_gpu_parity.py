"""GPU parity of the precompute passes against the CPU oracle.

Every comparison calls the HIP path through the C ABI (libfuzzyblue_hip.so)
and the oracle (oracle/libfb_oracle.so) on the same parameters.

Tolerances (relative error |x-y| / max(|y|, floor), floor = FLOOR * max|y|):
  * FP32 mode vs the FP32 restatement, pass by pass on identical inputs:
    transmittance 1e-5 (the reference's bar), other LUTs 1e-4 (libm-vs-GPU
    transcendental ulps and summation order).
  * Precision mode (FB_FLAG_PRECISE) vs the FP64 oracle, end to end:
    transmittance 1e-5, irradiance and scattering 1e-3 (the bar the metric
    is quoted at).
"""
import ctypes as C

import numpy as np
import pytest

from fuzzyblue_amd import Atmosphere, Parameters, abi
from fuzzyblue_amd.lib import check, load
from fuzzyblue_amd.params import DUMP_SIZES, SMOKE_SIZES
from oracle import Oracle

pytestmark = pytest.mark.gpu

FLOOR = 1e-6


def rel_err(x, y, floor=FLOOR):
    x = np.asarray(x, np.float64)[..., :3]
    y = np.asarray(y, np.float64)[..., :3]
    den = np.maximum(np.abs(y), floor * max(np.abs(y).max(), 1e-300))
    return np.abs(x - y) / den


def run(builder, p):
    pa = Atmosphere.build(builder, None, p)
    pa.wait()
    return pa


def outputs(pa):
    a = pa.atmosphere()
    return {k: a.download(k) for k in ("transmittance", "irradiance", "scattering")}


def scratch(pa):
    return {k: pa.download_scratch(k) for k in (
        "delta_irradiance", "delta_rayleigh", "delta_mie", "delta_multiple_scattering",
        "scattering_density")}


SMOKE = Parameters().replace(**SMOKE_SIZES)


# Rows of 32 texels: multiple scattering in 8 chunks of 6-7 samples.
NARROW32 = Parameters().replace(scattering_r_size=8, scattering_mu_size=32,
                                scattering_mu_s_size=8, scattering_nu_size=4)


@pytest.mark.parametrize("p", [SMOKE, NARROW32, Parameters().replace(**DUMP_SIZES)],
                         ids=["smoke", "narrow32", "dump"])
def test_passes_fp32_vs_restatement(builder, p):
    """Each pass on the GPU's own inputs against the FP32 restatement (at
    the smoke and narrow32 sizes multiple scattering runs chunked: its
    rows are at most 32 texels wide)."""
    o = Oracle("f32")
    print(f"\n[parity fp32] sizes {p.scattering_extent()}", flush=True)
    # order 1: transmittance, direct irradiance, single scattering
    pa1 = run(builder, p.replace(order=1))
    out1, scr1 = outputs(pa1), scratch(pa1)
    T = out1["transmittance"]
    T_ref = o.pass_transmittance(p)
    assert rel_err(T, T_ref).max() < 1e-5
    dE_dir = scr1["delta_irradiance"]
    assert rel_err(dE_dir, o.pass_direct_irradiance(p, T)).max() < 1e-4
    dR, dM, S1 = o.pass_single_scattering(p, T)
    assert rel_err(scr1["delta_rayleigh"], dR).max() < 1e-4
    assert rel_err(scr1["delta_mie"], dM).max() < 1e-4
    assert rel_err(out1["scattering"], S1).max() < 1e-4
    assert np.all(out1["irradiance"] == 0.0)  # order 1: irradiance stays cleared
    dR_g, dM_g = scr1["delta_rayleigh"], scr1["delta_mie"]
    # order 2: density(2), indirect(1), multiple
    pa2 = run(builder, p.replace(order=2))
    scr2 = scratch(pa2)
    zeros = np.zeros_like(dR_g)
    J2 = o.pass_density(p, 2, T, dR_g, dM_g, zeros, dE_dir)
    assert rel_err(scr2["scattering_density"], J2).max() < 1e-4
    dE1, E1 = o.pass_indirect(p, 1, dR_g, dM_g, zeros, np.zeros_like(dE_dir))
    assert rel_err(scr2["delta_irradiance"], dE1).max() < 1e-4
    dMS2, _ = o.pass_multiple(p, T, scr2["scattering_density"], out1["scattering"])
    assert rel_err(scr2["delta_multiple_scattering"], dMS2).max() < 1e-4
    # order 3: density(3) reads the order-2 multiple scattering
    pa3 = run(builder, p.replace(order=3))
    scr3 = scratch(pa3)
    J3 = o.pass_density(p, 3, T, dR_g, dM_g, scr2["delta_multiple_scattering"],
                        scr2["delta_irradiance"])
    assert rel_err(scr3["scattering_density"], J3).max() < 1e-4
    dE2, _ = o.pass_indirect(p, 2, dR_g, dM_g, scr2["delta_multiple_scattering"],
                             np.zeros_like(dE_dir))
    assert rel_err(scr3["delta_irradiance"], dE2).max() < 1e-4
    for pa in (pa1, pa2, pa3):
        pa.close()


def test_passes_fp32_default_size_sampled(builder):
    """FP32 mode at the default LUT sizes (256x64, 64x16, 32x128x32x8), pass
    by pass on the GPU's own inputs against the FP32 restatement: the 2D
    passes in full, the 3D passes (single scattering, density of orders 2
    and 3, multiple scattering) on 4,096 seeded texels each."""
    p = Parameters()
    assert p.flags == 0
    o = Oracle("f32")
    rng = np.random.default_rng(2024)
    sw, sh, sd = p.scattering_extent()
    idx = np.sort(rng.choice(sw * sh * sd, 4096, replace=False))
    at = lambda a: a.reshape(-1, 4)[idx]  # noqa: E731
    print(f"\n[parity fp32 default] {len(idx)} texels of {sw}x{sh}x{sd}", flush=True)
    pa1 = run(builder, p.replace(order=1))
    out1, scr1 = outputs(pa1), scratch(pa1)
    T = out1["transmittance"]
    assert rel_err(T, o.pass_transmittance(p)).max() < 1e-5
    dE_dir = scr1["delta_irradiance"]
    assert rel_err(dE_dir, o.pass_direct_irradiance(p, T)).max() < 1e-4
    ray, mie = o.single_texels(p, T, idx)
    assert rel_err(at(scr1["delta_rayleigh"]), ray).max() < 1e-4
    assert rel_err(at(scr1["delta_mie"]), mie).max() < 1e-4
    dR_g, dM_g = scr1["delta_rayleigh"], scr1["delta_mie"]
    zeros = np.zeros_like(dR_g)
    pa2 = run(builder, p.replace(order=2))
    scr2 = scratch(pa2)
    J2 = o.density_texels(p, 2, T, dR_g, dM_g, zeros, dE_dir, idx)
    assert rel_err(at(scr2["scattering_density"]), J2).max() < 1e-4
    dE1, _ = o.pass_indirect(p, 1, dR_g, dM_g, zeros, np.zeros_like(dE_dir))
    assert rel_err(scr2["delta_irradiance"], dE1).max() < 1e-4
    dMS2, _ = o.multiple_texels(p, T, scr2["scattering_density"], idx)
    assert rel_err(at(scr2["delta_multiple_scattering"]), dMS2).max() < 1e-4
    pa3 = run(builder, p.replace(order=3))
    scr3 = scratch(pa3)
    J3 = o.density_texels(p, 3, T, dR_g, dM_g, scr2["delta_multiple_scattering"],
                          scr2["delta_irradiance"], idx)
    assert rel_err(at(scr3["scattering_density"]), J3).max() < 1e-4
    for pa in (pa1, pa2, pa3):
        pa.close()


def test_full_precise_vs_fp64(builder):
    """Precision mode end to end against the FP64 oracle at the smoke sizes."""
    p = SMOKE.replace(flags=abi.FB_FLAG_PRECISE)
    print("\n[parity precise] smoke order 4", flush=True)
    pa = run(builder, p)
    out = outputs(pa)
    ref = Oracle("f64").precompute(p)
    eT = rel_err(out["transmittance"], ref["T"]).max()
    eE = rel_err(out["irradiance"], ref["E"]).max()
    eS = rel_err(out["scattering"], ref["S"]).max()
    print(f"precise vs fp64: T {eT:.3e} E {eE:.3e} S {eS:.3e}", flush=True)
    assert eT < 1e-5
    assert eE < 1e-3
    assert eS < 1e-3
    pa.close()


def test_full_fp32_vs_restatement(builder):
    """FP32 mode end to end against the FP32 restatement (smoke sizes)."""
    pa = run(builder, SMOKE)
    out = outputs(pa)
    ref = Oracle("f32").precompute(SMOKE)
    eT = rel_err(out["transmittance"], ref["T"]).max()
    eE = rel_err(out["irradiance"], ref["E"]).max()
    eS = rel_err(out["scattering"], ref["S"]).max()
    print(f"\nfp32 vs restatement: T {eT:.3e} E {eE:.3e} S {eS:.3e}", flush=True)
    assert eT < 1e-5
    assert eE < 1e-3
    assert eS < 1e-3
    pa.close()


def test_relaunch_is_deterministic(builder):
    """The captured graph replays to the same bits as the first build."""
    pa = Atmosphere.build(builder, None, SMOKE)
    pa.wait()
    first = outputs(pa)
    for _ in range(3):
        pa.relaunch(None)
        pa.wait()
        again = outputs(pa)
        for k in first:
            assert np.array_equal(first[k], again[k]), k
    pa.close()


def test_ready_event_hand_over(builder):
    """fb_pending_record_ready_event (acquire_ownership as an event): the
    caller's event completes once the precompute has, and a consumer that
    waited on it reads the finished LUTs."""
    lib = load()
    ev = C.c_void_p()
    check(lib.fb_event_create(C.byref(ev)))
    pa = Atmosphere.build(builder, None, SMOKE)
    pa.record_ready_event(ev)
    check(lib.fb_event_synchronize(ev))
    assert pa.is_ready()
    ref = Atmosphere.build(builder, None, SMOKE)
    ref.wait()
    assert np.array_equal(pa.atmosphere().download("scattering"),
                          ref.atmosphere().download("scattering"))
    check(lib.fb_event_destroy(ev))
    pa.close()
    ref.close()
