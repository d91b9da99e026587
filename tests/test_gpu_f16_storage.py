"""Binary16 storage (FB_FLAG_F16_STORAGE): the scattering LUT and the four
scattering-sized scratch LUTs (delta Rayleigh, delta Mie, delta multiple
scattering, scattering density) are RGBA binary16, every store rounded to
nearest even, as the reference's R16G16B16A16_SFLOAT images
(src/precompute.rs:1215, :1229-1234).  Transmittance and irradiance stay FP32.

Tolerances:
  * pass by pass (FP32 arithmetic) on the GPU's own inputs against the FP32
    restatement in binary16 mode: every stored texel is a binary16 value at
    most one binary16 ulp from the restatement's (the two FP32 values before
    rounding differ by up to 1e-4 relative, the FP32-mode bar, so they can
    straddle a rounding boundary), and at least 95% are bit-identical;
  * end to end: 3e-3 relative with a floor of 1e-3 x max (rounding flips
    propagate through the orders; the density LUT's smallest values fall
    below the binary16 subnormal range, as in the reference);
  * views, downloads, batch members and the render kernels read the
    binary16 LUTs exactly (bit-identical comparisons).
"""
import ctypes as C

import numpy as np
import pytest

from fuzzyblue_amd import (Atmosphere, DeviceBuffer, DrawParameters, Parameters, Renderer, abi,
                           synchronize)
from fuzzyblue_amd.lib import check, load
from fuzzyblue_amd.params import DUMP_SIZES, SMOKE_SIZES
from fuzzyblue_amd.workloads import random_atmosphere
from oracle import Oracle

pytestmark = pytest.mark.gpu

F16 = abi.FB_FLAG_F16_STORAGE
SMOKE = Parameters().replace(**SMOKE_SIZES)
DUMP = Parameters().replace(**DUMP_SIZES)
KEYS = ("transmittance", "irradiance", "scattering")
SCRATCH3D = ("delta_rayleigh", "delta_mie", "delta_multiple_scattering", "scattering_density")


def f16(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def rel_err(x, y, floor):
    x = np.asarray(x, np.float64)[..., :3]
    y = np.asarray(y, np.float64)[..., :3]
    return np.abs(x - y) / np.maximum(np.abs(y), floor * max(np.abs(y).max(), 1e-300))


def check_ulp16(name, got, want):
    """got (GPU) and want (restatement) are binary16 values at most one
    binary16 ulp apart, and mostly identical."""
    got = np.asarray(got, np.float32)
    want = np.asarray(want, np.float32)
    assert np.array_equal(got, f16(got)), name
    assert np.array_equal(want, f16(want)), name
    big = np.maximum(np.abs(got), np.abs(want)).astype(np.float16)
    ulp = np.spacing(big).astype(np.float32)
    d = np.abs(got - want)
    same = float(np.mean(got == want))
    print(f"  {name}: {same:.4f} identical, max {float((d / ulp).max()):.1f} ulp", flush=True)
    assert np.all(d <= ulp), name
    assert same >= 0.95, name


def build(builder, p):
    pa = Atmosphere.build(builder, None, p)
    pa.wait()
    return pa


def test_views_and_downloads(builder):
    """Binary16 LUTs report 8-byte texels; the float32 download is the exact
    widening of the stored binary16 values."""
    pa = build(builder, DUMP.replace(flags=F16, order=2))
    a = pa.atmosphere()
    ns = DUMP.scattering_nu_size * DUMP.scattering_mu_s_size * DUMP.scattering_mu_size * \
        DUMP.scattering_r_size
    v = a.lut("scattering")
    assert v.texel_bytes == 8 and v.bytes == ns * 8
    assert a.lut("transmittance").texel_bytes == 16
    assert a.lut("irradiance").texel_bytes == 16
    for k in SCRATCH3D:
        assert pa.scratch(k).texel_bytes == 8, k
    assert pa.scratch("delta_irradiance").texel_bytes == 16
    h = a.download("scattering", fp16=True)
    w = a.download("scattering")
    assert h.dtype == np.float16 and np.array_equal(h.astype(np.float32), w)
    # the captured schedule replays to the same bits
    pa.relaunch(None)
    pa.wait()
    assert np.array_equal(a.download("scattering", fp16=True), h)
    pa.close()


@pytest.mark.parametrize("p", [SMOKE, DUMP], ids=["smoke", "dump"])
def test_passes_f16_vs_restatement(builder, p):
    """Each pass on the GPU's own inputs against the FP32 restatement with
    binary16 stores."""
    o = Oracle("f32")
    p = p.replace(flags=F16)
    print(f"\n[f16 storage] sizes {p.scattering_extent()}", flush=True)
    pa1 = build(builder, p.replace(order=1))
    a1 = pa1.atmosphere()
    T = a1.download("transmittance")
    assert rel_err(T, o.pass_transmittance(p), 1e-6).max() < 1e-5
    dE_dir = pa1.download_scratch("delta_irradiance")
    dR, dM, S1 = o.pass_single_scattering(p, T, fp16=True)
    dR_g = pa1.download_scratch("delta_rayleigh")
    dM_g = pa1.download_scratch("delta_mie")
    S1_g = a1.download("scattering")
    check_ulp16("single rayleigh", dR_g, dR)
    check_ulp16("single mie", dM_g, dM)
    check_ulp16("scattering, order 1", S1_g, S1)
    zeros = np.zeros_like(dR_g)
    pa2 = build(builder, p.replace(order=2))
    J2_g = pa2.download_scratch("scattering_density")
    check_ulp16("density, order 2", J2_g,
                o.pass_density(p, 2, T, dR_g, dM_g, zeros, dE_dir, fp16=True))
    dE1, _ = o.pass_indirect(p, 1, dR_g, dM_g, zeros, np.zeros_like(dE_dir))
    dE1_g = pa2.download_scratch("delta_irradiance")
    assert rel_err(dE1_g, dE1, 1e-6).max() < 1e-4  # FP32 storage
    dMS2, S2 = o.pass_multiple(p, T, J2_g, S1_g, fp16=True)
    dMS2_g = pa2.download_scratch("delta_multiple_scattering")
    check_ulp16("multiple, order 2", dMS2_g, dMS2)
    check_ulp16("scattering, order 2", pa2.atmosphere().download("scattering"), S2)
    pa3 = build(builder, p.replace(order=3))
    check_ulp16("density, order 3", pa3.download_scratch("scattering_density"),
                o.pass_density(p, 3, T, dR_g, dM_g, dMS2_g, dE1_g, fp16=True))
    for pa in (pa1, pa2, pa3):
        pa.close()


@pytest.mark.parametrize("precise", [False, True], ids=["fp32", "precise"])
def test_full_f16_vs_oracle(builder, precise):
    """End to end (order 4) against the oracle in binary16 mode: the FP32
    restatement for FP32 arithmetic, the FP64 oracle for precision mode."""
    p = SMOKE.replace(flags=F16 | (abi.FB_FLAG_PRECISE if precise else 0))
    pa = build(builder, p)
    a = pa.atmosphere()
    ref = Oracle("f64" if precise else "f32").precompute(p, fp16=True)
    eT = rel_err(a.download("transmittance"), ref["T"], 1e-6).max()
    eE = rel_err(a.download("irradiance"), ref["E"], 1e-3).max()
    eS = rel_err(a.download("scattering"), ref["S"], 1e-3).max()
    print(f"\nf16 storage {'precise' if precise else 'fp32'}: T {eT:.2e} E {eE:.2e} "
          f"S {eS:.2e}", flush=True)
    assert eT < 1e-5
    assert eE < 3e-3
    assert eS < 3e-3
    pa.close()


def test_batch_members_equal_single_builds(builder):
    """A batch mixing binary16 and FP32 storage (and both precision modes)
    builds every member bit-identical to the same atmosphere built alone."""
    rng = np.random.default_rng(31)
    ps = [
        SMOKE.replace(flags=F16, order=3),
        SMOKE.replace(order=3),
        DUMP.replace(flags=F16 | abi.FB_FLAG_PRECISE, order=5),
        random_atmosphere(rng, False, resolution="half", order=4),
    ]
    ps[3].flags = F16
    pend = Atmosphere.build_batch(builder, None, ps)
    for pa in pend:
        pa.wait()
    for p, pb in zip(ps, pend):
        ps_ = build(builder, p)
        for k in KEYS:
            assert np.array_equal(pb.atmosphere().download(k), ps_.atmosphere().download(k)), \
                (k, p.flags)
        ps_.close()
    for pa in pend:
        pa.close()


def random_rays(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    alt = rng.uniform(0.0, 10.0, n)
    cams = np.stack([np.zeros(n), np.zeros(n), 6360.0 + alt], axis=1)
    cams[: n // 16, 2] = 6500.0  # in space
    zs = np.radians(rng.uniform(0.0, 108.0, n))
    suns = np.stack([np.sin(zs), np.zeros(n), np.cos(zs)], axis=1)
    return cams.astype(np.float32), v.astype(np.float32), suns.astype(np.float32)


def test_render_reads_binary16_scattering(builder):
    """Every render entry point on a binary16 LUT set equals the same entry
    point on an FP32 LUT set holding the widened values (the render
    arithmetic is FP32 either way)."""
    a16 = build(builder, DUMP.replace(flags=F16)).assert_ready()
    a32 = build(builder, DUMP).assert_ready()
    assert np.array_equal(a16.download("transmittance"), a32.download("transmittance"))
    S = np.ascontiguousarray(a16.download("scattering"))
    v = a32.lut("scattering")
    assert v.bytes == S.nbytes
    check(load().fb_memcpy_h2d(C.c_void_p(v.device_ptr), S.ctypes.data_as(C.c_void_p), S.nbytes))
    r = Renderer(builder)
    # rays, with and without points
    n = 4096
    cams, views, suns = random_rays(n, 41)
    pts = (cams + views * np.float32(20.0)).astype(np.float32)
    bufs = [DeviceBuffer.from_array(x) for x in (cams, views, suns, pts)]
    for points in (None, bufs[3]):
        out = []
        for a in (a16, a32):
            rad, tr = DeviceBuffer(n * 16), DeviceBuffer(n * 16)
            r.rays(None, a, bufs[0], 1, bufs[1], bufs[2], 1, n, rad, tr, points=points)
            synchronize()
            out.append((rad.to_array(np.float32, (n, 4)), tr.to_array(np.float32, (n, 4))))
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    # frames: the per-frame slice (inside the atmosphere) and the per-ray path
    alt = np.array([0.0, 59.9, 75.0], np.float32)
    zen = np.radians(np.array([30.0, 95.0, 60.0])).astype(np.float32)
    d_alt, d_zen = DeviceBuffer.from_array(alt), DeviceBuffer.from_array(zen)
    px = 9000
    out = []
    for a in (a16, a32):
        rad, tr = DeviceBuffer(len(alt) * px * 16), DeviceBuffer(len(alt) * px * 16)
        r.frames(None, a, d_alt, d_zen, len(alt), px, 11, rad, tr)
        synchronize()
        out.append((rad.to_array(np.float32, (len(alt), px, 4)),
                    tr.to_array(np.float32, (len(alt), px, 4))))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    # draw with a depth buffer
    w, h = 32, 24
    ivp = np.zeros((4, 4), np.float32)
    ivp[0] = [0.0, -0.75, 0.0, 0.0]
    ivp[1] = [0.0, 0.0, -0.5, 0.0]
    ivp[2] = [0.0, 0.0, 6362e3 / 0.1, 1.0 / 0.1]
    ivp[3] = [1.0, 0.0, 0.0, 0.0]
    draw = DrawParameters(inverse_viewproj=ivp.tolist(), camera_position=[0.0, 0.0, 6362.0],
                          sun_direction=[0.6, 0.0, 0.8])
    depth = np.random.default_rng(3).uniform(0.1 / 50e3, 0.1 / 500.0, (h, w)).astype(np.float32)
    depth[: h // 2] = 0.0
    d_depth = DeviceBuffer.from_array(depth)
    out = []
    for a in (a16, a32):
        color, tr = DeviceBuffer(w * h * 16), DeviceBuffer(w * h * 16)
        r.draw(None, a, draw, w, h, color, tr, depth=d_depth)
        synchronize()
        out.append((color.to_array(np.float32, (h, w, 4)), tr.to_array(np.float32, (h, w, 4))))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    a16.close()
    a32.close()
