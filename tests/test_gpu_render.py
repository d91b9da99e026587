"""GPU parity of the sky / lighting lookups (shaders/render_sky.h,
render_sky.frag, render_lighting.h) against the oracle, on the same LUT set.

The GPU kernels read the GPU-built LUTs; the oracle evaluates the same
lookups on the same LUT values (downloaded), so the comparison isolates the
render arithmetic.  Tolerance: 1e-4 relative with an absolute floor of 1e-6
x max.  The reference's fragment shader works in FP32; where its arithmetic
cancels (the distance |point - camera| of two ~6360 km positions, and the
S - T*S_p difference of GetSkyRadianceToPoint) the GPU is compared with the
FP32 restatement, which performs the same binary32 operations.
"""
import math

import numpy as np
import pytest

from fuzzyblue_amd import (Atmosphere, DeviceBuffer, DrawParameters, Parameters, Renderer,
                           ray_direction, synchronize)
from fuzzyblue_amd.params import DUMP_SIZES
from oracle import Oracle

pytestmark = pytest.mark.gpu

P = Parameters().replace(**DUMP_SIZES)


def rel_err(x, y, floor):
    x = np.asarray(x, np.float64)[..., :3]
    y = np.asarray(y, np.float64)[..., :3]
    return np.abs(x - y) / np.maximum(np.abs(y), floor * max(np.abs(y).max(), 1e-30))


@pytest.fixture(scope="module")
def atm(builder):
    pa = Atmosphere.build(builder, None, P)
    pa.wait()
    a = pa.assert_ready()
    T = a.download("transmittance").astype(np.float64)
    S = a.download("scattering").astype(np.float64)
    E = a.download("irradiance").astype(np.float64)
    yield a, T, S, E
    a.close()


def random_rays(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    alt = rng.uniform(0.0, 10.0, n)
    cams = np.stack([np.zeros(n), np.zeros(n), 6360.0 + alt], axis=1)
    zs = np.radians(rng.uniform(0.0, 108.0, n))
    suns = np.stack([np.sin(zs), np.zeros(n), np.cos(zs)], axis=1)
    return cams.astype(np.float32), v.astype(np.float32), suns.astype(np.float32)


def test_sky_rays_vs_oracle(builder, atm):
    a, T, S, _ = atm
    n = 4096
    cams, views, suns = random_rays(n, 7)
    # include cameras in space, some of whose rays miss the atmosphere
    cams[:256, 2] = 6500.0
    r = Renderer(builder)
    d_c, d_v, d_s = (DeviceBuffer.from_array(x) for x in (cams, views, suns))
    rad, tr = DeviceBuffer(n * 16), DeviceBuffer(n * 16)
    r.rays(None, a, d_c, 1, d_v, d_s, 1, n, rad, tr)
    synchronize()
    g_rad = rad.to_array(np.float32, (n, 4))
    g_tr = tr.to_array(np.float32, (n, 4))
    o_rad, o_tr = Oracle("f32").sky_rays(P, T, S, cams, views, suns)
    e_rad = rel_err(g_rad, o_rad, 1e-6)
    e_tr = rel_err(g_tr, o_tr, 1e-6)
    print(f"\nsky rays vs FP32: radiance {e_rad.max():.3e} transmittance {e_tr.max():.3e}",
          flush=True)
    assert e_rad.max() < 1e-4
    assert e_tr.max() < 1e-4
    assert np.all(g_tr[:, 3] == 1.0)
    # Against FP64 evaluation of the same lookups: FP32 geometry flips the
    # ground/sky half of near-horizon rays (render_sky.h:74), as in the
    # reference; everywhere else the two agree.
    d_rad, d_tr = Oracle("f64").sky_rays(P, T, S, cams.astype(np.float64),
                                          views.astype(np.float64), suns.astype(np.float64))
    e64 = rel_err(g_rad, d_rad, 1e-6)
    print(f"sky rays vs FP64: median {np.median(e64):.2e} p99 {np.quantile(e64, 0.99):.2e}",
          flush=True)
    assert np.quantile(e64, 0.99) < 1e-3


def test_sky_rays_to_point_vs_oracle(builder, atm):
    a, T, S, _ = atm
    n = 4096
    cams, views, suns = random_rays(n, 9)
    dist = np.random.default_rng(10).uniform(0.5, 50.0, n).astype(np.float32)
    pts = (cams + views * dist[:, None]).astype(np.float32)
    r = Renderer(builder)
    bufs = [DeviceBuffer.from_array(x) for x in (cams, views, suns, pts)]
    rad, tr = DeviceBuffer(n * 16), DeviceBuffer(n * 16)
    r.rays(None, a, bufs[0], 1, bufs[1], bufs[2], 1, n, rad, tr, points=bufs[3])
    synchronize()
    g_rad = rad.to_array(np.float32, (n, 4))
    g_tr = tr.to_array(np.float32, (n, 4))
    o_rad, o_tr = Oracle("f32").sky_rays(P, T, S, cams, views, suns, points=pts)
    e_rad = rel_err(g_rad, o_rad, 1e-4)
    e_tr = rel_err(g_tr, o_tr, 1e-6)
    print(f"\nsky rays to point: radiance {e_rad.max():.3e} transmittance {e_tr.max():.3e}",
          flush=True)
    assert e_tr.max() < 1e-4
    assert np.quantile(e_rad, 0.999) < 1e-4
    assert e_rad.max() < 1e-3


def test_light_shafts_vs_oracle(builder, atm):
    """fb_render_sky_rays_shadowed (the light-shaft term the reference keeps
    commented out, render_sky.h:81-106, :158-160, :172-177) against the FP32
    restatement, with and without points; zero lengths reproduce
    fb_render_sky_rays bit for bit."""
    a, T, S, _ = atm
    n = 4096
    cams, views, suns = random_rays(n, 13)
    cams[:256, 2] = 6500.0
    rng = np.random.default_rng(14)
    ell = rng.uniform(0.0, 20.0, n).astype(np.float32)
    ell[::7] = 0.0
    pts = (cams + views * rng.uniform(0.5, 50.0, n).astype(np.float32)[:, None]).astype(
        np.float32)
    r = Renderer(builder)
    bufs = [DeviceBuffer.from_array(x) for x in (cams, views, suns, pts, ell)]
    zero = DeviceBuffer.from_array(np.zeros(n, np.float32))
    for points, floor in ((None, 1e-6), (bufs[3], 1e-4)):
        outs = {}
        for name, sh in (("plain", None), ("zero", zero), ("shadow", bufs[4])):
            rad, tr = DeviceBuffer(n * 16), DeviceBuffer(n * 16)
            r.rays(None, a, bufs[0], 1, bufs[1], bufs[2], 1, n, rad, tr, points=points,
                   shadow_lengths=sh)
            synchronize()
            outs[name] = (rad.to_array(np.float32, (n, 4)), tr.to_array(np.float32, (n, 4)))
        assert np.array_equal(outs["plain"][0], outs["zero"][0])
        assert np.array_equal(outs["plain"][1], outs["zero"][1])
        assert np.array_equal(outs["plain"][1], outs["shadow"][1])  # transmittance unchanged
        o_rad, _ = Oracle("f32").sky_rays(P, T, S, cams, views, suns,
                                          points=pts if points is not None else None,
                                          shadow=ell)
        e = rel_err(outs["shadow"][0], o_rad, floor)
        print(f"\nlight shafts ({'to point' if points is not None else 'sky'}): "
              f"max {e.max():.3e} p99.9 {np.quantile(e, 0.999):.3e}", flush=True)
        assert np.quantile(e, 0.999) < 1e-4
        assert e.max() < 1e-3


def test_frames_match_rays(builder, atm):
    """fb_render_sky_frames == fb_render_sky_rays on the same hashed rays."""
    a, T, S, _ = atm
    frames, px = 3, 5000
    alt = np.array([0.5, 4.0, 9.5], np.float32)
    zen = np.radians(np.array([10.0, 85.0, 104.0])).astype(np.float32)
    r = Renderer(builder)
    d_alt, d_zen = DeviceBuffer.from_array(alt), DeviceBuffer.from_array(zen)
    rad = DeviceBuffer(frames * px * 16)
    r.frames(None, a, d_alt, d_zen, frames, px, 2, rad)
    synchronize()
    g = rad.to_array(np.float32, (frames, px, 4))
    for f in range(frames):
        views = np.stack([ray_direction(2, f, i) for i in range(px)])
        cams = np.array([[0.0, 0.0, np.float32(6360.0) + alt[f]]], np.float32)
        sun = np.array([[math.sin(zen[f]), 0.0, math.cos(zen[f])]], np.float32)
        o_rad, _ = Oracle("f32").sky_rays(P, T, S, cams, views, sun)
        assert rel_err(g[f], o_rad, 1e-6).max() < 1e-4


def test_frames_slice_and_general_paths(builder, atm):
    """Frame kernels evaluate the LUTs' r / mu_s interpolation once per frame
    when the camera is inside the atmosphere; a camera above the top takes
    the per-ray path.  Both agree with the oracle on the same hashed rays,
    over more pixels than one block handles (4096)."""
    a, T, S, _ = atm
    px = 9000
    alt = np.array([0.0, 59.9, 75.0], np.float32)  # bottom, just below top, above
    zen = np.radians(np.array([30.0, 95.0, 60.0])).astype(np.float32)
    r = Renderer(builder)
    d_alt, d_zen = DeviceBuffer.from_array(alt), DeviceBuffer.from_array(zen)
    rad = DeviceBuffer(len(alt) * px * 16)
    tr = DeviceBuffer(len(alt) * px * 16)
    r.frames(None, a, d_alt, d_zen, len(alt), px, 11, rad, tr)
    synchronize()
    g = rad.to_array(np.float32, (len(alt), px, 4))
    gt = tr.to_array(np.float32, (len(alt), px, 4))
    for f in range(len(alt)):
        views = np.stack([ray_direction(11, f, i) for i in range(px)])
        cams = np.array([[0.0, 0.0, np.float32(6360.0) + alt[f]]], np.float32)
        sun = np.array([[math.sin(zen[f]), 0.0, math.cos(zen[f])]], np.float32)
        o_rad, o_tr = Oracle("f32").sky_rays(P, T, S, cams, views, sun)
        assert rel_err(g[f], o_rad, 1e-6).max() < 1e-4, f
        assert rel_err(gt[f], o_tr, 1e-6).max() < 1e-4, f


def test_frames_and_rays_odd_extents(builder):
    """LUT extents that are no powers of two (transmittance 250 x 60,
    scattering 4 x 36 x 26 x 18, irradiance 60 x 14): the frames call's
    per-frame LUT slice (19 nu slices x 36 mu entries and the transmittance
    row) and the per-ray kernel both give the FP32 oracle's radiance and
    transmittance on the same hashed rays (1e-4, floor 1e-6 x max)."""
    p = Parameters().replace(transmittance_mu_size=250, transmittance_r_size=60,
                             scattering_r_size=4, scattering_mu_size=36,
                             scattering_mu_s_size=26, scattering_nu_size=18,
                             irradiance_mu_s_size=60, irradiance_r_size=14)
    pa = Atmosphere.build(builder, None, p)
    pa.wait()
    a = pa.assert_ready()
    T = a.download("transmittance").astype(np.float64)
    S = a.download("scattering").astype(np.float64)
    frames, px = 3, 6000
    alt = np.array([0.2, 3.0, 9.0], np.float32)
    zen = np.radians(np.array([20.0, 88.0, 100.0])).astype(np.float32)
    r = Renderer(builder)
    d_alt, d_zen = DeviceBuffer.from_array(alt), DeviceBuffer.from_array(zen)
    rad, tr = DeviceBuffer(frames * px * 16), DeviceBuffer(frames * px * 16)
    print(f"\n[render] odd extents {p.scattering_extent()}: {frames} frames x {px}",
          flush=True)
    r.frames(None, a, d_alt, d_zen, frames, px, 9, rad, tr)
    synchronize()
    g = rad.to_array(np.float32, (frames, px, 4))
    gt = tr.to_array(np.float32, (frames, px, 4))
    for f in range(frames):
        views = np.stack([ray_direction(9, f, i) for i in range(px)]).astype(np.float32)
        cams = np.array([[0.0, 0.0, np.float32(p.bottom_radius) + alt[f]]], np.float32)
        sun = np.array([[math.sin(zen[f]), 0.0, math.cos(zen[f])]], np.float32)
        d_c, d_v, d_s = (DeviceBuffer.from_array(x) for x in (cams, views, sun))
        rr, rt = DeviceBuffer(px * 16), DeviceBuffer(px * 16)
        r.rays(None, a, d_c, 0, d_v, d_s, 0, px, rr, rt)
        synchronize()
        g_rr = rr.to_array(np.float32, (px, 4))
        g_rt = rt.to_array(np.float32, (px, 4))
        o_rad, o_tr = Oracle("f32").sky_rays(p, T, S, cams, views, sun)
        errs = [rel_err(x, y, 1e-6).max()
                for x, y in ((g[f], o_rad), (gt[f], o_tr), (g_rr, o_rad), (g_rt, o_tr))]
        print(f"  frame {f}: frames {errs[0]:.2e}/{errs[1]:.2e} rays {errs[2]:.2e}/{errs[3]:.2e}",
              flush=True)
        assert max(errs) < 1e-4, (f, errs)
    a.close()


@pytest.mark.parametrize("w,h", [(64, 48), (67, 45)])
def test_draw_matches_fragment_shader(builder, atm, w, h):
    """render_sky.frag:24-35 over a frame with a depth buffer; composite
    mode reproduces the dual-source blend dst = color + dst * T.  67 x 45
    fills no block of the draw kernel evenly."""
    a, T, S, _ = atm
    # Inverse of a reverse-Z infinite perspective for a camera 2 km up looking
    # along +x (column-major, [column][row]); world units are metres
    # (render_sky.frag:27 scales by 1e-3).  Depth z unprojects to the point
    # at distance near/z; z = 0 is the sky at infinity.
    near = 0.1
    ivp = np.zeros((4, 4), np.float32)
    ivp[0] = [0.0, -0.75, 0.0, 0.0]                  # ndc x -> world -y
    ivp[1] = [0.0, 0.0, -0.5, 0.0]                   # ndc y (down) -> world -z
    ivp[2] = [0.0, 0.0, 6362e3 / near, 1.0 / near]   # depth -> camera / near
    ivp[3] = [1.0, 0.0, 0.0, 0.0]                    # forward +x
    draw = DrawParameters(inverse_viewproj=ivp.tolist(), camera_position=[0.0, 0.0, 6362.0],
                          sun_direction=[0.6, 0.0, 0.8])
    depth = np.random.default_rng(3).uniform(near / 50e3, near / 500.0, (h, w)).astype(np.float32)
    depth[: h // 2] = 0.0  # upper half: infinitely far
    r = Renderer(builder)
    d_depth = DeviceBuffer.from_array(depth)
    color, tr = DeviceBuffer(w * h * 16), DeviceBuffer(w * h * 16)
    r.draw(None, a, draw, w, h, color, tr, depth=d_depth)
    synchronize()
    g_c = color.to_array(np.float32, (h, w, 4))
    g_t = tr.to_array(np.float32, (h, w, 4))
    o_c, o_t = Oracle("f32").render_frame(P, T, S, ivp, draw.camera_position,
                                           draw.sun_direction, w, h, depth=depth)
    assert rel_err(g_t, o_t, 1e-6).max() < 1e-4
    assert rel_err(g_c, o_c, 1e-4).max() < 1e-3
    assert np.all(g_c[..., 3] == 0.0) and np.all(g_t[..., 3] == 1.0)
    # composite: dst = color + dst * T
    dst = np.random.default_rng(4).uniform(0, 1, (h, w, 4)).astype(np.float32)
    d_dst = DeviceBuffer.from_array(dst)
    r.draw(None, a, draw, w, h, d_dst, None, depth=d_depth, composite=True)
    synchronize()
    comp = d_dst.to_array(np.float32, (h, w, 4))
    want = g_c[..., :3] + dst[..., :3] * g_t[..., :3]
    assert np.allclose(comp[..., :3], want, rtol=1e-5, atol=1e-7)
    assert np.array_equal(comp[..., 3], dst[..., 3])


def test_sun_and_sky_irradiance_vs_oracle(builder, atm):
    a, T, _, E = atm
    n = 2048
    rng = np.random.default_rng(11)
    pts = np.stack([rng.uniform(-50, 50, n), rng.uniform(-50, 50, n),
                    6360.0 + rng.uniform(0, 30, n)], axis=1).astype(np.float32)
    nrm = rng.standard_normal((n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = nrm.astype(np.float32)
    zs = np.radians(rng.uniform(0, 100, n))
    suns = np.stack([np.sin(zs), np.zeros(n), np.cos(zs)], axis=1).astype(np.float32)
    r = Renderer(builder)
    bufs = [DeviceBuffer.from_array(x) for x in (pts, nrm, suns)]
    direct, sky = DeviceBuffer(n * 16), DeviceBuffer(n * 16)
    r.sun_and_sky_irradiance(None, a, bufs[0], bufs[1], bufs[2], 1, n, direct, sky)
    synchronize()
    g_d = direct.to_array(np.float32, (n, 4))
    g_s = sky.to_array(np.float32, (n, 4))
    o_d, o_s = Oracle("f32").sun_and_sky_irradiance(P, T, E, pts, nrm, suns)
    assert rel_err(g_d, o_d, 1e-6).max() < 1e-4
    assert rel_err(g_s, o_s, 1e-6).max() < 1e-4
    # FP64 evaluation: the sun-disc smoothstep near the horizon is FP32
    # sensitive (transmittance.h:71-73); elsewhere they agree.
    d_d, d_s = Oracle("f64").sun_and_sky_irradiance(P, T, E, pts, nrm, suns)
    assert np.quantile(rel_err(g_d, d_d, 1e-6), 0.99) < 1e-4
    assert rel_err(g_s, d_s, 1e-6).max() < 1e-4


def test_empty_inputs(builder, atm):
    """Zero rays, frames or points enqueue nothing and succeed; an empty batch
    builds nothing."""
    a, _, _, _ = atm
    r = Renderer(builder)
    one = DeviceBuffer.from_array(np.zeros(3, np.float32))
    out = DeviceBuffer(16)
    r.rays(None, a, one, 0, one, one, 0, 0, out)
    r.rays(None, a, one, 0, one, one, 0, 0, out, shadow_lengths=one)
    r.frames(None, a, one, one, 0, 100, 1, out)
    r.frames(None, a, one, one, 1, 0, 1, out)
    r.sun_and_sky_irradiance(None, a, one, one, one, 0, 0, out, out)
    synchronize()
    assert Atmosphere.build_batch(builder, None, []) == []


def test_default_size_frames_vs_oracle(builder):
    """The bench's render workload on the LUT set it is timed on: the
    default-size Earth set (scattering 32x128x32x8, precision mode) and
    1920x1080 frames from fb_render_sky_frames (cameras at U[0, 10] km, sun
    zeniths up to 108 degrees: the per-frame LDS slice of 22.5 KB), 20,000
    sampled pixels per frame against the FP32 restatement of the same
    lookups.  Also views exactly opposite the sun through the per-ray path,
    where |nu| rounds past 1."""
    p = Parameters().replace(flags=1)
    pa = Atmosphere.build(builder, None, p)
    pa.wait()
    a = pa.assert_ready()
    T = a.download("transmittance").astype(np.float64)
    S = a.download("scattering").astype(np.float64)
    frames, px = 4, 1920 * 1080
    alt = np.array([0.0, 2.5, 7.0, 10.0], np.float32)
    zen = np.radians(np.array([5.0, 60.0, 95.0, 108.0])).astype(np.float32)
    r = Renderer(builder)
    d_alt, d_zen = DeviceBuffer.from_array(alt), DeviceBuffer.from_array(zen)
    rad, tr = DeviceBuffer(frames * px * 16), DeviceBuffer(frames * px * 16)
    r.frames(None, a, d_alt, d_zen, frames, px, 2, rad, tr)
    synchronize()
    g_rad = rad.to_array(np.float32, (frames, px, 4))
    g_tr = tr.to_array(np.float32, (frames, px, 4))
    rng = np.random.default_rng(31)
    worst = 0.0
    for f in range(frames):
        idx = np.sort(rng.choice(px, 20000, replace=False))
        views = np.stack([ray_direction(2, f, int(i)) for i in idx])
        cams = np.array([[0.0, 0.0, np.float32(6360.0) + alt[f]]], np.float32)
        sun = np.array([[math.sin(zen[f]), 0.0, math.cos(zen[f])]], np.float32)
        o_rad, o_tr = Oracle("f32").sky_rays(p, T, S, cams, views, sun)
        e_rad = rel_err(g_rad[f, idx], o_rad, 1e-6).max()
        e_tr = rel_err(g_tr[f, idx], o_tr, 1e-6).max()
        worst = max(worst, e_rad, e_tr)
        assert e_rad < 1e-4 and e_tr < 1e-4, (f, e_rad, e_tr)
    print(f"\ndefault-size frames vs FP32 restatement: max rel err {worst:.3e}", flush=True)
    # views exactly opposite the sun (anti-solar point), per-ray path
    n = 256
    zs = np.radians(np.linspace(0.0, 108.0, n))
    suns = np.stack([np.sin(zs), np.zeros(n), np.cos(zs)], axis=1).astype(np.float32)
    views = (-suns).astype(np.float32)
    cams = np.stack([np.zeros(n), np.zeros(n), 6360.0 + np.linspace(0.0, 10.0, n)],
                    axis=1).astype(np.float32)
    bufs = [DeviceBuffer.from_array(x) for x in (cams, views, suns)]
    rr, rt = DeviceBuffer(n * 16), DeviceBuffer(n * 16)
    r.rays(None, a, bufs[0], 1, bufs[1], bufs[2], 1, n, rr, rt)
    synchronize()
    o_rad, _ = Oracle("f32").sky_rays(p, T, S, cams, views, suns)
    g = rr.to_array(np.float32, (n, 4))
    assert np.all(np.isfinite(g))
    assert rel_err(g, o_rad, 1e-6).max() < 1e-4
    a.close()
