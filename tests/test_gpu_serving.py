"""The serving loop (fb_server_*): continuous batching of precompute and
render requests, HBM-budget admission with LRU eviction, and the check
that each result is the same built alone and served in a mix."""
import numpy as np
import pytest

from fuzzyblue_amd import Atmosphere, Parameters, abi
from fuzzyblue_amd.params import DUMP_SIZES, RESOLUTIONS, SMOKE_SIZES
from fuzzyblue_amd.serving import Server
from fuzzyblue_amd.workloads import random_atmosphere

pytestmark = pytest.mark.gpu

GB = 1 << 30


def shapes(p):
    tw, th = p.transmittance_extent()
    iw, ih = p.irradiance_extent()
    sw, sh, sd = p.scattering_extent()
    return {abi.FB_LUT_TRANSMITTANCE: (th, tw, 4), abi.FB_LUT_IRRADIANCE: (ih, iw, 4),
            abi.FB_LUT_SCATTERING: (sd, sh, sw, 4)}


def fetch(server, lut, p):
    return {k: server.download(lut, k, s) for k, s in shapes(p).items()}


@pytest.mark.parametrize("share", [0, 60])
def test_alone_vs_mix(builder, share):
    """Alone: full-occupancy launches.  In the mix: the same (share 0) or
    persistent launches holding 60% of the resident-block slots (share 60);
    either way every texel is computed by the same code on the same inputs,
    so the results are bit-identical."""
    target = Parameters().replace(**DUMP_SIZES, order=4, flags=abi.FB_FLAG_PRECISE)
    # alone
    s = Server(builder, 8 * GB, keep_outputs=True)
    _, lut = s.submit_precompute(target)
    s.drain()
    alone = fetch(s, lut, target)
    rid = s.submit_render(lut, 3.0, 0.7, 96, 64, 12345)
    s.drain()
    alone_img = s.render_result(rid, 96, 64)
    infos = s.poll()
    assert all(i.status == abi.FB_OK for i in infos)
    s.close()
    # in a mix: other precomputes (batched with the target) and renders
    rng = np.random.default_rng(5)
    s = Server(builder, 8 * GB, keep_outputs=True, precompute_share_pct=share)
    others = [random_atmosphere(rng, True, resolution="half") for _ in range(3)]
    s.submit_precompute(others[0])
    _, lut = s.submit_precompute(target)
    for p in others[1:]:
        s.submit_precompute(p)
    rid = s.submit_render(lut, 3.0, 0.7, 96, 64, 12345)
    for k in range(6):
        s.submit_render(lut, float(k), 0.3 * k, 64, 32, k)
    s.drain()
    mixed = fetch(s, lut, target)
    mixed_img = s.render_result(rid, 96, 64)
    infos = s.poll()
    assert len(infos) == 11 and all(i.status == abi.FB_OK for i in infos)
    for k in alone:
        assert np.array_equal(alone[k], mixed[k]), k
    assert np.array_equal(alone_img, mixed_img)
    s.close()


def test_budget_rejection_and_unknown_lut(builder):
    p = Parameters().replace(**SMOKE_SIZES)
    need = int(p.scattering_extent()[0] * p.scattering_extent()[1] * p.scattering_extent()[2]
               * 16 * 6)
    s = Server(builder, need // 2)
    rid, _ = s.submit_precompute(p)
    bad = s.submit_render(987654, 1.0, 0.5, 8, 8, 1)
    s.drain()
    st = {i.id: i.status for i in s.poll()}
    assert st[rid] == abi.FB_EBUDGET
    assert st[bad] == abi.FB_ENOENT
    s.close()


def test_lru_eviction(builder):
    p = Parameters().replace(**RESOLUTIONS["half"])
    res, scratch = p.footprint()
    # room for two resident sets plus one build in flight
    s = Server(builder, 2 * res + res + scratch + 4096, max_inflight=1)
    luts = []
    for k in range(4):
        _, lut = s.submit_precompute(p.replace(order=2 + k))
        luts.append(lut)
        s.drain()
    infos = s.poll()
    assert all(i.status == abi.FB_OK for i in infos)
    st = s.stats()
    assert st.evictions >= 1
    resident = set(s.resident())
    assert luts[-1] in resident and luts[0] not in resident
    s.close()


def test_render_waits_for_inflight_precompute(builder):
    p = Parameters().replace(**DUMP_SIZES)
    s = Server(builder, 8 * GB, keep_outputs=True)
    _, lut = s.submit_precompute(p)
    rid = s.submit_render(lut, 1.0, 0.2, 32, 16, 3)
    s.drain()
    infos = {i.id: i for i in s.poll()}
    assert infos[rid].status == abi.FB_OK
    img = s.render_result(rid, 32, 16)
    assert np.all(np.isfinite(img)) and img[..., :3].max() > 0.0
    s.close()


def test_binary16_storage_requests(builder):
    """Binary16-storage sets (FB_FLAG_F16_STORAGE) served beside FP32 ones:
    each result equals the same set built directly, and the server holds
    the smaller resident footprint for the binary16 set."""
    p32 = Parameters().replace(**DUMP_SIZES, order=3, flags=abi.FB_FLAG_PRECISE)
    p16 = p32.replace(flags=abi.FB_FLAG_PRECISE | abi.FB_FLAG_F16_STORAGE)
    s = Server(builder, 8 * GB, keep_outputs=True)
    _, l32 = s.submit_precompute(p32)
    _, l16 = s.submit_precompute(p16)
    rid = s.submit_render(l16, 2.0, 0.6, 64, 32, 7)
    s.drain()
    assert all(i.status == abi.FB_OK for i in s.poll())
    for p, lut in ((p32, l32), (p16, l16)):
        pa = Atmosphere.build(builder, None, p)
        pa.wait()
        a = pa.atmosphere()
        got = fetch(s, lut, p)
        for k, name in ((abi.FB_LUT_TRANSMITTANCE, "transmittance"),
                        (abi.FB_LUT_IRRADIANCE, "irradiance"),
                        (abi.FB_LUT_SCATTERING, "scattering")):
            assert np.array_equal(got[k], a.download(name)), (p.flags, name)
        pa.close()
    img = s.render_result(rid, 64, 32)
    assert np.all(np.isfinite(img)) and img[..., :3].max() > 0.0
    assert s.stats().hbm_in_use == p32.footprint()[0] + p16.footprint()[0]
    s.close()


def test_queued_renders_pin_their_set(builder):
    """Renders pin their LUT set from submission: under a budget tight enough
    that renders queue for output memory while later precomputes need room,
    no render finds its set evicted (FB_ENOENT), and evicted sets are still
    freed so every request completes."""
    p = Parameters().replace(**RESOLUTIONS["half"])
    res, scratch = p.footprint()
    w, h = 640, 360
    out = w * h * 16
    s = Server(builder, 3 * res + scratch + 3 * out, max_inflight=1)
    luts = [s.submit_precompute(p.replace(order=2 + (k % 3)))[1] for k in range(5)]
    rids = []
    for k, lut in enumerate(luts):
        for j in range(3):
            rids.append(s.submit_render(lut, 1.0 + j, 0.2 * j, w, h, 100 * k + j))
    s.drain()
    infos = {i.id: i for i in s.poll()}
    assert len(infos) == len(luts) + len(rids)
    bad = {i.id: i.status for i in infos.values() if i.status != abi.FB_OK}
    assert not bad, bad
    st = s.stats()
    assert st.evictions >= 1
    assert st.hbm_peak <= 3 * res + scratch + 3 * out
    s.close()


def test_bulk_submission(builder):
    """fb_server_submit_precomputes: ids in submission order, every request
    completes, and each LUT set equals the set a plain build of the same
    parameters produces (a batch member is bit-identical to a plain build);
    an invalid parameter set in the list enqueues nothing."""
    rng = np.random.default_rng(11)
    ps = [random_atmosphere(rng, True, resolution="half") for _ in range(5)]
    ps.append(Parameters().replace(**DUMP_SIZES, order=3, flags=abi.FB_FLAG_PRECISE))
    s = Server(builder, 8 * GB, keep_outputs=True, n_streams=2)
    ids = s.submit_precomputes(ps)
    assert len(ids) == len(ps)
    assert [r for r, _ in ids] == sorted(r for r, _ in ids)
    s.drain()
    infos = s.poll()
    assert len(infos) == len(ps) and all(i.status == abi.FB_OK for i in infos)
    for p, (_, lut) in zip(ps[-2:], ids[-2:]):
        pa = Atmosphere.build(builder, None, p)
        pa.wait()
        a = pa.assert_ready()
        got = fetch(s, lut, p)
        for k, v in got.items():
            assert np.array_equal(v, a.download(
                {abi.FB_LUT_TRANSMITTANCE: "transmittance", abi.FB_LUT_IRRADIANCE: "irradiance",
                 abi.FB_LUT_SCATTERING: "scattering"}[k])), k
        a.close()
    bad = ps[0].replace(scattering_nu_size=1)
    before = s.stats().submitted
    with pytest.raises(Exception):
        s.submit_precomputes([ps[1], bad])
    assert s.stats().submitted == before
    s.close()


@pytest.mark.parametrize("admit_min_free", [3, 64])
def test_admission_hysteresis(builder, admit_min_free):
    """admit_min_free: with precomputes in flight the next group waits for
    that many free slots (64 > max_inflight: only when all are free); every
    request still completes, in more than one admission when the submission
    exceeds the slots, and the sets equal plain builds."""
    rng = np.random.default_rng(12)
    ps = [random_atmosphere(rng, True, resolution="half") for _ in range(13)]
    print(f"admit_min_free={admit_min_free}: 13 half sets, 4 slots", flush=True)
    s = Server(builder, 8 * GB, max_inflight=4, keep_outputs=True, n_streams=2,
               admit_min_free=admit_min_free)
    ids = s.submit_precomputes(ps)
    s.drain()
    infos = s.poll()
    assert len(infos) == len(ps) and all(i.status == abi.FB_OK for i in infos)
    starts = sorted({round(i.t_start, 6) for i in infos})
    assert len(starts) >= 4  # 13 requests through 4 slots
    for p, (_, lut) in zip(ps[:1] + ps[-1:], ids[:1] + ids[-1:]):
        pa = Atmosphere.build(builder, None, p)
        pa.wait()
        a = pa.assert_ready()
        got = fetch(s, lut, p)
        for k, v in got.items():
            assert np.array_equal(v, a.download(
                {abi.FB_LUT_TRANSMITTANCE: "transmittance", abi.FB_LUT_IRRADIANCE: "irradiance",
                 abi.FB_LUT_SCATTERING: "scattering"}[k])), k
        a.close()
    s.close()
