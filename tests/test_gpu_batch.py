"""Batched precompute (fb_atmosphere_build_batch): every atmosphere of a
mixed batch — uneven LUT sizes, orders and parameters in one launch per
pass and kernel variant — is bit-identical to the same atmosphere built alone (the batched
kernels run the same per-block code on the same inputs), the captured graph
replays to the same bits, and a batch member agrees with the oracle."""
import numpy as np
import pytest

from fuzzyblue_amd import Atmosphere, Parameters, abi
from fuzzyblue_amd.params import DUMP_SIZES, RESOLUTIONS, SMOKE_SIZES
from fuzzyblue_amd.workloads import random_atmosphere
from oracle import Oracle

pytestmark = pytest.mark.gpu

KEYS = ("transmittance", "irradiance", "scattering")


def small_mix(precise):
    rng = np.random.default_rng(21)
    ps = [
        Parameters().replace(**SMOKE_SIZES, order=2),
        Parameters().replace(**DUMP_SIZES, order=5),
        random_atmosphere(rng, precise, resolution="half", order=3),
        random_atmosphere(rng, precise, resolution="half", order=7),
        random_atmosphere(rng, precise, resolution="half", order=2),
        Parameters().replace(**RESOLUTIONS["half"], order=4, ground_albedo=[0.3, 0.2, 0.1]),
        # rows of 32 texels: chunked multiple scattering, as in a plain build
        Parameters().replace(scattering_r_size=8, scattering_mu_size=32,
                             scattering_mu_s_size=8, scattering_nu_size=4, order=4),
        # two density texels per thread, a tile's 324 threads ending in a
        # partial block
        Parameters().replace(scattering_r_size=4, scattering_mu_size=36,
                             scattering_mu_s_size=26, scattering_nu_size=18, order=3),
        # extents that fill no grid or tile evenly
        Parameters().replace(transmittance_mu_size=7, transmittance_r_size=5,
                             scattering_r_size=5, scattering_mu_size=6, scattering_mu_s_size=3,
                             scattering_nu_size=3, irradiance_mu_s_size=5, irradiance_r_size=3,
                             order=3),
    ]
    for p in ps:
        p.flags = abi.FB_FLAG_PRECISE if precise else 0
    return ps


@pytest.mark.parametrize("precise", [False, True], ids=["fp32", "precise"])
def test_batch_members_equal_single_builds(builder, precise):
    ps = small_mix(precise)
    print(f"\n[batch] {len(ps)} atmospheres, orders {[p.order for p in ps]}", flush=True)
    pend = Atmosphere.build_batch(builder, None, ps)
    for pa in pend:
        pa.wait()
    got = [{k: pa.atmosphere().download(k) for k in KEYS} for pa in pend]
    for p, g in zip(ps, got):
        pa = Atmosphere.build(builder, None, p)
        pa.wait()
        a = pa.atmosphere()
        for k in KEYS:
            assert np.array_equal(a.download(k), g[k]), (k, p.order)
        pa.close()
    # replaying the captured batch graph reproduces every member
    pend[0].relaunch(None)
    for pa in pend:
        pa.wait()
    for pa, g in zip(pend, got):
        for k in KEYS:
            assert np.array_equal(pa.atmosphere().download(k), g[k])
    for pa in pend:
        pa.close()


@pytest.mark.parametrize("precise", [False, True], ids=["fp32", "precise"])
def test_batch_default_size_members_equal_single_builds(builder, precise):
    """Default-size members (density with two texels per thread, beside
    half-resolution members with one) equal their plain builds bit for
    bit."""
    rng = np.random.default_rng(22)
    fl = abi.FB_FLAG_PRECISE if precise else 0
    ps = [Parameters().replace(order=3, flags=fl),
          random_atmosphere(rng, precise, resolution="default", order=4),
          random_atmosphere(rng, precise, resolution="half", order=3)]
    print(f"\n[batch] default-size members, precise={precise}", flush=True)
    pend = Atmosphere.build_batch(builder, None, ps)
    for pa in pend:
        pa.wait()
    for p, pb in zip(ps, pend):
        pa = Atmosphere.build(builder, None, p)
        pa.wait()
        for k in KEYS:
            assert np.array_equal(pa.atmosphere().download(k), pb.atmosphere().download(k)), k
        pa.close()
    for pa in pend:
        pa.close()


def test_batch_member_vs_fp64_oracle(builder):
    ps = small_mix(True)
    pend = Atmosphere.build_batch(builder, None, ps)
    pend[1].wait()
    out = {k: pend[1].atmosphere().download(k) for k in KEYS}
    ref = Oracle("f64").precompute(ps[1])
    for k, key, tol in (("transmittance", "T", 1e-5), ("irradiance", "E", 1e-3),
                        ("scattering", "S", 1e-3)):
        x = out[k][..., :3].astype(np.float64)
        y = ref[key][..., :3]
        e = (np.abs(x - y) / np.maximum(np.abs(y), 1e-6 * np.abs(y).max())).max()
        assert e < tol, (k, e)
    for pa in pend:
        pa.close()


def test_batch_relaunch_refused_after_member_released(builder):
    ps = small_mix(False)[:3]
    pend = Atmosphere.build_batch(builder, None, ps)
    for pa in pend:
        pa.wait()
    atm = pend[0].assert_ready()
    with pytest.raises(Exception):
        pend[1].relaunch(None)
    atm.close()
    for pa in pend[1:]:
        pa.close()


def test_mixed_half_members_vs_fp64_fixture(builder):
    """Eight random atmospheres of the mixed workload's ranges (half
    resolution, orders 2-8), built as one precision-mode batch, against their
    FP64-oracle LUTs (tests/golden/make_golden_mixed.py)."""
    import json
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "mixed_half_fp64.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    g = np.load(path)
    ps = [Parameters.from_dict(d) for d in json.loads(str(g["params"]))]
    pend = Atmosphere.build_batch(builder, None, ps)
    worst = {}
    for i, (p, pa) in enumerate(zip(ps, pend)):
        pa.wait()
        a = pa.atmosphere()
        got = {"T": a.download("transmittance")[..., :3],
               "E": a.download("irradiance")[..., :3],
               "S": a.download("scattering").reshape(-1, 4)[g["S_idx"]]}
        errs = {}
        for key in ("T", "E", "S"):
            x = got[key].astype(np.float64)
            y = g[f"{key}{i}"].astype(np.float64)
            scale = np.abs(y).reshape(-1, y.shape[-1]).max(axis=0)
            errs[key] = float((np.abs(x - y) / np.maximum(np.abs(y), 1e-6 * scale)).max())
            worst[key] = max(worst.get(key, 0.0), errs[key])
        print(f"\nmember {i} (order {p.order}, bottom {p.bottom_radius:.0f} km, thickness "
              f"{p.top_radius - p.bottom_radius:.0f} km): {errs}", flush=True)
    for pa in pend:
        pa.close()
    assert worst["T"] < 1e-5, worst
    assert worst["E"] < 1e-3, worst
    assert worst["S"] < 1e-3, worst


def test_mixed_members_all_resolutions_vs_fp64_fixture(builder):
    """The precision qualifier on the sizes that dominate the mixed set's
    time: six members of the seeded 256-atmosphere set itself at the
    "default" (32x128x32x8) and "double" (64x256x64x16) resolutions
    (tests/golden/make_golden_mixed_large.py), built as ONE precision-mode
    batch together with the eight "half" fixtures, each against its FP64
    oracle LUTs: max relative error <= 1e-3 (1e-5 for transmittance), floor
    1e-6 x max per channel."""
    import json
    import os

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    members = []
    for name in ("mixed_half_fp64.npz", "mixed_large_fp64.npz"):
        path = os.path.join(here, name)
        if not os.path.exists(path):
            pytest.skip(f"{name} not generated")
        g = np.load(path)
        for i, d in enumerate(json.loads(str(g["params"]))):
            idx = g[f"S_idx{i}"] if f"S_idx{i}" in g else g["S_idx"]
            members.append((Parameters.from_dict(d), g[f"T{i}"], g[f"E{i}"], g[f"S{i}"], idx))
    pend = Atmosphere.build_batch(builder, None, [m[0] for m in members])
    worst = {}
    for (p, T64, E64, S64, idx), pa in zip(members, pend):
        pa.wait()
        a = pa.atmosphere()
        got = {"T": (a.download("transmittance")[..., :3], T64),
               "E": (a.download("irradiance")[..., :3], E64),
               "S": (a.download("scattering").reshape(-1, 4)[idx], S64)}
        cls = {4: "half", 8: "default", 16: "double"}[p.scattering_nu_size]
        errs = {}
        for key, (x, y) in got.items():
            x = x.astype(np.float64)
            y = y.astype(np.float64)
            scale = np.abs(y).reshape(-1, y.shape[-1]).max(axis=0)
            errs[key] = float((np.abs(x - y) / np.maximum(np.abs(y), 1e-6 * scale)).max())
            w = worst.setdefault(cls, {})
            w[key] = max(w.get(key, 0.0), errs[key])
        print(f"\n{cls} member (order {p.order}, bottom {p.bottom_radius:.0f} km): {errs}",
              flush=True)
    for pa in pend:
        pa.close()
    print(f"\nworst by class: {worst}", flush=True)
    assert set(worst) == {"half", "default", "double"}
    for cls, w in worst.items():
        assert w["T"] < 1e-5, (cls, w)
        assert w["E"] < 1e-3, (cls, w)
        assert w["S"] < 1e-3, (cls, w)
