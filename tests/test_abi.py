"""C-ABI surface on the CPU: the library loads, exports every declared
symbol, and the host-only entry points (defaults, extents, validation,
footprint) behave like the reference's Parameters (src/precompute.rs)."""
import ctypes as C

import numpy as np
import pytest

from fuzzyblue_amd import Parameters, abi
from fuzzyblue_amd.lib import check, declared_functions, load, FuzzyblueError


def test_library_exports_every_declared_symbol():
    lib = load()
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version():
    assert load().fb_abi_version() == abi.FB_ABI_VERSION


def test_defaults_match_reference_default():
    """fb_parameters_default == Parameters::default (precompute.rs:849-935)."""
    c = abi.fb_parameters()
    load().fb_parameters_default(C.byref(c))
    assert bytes(c) == bytes(Parameters().to_c())
    assert c.order == 4 and c.scattering_nu_size == 8 and c.bottom_radius == np.float32(6360.0)
    assert c.mie_density.layers[1].exp_scale == np.float32(-0.833333)
    assert c.absorption_density.layers[0].width == np.float32(25.0)
    assert c.mu_s_min == np.float32(-0.207912)


def test_extents():
    """Parameters::*_extent (precompute.rs:771-793)."""
    p = Parameters()
    c = p.to_c()
    t = (C.c_uint32 * 2)()
    i = (C.c_uint32 * 2)()
    s = (C.c_uint32 * 3)()
    load().fb_parameters_extents(C.byref(c), t, i, s)
    assert tuple(t) == p.transmittance_extent() == (256, 64)
    assert tuple(i) == p.irradiance_extent() == (64, 16)
    assert tuple(s) == p.scattering_extent() == (256, 128, 32)


def test_footprint():
    c = Parameters().to_c()
    res = C.c_size_t()
    scr = C.c_size_t()
    tot = load().fb_parameters_footprint(C.byref(c), C.byref(res), C.byref(scr))
    ns = 256 * 128 * 32
    assert res.value == (256 * 64 + 64 * 16 + ns) * 16
    # optical depths (transmittance-sized), irradiance delta, four
    # scattering-sized scratch LUTs and the multiple-scattering row table
    # (4 float4 x 51 samples per (r, mu) row)
    # the density geometry table (33 float4 per r, NU + 2 per mu_s) and the
    # single-scattering row table (5 float4 x 51 samples per row)
    rows = 32 * 128
    assert scr.value == (256 * 64 + 64 * 16 + 4 * ns + rows * 4 * 51 + 32 * 33 + 32 * 10
                         + rows * 5 * 51) * 16
    assert tot == res.value + scr.value


def test_footprint_f16_storage():
    """FB_FLAG_F16_STORAGE: the scattering LUT and the four scattering-sized
    scratch LUTs take 8 bytes per texel (R16G16B16A16_SFLOAT,
    src/precompute.rs:1215); everything else is unchanged."""
    lib = load()
    full = [C.c_size_t(), C.c_size_t()]
    half = [C.c_size_t(), C.c_size_t()]
    p = Parameters()
    lib.fb_parameters_footprint(C.byref(p.to_c()), C.byref(full[0]), C.byref(full[1]))
    h = p.replace(flags=abi.FB_FLAG_F16_STORAGE | abi.FB_FLAG_PRECISE)
    check(lib.fb_parameters_validate(C.byref(h.to_c())))
    lib.fb_parameters_footprint(C.byref(h.to_c()), C.byref(half[0]), C.byref(half[1]))
    ns = 256 * 128 * 32
    assert full[0].value - half[0].value == ns * 8
    assert full[1].value - half[1].value == 4 * ns * 8


@pytest.mark.parametrize(
    "field,value",
    [
        ("scattering_nu_size", 1),
        ("scattering_mu_size", 2),
        ("transmittance_mu_size", 0),
        ("irradiance_r_size", 1),
        ("order", 0),
        ("top_radius", 6000.0),
        ("bottom_radius", -1.0),
        ("mie_phase_function_g", 1.0),
        ("mu_s_min", 0.1),
        ("sun_angular_radius", 0.2),
        ("flags", 0x80),
        ("flags", 0x4),
        ("scattering_nu_size", 1000),
    ],
)
def test_validation_rejects(field, value):
    p = Parameters().replace(**{field: value})
    st = load().fb_parameters_validate(C.byref(p.to_c()))
    assert st == abi.FB_EINVAL
    assert load().fb_last_error()


def test_validation_accepts_defaults_and_smoke_sizes():
    from fuzzyblue_amd.params import SMOKE_SIZES, RESOLUTIONS

    for p in [Parameters(), Parameters().replace(**SMOKE_SIZES)] + [
        Parameters().replace(**r) for r in RESOLUTIONS.values()
    ]:
        check(load().fb_parameters_validate(C.byref(p.to_c())))


def test_validation_rejects_luts_beyond_32bit_offsets():
    """The kernels address a scattering-sized LUT and the transmittance LUT
    with 32-bit byte offsets: sets whose LUTs reach 2 GiB are rejected
    (FB_EINVAL), in both storage modes; the largest set below the limit is
    accepted."""
    lib = load()
    # 64 x 512 x 64 x 16 texels = 2^25: 512 MiB FP32, accepted
    ok = Parameters().replace(scattering_r_size=64, scattering_mu_size=512,
                              scattering_mu_s_size=64, scattering_nu_size=16)
    check(lib.fb_parameters_validate(C.byref(ok.to_c())))
    # 2^27 texels: exactly 2 GiB as RGBA32F, 1 GiB as binary16
    big = ok.replace(scattering_mu_size=2048)
    assert lib.fb_parameters_validate(C.byref(big.to_c())) == abi.FB_EINVAL
    assert b"2 GiB" in lib.fb_last_error()
    check(lib.fb_parameters_validate(C.byref(big.replace(flags=abi.FB_FLAG_F16_STORAGE).to_c())))
    bigger = big.replace(scattering_mu_size=4096, flags=abi.FB_FLAG_F16_STORAGE)
    assert lib.fb_parameters_validate(C.byref(bigger.to_c())) == abi.FB_EINVAL
    # transmittance: 65536 x 2048 texels x 16 B = 2 GiB
    t = Parameters().replace(transmittance_mu_size=65536, transmittance_r_size=2048)
    assert lib.fb_parameters_validate(C.byref(t.to_c())) == abi.FB_EINVAL
    check(lib.fb_parameters_validate(C.byref(t.replace(transmittance_r_size=2047).to_c())))


def test_negative_coefficient_rejected():
    p = Parameters()
    p.rayleigh_scattering = [-1.0, 0.0, 0.0]
    assert load().fb_parameters_validate(C.byref(p.to_c())) == abi.FB_EINVAL


def test_status_strings():
    lib = load()
    for st in abi.STATUS_NAMES:
        assert lib.fb_status_string(st).startswith(abi.STATUS_NAMES[st].encode())


def test_params_roundtrip():
    p = Parameters().replace(order=7, mie_phase_function_g=0.7)
    q = Parameters.from_c(p.to_c())
    assert bytes(q.to_c()) == bytes(p.to_c())


def test_ray_direction_unit_and_deterministic():
    from fuzzyblue_amd import ray_direction

    for i in range(50):
        d = ray_direction(2, 3, i)
        assert abs(float(np.linalg.norm(d.astype(np.float64))) - 1.0) < 1e-6
        assert np.array_equal(d, ray_direction(2, 3, i))


def test_mixed_workload_matches_committed_fixture():
    """workloads.mixed_batch(256, seed 4) is pinned by a committed JSON copy."""
    import json
    import os

    from fuzzyblue_amd.workloads import mixed_batch

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "mixed256_seed4.json")
    with open(path) as f:
        fixed = [Parameters.from_dict(d) for d in json.load(f)]
    assert fixed == mixed_batch(256, 4, precise=True)


def test_large_fixture_members_are_mixed_set_members():
    """tests/golden/mixed_large_fp64.npz holds members of the committed
    256-atmosphere set: its parameters equal mixed256_seed4.json at
    mix_index, three at "default" and three at "double" resolution."""
    import json
    import os

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    g = np.load(os.path.join(here, "mixed_large_fp64.npz"))
    mix = json.load(open(os.path.join(here, "mixed256_seed4.json")))
    ps = json.loads(str(g["params"]))
    assert [mix[int(i)] for i in g["mix_index"]] == ps
    assert sorted(p["scattering_nu_size"] for p in ps) == [8, 8, 8, 16, 16, 16]
