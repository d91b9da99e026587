"""fb_atmosphere_build_into: the precompute into caller-owned output LUTs
with every transient LUT in the library's static device scratch (no device
allocation).  Called through the C ABI; its LUTs must equal
fb_atmosphere_build's bit for bit, on two calls in a row that share the
scratch, the second with smaller LUTs."""
import ctypes as C

import numpy as np
import pytest

from fuzzyblue_amd import Atmosphere, DeviceBuffer, Parameters, abi, synchronize
from fuzzyblue_amd.lib import load
from fuzzyblue_amd.params import DUMP_SIZES, RESOLUTIONS

pytestmark = pytest.mark.gpu


def lut_shapes(p):
    tw, th = p.transmittance_extent()
    iw, ih = p.irradiance_extent()
    sw, sh, sd = p.scattering_extent()
    return (th, tw, 4), (ih, iw, 4), (sd, sh, sw, 4)


def build_into(builder, p):
    ts, is_, ss = lut_shapes(p)
    f16 = bool(p.flags & abi.FB_FLAG_F16_STORAGE)
    sbytes = int(np.prod(ss)) * (2 if f16 else 4)
    bufs = [DeviceBuffer(int(np.prod(ts)) * 4), DeviceBuffer(int(np.prod(is_)) * 4),
            DeviceBuffer(sbytes)]
    for b in bufs:
        b.zero()
    c = p.to_c()
    st = load().fb_atmosphere_build_into(builder.handle, None, C.byref(c), bufs[0].ptr,
                                         bufs[0].nbytes, bufs[1].ptr, bufs[1].nbytes,
                                         bufs[2].ptr, bufs[2].nbytes)
    assert st == abi.FB_OK, load().fb_last_error()
    synchronize()
    out = (bufs[0].to_array(np.float32, ts), bufs[1].to_array(np.float32, is_),
           bufs[2].to_array(np.float16 if f16 else np.float32, ss))
    for b in bufs:
        b.close()
    return out


def build_plain(builder, p):
    f16 = bool(p.flags & abi.FB_FLAG_F16_STORAGE)
    pa = Atmosphere.build(builder, None, p)
    pa.wait()
    a = pa.assert_ready()
    out = (a.download("transmittance"), a.download("irradiance"),
           a.download("scattering", fp16=f16))
    a.close()
    return out


def test_build_into_matches_build(builder):
    precise = abi.FB_FLAG_PRECISE
    cases = [
        Parameters().replace(flags=precise),                       # default sizes, 93 MiB
        Parameters().replace(**DUMP_SIZES, order=5, flags=precise,
                             rayleigh_scattering=[0.007, 0.015, 0.03]),  # smaller, reuses it
        Parameters().replace(**DUMP_SIZES, order=3),               # FP32 mode
        # extents that are no powers of two (unaligned scratch offsets)
        Parameters().replace(transmittance_mu_size=250, transmittance_r_size=60,
                             scattering_r_size=4, scattering_mu_size=36,
                             scattering_mu_s_size=26, scattering_nu_size=18,
                             irradiance_mu_s_size=60, irradiance_r_size=14, order=4,
                             flags=precise),
        Parameters().replace(flags=precise | abi.FB_FLAG_F16_STORAGE),
    ]
    for p in cases:
        print(f"build_into {p.scattering_extent()} order {p.order} flags {p.flags}", flush=True)
        got = build_into(builder, p)
        want = build_plain(builder, p)
        for g, w, name in zip(got, want, ("transmittance", "irradiance", "scattering")):
            assert np.array_equal(g, w), (p.flags, name)


def test_build_into_rejects_oversize_and_short_outputs(builder):
    lib = load()
    limit = lib.fb_build_into_scratch_limit()
    assert limit == 96 << 20
    big = Parameters().replace(**RESOLUTIONS["double"]).to_c()
    one = DeviceBuffer(16)
    st = lib.fb_atmosphere_build_into(builder.handle, None, C.byref(big), one.ptr, 1 << 40,
                                      one.ptr, 1 << 40, one.ptr, 1 << 40)
    assert st == abi.FB_EINVAL and b"scratch" in lib.fb_last_error()
    small = Parameters().replace(**DUMP_SIZES).to_c()
    st = lib.fb_atmosphere_build_into(builder.handle, None, C.byref(small), one.ptr, 16,
                                      one.ptr, 16, one.ptr, 16)
    assert st == abi.FB_EINVAL
    synchronize()
    one.close()
