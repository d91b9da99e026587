"""LUT files (fb_atmosphere_save / fb_atmosphere_load): the reference dump's
OpenEXR images (examples/dump.rs:110-126, 175-193, 331-367), read here by an
independent OpenEXR parser written from the published file layout, and
sets reloaded from them render bit-identically."""
import ctypes as C
import math
import struct

import numpy as np
import pytest

from fuzzyblue_amd import Atmosphere, DeviceBuffer, Parameters, Renderer, abi, synchronize
from fuzzyblue_amd.lib import FuzzyblueError, load
from fuzzyblue_amd.params import DUMP_SIZES

pytestmark = pytest.mark.gpu


def read_exr(path):
    """Uncompressed single-part scanline OpenEXR: (attributes, {channel:
    (H, W) array})."""
    data = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", data, 0)
    assert magic == 20000630 and version == 2
    pos = 8
    attrs = {}

    def cstr(p):
        e = data.index(b"\0", p)
        return data[p:e].decode(), e + 1

    while True:
        name, pos = cstr(pos)
        if not name:
            break
        typ, pos = cstr(pos)
        (size,) = struct.unpack_from("<i", data, pos)
        pos += 4
        attrs[name] = (typ, data[pos:pos + size])
        pos += size
    chans = []
    typ, raw = attrs["channels"]
    assert typ == "chlist"
    p = 0
    while raw[p] != 0:
        e = raw.index(b"\0", p)
        cname = raw[p:e].decode()
        ptype, _, xs, ys = struct.unpack_from("<iIii", raw, e + 1)
        assert xs == 1 and ys == 1
        chans.append((cname, ptype))
        p = e + 1 + 16
    assert attrs["compression"][1] == b"\0"
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{h}Q", data, pos)
    out = {c: np.zeros((h, w), np.float32) for c, _ in chans}
    for y in range(h):
        q = offsets[y]
        yy, size = struct.unpack_from("<ii", data, q)
        assert yy == y
        q += 8
        for cname, ptype in chans:
            dt = np.float16 if ptype == 1 else np.float32
            n = w * np.dtype(dt).itemsize
            out[cname][y] = np.frombuffer(data, dt, w, q).astype(np.float32)
            q += n
        assert q == offsets[y] + 8 + size
    return attrs, out, chans


def stack_layers(chan, depth):
    """(depth, H, W, 4) from channels named like examples/dump.rs:358-367."""
    names = "RGBA"
    if depth == 1:
        return np.stack([chan[c] for c in names], axis=-1)[None]
    return np.stack([np.stack([chan[f"{z}.{c}"] for c in names], axis=-1)
                     for z in range(depth)])


def build(builder, p):
    pa = Atmosphere.build(builder, None, p)
    pa.wait()
    return pa.assert_ready()


def render(builder, a, frames=2, px=50000):
    alt = np.array([1.0, 8.0], np.float32)[:frames]
    zen = np.radians(np.array([30.0, 93.0], np.float32))[:frames]
    r = Renderer(builder)
    d_alt, d_zen = DeviceBuffer.from_array(alt), DeviceBuffer.from_array(zen)
    rad, tr = DeviceBuffer(frames * px * 16), DeviceBuffer(frames * px * 16)
    r.frames(None, a, d_alt, d_zen, frames, px, 5, rad, tr)
    synchronize()
    return rad.to_array(np.float32, (frames, px, 4)), tr.to_array(np.float32, (frames, px, 4))


# Extents that are no powers of two (the file's rows, layers and channel
# blocks then have odd strides).
ODD_SIZES = dict(transmittance_mu_size=250, transmittance_r_size=60, scattering_r_size=4,
                 scattering_mu_size=36, scattering_mu_s_size=26, scattering_nu_size=18,
                 irradiance_mu_s_size=60, irradiance_r_size=14)


@pytest.mark.parametrize("sizes", ["dump", "odd"])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_save_load_round_trip(builder, tmp_path, storage, sizes):
    flags = abi.FB_FLAG_PRECISE | (abi.FB_FLAG_F16_STORAGE if storage == "f16" else 0)
    p = Parameters().replace(**(DUMP_SIZES if sizes == "dump" else ODD_SIZES), order=3,
                             flags=flags)
    print(f"\n[lutio] {sizes} {storage} extents {p.scattering_extent()}", flush=True)
    a = build(builder, p)
    fp16 = storage == "f16"
    paths = a.save_all(str(tmp_path / "set_"), fp16=fp16)
    # the files, read independently
    tw, th = p.transmittance_extent()
    sw, sh, sd = p.scattering_extent()
    attrs, ch, chans = read_exr(paths[0])
    assert [c for c, _ in chans] == ["A", "B", "G", "R"] and {t for _, t in chans} == {2}
    assert np.array_equal(stack_layers(ch, 1)[0], a.download("transmittance"))
    assert attrs["fuzzyblueLut"][1] == b"transmittance"
    assert attrs["fuzzyblueParameters"][1] == bytes(p.to_c())
    _, ch, _ = read_exr(paths[1])
    assert np.array_equal(stack_layers(ch, 1)[0], a.download("irradiance"))
    _, ch, chans = read_exr(paths[2])
    assert len(chans) == 4 * sd and chans[:5] == [("0.A", 1 if fp16 else 2), ("0.B", 1 if fp16 else 2),
                                                   ("0.G", 1 if fp16 else 2), ("0.R", 1 if fp16 else 2),
                                                   ("1.A", 1 if fp16 else 2)]
    s_file = stack_layers(ch, sd)
    want = a.download("scattering", fp16=fp16).astype(np.float32)
    assert s_file.shape == (sd, sh, sw, 4)
    assert np.array_equal(s_file, want)
    # reload: the same LUTs, and renders bit-identical
    b_ = Atmosphere.load(builder, *paths)
    for name in ("transmittance", "irradiance", "scattering"):
        assert np.array_equal(b_.download(name), a.download(name)), name
    assert b_.parameters() == a.parameters()
    ra, ta = render(builder, a)
    rb, tb = render(builder, b_)
    assert np.array_equal(ra, rb) and np.array_equal(ta, tb)
    b_.close()
    a.close()


def test_half_file_into_float_storage_and_errors(builder, tmp_path):
    """A half scattering file loads into an FP32-storage set as the exact
    widening of the half values; files of another atmosphere or LUT are
    refused."""
    p = Parameters().replace(**DUMP_SIZES, order=2, flags=abi.FB_FLAG_PRECISE)
    a = build(builder, p)
    t, e, s = a.save_all(str(tmp_path / "h_"), fp16=True)
    b_ = Atmosphere.load(builder, t, e, s)
    assert np.array_equal(b_.download("scattering"),
                          a.download("scattering", fp16=True).astype(np.float32))
    b_.close()
    q = p.replace(order=3, ground_albedo=[0.3, 0.3, 0.3])
    other = build(builder, q)
    t2, _, _ = other.save_all(str(tmp_path / "o_"))
    with pytest.raises(FuzzyblueError) as exc:
        Atmosphere.load(builder, t2, e, s)
    assert exc.value.status == abi.FB_EINVAL
    with pytest.raises(FuzzyblueError):
        Atmosphere.load(builder, e, t, s)  # LUTs swapped
    with pytest.raises(FuzzyblueError):
        Atmosphere.load(builder, t, e, str(tmp_path / "missing.exr"))
    other.close()
    a.close()
