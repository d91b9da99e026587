"""Lifetime of a finished LUT set under renders on several streams.

fb_atmosphere_destroy must wait for every render enqueued against the set,
on any stream, before its LUTs return to the builder's pool: the next set of
the same sizes reuses those blocks at once."""
import numpy as np
import pytest

from fuzzyblue_amd import Atmosphere, DeviceBuffer, Parameters, Renderer, synchronize
from fuzzyblue_amd.api import Stream
from fuzzyblue_amd.params import DUMP_SIZES

pytestmark = pytest.mark.gpu

P = Parameters().replace(**DUMP_SIZES)


def build(builder, p):
    pa = Atmosphere.build(builder, None, p)
    pa.wait()
    return pa.assert_ready()


def test_destroy_waits_for_renders_on_every_stream(builder):
    frames, px = 8, 1920 * 1080
    alt = np.linspace(0.5, 9.5, frames).astype(np.float32)
    zen = np.radians(np.linspace(5.0, 100.0, frames)).astype(np.float32)
    d_alt, d_zen = DeviceBuffer.from_array(alt), DeviceBuffer.from_array(zen)
    r = Renderer(builder)
    # single-stream reference
    a = build(builder, P)
    want = DeviceBuffer(frames * px * 16)
    r.frames(None, a, d_alt, d_zen, frames, px, 21, want)
    synchronize()
    ref = want.to_array(np.float32, (frames, px, 4))
    a.close()
    # two streams, destroy right after enqueueing, then reuse the pool blocks
    a = build(builder, P)
    s1, s2 = Stream(), Stream()
    out1, out2 = DeviceBuffer(frames * px * 16), DeviceBuffer(frames * px * 16)
    r.frames(s1, a, d_alt, d_zen, frames, px, 21, out1)
    r.frames(s2, a, d_alt, d_zen, frames, px, 21, out2)
    a.close()
    other = build(builder, P.replace(rayleigh_scattering=[0.05, 0.05, 0.05], order=2))
    s1.synchronize()
    s2.synchronize()
    assert np.array_equal(out1.to_array(np.float32, (frames, px, 4)), ref)
    assert np.array_equal(out2.to_array(np.float32, (frames, px, 4)), ref)
    other.close()
    s1.close()
    s2.close()
