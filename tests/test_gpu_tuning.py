"""Block-size autotuner (fb_builder_autotune / get / set): every block shape
of the tunable passes must give bit-identical LUTs, plain and batched."""
import numpy as np
import pytest

from fuzzyblue_amd import Atmosphere, Builder, Parameters, abi
from fuzzyblue_amd.lib import FuzzyblueError
from fuzzyblue_amd.params import DUMP_SIZES, RESOLUTIONS

pytestmark = pytest.mark.gpu

LUTS = ("transmittance", "irradiance", "scattering")


def build(b, p):
    pa = Atmosphere.build(b, None, p)
    pa.wait()
    out = {k: pa.atmosphere().download(k) for k in LUTS}
    pa.close()
    return out


def same(a, b):
    for k in LUTS:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("precise", [True, False])
def test_every_block_shape_gives_identical_luts(precise):
    b = Builder(0)
    p = Parameters().replace(**DUMP_SIZES, order=3,
                             flags=abi.FB_FLAG_PRECISE if precise else 0)
    tuned, t = b.tuning(p)
    assert not tuned
    base = build(b, p)
    remap_all = (1 << abi.FB_PASS_SINGLE_SCATTERING) | (1 << abi.FB_PASS_SCATTERING_DENSITY) \
        | (1 << abi.FB_PASS_MULTIPLE_SCATTERING)
    for threads, waves, remap in ((64, 1, 0), (128, 2, remap_all), (256, 4, remap_all)):
        b.set_tuning(p, single_scattering_threads=threads, density_order2_threads=threads,
                     density_threads=threads, indirect_order1_waves=waves,
                     indirect_waves=waves, multiple_scattering_threads=threads,
                     xcd_remap=remap)
        tuned, t = b.tuning(p)
        assert tuned and t["density_threads"] == threads and t["indirect_waves"] == waves
        same(base, build(b, p))
    chosen = b.autotune(p)
    print(f"\nautotune {'precise' if precise else 'fp32'}: {chosen}", flush=True)
    assert chosen["multiple_scattering_threads"] in (64, 128, 256)
    assert all(ms > 0 for ms in chosen["best_ms"])
    same(base, build(b, p))
    b.close()


def test_set_tuning_rejects_invalid_shapes():
    b = Builder(0)
    p = Parameters().replace(**RESOLUTIONS["double"])  # 1024 texels per row
    with pytest.raises(FuzzyblueError):
        b.set_tuning(p, density_threads=96)
    with pytest.raises(FuzzyblueError):
        b.set_tuning(p, multiple_scattering_threads=128)  # > 4 texels per thread
    with pytest.raises(FuzzyblueError):
        b.set_tuning(p, indirect_waves=8)
    with pytest.raises(FuzzyblueError):
        b.set_tuning(p, xcd_remap=1 << abi.FB_PASS_TRANSMITTANCE)
    with pytest.raises(FuzzyblueError):
        b.set_tuning(p, multiple_scattering_threads=2048)
    with pytest.raises(FuzzyblueError):
        b.set_tuning(p, single_scattering_threads=512)  # only multiple scattering goes past 256
    b.set_tuning(p, multiple_scattering_threads=256)
    b.set_tuning(p, multiple_scattering_threads=1024)  # one texel per thread
    small = Parameters().replace(**DUMP_SIZES)  # 64 texels per row
    with pytest.raises(FuzzyblueError):
        b.set_tuning(small, multiple_scattering_threads=512)
    b.close()


def test_batch_uses_tuning_and_matches_plain():
    b = Builder(0)
    ps = [Parameters().replace(**RESOLUTIONS["half"], order=o, flags=abi.FB_FLAG_PRECISE)
          for o in (2, 3)]
    ps.append(Parameters().replace(**DUMP_SIZES, order=3, flags=abi.FB_FLAG_PRECISE))
    plain = [build(b, p) for p in ps]
    for p in ps:
        b.set_tuning(p, density_threads=64, density_order2_threads=128,
                     multiple_scattering_threads=64, single_scattering_threads=128,
                     indirect_waves=1, indirect_order1_waves=2)
    pend = Atmosphere.build_batch(b, None, ps)
    for pa, ref in zip(pend, plain):
        pa.wait()
        same(ref, {k: pa.atmosphere().download(k) for k in LUTS})
    for pa in pend:
        pa.close()
    b.close()
