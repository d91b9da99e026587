"""Edge shapes of the LUTs on the GPU, end to end against the oracle:

  * ragged: every extent small and none a multiple of a block or tile size
    (transmittance 7x5, irradiance 5x3, scattering 5x6x3x3), so every grid's
    tail blocks and partial tiles run (the reference's extent/WG dispatches
    would leave such texels unwritten, src/precompute.rs:1742-1744; the
    build uses ceil-div grids);
  * widenu: the largest nu extent the kernels accept (32 slices);
  * widerow: the widest scattering row (NU x MUS = 1024 texels);
  * pairs: enough texels (67,392) for one thread per texel in the density
    pass, with NU even (18, not a power of two), so each thread takes two
    texels (mu, nu) and (mu, nu + 9); an (r, mu_s) tile's 324 threads end
    in a partial block.

FP32 mode against the FP32 restatement and precision mode against the FP64
oracle, both within 1e-3 (floor 1e-6 x max; transmittance 1e-5), and the
binary16 storage mode against the restatement in binary16 mode within 3e-3
(floor 1e-3 x max).

Except for "pairs" in FP32 and binary16 modes, which are held to an error
bound relative to the LUT's maximum (1e-5 x max FP32, 1e-3 x max binary16;
measured 3.2e-6 and 5.2e-4).  With 18 nu slices of 26 mu_s texels the
binary32 texture coordinate ((slice + u_mu_s) / NU) * W - 0.5
(scattering.h:149-154) of a lookup at mu_s_min rounds up to a few 1e-5 of
a texel into the previous nu slice, whose last texel (mu_s_max) is up to
1e7 times brighter: the darkest texels then differ between any two binary32
evaluations that round differently.  The FP32 restatement itself is 2.4e-2
x max from the FP64 oracle at this shape, and precision mode, which
computes the coordinates in FP64, meets the relative bar.
"""
import numpy as np
import pytest

from fuzzyblue_amd import Atmosphere, Parameters, abi
from oracle import Oracle

pytestmark = pytest.mark.gpu

SHAPES = {
    "ragged": dict(transmittance_mu_size=7, transmittance_r_size=5, scattering_r_size=5,
                   scattering_mu_size=6, scattering_mu_s_size=3, scattering_nu_size=3,
                   irradiance_mu_s_size=5, irradiance_r_size=3),
    "widenu": dict(scattering_r_size=4, scattering_mu_size=16, scattering_mu_s_size=4,
                   scattering_nu_size=32),
    "widerow": dict(scattering_r_size=3, scattering_mu_size=8, scattering_mu_s_size=64,
                    scattering_nu_size=16),
    "pairs": dict(scattering_r_size=4, scattering_mu_size=36, scattering_mu_s_size=26,
                  scattering_nu_size=18),
}


# Bounds relative to the LUT's maximum for (shape, mode) pairs whose binary32
# texture coordinates leak across nu slices (see the module docstring).
MAX_BOUND = {("pairs", "fp32"): 1e-5, ("pairs", "f16"): 1e-3}


def rel_err(x, y, floor):
    x = np.asarray(x, np.float64)[..., :3]
    y = np.asarray(y, np.float64)[..., :3]
    return np.abs(x - y) / np.maximum(np.abs(y), floor * max(np.abs(y).max(), 1e-300))


def outputs(builder, p):
    pa = Atmosphere.build(builder, None, p)
    pa.wait()
    a = pa.atmosphere()
    out = {k: a.download(k) for k in ("transmittance", "irradiance", "scattering")}
    pa.close()
    return out


@pytest.mark.parametrize("name", sorted(SHAPES))
@pytest.mark.parametrize("mode", ["fp32", "precise", "f16"])
def test_shape_vs_oracle(builder, name, mode):
    flags = {"fp32": 0, "precise": abi.FB_FLAG_PRECISE, "f16": abi.FB_FLAG_F16_STORAGE}[mode]
    p = Parameters().replace(**SHAPES[name], order=4, flags=flags)
    print(f"\n[shape] {name} {mode} extents {p.scattering_extent()}", flush=True)
    got = outputs(builder, p)
    ref = Oracle("f64" if mode == "precise" else "f32").precompute(p, fp16=mode == "f16")
    floor, tol = (1e-3, 3e-3) if mode == "f16" else (1e-6, 1e-3)
    eT = rel_err(got["transmittance"], ref["T"], 1e-6).max()
    if (name, mode) in MAX_BOUND:
        # floor 1: the error relative to the maximum
        floor, tol = 1.0, MAX_BOUND[(name, mode)]
    eE = rel_err(got["irradiance"], ref["E"], floor).max()
    eS = rel_err(got["scattering"], ref["S"], floor).max()
    print(f"  T {eT:.2e} E {eE:.2e} S {eS:.2e}", flush=True)
    assert eT < 1e-5
    assert eE < tol
    assert eS < tol
    for v in got.values():
        assert np.all(np.isfinite(v))
