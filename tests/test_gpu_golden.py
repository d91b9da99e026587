"""The default Earth LUT set on the GPU against the committed FP64-oracle
fixture (tests/golden/earth_default_fp64.npz, from tests/golden/make_golden.py):
precision mode must stay within 1e-5 (transmittance) and 1e-3 (irradiance,
scattering) relative error, floor 1e-6 x max."""
import os

import numpy as np
import pytest

from fuzzyblue_amd import Atmosphere, Parameters, abi

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def rel(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return np.abs(x - y) / np.maximum(np.abs(y), 1e-6 * np.abs(y).max())


def test_default_earth_precise_vs_fp64_fixture(builder):
    g = np.load(os.path.join(HERE, "golden", "earth_default_fp64.npz"))
    pa = Atmosphere.build(builder, None, Parameters().replace(flags=abi.FB_FLAG_PRECISE))
    pa.wait()
    a = pa.atmosphere()
    T = a.download("transmittance")[..., :3]
    E = a.download("irradiance")[..., :3]
    S = a.download("scattering").reshape(-1, 4)[g["S_idx"]]
    eT, eE, eS = rel(T, g["T"]).max(), rel(E, g["E"]).max(), rel(S[:, :3], g["S"][:, :3]).max()
    eA = rel(S[:, 3], g["S"][:, 3]).max()
    print(f"\ndefault precise vs FP64: T {eT:.2e} E {eE:.2e} S {eS:.2e} S.a {eA:.2e}", flush=True)
    assert eT < 1e-5
    assert eE < 1e-3
    assert eS < 1e-3
    assert eA < 1e-3
    pa.close()


def test_default_earth_per_order_vs_fp64_fixture(builder):
    """Per-order irradiance deltas: after an order-n build the delta
    irradiance holds the indirect irradiance of order n-1."""
    g = np.load(os.path.join(HERE, "golden", "earth_default_fp64.npz"))
    for order in (2, 3, 4):
        pa = Atmosphere.build(builder, None,
                              Parameters().replace(order=order, flags=abi.FB_FLAG_PRECISE))
        pa.wait()
        dE = pa.download_scratch("delta_irradiance")[..., :3]
        e = rel(dE, g["orders_E"][order - 1]).max()
        print(f"order {order}: delta irradiance vs FP64 {e:.2e}", flush=True)
        assert e < 1e-3
        pa.close()
