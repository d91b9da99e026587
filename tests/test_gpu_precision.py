"""Precision workload: the Earth atmosphere at twice the default resolution
(transmittance 512x128, irradiance 128x32, scattering 64x256x64x16) and
scattering order 8, in precision mode, against the FP64-oracle fixture
tests/golden/earth_double_order8_fp64.npz (tests/golden/make_golden_double.py).

Relative error |x - x64| / max(|x64|, 1e-6 max|LUT64|) per channel, reported
per LUT and per scattering order, separately for texels whose view cosine
lies within 1e-3 of the horizon and for the rest.  Bounds: 1e-5 on
transmittance, 1e-3 on irradiance and scattering, every order, both strata.
"""
import os

import numpy as np
import pytest

from fuzzyblue_amd import Atmosphere, Parameters, abi
from fuzzyblue_amd.params import RESOLUTIONS

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "earth_double_order8_fp64.npz")


def rel(x, y, scale):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return np.abs(x - y) / np.maximum(np.abs(y), 1e-6 * scale)


def stats(e, band):
    return {"band_max": float(e[band].max()), "band_mean": float(e[band].mean()),
            "else_max": float(e[~band].max()), "else_mean": float(e[~band].mean())}


def build_order(builder, order):
    p = Parameters().replace(**RESOLUTIONS["double"], order=order, flags=abi.FB_FLAG_PRECISE)
    pa = Atmosphere.build(builder, None, p)
    pa.wait()
    return pa


@pytest.fixture(scope="module")
def golden():
    if not os.path.exists(FIXTURE):
        pytest.skip("fixture not generated")
    return np.load(FIXTURE)


def test_double_order8_luts_vs_fp64(builder, golden):
    g = golden
    pa = build_order(builder, 8)
    a = pa.atmosphere()
    T = a.download("transmittance")[..., :3]
    E = a.download("irradiance")[..., :3]
    S = a.download("scattering").reshape(-1, 4)[g["S_idx"]]
    pa.close()
    band = g["S_band"]
    eT = rel(T, g["T"], np.abs(g["T"]).max()).max()
    eE = rel(E, g["E"], np.abs(g["E"]).max()).max()
    sc = np.abs(g["S"][:, :3]).max(axis=0)
    eS = rel(S[:, :3], g["S"][:, :3], sc).max(axis=1)
    eA = rel(S[:, 3], g["S"][:, 3], np.abs(g["S"][:, 3]).max())
    print(f"\ndouble/order 8 vs FP64: T {eT:.2e}  E {eE:.2e}  S {stats(eS, band)}  "
          f"S.a {stats(eA, band)}", flush=True)
    assert eT < 1e-5
    assert eE < 1e-3
    assert eS.max() < 1e-3
    assert eA.max() < 1e-3


def test_double_per_order_vs_fp64(builder, golden):
    """After an order-n build the scratch holds the order-n terms: delta
    multiple scattering (order 1: single Rayleigh) and delta irradiance
    (order 1: direct; n >= 2: the indirect irradiance of order n-1)."""
    g = golden
    band = g["S_band"][g["O_sel"]]
    idx = g["S_idx"][g["O_sel"]]
    for order in range(1, 9):
        pa = build_order(builder, order)
        which = "delta_rayleigh" if order == 1 else "delta_multiple_scattering"
        s = pa.download_scratch(which).reshape(-1, 4)[idx, :3]
        e_irr = pa.download_scratch("delta_irradiance")[..., :3]
        pa.close()
        ref_s = g["orders_S"][order - 1]
        ref_e = g["orders_E"][order - 1]
        es = rel(s, ref_s, np.abs(ref_s).max(axis=0)).max(axis=1)
        ee = rel(e_irr, ref_e, np.abs(ref_e).max()).max()
        print(f"order {order}: scattering {stats(es, band)}  irradiance max {ee:.2e}",
              flush=True)
        assert es.max() < 1e-3, order
        assert ee < 1e-3, order
