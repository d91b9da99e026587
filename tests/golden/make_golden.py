"""Generate the committed FP64-oracle fixtures under tests/golden/.

The reference ships no golden vectors (its only test checks for Vulkan
validation errors, tests/smoke.rs:169-171), so the fixtures come from this
repository's FP64 oracle (oracle/atmo_oracle.c), itself pinned by the
closed-form and invariant tests in tests/test_oracle_pinning.py.

earth_default_fp64.npz — Parameters::default() (src/precompute.rs:849-935),
order 4, evaluated in FP64:
  T          full transmittance LUT (64, 256, 3), stored as float32
  E          full irradiance LUT (16, 64, 3), float32
  S_idx      16384 seeded texel indices into the packed scattering LUT
  S          scattering RGBA at those texels (16384, 4), float32
  orders_E   per-order irradiance deltas (4, 16, 64, 3)
  orders_S   per-order scattering deltas at S_idx[:4096] (4, 4096, 3)
Run: python tests/golden/make_golden.py   (about 80 s on 8 cores)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fuzzyblue_amd.params import Parameters  # noqa: E402
from oracle import Oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    p = Parameters()
    r = Oracle("f64").precompute(p, per_order=True)
    S = r["S"].reshape(-1, 4)
    rng = np.random.default_rng(8)
    idx = np.sort(rng.choice(S.shape[0], size=16384, replace=False)).astype(np.int64)
    orders_S = r["orders_S"].reshape(p.order, -1, 4)[:, idx[:4096], :3]
    np.savez_compressed(
        os.path.join(HERE, "earth_default_fp64.npz"),
        T=r["T"][..., :3].astype(np.float32),
        E=r["E"][..., :3].astype(np.float32),
        S_idx=idx,
        S=S[idx].astype(np.float32),
        orders_E=r["orders_E"][..., :3].astype(np.float32),
        orders_S=orders_S.astype(np.float32),
    )
    print("wrote earth_default_fp64.npz")


if __name__ == "__main__":
    main()
