"""Generate the fixtures of the mixed 256-atmosphere workload.

mixed256_seed4.json — the parameter sets of workloads.mixed_batch(256,
seed=4) (precision mode), committed so the workload is pinned as data.

mixed_half_fp64.npz — FP64-oracle LUTs (oracle/atmo_oracle.c) of the first
8 "half"-resolution members of the same generator stream (seed 14), whose
parameters span the mix's ranges (bottom radius 2000-10000 km, thickness
20-200 km, Rayleigh / Mie coefficients 1e-4 - 1e-1 /km, orders 2-8):
  params     JSON of the 8 parameter sets
  T<i>, E<i> full transmittance / irradiance LUTs (rgb), float32
  S<i>       scattering RGBA at the 4096 texels S_idx, float32
  S_idx      seeded texel indices into the packed half-size scattering LUT
Run: python tests/golden/make_golden_mixed.py   (about 2 min on 8 cores)
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fuzzyblue_amd.workloads import mixed_batch, random_atmosphere  # noqa: E402
from oracle import Oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def half_members(n=8, seed=14):
    rng = np.random.default_rng(seed)
    return [random_atmosphere(rng, True, resolution="half") for _ in range(n)]


def main():
    with open(os.path.join(HERE, "mixed256_seed4.json"), "w") as f:
        json.dump([p.to_dict() for p in mixed_batch(256, 4, precise=True)], f,
                  separators=(",", ":"))
    ps = half_members()
    o = Oracle("f64")
    rng = np.random.default_rng(9)
    ns = 16 * 64 * 16 * 4
    idx = np.sort(rng.choice(ns, 4096, replace=False)).astype(np.int64)
    out = {"params": json.dumps([p.to_dict() for p in ps]), "S_idx": idx}
    for i, p in enumerate(ps):
        r = o.precompute(p)
        out[f"T{i}"] = r["T"][..., :3].astype(np.float32)
        out[f"E{i}"] = r["E"][..., :3].astype(np.float32)
        out[f"S{i}"] = r["S"].reshape(-1, 4)[idx].astype(np.float32)
        print(f"member {i}: order {p.order}, bottom {p.bottom_radius:.0f} km", flush=True)
    np.savez_compressed(os.path.join(HERE, "mixed_half_fp64.npz"), **out)
    print("wrote mixed256_seed4.json, mixed_half_fp64.npz")


if __name__ == "__main__":
    main()
