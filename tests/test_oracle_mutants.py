"""Mutation tests of the oracle's pins.

Each case copies oracle/atmo_oracle.c, plants one plausible mistake by a
textual substitution, builds the copy (FP64 only) into a temporary library
and requires the named closed-form check of tests/oracle_checks.py to FAIL
on it.  The unmodified oracle passes every one of those checks
(tests/test_oracle_pinning.py), so each planted mistake is caught by a
check that the mathematics fixes, not by comparing the oracle with itself
in another precision.
"""
import os
import shutil
import subprocess

import pytest

import oracle_checks as chk
from oracle import Oracle

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
SRC = os.path.join(ORACLE_DIR, "atmo_oracle.c")

# (name, text in atmo_oracle.c, replacement, check that must fail)
MUTANTS = [
    (
        # GetScattering order 1 (scattering.h:172-175) with P_M(g, -nu)
        "order1_mie_phase_sign",
        "get_scattering_tex(a, dM, r, mu, mu_s, nu, ground, m);\n"
        "        real pr = phase_rayleigh(nu), pm = phase_mie(a->g, nu);",
        "get_scattering_tex(a, dM, r, mu, mu_s, nu, ground, m);\n"
        "        real pr = phase_rayleigh(nu), pm = phase_mie(a->g, -nu);",
        ("check_order1_indirect", "check_order1_density"),
    ),
    (
        # GetExtrapolatedSingleMieScattering (render_sky.h:17-18) with the
        # beta ratios inverted
        "mie_extrapolation_ratio_inverted",
        "(a->rayleigh_scattering[0] / a->mie_scattering[0]) *\n"
        "                 (a->mie_scattering[c] / a->rayleigh_scattering[c]);",
        "(a->mie_scattering[0] / a->rayleigh_scattering[0]) *\n"
        "                 (a->rayleigh_scattering[c] / a->mie_scattering[c]);",
        ("check_mie_extrapolation",),
    ),
    (
        # GetTransmittance (transmittance.h:43-60) with the ground branch
        # replaced by the sky form
        "transmittance_ground_branch_as_sky",
        "    if (ground) {\n        transmittance_to_top(a, T, r_d, -mu_d, n);",
        "    if (0) {\n        transmittance_to_top(a, T, r_d, -mu_d, n);",
        ("check_transmittance_segments",),
    ),
    (
        # the manual nu lerp (scattering.h:146-154) with swapped weights
        "nu_lerp_weights_swapped",
        "out[c] = s0[c] * (1.0 - lerp) + s1[c] * lerp;",
        "out[c] = s0[c] * lerp + s1[c] * (1.0 - lerp);",
        ("check_nu_lerp_affine",),
    ),
]


def _build_mutant(tmp_path, old, new):
    text = open(SRC).read()
    assert text.count(old) == 1, f"mutation site not unique in atmo_oracle.c: {old!r}"
    src = tmp_path / "atmo_oracle.c"
    src.write_text(text.replace(old, new))
    lib = tmp_path / "libfb_oracle_mutant.so"
    cc = shutil.which("gcc") or "cc"
    # -I oracle/ resolves the source's "../include/fuzzyblue_hip.h"
    subprocess.check_call([
        cc, "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-math-errno",
        "-std=c11", "-DORACLE_REAL=double", "-DORACLE_SFX=_f64", "-I", ORACLE_DIR,
        str(src), "-o", str(lib), "-lm",
    ])
    return str(lib)


@pytest.mark.parametrize("name,old,new,checks", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_fails_its_check(tmp_path, name, old, new, checks):
    mutant = Oracle("f64", lib_path=_build_mutant(tmp_path, old, new))
    for check in checks:
        with pytest.raises(AssertionError):
            getattr(chk, check)(mutant)
