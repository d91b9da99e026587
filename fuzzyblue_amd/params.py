"""Host-side ``Parameters`` / ``DensityProfile`` / ``DensityProfileLayer``.

Mirrors the reference's public parameter set (src/precompute.rs:655-769) and
its ``Default`` (src/precompute.rs:849-935): the same field names (the
reference spells ``absorbtion_*``; both spellings are accepted here), the
same units (kilometres, src/precompute.rs:680) and the same default Earth
atmosphere.  ``to_c()`` produces the C-ABI ``fb_parameters``.
"""
from __future__ import annotations

import copy
import dataclasses
from dataclasses import dataclass, field
from typing import List

from . import abi


@dataclass
class DensityProfileLayer:
    """src/precompute.rs:655-665: density = clamp(exp_term*exp(exp_scale*h)
    + linear_term*h + constant_term, 0, 1)."""

    width: float = 0.0
    exp_term: float = 0.0
    exp_scale: float = 0.0
    linear_term: float = 0.0
    constant_term: float = 0.0


@dataclass
class DensityProfile:
    """src/precompute.rs:667-676: two layers, bottom to top."""

    layers: List[DensityProfileLayer] = field(
        default_factory=lambda: [DensityProfileLayer(), DensityProfileLayer()]
    )


def _exp_profile(scale: float) -> DensityProfile:
    return DensityProfile(
        [DensityProfileLayer(), DensityProfileLayer(0.0, 1.0, scale, 0.0, 0.0)]
    )


@dataclass
class Parameters:
    """Parameters governing generated skies (src/precompute.rs:677-769).

    Field defaults are ``Parameters::default()`` (src/precompute.rs:849-935).
    """

    order: int = 4
    transmittance_mu_size: int = 256
    transmittance_r_size: int = 64
    scattering_r_size: int = 32
    scattering_mu_size: int = 128
    scattering_mu_s_size: int = 32
    scattering_nu_size: int = 8
    irradiance_mu_s_size: int = 64
    irradiance_r_size: int = 16
    solar_irradiance: List[float] = field(default_factory=lambda: [1.474, 1.850, 1.91198])
    sun_angular_radius: float = 0.004675
    bottom_radius: float = 6360.0
    top_radius: float = 6420.0
    rayleigh_density: DensityProfile = field(default_factory=lambda: _exp_profile(-0.125))
    rayleigh_scattering: List[float] = field(
        default_factory=lambda: [0.005802, 0.013558, 0.033100]
    )
    mie_density: DensityProfile = field(default_factory=lambda: _exp_profile(-0.833333))
    mie_scattering: List[float] = field(default_factory=lambda: [0.003996, 0.003996, 0.003996])
    mie_extinction: List[float] = field(default_factory=lambda: [0.004440, 0.004440, 0.004440])
    mie_phase_function_g: float = 0.8
    absorption_density: DensityProfile = field(
        default_factory=lambda: DensityProfile(
            [
                DensityProfileLayer(25.0, 0.0, 0.0, 0.066667, -0.666667),
                DensityProfileLayer(0.0, 0.0, 0.0, -0.066667, 2.666667),
            ]
        )
    )
    absorption_extinction: List[float] = field(default_factory=lambda: [6.5e-4, 1.881e-3, 8.5e-5])
    ground_albedo: List[float] = field(default_factory=lambda: [0.1, 0.1, 0.1])
    mu_s_min: float = -0.207912
    flags: int = 0

    # The reference's spelling (src/precompute.rs:756,761).
    @property
    def absorbtion_density(self) -> DensityProfile:
        return self.absorption_density

    @absorbtion_density.setter
    def absorbtion_density(self, v: DensityProfile) -> None:
        self.absorption_density = v

    @property
    def absorbtion_extinction(self) -> List[float]:
        return self.absorption_extinction

    @absorbtion_extinction.setter
    def absorbtion_extinction(self, v: List[float]) -> None:
        self.absorption_extinction = v

    # src/precompute.rs:771-793
    def transmittance_extent(self):
        return (self.transmittance_mu_size, self.transmittance_r_size)

    def irradiance_extent(self):
        return (self.irradiance_mu_s_size, self.irradiance_r_size)

    def scattering_extent(self):
        return (
            self.scattering_nu_size * self.scattering_mu_s_size,
            self.scattering_mu_size,
            self.scattering_r_size,
        )

    def to_dict(self) -> dict:
        """Plain-data form (JSON fixtures)."""
        return dataclasses.asdict(self)

    @staticmethod
    def from_dict(d: dict) -> "Parameters":
        d = dict(d)
        for k in ("rayleigh_density", "mie_density", "absorption_density"):
            d[k] = DensityProfile([DensityProfileLayer(**l) for l in d[k]["layers"]])
        return Parameters(**d)

    def footprint(self):
        """(resident, scratch) device bytes of one LUT set, as the library
        allocates them (fb_parameters_footprint)."""
        import ctypes as C

        from .lib import load

        res, scr = C.c_size_t(), C.c_size_t()
        c = self.to_c()
        load().fb_parameters_footprint(C.byref(c), C.byref(res), C.byref(scr))
        return res.value, scr.value

    def replace(self, **kw) -> "Parameters":
        p = copy.deepcopy(self)
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        return p

    def to_c(self) -> abi.fb_parameters:
        c = abi.fb_parameters()
        for name in (
            "order",
            "transmittance_mu_size",
            "transmittance_r_size",
            "scattering_r_size",
            "scattering_mu_size",
            "scattering_mu_s_size",
            "scattering_nu_size",
            "irradiance_mu_s_size",
            "irradiance_r_size",
            "flags",
        ):
            setattr(c, name, int(getattr(self, name)))
        for name in (
            "sun_angular_radius",
            "bottom_radius",
            "top_radius",
            "mie_phase_function_g",
            "mu_s_min",
        ):
            setattr(c, name, float(getattr(self, name)))
        for name in (
            "solar_irradiance",
            "rayleigh_scattering",
            "mie_scattering",
            "mie_extinction",
            "absorption_extinction",
            "ground_albedo",
        ):
            arr = getattr(c, name)
            vals = getattr(self, name)
            for i in range(3):
                arr[i] = float(vals[i])
        for name in ("rayleigh_density", "mie_density", "absorption_density"):
            src = getattr(self, name)
            dst = getattr(c, name)
            for i in range(2):
                l = src.layers[i]
                dst.layers[i].width = l.width
                dst.layers[i].exp_term = l.exp_term
                dst.layers[i].exp_scale = l.exp_scale
                dst.layers[i].linear_term = l.linear_term
                dst.layers[i].constant_term = l.constant_term
        return c

    @staticmethod
    def from_c(c: abi.fb_parameters) -> "Parameters":
        p = Parameters()
        for name in (
            "order",
            "transmittance_mu_size",
            "transmittance_r_size",
            "scattering_r_size",
            "scattering_mu_size",
            "scattering_mu_s_size",
            "scattering_nu_size",
            "irradiance_mu_s_size",
            "irradiance_r_size",
            "flags",
        ):
            setattr(p, name, int(getattr(c, name)))
        for name in (
            "sun_angular_radius",
            "bottom_radius",
            "top_radius",
            "mie_phase_function_g",
            "mu_s_min",
        ):
            setattr(p, name, float(getattr(c, name)))
        for name in (
            "solar_irradiance",
            "rayleigh_scattering",
            "mie_scattering",
            "mie_extinction",
            "absorption_extinction",
            "ground_albedo",
        ):
            setattr(p, name, [float(x) for x in getattr(c, name)])
        for name in ("rayleigh_density", "mie_density", "absorption_density"):
            src = getattr(c, name)
            setattr(
                p,
                name,
                DensityProfile(
                    [
                        DensityProfileLayer(
                            l.width, l.exp_term, l.exp_scale, l.linear_term, l.constant_term
                        )
                        for l in src.layers
                    ]
                ),
            )
        return p


# Resolution classes of the mixed-atmosphere batch: "half", "default",
# "double" of the default LUT sizes.
RESOLUTIONS = {
    "half": dict(
        transmittance_mu_size=128,
        transmittance_r_size=32,
        scattering_r_size=16,
        scattering_mu_size=64,
        scattering_mu_s_size=16,
        scattering_nu_size=4,
        irradiance_mu_s_size=32,
        irradiance_r_size=8,
    ),
    "default": dict(
        transmittance_mu_size=256,
        transmittance_r_size=64,
        scattering_r_size=32,
        scattering_mu_size=128,
        scattering_mu_s_size=32,
        scattering_nu_size=8,
        irradiance_mu_s_size=64,
        irradiance_r_size=16,
    ),
    "double": dict(
        transmittance_mu_size=512,
        transmittance_r_size=128,
        scattering_r_size=64,
        scattering_mu_size=256,
        scattering_mu_s_size=64,
        scattering_nu_size=16,
        irradiance_mu_s_size=128,
        irradiance_r_size=32,
    ),
}

# The reference's test and bench use reduced scattering sizes "for speed"
# (tests/smoke.rs:136-142, benches/precompute.rs:127-133).
SMOKE_SIZES = dict(
    scattering_r_size=8, scattering_mu_size=32, scattering_mu_s_size=8, scattering_nu_size=2
)
# examples/dump.rs:96-108
DUMP_SIZES = dict(
    scattering_r_size=16, scattering_mu_size=64, scattering_mu_s_size=16, scattering_nu_size=4
)
