"""ctypes loader for the CPU oracle (oracle/libfb_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the reference the GPU results are
compared with.  The product never imports this package.

Two precisions are exposed with the same API: ``Oracle("f64")`` — the FP64
oracle (the reference's discretisation evaluated in double) — and
``Oracle("f32")`` — the FP32 restatement of the reference GLSL.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from fuzzyblue_amd import abi
from fuzzyblue_amd.params import Parameters

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfb_oracle.so")

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "all"], cwd=HERE)
    return LIB_PATH


def load(path: str | None = None):
    """The oracle library; ``path`` loads another build of the same source
    (the mutation tests build copies with one planted mistake each)."""
    global _lib
    if path is not None:
        return C.CDLL(path)
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One precision of the oracle.  Arrays are numpy float64/float32 RGBA."""

    def __init__(self, precision: str = "f64", lib_path: str | None = None):
        assert precision in ("f64", "f32")
        self.precision = precision
        self.dtype = np.float64 if precision == "f64" else np.float32
        self.ctype = C.c_double if precision == "f64" else C.c_float
        self.lib = load(lib_path)
        self._sig()

    def _fn(self, name):
        return getattr(self.lib, f"{name}_{self.precision}")

    def _sig(self):
        R = self.ctype
        P = C.POINTER(abi.fb_parameters)
        V = C.c_void_p
        sigs = {
            "oracle_phase_rayleigh": (R, [R]),
            "oracle_phase_mie": (R, [R, R]),
            "oracle_tex_from_unit": (R, [R, C.c_int]),
            "oracle_unit_from_tex": (R, [R, C.c_int]),
            "oracle_frag_from_texel": (R, [C.c_uint, C.c_uint]),
            "oracle_round_half": (R, [R]),
            "oracle_profile_density": (R, [P, C.c_int, R]),
            "oracle_distance_to_top": (R, [P, R, R]),
            "oracle_distance_to_bottom": (R, [P, R, R]),
            "oracle_ray_intersects_ground": (C.c_int, [P, R, R]),
            "oracle_transmittance_uv": (None, [P, R, R, V]),
            "oracle_transmittance_rmu": (None, [P, R, R, V]),
            "oracle_optical_length": (R, [P, C.c_int, R, R]),
            "oracle_irradiance_uv": (None, [P, R, R, V]),
            "oracle_scattering_uvwz": (None, [P, R, R, R, R, C.c_int, V]),
            "oracle_scattering_from_uvwz": (C.c_int, [P, V, V]),
            "oracle_scattering_texel": (C.c_int, [P, C.c_uint, C.c_uint, C.c_uint, V]),
            "oracle_sample2d": (None, [V, C.c_int, C.c_int, R, R, V]),
            "oracle_sample3d": (None, [V, C.c_int, C.c_int, C.c_int, R, R, R, V]),
            "oracle_get_transmittance": (None, [P, V, R, R, R, C.c_int, V]),
            "oracle_get_scattering": (None, [P, V, R, R, R, R, C.c_int, V]),
            "oracle_density_at": (None, [P, C.c_int, V, V, V, V, V, R, R, R, R, V]),
            "oracle_extrapolated_single_mie": (None, [P, V, V]),
            "oracle_pass_transmittance": (None, [P, V]),
            "oracle_pass_direct_irradiance": (None, [P, V, V]),
            "oracle_pass_single_scattering": (None, [P, C.c_int, V, V, V, V]),
            "oracle_pass_density": (None, [P, C.c_int, C.c_int, V, V, V, V, V, V]),
            "oracle_pass_indirect": (None, [P, C.c_int, V, V, V, V, V]),
            "oracle_pass_multiple": (None, [P, C.c_int, V, V, V, V]),
            "oracle_density_texels": (None, [P, C.c_int, V, V, V, V, V, V, C.c_long, V]),
            "oracle_multiple_texels": (None, [P, V, V, V, C.c_long, V, V]),
            "oracle_single_texels": (None, [P, V, V, C.c_long, V, V]),
            "oracle_precompute": (C.c_int, [P, C.c_int, V, V, V, V, V]),
            "oracle_sky_rays": (
                None,
                [P, V, V, V, C.c_int, V, V, V, C.c_int, V, C.c_long, V, V],
            ),
            "oracle_render_frame": (None, [P, V, V, V, V, V, V, C.c_int, C.c_int, V, V]),
            "oracle_sun_and_sky_irradiance": (None, [P, V, V, V, V, V, C.c_int, C.c_long, V, V]),
        }
        for name, (res, args) in sigs.items():
            f = self._fn(name)
            f.restype = res
            f.argtypes = args

    # ------------------------------------------------------------ helpers
    @staticmethod
    def cparams(p) -> abi.fb_parameters:
        return p.to_c() if isinstance(p, Parameters) else p

    def arr(self, *shape):
        return np.zeros(shape, dtype=self.dtype)

    def asarr(self, a):
        return np.ascontiguousarray(a, dtype=self.dtype)

    @staticmethod
    def shapes(p):
        p = p if isinstance(p, Parameters) else Parameters.from_c(p)
        tw, th = p.transmittance_extent()
        iw, ih = p.irradiance_extent()
        sw, sh, sd = p.scattering_extent()
        return (th, tw, 4), (ih, iw, 4), (sd, sh, sw, 4)

    # ------------------------------------------------------------ scalars
    def phase_rayleigh(self, nu):
        return self._fn("oracle_phase_rayleigh")(nu)

    def phase_mie(self, g, nu):
        return self._fn("oracle_phase_mie")(g, nu)

    def tex_from_unit(self, x, n):
        return self._fn("oracle_tex_from_unit")(x, n)

    def unit_from_tex(self, u, n):
        return self._fn("oracle_unit_from_tex")(u, n)

    def frag_from_texel(self, x, n):
        return self._fn("oracle_frag_from_texel")(x, n)

    def round_half(self, v):
        return self._fn("oracle_round_half")(v)

    def profile_density(self, p, which, h):
        c = self.cparams(p)
        return self._fn("oracle_profile_density")(C.byref(c), which, h)

    def distance_to_top(self, p, r, mu):
        c = self.cparams(p)
        return self._fn("oracle_distance_to_top")(C.byref(c), r, mu)

    def distance_to_bottom(self, p, r, mu):
        c = self.cparams(p)
        return self._fn("oracle_distance_to_bottom")(C.byref(c), r, mu)

    def ray_intersects_ground(self, p, r, mu):
        c = self.cparams(p)
        return bool(self._fn("oracle_ray_intersects_ground")(C.byref(c), r, mu))

    def transmittance_uv(self, p, r, mu):
        c = self.cparams(p)
        out = self.arr(2)
        self._fn("oracle_transmittance_uv")(C.byref(c), r, mu, _ptr(out))
        return out

    def transmittance_rmu(self, p, x_mu, x_r):
        c = self.cparams(p)
        out = self.arr(2)
        self._fn("oracle_transmittance_rmu")(C.byref(c), x_mu, x_r, _ptr(out))
        return out

    def optical_length(self, p, which, r, mu):
        c = self.cparams(p)
        return self._fn("oracle_optical_length")(C.byref(c), which, r, mu)

    def irradiance_uv(self, p, r, mu_s):
        c = self.cparams(p)
        out = self.arr(2)
        self._fn("oracle_irradiance_uv")(C.byref(c), r, mu_s, _ptr(out))
        return out

    def scattering_uvwz(self, p, r, mu, mu_s, nu, ground):
        c = self.cparams(p)
        out = self.arr(4)
        self._fn("oracle_scattering_uvwz")(C.byref(c), r, mu, mu_s, nu, int(ground), _ptr(out))
        return out

    def scattering_from_uvwz(self, p, uvwz):
        c = self.cparams(p)
        u = self.asarr(uvwz)
        out = self.arr(4)
        g = self._fn("oracle_scattering_from_uvwz")(C.byref(c), _ptr(u), _ptr(out))
        return out, bool(g)

    def scattering_texel(self, p, x, y, z):
        c = self.cparams(p)
        out = self.arr(4)
        g = self._fn("oracle_scattering_texel")(C.byref(c), x, y, z, _ptr(out))
        return out, bool(g)

    def sample2d(self, tex, u, v):
        t = self.asarr(tex)
        out = self.arr(4)
        self._fn("oracle_sample2d")(_ptr(t), t.shape[1], t.shape[0], u, v, _ptr(out))
        return out

    def sample3d(self, tex, u, v, s):
        t = self.asarr(tex)
        out = self.arr(4)
        self._fn("oracle_sample3d")(
            _ptr(t), t.shape[2], t.shape[1], t.shape[0], u, v, s, _ptr(out)
        )
        return out

    def get_transmittance(self, p, T, r, mu, d, ground):
        c = self.cparams(p)
        t = self.asarr(T)
        out = self.arr(3)
        self._fn("oracle_get_transmittance")(C.byref(c), _ptr(t), r, mu, d, int(ground), _ptr(out))
        return out

    def get_scattering(self, p, tex, r, mu, mu_s, nu, ground):
        c = self.cparams(p)
        t = self.asarr(tex)
        out = self.arr(4)
        self._fn("oracle_get_scattering")(
            C.byref(c), _ptr(t), r, mu, mu_s, nu, int(ground), _ptr(out)
        )
        return out

    def density_at(self, p, order, T, dR, dM, dMS, dE, r, mu, mu_s, nu):
        c = self.cparams(p)
        args = [self.asarr(x) for x in (T, dR, dM, dMS, dE)]
        out = self.arr(3)
        self._fn("oracle_density_at")(
            C.byref(c), int(order), *[_ptr(a) for a in args], r, mu, mu_s, nu, _ptr(out)
        )
        return out

    def extrapolated_single_mie(self, p, s):
        """render_sky.h:9-19 over an (..., 4) array of scattering values."""
        c = self.cparams(p)
        s = self.asarr(s)
        flat = s.reshape(-1, 4)
        out = self.arr(flat.shape[0], 3)
        f = self._fn("oracle_extrapolated_single_mie")
        for i in range(flat.shape[0]):
            f(C.byref(c), _ptr(flat[i]), _ptr(out[i]))
        return out.reshape(s.shape[:-1] + (3,))

    # ------------------------------------------------------------ passes
    def pass_transmittance(self, p):
        c = self.cparams(p)
        ts, _, _ = self.shapes(c)
        T = self.arr(*ts)
        self._fn("oracle_pass_transmittance")(C.byref(c), _ptr(T))
        return T

    def pass_direct_irradiance(self, p, T):
        c = self.cparams(p)
        _, is_, _ = self.shapes(c)
        T = self.asarr(T)
        dE = self.arr(*is_)
        self._fn("oracle_pass_direct_irradiance")(C.byref(c), _ptr(T), _ptr(dE))
        return dE

    def pass_single_scattering(self, p, T, fp16=False):
        c = self.cparams(p)
        _, _, ss = self.shapes(c)
        T = self.asarr(T)
        dR, dM, S = self.arr(*ss), self.arr(*ss), self.arr(*ss)
        self._fn("oracle_pass_single_scattering")(
            C.byref(c), int(fp16), _ptr(T), _ptr(dR), _ptr(dM), _ptr(S)
        )
        return dR, dM, S

    def pass_density(self, p, order, T, dR, dM, dMS, dE, fp16=False):
        c = self.cparams(p)
        _, _, ss = self.shapes(c)
        args = [self.asarr(x) for x in (T, dR, dM, dMS, dE)]
        J = self.arr(*ss)
        self._fn("oracle_pass_density")(
            C.byref(c), int(fp16), int(order), *[_ptr(a) for a in args], _ptr(J)
        )
        return J

    def pass_indirect(self, p, order, dR, dM, dMS, E):
        c = self.cparams(p)
        _, is_, _ = self.shapes(c)
        args = [self.asarr(x) for x in (dR, dM, dMS)]
        dE = self.arr(*is_)
        E = self.asarr(E).copy()
        self._fn("oracle_pass_indirect")(
            C.byref(c), int(order), *[_ptr(a) for a in args], _ptr(dE), _ptr(E)
        )
        return dE, E

    def pass_multiple(self, p, T, J, S, fp16=False):
        c = self.cparams(p)
        _, _, ss = self.shapes(c)
        T, J = self.asarr(T), self.asarr(J)
        dMS = self.arr(*ss)
        S = self.asarr(S).copy()
        self._fn("oracle_pass_multiple")(
            C.byref(c), int(fp16), _ptr(T), _ptr(J), _ptr(dMS), _ptr(S)
        )
        return dMS, S

    def density_texels(self, p, order, T, dR, dM, dMS, dE, idx):
        c = self.cparams(p)
        args = [self.asarr(x) for x in (T, dR, dM, dMS, dE)]
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        out = self.arr(len(idx), 4)
        self._fn("oracle_density_texels")(
            C.byref(c), int(order), *[_ptr(a) for a in args], _ptr(idx), len(idx), _ptr(out)
        )
        return out

    def multiple_texels(self, p, T, J, idx):
        c = self.cparams(p)
        T, J = self.asarr(T), self.asarr(J)
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        out = self.arr(len(idx), 4)
        nu = self.arr(len(idx))
        self._fn("oracle_multiple_texels")(
            C.byref(c), _ptr(T), _ptr(J), _ptr(idx), len(idx), _ptr(out), _ptr(nu)
        )
        return out, nu

    def single_texels(self, p, T, idx):
        c = self.cparams(p)
        T = self.asarr(T)
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        ray = self.arr(len(idx), 3)
        mie = self.arr(len(idx), 3)
        self._fn("oracle_single_texels")(
            C.byref(c), _ptr(T), _ptr(idx), len(idx), _ptr(ray), _ptr(mie)
        )
        return ray, mie

    def precompute(self, p, fp16=False, per_order=False):
        """The whole precompute.  Returns dict T, E (irradiance), S
        (scattering) and, with per_order, 'orders_S' and 'orders_E'."""
        c = self.cparams(p)
        ts, is_, ss = self.shapes(c)
        T, E, S = self.arr(*ts), self.arr(*is_), self.arr(*ss)
        po_s = po_e = None
        if per_order:
            po_s = self.arr(c.order, *ss)
            po_e = self.arr(c.order, *is_)
        rc = self._fn("oracle_precompute")(
            C.byref(c),
            int(fp16),
            _ptr(T),
            _ptr(E),
            _ptr(S),
            _ptr(po_s) if po_s is not None else None,
            _ptr(po_e) if po_e is not None else None,
        )
        if rc != 0:
            raise MemoryError("oracle_precompute failed")
        out = {"T": T, "E": E, "S": S}
        if per_order:
            out["orders_S"] = po_s
            out["orders_E"] = po_e
        return out

    def sky_rays(self, p, T, S, cameras, views, suns, points=None, shadow=None):
        c = self.cparams(p)
        T, S = self.asarr(T), self.asarr(S)
        cams = self.asarr(np.atleast_2d(cameras))
        suns_ = self.asarr(np.atleast_2d(suns))
        views = self.asarr(np.atleast_2d(views))
        n = views.shape[0]
        pts = self.asarr(np.atleast_2d(points)) if points is not None else None
        sh = self.asarr(np.asarray(shadow).reshape(n)) if shadow is not None else None
        rad, trans = self.arr(n, 3), self.arr(n, 3)
        self._fn("oracle_sky_rays")(
            C.byref(c),
            _ptr(T),
            _ptr(S),
            _ptr(cams),
            1 if cams.shape[0] > 1 else 0,
            _ptr(views),
            _ptr(pts) if pts is not None else None,
            _ptr(suns_),
            1 if suns_.shape[0] > 1 else 0,
            _ptr(sh) if sh is not None else None,
            n,
            _ptr(rad),
            _ptr(trans),
        )
        return rad, trans

    def render_frame(self, p, T, S, inverse_viewproj, camera, sun, width, height, depth=None):
        c = self.cparams(p)
        T, S = self.asarr(T), self.asarr(S)
        ivp = np.ascontiguousarray(inverse_viewproj, dtype=np.float32).reshape(16)
        cam = np.ascontiguousarray(camera, dtype=np.float32)
        sn = np.ascontiguousarray(sun, dtype=np.float32)
        dep = np.ascontiguousarray(depth, dtype=np.float32) if depth is not None else None
        color, trans = self.arr(height, width, 4), self.arr(height, width, 4)
        self._fn("oracle_render_frame")(
            C.byref(c),
            _ptr(T),
            _ptr(S),
            _ptr(ivp),
            _ptr(cam),
            _ptr(sn),
            _ptr(dep) if dep is not None else None,
            width,
            height,
            _ptr(color),
            _ptr(trans),
        )
        return color, trans

    def sun_and_sky_irradiance(self, p, T, E, points, normals, suns):
        c = self.cparams(p)
        T, E = self.asarr(T), self.asarr(E)
        pts = self.asarr(np.atleast_2d(points))
        nrm = self.asarr(np.atleast_2d(normals))
        sn = self.asarr(np.atleast_2d(suns))
        n = pts.shape[0]
        direct, sky = self.arr(n, 3), self.arr(n, 3)
        self._fn("oracle_sun_and_sky_irradiance")(
            C.byref(c),
            _ptr(T),
            _ptr(E),
            _ptr(pts),
            _ptr(nrm),
            _ptr(sn),
            1 if sn.shape[0] > 1 else 0,
            n,
            _ptr(direct),
            _ptr(sky),
        )
        return direct, sky
