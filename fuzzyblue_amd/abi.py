"""ctypes mirror of include/fuzzyblue_hip.h (data layouts and constants only).

Nothing here loads a shared library; :mod:`fuzzyblue_amd.lib` does that.
The structures follow the reference's ``Parameters`` / ``DensityProfile`` /
``DensityProfileLayer`` (src/precompute.rs:655-769) and ``DrawParameters``
(src/render.rs:243-252), minus the Vulkan synchronisation fields.
"""
import ctypes as C

FB_ABI_VERSION = 1

# fb_status
FB_OK = 0
FB_EINVAL = 1
FB_ENOMEM = 2
FB_EBUDGET = 3
FB_EHIP = 4
FB_EBUSY = 5
FB_ENOTREADY = 6
FB_ENOENT = 7
FB_ENODEV = 8

STATUS_NAMES = {
    FB_OK: "FB_OK",
    FB_EINVAL: "FB_EINVAL",
    FB_ENOMEM: "FB_ENOMEM",
    FB_EBUDGET: "FB_EBUDGET",
    FB_EHIP: "FB_EHIP",
    FB_EBUSY: "FB_EBUSY",
    FB_ENOTREADY: "FB_ENOTREADY",
    FB_ENOENT: "FB_ENOENT",
    FB_ENODEV: "FB_ENODEV",
}

FB_FLAG_PRECISE = 0x1
FB_FLAG_F16_STORAGE = 0x2

# fb_lut
FB_LUT_TRANSMITTANCE = 0
FB_LUT_IRRADIANCE = 1
FB_LUT_SCATTERING = 2
FB_LUT_DELTA_IRRADIANCE = 3
FB_LUT_DELTA_RAYLEIGH = 4
FB_LUT_DELTA_MIE = 5
FB_LUT_DELTA_MULTIPLE_SCATTERING = 6
FB_LUT_SCATTERING_DENSITY = 7

FB_FORMAT_F32 = 0
FB_FORMAT_F16 = 1

FB_RENDER_COMPOSITE = 0x1

FB_REQUEST_PRECOMPUTE = 0
FB_REQUEST_RENDER = 1


class fb_density_profile_layer(C.Structure):
    _fields_ = [
        ("width", C.c_float),
        ("exp_term", C.c_float),
        ("exp_scale", C.c_float),
        ("linear_term", C.c_float),
        ("constant_term", C.c_float),
    ]


class fb_density_profile(C.Structure):
    _fields_ = [("layers", fb_density_profile_layer * 2)]


class fb_parameters(C.Structure):
    _fields_ = [
        ("order", C.c_uint32),
        ("transmittance_mu_size", C.c_uint32),
        ("transmittance_r_size", C.c_uint32),
        ("scattering_r_size", C.c_uint32),
        ("scattering_mu_size", C.c_uint32),
        ("scattering_mu_s_size", C.c_uint32),
        ("scattering_nu_size", C.c_uint32),
        ("irradiance_mu_s_size", C.c_uint32),
        ("irradiance_r_size", C.c_uint32),
        ("solar_irradiance", C.c_float * 3),
        ("sun_angular_radius", C.c_float),
        ("bottom_radius", C.c_float),
        ("top_radius", C.c_float),
        ("rayleigh_density", fb_density_profile),
        ("rayleigh_scattering", C.c_float * 3),
        ("mie_density", fb_density_profile),
        ("mie_scattering", C.c_float * 3),
        ("mie_extinction", C.c_float * 3),
        ("mie_phase_function_g", C.c_float),
        ("absorption_density", fb_density_profile),
        ("absorption_extinction", C.c_float * 3),
        ("ground_albedo", C.c_float * 3),
        ("mu_s_min", C.c_float),
        ("flags", C.c_uint32),
    ]


FB_PASS_TRANSMITTANCE = 0
FB_PASS_DIRECT_IRRADIANCE = 1
FB_PASS_SINGLE_SCATTERING = 2
FB_PASS_CLEAR_IRRADIANCE = 3
FB_PASS_SCATTERING_DENSITY = 4
FB_PASS_INDIRECT_IRRADIANCE = 5
FB_PASS_MULTIPLE_SCATTERING = 6
FB_PASS_GEOMETRY = 7

FB_PASS_NAMES = {
    0: "transmittance",
    1: "direct_irradiance",
    2: "single_scattering",
    3: "clear_irradiance",
    4: "scattering_density",
    5: "indirect_irradiance",
    6: "multiple_scattering",
    7: "geometry",
}


class fb_pass_timing(C.Structure):
    _fields_ = [("pass_", C.c_uint32), ("order", C.c_uint32), ("ms", C.c_float)]


class fb_tuning(C.Structure):
    _fields_ = [
        ("single_scattering_threads", C.c_uint32),
        ("density_order2_threads", C.c_uint32),
        ("density_threads", C.c_uint32),
        ("indirect_order1_waves", C.c_uint32),
        ("indirect_waves", C.c_uint32),
        ("multiple_scattering_threads", C.c_uint32),
        ("xcd_remap", C.c_uint32),
        ("best_ms", C.c_float * 6),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "best_ms"}
        d["best_ms"] = list(self.best_ms)
        return d


class fb_lut_view(C.Structure):
    _fields_ = [
        ("device_ptr", C.c_void_p),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("depth", C.c_uint32),
        ("texel_bytes", C.c_uint32),
        ("bytes", C.c_size_t),
    ]


class fb_lut_export(C.Structure):
    _fields_ = [
        ("fd", C.c_int),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("depth", C.c_uint32),
        ("texel_bytes", C.c_uint32),
        ("vk_format", C.c_uint32),
        ("bytes", C.c_size_t),
        ("alloc_bytes", C.c_size_t),
        ("offset", C.c_size_t),
        ("row_pitch", C.c_size_t),
        ("slice_pitch", C.c_size_t),
        ("device_ptr", C.c_void_p),
        ("opaque", C.c_void_p),
    ]


class fb_draw_parameters(C.Structure):
    _fields_ = [
        ("inverse_viewproj", (C.c_float * 4) * 4),
        ("camera_position", C.c_float * 3),
        ("sun_direction", C.c_float * 3),
    ]


class fb_server_config(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_size_t),
        ("max_inflight", C.c_uint32),
        ("n_streams", C.c_uint32),
        ("keep_outputs", C.c_uint32),
        ("precompute_share_pct", C.c_uint32),
        ("admit_min_free", C.c_uint32),
    ]


class fb_request_info(C.Structure):
    _fields_ = [
        ("id", C.c_uint64),
        ("kind", C.c_uint32),
        ("status", C.c_int32),
        ("lut_id", C.c_uint64),
        ("t_submit", C.c_double),
        ("t_start", C.c_double),
        ("t_done", C.c_double),
        ("gpu_ms", C.c_float),
        ("gpu_wait_ms", C.c_float),
    ]


class fb_server_stats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("rejected", C.c_uint64),
        ("evictions", C.c_uint64),
        ("hbm_in_use", C.c_size_t),
        ("hbm_peak", C.c_size_t),
        ("resident_sets", C.c_uint32),
        ("step_max_ms", C.c_float * 3),
        ("slow_steps", C.c_uint64),
    ]
