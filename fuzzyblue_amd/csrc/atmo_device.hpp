// atmo_device.hpp — device-side atmosphere math for gfx950.
//
// Restates the reference's GLSL device library (shaders/util.h, params.h,
// transmittance.h, irradiance.h, scattering.h) as HIP device functions.
// Geometry (ray–sphere distances, the r/mu/mu_s texture parameterisations,
// sample positions along rays) is templated on a scalar type G:
//   G = float  reproduces the reference's binary32 arithmetic operation by
//              operation (the library is compiled with -ffp-contract=off and
//              IEEE division / square root, so no operation is fused or
//              approximated behind the source's back);
//   G = double is the precision mode (FB_FLAG_PRECISE): FP64 where FP32
//              cancellation breaks — r^2 - b^2 near the ground, the
//              ray–sphere discriminants near the horizon, r_d along rays.
// Radiance values, LUT texels and the quadrature accumulations stay FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbk {

struct DevLayer {
    float width, exp_term, exp_scale, linear_term, constant_term;
};

// Mirrors shaders/params.h:26-87 (AtmosphereParameters), field for field,
// plus the texture sizes.  Passed by value as a kernel argument (252 B) or
// read from an HBM array in the batched passes.
struct DevParams {
    float solar[3];
    float sun_angular_radius;
    float rayleigh_scattering[3];
    float bottom;
    float mie_scattering[3];
    float top;
    float mie_extinction[3];
    float g;
    float albedo[3];
    float mu_s_min;
    float absorption_extinction[3];
    int flags;
    int t_mu, t_r, s_r, s_mu, s_mus, s_nu, i_mus, i_r;
    int den_split;        // scattering density: threads per texel (1, or kDenTheta)
    DevLayer prof[3][2];  // rayleigh, mie, absorption
};

constexpr float kPi = 3.14159265358979323846f;

// ---------------------------------------------------------------- util.h
template <typename G>
__device__ __forceinline__ G clampg(G x, G lo, G hi) {
    return fmin(fmax(x, lo), hi);
}
// util.h:6-8
template <typename G>
__device__ __forceinline__ G clamp_cosine(G mu) {
    return clampg<G>(mu, G(-1), G(1));
}
// util.h:10-12
template <typename G>
__device__ __forceinline__ G clamp_distance(G d) {
    return fmax(d, G(0));
}
// util.h:14-16
template <typename G>
__device__ __forceinline__ G safe_sqrt(G a) {
    return sqrt(fmax(a, G(0)));
}
// util.h:18-20
template <typename G>
__device__ __forceinline__ G tex_from_unit(G x, int n) {
    return G(0.5) / G(n) + x * (G(1) - G(1) / G(n));
}
// util.h:22-24
template <typename G>
__device__ __forceinline__ G unit_from_tex(G u, int n) {
    return (u - G(0.5) / G(n)) / (G(1) - G(1) / G(n));
}
// util.h:26-29
__device__ __forceinline__ float phase_rayleigh(float nu) {
    float k = 3.0f / (16.0f * kPi);
    return k * (1.0f + nu * nu);
}
// util.h:31-34.  pow(x, 1.5) evaluated as x*sqrt(x).
__device__ __forceinline__ float phase_mie(float g, float nu) {
    float k = 3.0f / (8.0f * kPi) * (1.0f - g * g) / (2.0f + g * g);
    float x = 1.0f + g * g - 2.0f * g * nu;
    return k * (1.0f + nu * nu) / (x * sqrtf(x));
}
// util.h:36-38 (uint texel -> fragment coordinate)
template <typename G>
__device__ __forceinline__ G frag_from_texel(unsigned x, unsigned n) {
    return G(n) * tex_from_unit<G>(G(x) / G(n - 1), int(n));
}
template <typename G>
__device__ __forceinline__ G smoothstep_g(G e0, G e1, G x) {
    G t = clampg<G>((x - e0) / (e1 - e0), G(0), G(1));
    return t * t * (G(3) - G(2) * t);
}

// -------------------------------------------------------------- params.h
// params.h:89-93 (altitude in FP32, as the density value is FP32)
__device__ __forceinline__ float layer_density(const DevLayer& l, float altitude) {
    float density = l.exp_term * expf(l.exp_scale * altitude) + l.linear_term * altitude +
                    l.constant_term;
    return fminf(fmaxf(density, 0.0f), 1.0f);
}
// params.h:95-99
__device__ __forceinline__ float profile_density(const DevLayer* p, float altitude) {
    return altitude < p[0].width ? layer_density(p[0], altitude)
                                 : layer_density(p[1], altitude);
}
// params.h:101-103
template <typename G>
__device__ __forceinline__ G clamp_radius(const DevParams& P, G r) {
    return clampg<G>(r, G(P.bottom), G(P.top));
}
// params.h:105-110
template <typename G>
__device__ __forceinline__ G distance_to_top(const DevParams& P, G r, G mu) {
    G discriminant = r * r * (mu * mu - G(1)) + G(P.top) * G(P.top);
    return clamp_distance<G>(-r * mu + safe_sqrt<G>(discriminant));
}
// params.h:112-117
template <typename G>
__device__ __forceinline__ G distance_to_bottom(const DevParams& P, G r, G mu) {
    G discriminant = r * r * (mu * mu - G(1)) + G(P.bottom) * G(P.bottom);
    return clamp_distance<G>(-r * mu - safe_sqrt<G>(discriminant));
}
// params.h:119-124
template <typename G>
__device__ __forceinline__ bool ray_intersects_ground(const DevParams& P, G r, G mu) {
    return mu < G(0) && r * r * (mu * mu - G(1)) + G(P.bottom) * G(P.bottom) >= G(0);
}
// params.h:126-133
template <typename G>
__device__ __forceinline__ G distance_to_nearest(const DevParams& P, G r, G mu, bool ground) {
    return ground ? distance_to_bottom<G>(P, r, mu) : distance_to_top<G>(P, r, mu);
}

// ------------------------------------------------------ software sampler
// VkSampler LINEAR + CLAMP_TO_EDGE (src/precompute.rs:85-98): texel
// coordinate u*N - 0.5, taps floor and floor+1 clamped to [0, N-1].  The
// coordinate is clamped to [-1, N] first so NaN / out-of-range inputs index
// in bounds (fmax/fmin drop a NaN operand).
struct Taps {
    int i0, i1;
    float w;  // weight of i1
};
template <typename G>
__device__ __forceinline__ Taps linear_taps(G u, int n) {
    G x = u * G(n) - G(0.5);
    x = fmin(fmax(x, G(-1)), G(n));
    G f = floor(x);
    Taps t;
    t.w = float(x - f);
    int i = int(f);
    t.i0 = min(max(i, 0), n - 1);
    t.i1 = min(max(i + 1, 0), n - 1);
    return t;
}

__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float w) {
    float v = 1.0f - w;
    return make_float4(a.x * v + b.x * w, a.y * v + b.y * w, a.z * v + b.z * w,
                       a.w * v + b.w * w);
}
__device__ __forceinline__ float4 ldg4(const float4* p) { return __ldg(p); }

// Binary16 storage of the scattering LUT and the 3D scratch LUTs (the
// reference's R16G16B16A16_SFLOAT images, src/precompute.rs:1215): DevParams
// .flags bit kFlagF16Storage (= FB_FLAG_F16_STORAGE).  Those LUTs keep their
// float4* type in every signature; with the bit set they hold 8-byte
// binary16 texels at the same texel indices.
constexpr int kFlagF16Storage = 0x2;
typedef _Float16 Half4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 widen4(Half4 v) {
    return make_float4(float(v.x), float(v.y), float(v.z), float(v.w));
}
// Round to nearest even, as an image store into a 16F image.
typedef float Float4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ Half4 narrow4(float4 v) {
    const Float4v f = {v.x, v.y, v.z, v.w};
    return __builtin_convertvector(f, Half4);
}
// Texel i of a scattering-type LUT, binary16 when `h`.
__device__ __forceinline__ float4 ld_tex(const float4* __restrict__ t, size_t i, bool h) {
    if (h) return widen4(reinterpret_cast<const Half4*>(t)[i]);
    return ldg4(t + i);
}
__device__ __forceinline__ void st_tex(float4* __restrict__ t, size_t i, float4 v, bool h) {
    if (h)
        reinterpret_cast<Half4*>(t)[i] = narrow4(v);
    else
        t[i] = v;
}

template <typename G>
__device__ __forceinline__ float4 sample2d(const float4* __restrict__ tex, int w, int h, G u,
                                           G v) {
    Taps tx = linear_taps<G>(u, w), ty = linear_taps<G>(v, h);
    float4 t00 = ldg4(tex + ty.i0 * w + tx.i0);
    float4 t01 = ldg4(tex + ty.i0 * w + tx.i1);
    float4 t10 = ldg4(tex + ty.i1 * w + tx.i0);
    float4 t11 = ldg4(tex + ty.i1 * w + tx.i1);
    return lerp4(lerp4(t00, t01, tx.w), lerp4(t10, t11, tx.w), ty.w);
}

// ------------------------------------------------------- transmittance.h
// transmittance.h:7-24
template <typename G>
__device__ __forceinline__ void transmittance_uv(const DevParams& P, G r, G mu, G& u, G& v) {
    G H = sqrt(G(P.top) * G(P.top) - G(P.bottom) * G(P.bottom));
    G rho = safe_sqrt<G>(r * r - G(P.bottom) * G(P.bottom));
    G d = distance_to_top<G>(P, r, mu);
    G d_min = G(P.top) - r;
    G d_max = rho + H;
    G x_mu = (d - d_min) / (d_max - d_min);
    G x_r = rho / H;
    u = tex_from_unit<G>(x_mu, P.t_mu);
    v = tex_from_unit<G>(x_r, P.t_r);
}
// transmittance.h:26-33
template <typename G>
__device__ __forceinline__ float3 transmittance_to_top(const DevParams& P,
                                                       const float4* __restrict__ T, G r,
                                                       G mu) {
    G u, v;
    transmittance_uv<G>(P, r, mu, u, v);
    float4 s = sample2d<G>(T, P.t_mu, P.t_r, u, v);
    return make_float3(s.x, s.y, s.z);
}
// transmittance.h:35-61
template <typename G>
__device__ __forceinline__ float3 get_transmittance(const DevParams& P,
                                                    const float4* __restrict__ T, G r, G mu,
                                                    G d, bool ground) {
    G r_d = clamp_radius<G>(P, sqrt(d * d + G(2) * r * mu * d + r * r));
    G mu_d = clamp_cosine<G>((r * mu + d) / r_d);
    float3 n, q;
    if (ground) {
        n = transmittance_to_top<G>(P, T, r_d, -mu_d);
        q = transmittance_to_top<G>(P, T, r, -mu);
    } else {
        n = transmittance_to_top<G>(P, T, r, mu);
        q = transmittance_to_top<G>(P, T, r_d, mu_d);
    }
    return make_float3(fminf(n.x / q.x, 1.0f), fminf(n.y / q.y, 1.0f), fminf(n.z / q.z, 1.0f));
}
// Precision mode: transmittance lookups in extended range.  The bilinear
// lookup of transmittance.h:26-33 interpolates T_k = exp(-tau_k) over four
// texels; from the optical depths tau_k (the transmittance pass stores
// them beside T) the same sum is exp(-tau_0) * sum_k w_k exp(tau_0 - tau_k),
// tau_0 the smallest depth among taps of nonzero weight, evaluated in FP64.
// Long horizontal paths (tau > 87) have T below binary32's range; the
// quotient of two such lookups in get_transmittance then loses every digit,
// while this form keeps them.  Returns log(T) per channel.
__device__ __forceinline__ void log_transmittance_to_top(const DevParams& P,
                                                         const float4* __restrict__ OD,
                                                         double r, double mu, double out[3]) {
    double u, v;
    transmittance_uv<double>(P, r, mu, u, v);
    const Taps tx = linear_taps<double>(u, P.t_mu), ty = linear_taps<double>(v, P.t_r);
    const float4 q[4] = {OD[ty.i0 * P.t_mu + tx.i0], OD[ty.i0 * P.t_mu + tx.i1],
                         OD[ty.i1 * P.t_mu + tx.i0], OD[ty.i1 * P.t_mu + tx.i1]};
    const double wx = tx.w, wy = ty.w;
    const double w[4] = {(1.0 - wx) * (1.0 - wy), wx * (1.0 - wy), (1.0 - wx) * wy, wx * wy};
    for (int c = 0; c < 3; ++c) {
        double t[4];
        for (int k = 0; k < 4; ++k) t[k] = c == 0 ? q[k].x : (c == 1 ? q[k].y : q[k].z);
        double t0 = 1e300;
        for (int k = 0; k < 4; ++k)
            if (w[k] > 0.0) t0 = fmin(t0, t[k]);
        double s = 0.0;
        for (int k = 0; k < 4; ++k)
            if (w[k] > 0.0) s += w[k] * exp(t0 - t[k]);
        out[c] = log(s) - t0;
    }
}
// get_transmittance (transmittance.h:35-61) in extended range.
__device__ __forceinline__ float3 get_transmittance_ext(const DevParams& P,
                                                       const float4* __restrict__ OD, double r,
                                                       double mu, double d, bool ground) {
    const double r_d = clamp_radius<double>(P, sqrt(d * d + 2.0 * r * mu * d + r * r));
    const double mu_d = clamp_cosine<double>((r * mu + d) / r_d);
    double n[3], q[3];
    if (ground) {
        log_transmittance_to_top(P, OD, r_d, -mu_d, n);
        log_transmittance_to_top(P, OD, r, -mu, q);
    } else {
        log_transmittance_to_top(P, OD, r, mu, n);
        log_transmittance_to_top(P, OD, r_d, mu_d, q);
    }
    return make_float3(float(fmin(exp(n[0] - q[0]), 1.0)), float(fmin(exp(n[1] - q[1]), 1.0)),
                       float(fmin(exp(n[2] - q[2]), 1.0)));
}

// Precision mode's get_transmittance: the binary32 LUT quotient while its
// denominator is a normal number well above binary32's range limit (the
// quotient then keeps full precision), else the extended-range form.
__device__ __forceinline__ float3 get_transmittance_precise(const DevParams& P,
                                                           const float4* __restrict__ T,
                                                           const float4* __restrict__ OD,
                                                           double r, double mu, double d,
                                                           bool ground) {
    const double r_d = clamp_radius<double>(P, sqrt(d * d + 2.0 * r * mu * d + r * r));
    const double mu_d = clamp_cosine<double>((r * mu + d) / r_d);
    float3 n, q;
    if (ground) {
        n = transmittance_to_top<double>(P, T, r_d, -mu_d);
        q = transmittance_to_top<double>(P, T, r, -mu);
    } else {
        n = transmittance_to_top<double>(P, T, r, mu);
        q = transmittance_to_top<double>(P, T, r_d, mu_d);
    }
    if (fminf(q.x, fminf(q.y, q.z)) > 1e-30f)
        return make_float3(fminf(n.x / q.x, 1.0f), fminf(n.y / q.y, 1.0f), fminf(n.z / q.z, 1.0f));
    return get_transmittance_ext(P, OD, r, mu, d, ground);
}

// transmittance.h:63-74
template <typename G>
__device__ __forceinline__ float3 transmittance_to_sun(const DevParams& P,
                                                       const float4* __restrict__ T, G r,
                                                       G mu_s) {
    G sin_theta_h = G(P.bottom) / r;
    G cos_theta_h = -sqrt(fmax(G(1) - sin_theta_h * sin_theta_h, G(0)));
    float3 t = transmittance_to_top<G>(P, T, r, mu_s);
    float s = float(smoothstep_g<G>(-sin_theta_h * G(P.sun_angular_radius),
                                    sin_theta_h * G(P.sun_angular_radius), mu_s - cos_theta_h));
    return make_float3(t.x * s, t.y * s, t.z * s);
}
// transmittance.comp:48-64
template <typename G>
__device__ __forceinline__ void rmu_from_transmittance_unit(const DevParams& P, G x_mu, G x_r,
                                                            G& r_out, G& mu_out) {
    G H = sqrt(G(P.top) * G(P.top) - G(P.bottom) * G(P.bottom));
    G rho = H * x_r;
    G r = sqrt(rho * rho + G(P.bottom) * G(P.bottom));
    G d_min = G(P.top) - r;
    G d_max = rho + H;
    G d = d_min + x_mu * (d_max - d_min);
    G mu = d == G(0) ? G(1) : (H * H - rho * rho - d * d) / (G(2) * r * d);
    r_out = r;
    mu_out = clamp_cosine<G>(mu);
}

// ---------------------------------------------------------- irradiance.h
// irradiance.h:7-18
template <typename G>
__device__ __forceinline__ void rmus_from_irradiance_unit(const DevParams& P, G x_mu_s, G x_r,
                                                          G& r, G& mu_s) {
    r = G(P.bottom) + x_r * (G(P.top) - G(P.bottom));
    mu_s = clamp_cosine<G>(G(2) * x_mu_s - G(1));
}
// irradiance.h:20-30
template <typename G>
__device__ __forceinline__ void irradiance_uv(const DevParams& P, G r, G mu_s, G& u, G& v) {
    G x_r = (r - G(P.bottom)) / (G(P.top) - G(P.bottom));
    G x_mu_s = mu_s * G(0.5) + G(0.5);
    u = tex_from_unit<G>(x_mu_s, P.i_mus);
    v = tex_from_unit<G>(x_r, P.i_r);
}
// irradiance.h:32-38
template <typename G>
__device__ __forceinline__ float3 get_irradiance(const DevParams& P,
                                                 const float4* __restrict__ E, G r, G mu_s) {
    G u, v;
    irradiance_uv<G>(P, r, mu_s, u, v);
    float4 s = sample2d<G>(E, P.i_mus, P.i_r, u, v);
    return make_float3(s.x, s.y, s.z);
}

// ---------------------------------------------------------- scattering.h
// Per-component pieces of GetScatteringTextureUvwzFromRMuMuSNu
// (scattering.h:7-60), split so that callers can hoist the parts that are
// constant over a loop.
template <typename G>
__device__ __forceinline__ G scattering_u_r(const DevParams& P, G r) {
    G H = sqrt(G(P.top) * G(P.top) - G(P.bottom) * G(P.bottom));
    G rho = safe_sqrt<G>(r * r - G(P.bottom) * G(P.bottom));
    return tex_from_unit<G>(rho / H, P.s_r);
}
template <typename G>
__device__ __forceinline__ G scattering_u_mu(const DevParams& P, G r, G mu, bool ground) {
    G H = sqrt(G(P.top) * G(P.top) - G(P.bottom) * G(P.bottom));
    G rho = safe_sqrt<G>(r * r - G(P.bottom) * G(P.bottom));
    G r_mu = r * mu;
    G discriminant = r_mu * r_mu - r * r + G(P.bottom) * G(P.bottom);
    if (ground) {
        G d = -r_mu - safe_sqrt<G>(discriminant);
        G d_min = r - G(P.bottom);
        G d_max = rho;
        return G(0.5) - G(0.5) * tex_from_unit<G>(
                                     d_max == d_min ? G(0) : (d - d_min) / (d_max - d_min),
                                     P.s_mu / 2);
    } else {
        G d = -r_mu + safe_sqrt<G>(discriminant + H * H);
        G d_min = G(P.top) - r;
        G d_max = rho + H;
        return G(0.5) + G(0.5) * tex_from_unit<G>((d - d_min) / (d_max - d_min), P.s_mu / 2);
    }
}
template <typename G>
__device__ __forceinline__ G scattering_u_mu_s(const DevParams& P, G mu_s) {
    G H = sqrt(G(P.top) * G(P.top) - G(P.bottom) * G(P.bottom));
    G d = distance_to_top<G>(P, G(P.bottom), mu_s);
    G d_min = G(P.top) - G(P.bottom);
    G d_max = H;
    G a = (d - d_min) / (d_max - d_min);
    G A = G(-2) * G(P.mu_s_min) * G(P.bottom) / (d_max - d_min);
    return tex_from_unit<G>(fmax(G(1) - a / A, G(0)) / (G(1) + a), P.s_mus);
}

// scattering.h:62-114 and :116-137 — the (r, mu, mu_s, nu) of a scattering
// texel (x, y, z) of the packed (NU*MUS) x MU x R extent.
template <typename G>
struct TexelCoord {
    G r, mu, mu_s, nu;
    bool ground;
};
template <typename G>
__device__ __forceinline__ TexelCoord<G> scattering_texel_coord(const DevParams& P, unsigned x,
                                                                unsigned y, unsigned z) {
    unsigned W = unsigned(P.s_nu * P.s_mus);
    G fx = frag_from_texel<G>(x, W);
    G fy = frag_from_texel<G>(y, unsigned(P.s_mu));
    G fz = frag_from_texel<G>(z, unsigned(P.s_r));
    G frag_nu = floor(fx / G(P.s_mus));
    G frag_mu_s = fx - G(P.s_mus) * floor(fx / G(P.s_mus));  // GLSL mod
    G ux = frag_nu / G(P.s_nu - 1), uy = frag_mu_s / G(P.s_mus), uz = fy / G(P.s_mu),
      uw = fz / G(P.s_r);
    TexelCoord<G> c;
    G H = sqrt(G(P.top) * G(P.top) - G(P.bottom) * G(P.bottom));
    G rho = H * unit_from_tex<G>(uw, P.s_r);
    G r = sqrt(rho * rho + G(P.bottom) * G(P.bottom));
    G mu;
    if (uz < G(0.5)) {
        G d_min = r - G(P.bottom);
        G d_max = rho;
        G d = d_min + (d_max - d_min) * unit_from_tex<G>(G(1) - G(2) * uz, P.s_mu / 2);
        mu = d == G(0) ? G(-1) : clamp_cosine<G>(-(rho * rho + d * d) / (G(2) * r * d));
        c.ground = true;
    } else {
        G d_min = G(P.top) - r;
        G d_max = rho + H;
        G d = d_min + (d_max - d_min) * unit_from_tex<G>(G(2) * uz - G(1), P.s_mu / 2);
        mu = d == G(0) ? G(1) : clamp_cosine<G>((H * H - rho * rho - d * d) / (G(2) * r * d));
        c.ground = false;
    }
    G x_mu_s = unit_from_tex<G>(uy, P.s_mus);
    G d_min = G(P.top) - G(P.bottom);
    G d_max = H;
    G A = G(-2) * G(P.mu_s_min) * G(P.bottom) / (d_max - d_min);
    G a = (A - x_mu_s * A) / (G(1) + x_mu_s * A);
    G d = d_min + fmin(a, A) * (d_max - d_min);
    G mu_s = d == G(0) ? G(1) : clamp_cosine<G>((H * H - d * d) / (G(2) * G(P.bottom) * d));
    G nu = clamp_cosine<G>(ux * G(2) - G(1));
    G s = sqrt((G(1) - mu * mu) * (G(1) - mu_s * mu_s));
    nu = clampg<G>(nu, mu * mu_s - s, mu * mu_s + s);
    c.r = r;
    c.mu = mu;
    c.mu_s = mu_s;
    c.nu = nu;
    return c;
}

// Trilinear sample of a packed scattering LUT at texel coordinate
// (x within [0, W), u_mu, u_r) — the (y, z) taps are shared by the caller.
struct Taps3 {
    Taps y, z;
};

// scattering.h:139-155: GetScattering on one texture, nu lerped manually
// between two trilinear samples at slices floor(u_nu*(NU-1)) and +1.  H: the
// texture is binary16 (FB_FLAG_F16_STORAGE).
template <typename G, bool H>
__device__ __forceinline__ float4 get_scattering_tex(const DevParams& P,
                                                     const float4* __restrict__ tex, G r, G mu,
                                                     G mu_s, G nu, bool ground) {
    G u_r = scattering_u_r<G>(P, r);
    G u_mu = scattering_u_mu<G>(P, r, mu, ground);
    G u_mu_s = scattering_u_mu_s<G>(P, mu_s);
    G u_nu = (nu + G(1)) / G(2);
    G tex_coord_x = u_nu * G(P.s_nu - 1);
    G tex_x = floor(tex_coord_x);
    float lerp = float(tex_coord_x - tex_x);
    G x0 = (tex_x + u_mu_s) / G(P.s_nu);
    G x1 = (tex_x + G(1) + u_mu_s) / G(P.s_nu);
    int W = P.s_nu * P.s_mus;
    Taps ty = linear_taps<G>(u_mu, P.s_mu), tz = linear_taps<G>(u_r, P.s_r);
    Taps tx0 = linear_taps<G>(x0, W), tx1 = linear_taps<G>(x1, W);
    const size_t p00 = (size_t(tz.i0) * P.s_mu + ty.i0) * W;
    const size_t p01 = (size_t(tz.i0) * P.s_mu + ty.i1) * W;
    const size_t p10 = (size_t(tz.i1) * P.s_mu + ty.i0) * W;
    const size_t p11 = (size_t(tz.i1) * P.s_mu + ty.i1) * W;
    auto at = [&](size_t row, int x) { return ld_tex(tex, row + x, H); };
    float4 a0 = lerp4(lerp4(lerp4(at(p00, tx0.i0), at(p00, tx0.i1), tx0.w),
                            lerp4(at(p01, tx0.i0), at(p01, tx0.i1), tx0.w), ty.w),
                      lerp4(lerp4(at(p10, tx0.i0), at(p10, tx0.i1), tx0.w),
                            lerp4(at(p11, tx0.i0), at(p11, tx0.i1), tx0.w), ty.w),
                      tz.w);
    float4 a1 = lerp4(lerp4(lerp4(at(p00, tx1.i0), at(p00, tx1.i1), tx1.w),
                            lerp4(at(p01, tx1.i0), at(p01, tx1.i1), tx1.w), ty.w),
                      lerp4(lerp4(at(p10, tx1.i0), at(p10, tx1.i1), tx1.w),
                            lerp4(at(p11, tx1.i0), at(p11, tx1.i1), tx1.w), ty.w),
                      tz.w);
    return make_float4(a0.x * (1.0f - lerp) + a1.x * lerp, a0.y * (1.0f - lerp) + a1.y * lerp,
                       a0.z * (1.0f - lerp) + a1.z * lerp, a0.w * (1.0f - lerp) + a1.w * lerp);
}

}  // namespace fbk
