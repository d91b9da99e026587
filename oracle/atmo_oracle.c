/*
 * atmo_oracle.c — CPU restatement of fuzzyblue's precompute passes and sky /
 * lighting lookups.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU reference the GPU results are compared
 * with.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (fuzzyblue_amd/, libfuzzyblue_hip.so) never
 * links or calls it.
 *
 * It is compiled twice from this one source:
 *   ORACLE_REAL=double  -> *_f64 symbols: the FP64 oracle.  It evaluates the
 *                          SAME discretisation as the reference (sample
 *                          counts, trapezoid/midpoint weights, texture
 *                          parameterisations, Vulkan LINEAR/CLAMP_TO_EDGE
 *                          sampling) in double precision.
 *   ORACLE_REAL=float   -> *_f32 symbols: the FP32 restatement, built with
 *                          -fsingle-precision-constant -ffp-contract=off so
 *                          every operation rounds like the reference GLSL
 *                          written out in IEEE binary32 (modulo libm ulps).
 * Optional FP16 storage (fp16 != 0) rounds every store into a 3D texture to
 * binary16, as the reference's R16G16B16A16_SFLOAT images do
 * (src/precompute.rs:1215,1229-1234).
 *
 * Each function cites the reference file:line it restates.  Sampler
 * semantics follow the reference's VkSampler (src/precompute.rs:85-98):
 * LINEAR min/mag filter, CLAMP_TO_EDGE, one mip level, with exact (not
 * fixed-point) filter weights.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

#include "../include/fuzzyblue_hip.h"

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#define ORACLE_SFX _f64
#endif

typedef ORACLE_REAL real;

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, ORACLE_SFX)

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* Atmosphere parameters in working precision — shaders/params.h:26-87  */
/* ------------------------------------------------------------------ */
typedef struct {
    real width, exp_term, exp_scale, linear_term, constant_term;
} layer_t;
typedef struct {
    layer_t layers[2];
} profile_t;

typedef struct {
    real solar[3];
    real sun_angular_radius;
    real rayleigh_scattering[3];
    real bottom;
    real mie_scattering[3];
    real top;
    real mie_extinction[3];
    real g;
    real albedo[3];
    real mu_s_min;
    real absorption_extinction[3];
    int t_mu, t_r, s_r, s_mu, s_mus, s_nu, i_mus, i_r;
    profile_t rayleigh, mie, absorption;
    int fp16;
} atmo_t;

static void load_profile(profile_t *dst, const fb_density_profile *src) {
    for (int i = 0; i < 2; ++i) {
        dst->layers[i].width = src->layers[i].width;
        dst->layers[i].exp_term = src->layers[i].exp_term;
        dst->layers[i].exp_scale = src->layers[i].exp_scale;
        dst->layers[i].linear_term = src->layers[i].linear_term;
        dst->layers[i].constant_term = src->layers[i].constant_term;
    }
}

/* ParamsRaw::new (src/precompute.rs:966-993): a field-by-field copy. */
static void load_atmo(atmo_t *a, const fb_parameters *p, int fp16) {
    for (int c = 0; c < 3; ++c) {
        a->solar[c] = p->solar_irradiance[c];
        a->rayleigh_scattering[c] = p->rayleigh_scattering[c];
        a->mie_scattering[c] = p->mie_scattering[c];
        a->mie_extinction[c] = p->mie_extinction[c];
        a->albedo[c] = p->ground_albedo[c];
        a->absorption_extinction[c] = p->absorption_extinction[c];
    }
    a->sun_angular_radius = p->sun_angular_radius;
    a->bottom = p->bottom_radius;
    a->top = p->top_radius;
    a->g = p->mie_phase_function_g;
    a->mu_s_min = p->mu_s_min;
    a->t_mu = (int)p->transmittance_mu_size;
    a->t_r = (int)p->transmittance_r_size;
    a->s_r = (int)p->scattering_r_size;
    a->s_mu = (int)p->scattering_mu_size;
    a->s_mus = (int)p->scattering_mu_s_size;
    a->s_nu = (int)p->scattering_nu_size;
    a->i_mus = (int)p->irradiance_mu_s_size;
    a->i_r = (int)p->irradiance_r_size;
    load_profile(&a->rayleigh, &p->rayleigh_density);
    load_profile(&a->mie, &p->mie_density);
    load_profile(&a->absorption, &p->absorption_density);
    a->fp16 = fp16;
}

/* ------------------------------------------------------------------ */
/* Scalar utilities — shaders/util.h                                    */
/* ------------------------------------------------------------------ */
static const real PI = 3.14159265358979323846;

static real clampr(real x, real lo, real hi) { return fmin(fmax(x, lo), hi); }

/* util.h:6-8 */
static real clamp_cosine(real mu) { return clampr(mu, -1.0, 1.0); }
/* util.h:10-12 */
static real clamp_distance(real d) { return fmax(d, 0.0); }
/* util.h:14-16 */
static real safe_sqrt(real a) { return sqrt(fmax(a, 0.0)); }
/* util.h:18-20 */
static real tex_from_unit(real x, int n) {
    return 0.5 / (real)n + x * (1.0 - 1.0 / (real)n);
}
/* util.h:22-24 */
static real unit_from_tex(real u, int n) {
    return (u - 0.5 / (real)n) / (1.0 - 1.0 / (real)n);
}
/* util.h:26-29 */
static real phase_rayleigh(real nu) {
    real k = 3.0 / (16.0 * PI);
    return k * (1.0 + nu * nu);
}
/* util.h:31-34 */
static real phase_mie(real g, real nu) {
    real k = 3.0 / (8.0 * PI) * (1.0 - g * g) / (2.0 + g * g);
    return k * (1.0 + nu * nu) / pow(1.0 + g * g - 2.0 * g * nu, 1.5);
}
/* util.h:36-38: uint x converted to float, divided by float(size-1). */
static real frag_from_texel(unsigned x, unsigned n) {
    return (real)n * tex_from_unit((real)x / (real)(n - 1), (int)n);
}
/* GLSL smoothstep */
static real smoothstep_r(real e0, real e1, real x) {
    real t = clampr((x - e0) / (e1 - e0), 0.0, 1.0);
    return t * t * (3.0 - 2.0 * t);
}
/* GLSL mod */
static real glsl_mod(real x, real y) { return x - y * floor(x / y); }

/* binary16 round-to-nearest-even of a value, returned widened. */
static real round_half(real v) {
    float f = (float)v;
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = x & 0x80000000u;
    uint32_t absx = x & 0x7fffffffu;
    uint16_t h;
    if (absx >= 0x7f800000u) {
        h = (uint16_t)((sign >> 16) | (absx > 0x7f800000u ? 0x7e00u : 0x7c00u));
    } else if (absx >= 0x477ff000u) { /* >= 65520 rounds to inf */
        h = (uint16_t)((sign >> 16) | 0x7c00u);
    } else if (absx < 0x38800000u) { /* subnormal half or zero */
        if (absx < 0x33000000u) {
            h = (uint16_t)(sign >> 16);
        } else {
            uint32_t e = absx >> 23;
            uint32_t m = (absx & 0x7fffffu) | 0x800000u;
            uint32_t shift = 126u - e; /* 14..24 */
            uint32_t q = m >> shift;
            uint32_t rem = m & ((1u << shift) - 1u);
            uint32_t half = 1u << (shift - 1u);
            if (rem > half || (rem == half && (q & 1u))) q++;
            h = (uint16_t)((sign >> 16) | q);
        }
    } else {
        uint32_t q = ((absx - 0x38000000u) >> 13);
        uint32_t rem = absx & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
        h = (uint16_t)((sign >> 16) | q);
    }
    /* widen back */
    uint32_t hs = (uint32_t)(h & 0x8000u) << 16;
    uint32_t he = (h >> 10) & 0x1fu;
    uint32_t hm = h & 0x3ffu;
    uint32_t out;
    if (he == 0) {
        if (hm == 0) {
            out = hs;
        } else {
            float m = (float)hm * (1.0f / 16777216.0f); /* 2^-24 */
            memcpy(&out, &m, 4);
            out |= hs;
        }
    } else if (he == 31) {
        out = hs | 0x7f800000u | (hm << 13);
    } else {
        out = hs | ((he + 112u) << 23) | (hm << 13);
    }
    float r;
    memcpy(&r, &out, 4);
    return (real)r;
}

/* ------------------------------------------------------------------ */
/* Density profiles and geometry — shaders/params.h:89-133              */
/* ------------------------------------------------------------------ */
/* params.h:89-93 */
static real layer_density(const layer_t *l, real altitude) {
    real density = l->exp_term * exp(l->exp_scale * altitude) +
                   l->linear_term * altitude + l->constant_term;
    return clampr(density, 0.0, 1.0);
}
/* params.h:95-99 */
static real profile_density(const profile_t *p, real altitude) {
    return altitude < p->layers[0].width ? layer_density(&p->layers[0], altitude)
                                         : layer_density(&p->layers[1], altitude);
}
/* params.h:101-103 */
static real clamp_radius(const atmo_t *a, real r) {
    return clampr(r, a->bottom, a->top);
}
/* params.h:105-110 */
static real distance_to_top(const atmo_t *a, real r, real mu) {
    real discriminant = r * r * (mu * mu - 1.0) + a->top * a->top;
    return clamp_distance(-r * mu + safe_sqrt(discriminant));
}
/* params.h:112-117 */
static real distance_to_bottom(const atmo_t *a, real r, real mu) {
    real discriminant = r * r * (mu * mu - 1.0) + a->bottom * a->bottom;
    return clamp_distance(-r * mu - safe_sqrt(discriminant));
}
/* params.h:119-124 */
static int ray_intersects_ground(const atmo_t *a, real r, real mu) {
    return mu < 0.0 && r * r * (mu * mu - 1.0) + a->bottom * a->bottom >= 0.0;
}
/* params.h:126-133 */
static real distance_to_nearest(const atmo_t *a, real r, real mu, int ground) {
    return ground ? distance_to_bottom(a, r, mu) : distance_to_top(a, r, mu);
}

/* ------------------------------------------------------------------ */
/* Software sampler: VkSampler LINEAR + CLAMP_TO_EDGE                   */
/* (src/precompute.rs:85-98).  Texel coordinate = u*N - 0.5; taps       */
/* floor() and floor()+1 clamped to [0, N-1]; exact weights.            */
/* ------------------------------------------------------------------ */
typedef struct {
    int i0, i1;
    real w; /* weight of i1 */
} taps_t;

static taps_t linear_taps(real u, int n) {
    real x = u * (real)n - 0.5;
    /* Clamp before converting so that out-of-range or NaN coordinates
     * behave like CLAMP_TO_EDGE (fmax/fmin drop a NaN operand). */
    x = fmin(fmax(x, -1.0), (real)n);
    real f = floor(x);
    taps_t t;
    t.w = x - f;
    int i = (int)f;
    t.i0 = i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    t.i1 = i + 1 < 0 ? 0 : (i + 1 > n - 1 ? n - 1 : i + 1);
    return t;
}

static void sample2d(const real *tex, int w, int h, real u, real v, real out[4]) {
    taps_t tx = linear_taps(u, w), ty = linear_taps(v, h);
    const real *t00 = tex + 4 * ((size_t)ty.i0 * w + tx.i0);
    const real *t01 = tex + 4 * ((size_t)ty.i0 * w + tx.i1);
    const real *t10 = tex + 4 * ((size_t)ty.i1 * w + tx.i0);
    const real *t11 = tex + 4 * ((size_t)ty.i1 * w + tx.i1);
    for (int c = 0; c < 4; ++c) {
        out[c] = (1.0 - tx.w) * (1.0 - ty.w) * t00[c] + tx.w * (1.0 - ty.w) * t01[c] +
                 (1.0 - tx.w) * ty.w * t10[c] + tx.w * ty.w * t11[c];
    }
}

static void sample3d(const real *tex, int w, int h, int d, real u, real v, real s,
                     real out[4]) {
    taps_t tx = linear_taps(u, w), ty = linear_taps(v, h), tz = linear_taps(s, d);
    int xs[2] = {tx.i0, tx.i1}, ys[2] = {ty.i0, ty.i1}, zs[2] = {tz.i0, tz.i1};
    real wx[2] = {1.0 - tx.w, tx.w}, wy[2] = {1.0 - ty.w, ty.w},
         wz[2] = {1.0 - tz.w, tz.w};
    for (int c = 0; c < 4; ++c) out[c] = 0.0;
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) {
                const real *t = tex + 4 * (((size_t)zs[k] * h + ys[j]) * w + xs[i]);
                real wt = wx[i] * wy[j] * wz[k];
                for (int c = 0; c < 4; ++c) out[c] += wt * t[c];
            }
}

/* ------------------------------------------------------------------ */
/* Transmittance — shaders/transmittance.h, transmittance.comp          */
/* ------------------------------------------------------------------ */
/* transmittance.h:7-24 */
static void transmittance_uv_from_rmu(const atmo_t *a, real r, real mu, real uv[2]) {
    real H = sqrt(a->top * a->top - a->bottom * a->bottom);
    real rho = safe_sqrt(r * r - a->bottom * a->bottom);
    real d = distance_to_top(a, r, mu);
    real d_min = a->top - r;
    real d_max = rho + H;
    real x_mu = (d - d_min) / (d_max - d_min);
    real x_r = rho / H;
    uv[0] = tex_from_unit(x_mu, a->t_mu);
    uv[1] = tex_from_unit(x_r, a->t_r);
}
/* transmittance.h:26-33 */
static void transmittance_to_top(const atmo_t *a, const real *T, real r, real mu,
                                 real out[3]) {
    real uv[2], s[4];
    transmittance_uv_from_rmu(a, r, mu, uv);
    sample2d(T, a->t_mu, a->t_r, uv[0], uv[1], s);
    out[0] = s[0];
    out[1] = s[1];
    out[2] = s[2];
}
/* transmittance.h:35-61 */
static void get_transmittance(const atmo_t *a, const real *T, real r, real mu, real d,
                              int ground, real out[3]) {
    real r_d = clamp_radius(a, sqrt(d * d + 2.0 * r * mu * d + r * r));
    real mu_d = clamp_cosine((r * mu + d) / r_d);
    real n[3], q[3];
    if (ground) {
        transmittance_to_top(a, T, r_d, -mu_d, n);
        transmittance_to_top(a, T, r, -mu, q);
    } else {
        transmittance_to_top(a, T, r, mu, n);
        transmittance_to_top(a, T, r_d, mu_d, q);
    }
    for (int c = 0; c < 3; ++c) out[c] = fmin(n[c] / q[c], 1.0);
}
/* transmittance.h:63-74 */
static void transmittance_to_sun(const atmo_t *a, const real *T, real r, real mu_s,
                                 real out[3]) {
    real sin_theta_h = a->bottom / r;
    real cos_theta_h = -sqrt(fmax(1.0 - sin_theta_h * sin_theta_h, 0.0));
    real t[3];
    transmittance_to_top(a, T, r, mu_s, t);
    real s = smoothstep_r(-sin_theta_h * a->sun_angular_radius,
                          sin_theta_h * a->sun_angular_radius, mu_s - cos_theta_h);
    for (int c = 0; c < 3; ++c) out[c] = t[c] * s;
}
/* transmittance.comp:8-32 */
static real optical_length_to_top(const atmo_t *a, const profile_t *p, real r, real mu) {
    const int SAMPLE_COUNT = 500;
    real dx = distance_to_top(a, r, mu) / SAMPLE_COUNT;
    real result = 0.0;
    for (int i = 0; i <= SAMPLE_COUNT; ++i) {
        real d_i = (real)i * dx;
        real r_i = sqrt(d_i * d_i + 2.0 * r * mu * d_i + r * r);
        real y_i = profile_density(p, r_i - a->bottom);
        real weight_i = (i == 0 || i == SAMPLE_COUNT) ? 0.5 : 1.0;
        result += y_i * weight_i * dx;
    }
    return result;
}
/* transmittance.comp:34-46 */
static void compute_transmittance_to_top(const atmo_t *a, real r, real mu, real out[3]) {
    real lr = optical_length_to_top(a, &a->rayleigh, r, mu);
    real lm = optical_length_to_top(a, &a->mie, r, mu);
    real lo = optical_length_to_top(a, &a->absorption, r, mu);
    for (int c = 0; c < 3; ++c)
        out[c] = exp(-(a->rayleigh_scattering[c] * lr + a->mie_extinction[c] * lm +
                       a->absorption_extinction[c] * lo));
}
/* transmittance.comp:48-64 */
static void rmu_from_unit_ranges(const atmo_t *a, real x_mu, real x_r, real *r_out,
                                 real *mu_out) {
    real H = sqrt(a->top * a->top - a->bottom * a->bottom);
    real rho = H * x_r;
    real r = sqrt(rho * rho + a->bottom * a->bottom);
    real d_min = a->top - r;
    real d_max = rho + H;
    real d = d_min + x_mu * (d_max - d_min);
    real mu = d == 0.0 ? 1.0 : (H * H - rho * rho - d * d) / (2.0 * r * d);
    *r_out = r;
    *mu_out = clamp_cosine(mu);
}

/* ------------------------------------------------------------------ */
/* Irradiance parameterisation — shaders/irradiance.h                   */
/* ------------------------------------------------------------------ */
/* irradiance.h:7-18 */
static void rmus_from_irradiance_unit(const atmo_t *a, real x_mu_s, real x_r, real *r,
                                      real *mu_s) {
    *r = a->bottom + x_r * (a->top - a->bottom);
    *mu_s = clamp_cosine(2.0 * x_mu_s - 1.0);
}
/* irradiance.h:20-30 */
static void irradiance_uv_from_rmus(const atmo_t *a, real r, real mu_s, real uv[2]) {
    real x_r = (r - a->bottom) / (a->top - a->bottom);
    real x_mu_s = mu_s * 0.5 + 0.5;
    uv[0] = tex_from_unit(x_mu_s, a->i_mus);
    uv[1] = tex_from_unit(x_r, a->i_r);
}
/* irradiance.h:32-38 */
static void get_irradiance(const atmo_t *a, const real *E, real r, real mu_s,
                           real out[3]) {
    real uv[2], s[4];
    irradiance_uv_from_rmus(a, r, mu_s, uv);
    sample2d(E, a->i_mus, a->i_r, uv[0], uv[1], s);
    out[0] = s[0];
    out[1] = s[1];
    out[2] = s[2];
}

/* ------------------------------------------------------------------ */
/* Scattering parameterisation — shaders/scattering.h                   */
/* ------------------------------------------------------------------ */
/* scattering.h:7-60 */
static void scattering_uvwz(const atmo_t *a, real r, real mu, real mu_s, real nu,
                            int ground, real out[4]) {
    real H = sqrt(a->top * a->top - a->bottom * a->bottom);
    real rho = safe_sqrt(r * r - a->bottom * a->bottom);
    real u_r = tex_from_unit(rho / H, a->s_r);
    real r_mu = r * mu;
    real discriminant = r_mu * r_mu - r * r + a->bottom * a->bottom;
    real u_mu;
    if (ground) {
        real d = -r_mu - safe_sqrt(discriminant);
        real d_min = r - a->bottom;
        real d_max = rho;
        u_mu = 0.5 - 0.5 * tex_from_unit(d_max == d_min ? 0.0
                                                        : (d - d_min) / (d_max - d_min),
                                         a->s_mu / 2);
    } else {
        real d = -r_mu + safe_sqrt(discriminant + H * H);
        real d_min = a->top - r;
        real d_max = rho + H;
        u_mu = 0.5 + 0.5 * tex_from_unit((d - d_min) / (d_max - d_min), a->s_mu / 2);
    }
    real d = distance_to_top(a, a->bottom, mu_s);
    real d_min = a->top - a->bottom;
    real d_max = H;
    real aa = (d - d_min) / (d_max - d_min);
    real A = -2.0 * a->mu_s_min * a->bottom / (d_max - d_min);
    real u_mu_s = tex_from_unit(fmax(1.0 - aa / A, 0.0) / (1.0 + aa), a->s_mus);
    real u_nu = (nu + 1.0) / 2.0;
    out[0] = u_nu;
    out[1] = u_mu_s;
    out[2] = u_mu;
    out[3] = u_r;
}
/* scattering.h:62-114 */
static void rmumusnu_from_uvwz(const atmo_t *a, const real uvwz[4], real *r_out,
                               real *mu_out, real *mu_s_out, real *nu_out,
                               int *ground_out) {
    real H = sqrt(a->top * a->top - a->bottom * a->bottom);
    real rho = H * unit_from_tex(uvwz[3], a->s_r);
    real r = sqrt(rho * rho + a->bottom * a->bottom);
    real mu;
    if (uvwz[2] < 0.5) {
        real d_min = r - a->bottom;
        real d_max = rho;
        real d = d_min + (d_max - d_min) * unit_from_tex(1.0 - 2.0 * uvwz[2], a->s_mu / 2);
        mu = d == 0.0 ? -1.0 : clamp_cosine(-(rho * rho + d * d) / (2.0 * r * d));
        *ground_out = 1;
    } else {
        real d_min = a->top - r;
        real d_max = rho + H;
        real d = d_min + (d_max - d_min) * unit_from_tex(2.0 * uvwz[2] - 1.0, a->s_mu / 2);
        mu = d == 0.0 ? 1.0 : clamp_cosine((H * H - rho * rho - d * d) / (2.0 * r * d));
        *ground_out = 0;
    }
    real x_mu_s = unit_from_tex(uvwz[1], a->s_mus);
    real d_min = a->top - a->bottom;
    real d_max = H;
    real A = -2.0 * a->mu_s_min * a->bottom / (d_max - d_min);
    real aa = (A - x_mu_s * A) / (1.0 + x_mu_s * A);
    real d = d_min + fmin(aa, A) * (d_max - d_min);
    real mu_s = d == 0.0 ? 1.0 : clamp_cosine((H * H - d * d) / (2.0 * a->bottom * d));
    *r_out = r;
    *mu_out = mu;
    *mu_s_out = mu_s;
    *nu_out = clamp_cosine(uvwz[0] * 2.0 - 1.0);
}
/* scattering.h:116-137 */
static void rmumusnu_from_frag(const atmo_t *a, const real frag[3], real *r, real *mu,
                               real *mu_s, real *nu, int *ground) {
    real size[4] = {(real)(a->s_nu - 1), (real)a->s_mus, (real)a->s_mu, (real)a->s_r};
    real frag_nu = floor(frag[0] / (real)a->s_mus);
    real frag_mu_s = glsl_mod(frag[0], (real)a->s_mus);
    real uvwz[4] = {frag_nu / size[0], frag_mu_s / size[1], frag[1] / size[2],
                    frag[2] / size[3]};
    rmumusnu_from_uvwz(a, uvwz, r, mu, mu_s, nu, ground);
    real s = sqrt((1.0 - *mu * *mu) * (1.0 - *mu_s * *mu_s));
    *nu = clampr(*nu, *mu * *mu_s - s, *mu * *mu_s + s);
}
/* scattering.h:181-190 with util.h:40-45 */
static void scattering_frag_from_texel(const atmo_t *a, unsigned x, unsigned y,
                                       unsigned z, real frag[3]) {
    unsigned w = (unsigned)(a->s_nu * a->s_mus), h = (unsigned)a->s_mu,
             d = (unsigned)a->s_r;
    frag[0] = frag_from_texel(x, w);
    frag[1] = frag_from_texel(y, h);
    frag[2] = frag_from_texel(z, d);
}

/* scattering.h:139-155: manual nu lerp between two trilinear samples. */
static void get_scattering_tex(const atmo_t *a, const real *tex, real r, real mu,
                               real mu_s, real nu, int ground, real out[4]) {
    real uvwz[4];
    scattering_uvwz(a, r, mu, mu_s, nu, ground, uvwz);
    real tex_coord_x = uvwz[0] * (real)(a->s_nu - 1);
    real tex_x = floor(tex_coord_x);
    real lerp = tex_coord_x - tex_x;
    real x0 = (tex_x + uvwz[1]) / (real)a->s_nu;
    real x1 = (tex_x + 1.0 + uvwz[1]) / (real)a->s_nu;
    real s0[4], s1[4];
    int w = a->s_nu * a->s_mus;
    sample3d(tex, w, a->s_mu, a->s_r, x0, uvwz[2], uvwz[3], s0);
    sample3d(tex, w, a->s_mu, a->s_r, x1, uvwz[2], uvwz[3], s1);
    for (int c = 0; c < 4; ++c) out[c] = s0[c] * (1.0 - lerp) + s1[c] * lerp;
}
/* scattering.h:157-179 */
static void get_scattering_order(const atmo_t *a, const real *dR, const real *dM,
                                 const real *dMS, real r, real mu, real mu_s, real nu,
                                 int ground, int order, real out[3]) {
    real s[4];
    if (order == 1) {
        real m[4];
        get_scattering_tex(a, dR, r, mu, mu_s, nu, ground, s);
        get_scattering_tex(a, dM, r, mu, mu_s, nu, ground, m);
        real pr = phase_rayleigh(nu), pm = phase_mie(a->g, nu);
        for (int c = 0; c < 3; ++c) out[c] = s[c] * pr + m[c] * pm;
    } else {
        get_scattering_tex(a, dMS, r, mu, mu_s, nu, ground, s);
        for (int c = 0; c < 3; ++c) out[c] = s[c];
    }
}

static void store4(const atmo_t *a, real *dst, const real v[4], int is3d) {
    for (int c = 0; c < 4; ++c) dst[c] = (is3d && a->fp16) ? round_half(v[c]) : v[c];
}

/* ------------------------------------------------------------------ */
/* Pass 1: transmittance — transmittance.comp:71-80                      */
/* ------------------------------------------------------------------ */
static void pass_transmittance(const atmo_t *a, real *T) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int j = 0; j < a->t_r; ++j)
        for (int i = 0; i < a->t_mu; ++i) {
            real x_mu = (real)i / (real)(a->t_mu - 1);
            real x_r = (real)j / (real)(a->t_r - 1);
            real r, mu, t[3];
            rmu_from_unit_ranges(a, x_mu, x_r, &r, &mu);
            compute_transmittance_to_top(a, r, mu, t);
            real v[4] = {t[0], t[1], t[2], 1.0};
            store4(a, T + 4 * ((size_t)j * a->t_mu + i), v, 0);
        }
}

/* ------------------------------------------------------------------ */
/* Pass 2: direct irradiance — direct_irradiance.comp:10-46             */
/* ------------------------------------------------------------------ */
static void compute_direct_irradiance(const atmo_t *a, const real *T, real r, real mu_s,
                                      real out[3]) {
    real alpha_s = a->sun_angular_radius;
    real avg = mu_s < -alpha_s
                   ? 0.0
                   : (mu_s > alpha_s ? mu_s
                                     : (mu_s + alpha_s) * (mu_s + alpha_s) / (4.0 * alpha_s));
    real t[3];
    transmittance_to_top(a, T, r, mu_s, t);
    for (int c = 0; c < 3; ++c) out[c] = a->solar[c] * t[c] * avg;
}
static void pass_direct_irradiance(const atmo_t *a, const real *T, real *dE) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < a->i_r; ++j)
        for (int i = 0; i < a->i_mus; ++i) {
            real x_mu_s = (real)i / (real)(a->i_mus - 1);
            real x_r = (real)j / (real)(a->i_r - 1);
            real r, mu_s, e[3];
            rmus_from_irradiance_unit(a, x_mu_s, x_r, &r, &mu_s);
            compute_direct_irradiance(a, T, r, mu_s, e);
            real v[4] = {e[0], e[1], e[2], 0.0};
            store4(a, dE + 4 * ((size_t)j * a->i_mus + i), v, 0);
        }
}

/* ------------------------------------------------------------------ */
/* Pass 3: single scattering — single_scattering.comp:10-101            */
/* ------------------------------------------------------------------ */
static void single_scattering_integrand(const atmo_t *a, const real *T, real r, real mu,
                                        real mu_s, real nu, real d, int ground,
                                        real rayleigh[3], real mie[3]) {
    real r_d = clamp_radius(a, sqrt(d * d + 2.0 * r * mu * d + r * r));
    real mu_s_d = clamp_cosine((r * mu_s + d * nu) / r_d);
    real tv[3], ts[3];
    get_transmittance(a, T, r, mu, d, ground, tv);
    transmittance_to_sun(a, T, r_d, mu_s_d, ts);
    real rd = profile_density(&a->rayleigh, r_d - a->bottom);
    real md = profile_density(&a->mie, r_d - a->bottom);
    for (int c = 0; c < 3; ++c) {
        real t = tv[c] * ts[c];
        rayleigh[c] = t * rd;
        mie[c] = t * md;
    }
}
static void compute_single_scattering(const atmo_t *a, const real *T, real r, real mu,
                                      real mu_s, real nu, int ground, real rayleigh[3],
                                      real mie[3]) {
    const int SAMPLE_COUNT = 50;
    real dx = distance_to_nearest(a, r, mu, ground) / (real)SAMPLE_COUNT;
    real rs[3] = {0, 0, 0}, ms[3] = {0, 0, 0};
    for (int i = 0; i <= SAMPLE_COUNT; ++i) {
        real d_i = (real)i * dx;
        real ri[3], mi[3];
        single_scattering_integrand(a, T, r, mu, mu_s, nu, d_i, ground, ri, mi);
        real w = (i == 0 || i == SAMPLE_COUNT) ? 0.5 : 1.0;
        for (int c = 0; c < 3; ++c) {
            rs[c] += ri[c] * w;
            ms[c] += mi[c] * w;
        }
    }
    for (int c = 0; c < 3; ++c) {
        rayleigh[c] = rs[c] * dx * a->solar[c] * a->rayleigh_scattering[c];
        mie[c] = ms[c] * dx * a->solar[c] * a->mie_scattering[c];
    }
}
static void texel_rmumusnu(const atmo_t *a, unsigned x, unsigned y, unsigned z, real *r,
                           real *mu, real *mu_s, real *nu, int *ground) {
    real frag[3];
    scattering_frag_from_texel(a, x, y, z, frag);
    rmumusnu_from_frag(a, frag, r, mu, mu_s, nu, ground);
}
static void pass_single_scattering(const atmo_t *a, const real *T, real *dR, real *dM,
                                   real *S) {
    int w = a->s_nu * a->s_mus;
    long n = (long)w * a->s_mu * a->s_r;
#pragma omp parallel for schedule(dynamic, 64)
    for (long idx = 0; idx < n; ++idx) {
        unsigned x = (unsigned)(idx % w), y = (unsigned)((idx / w) % a->s_mu),
                 z = (unsigned)(idx / ((long)w * a->s_mu));
        real r, mu, mu_s, nu, ray[3], mie[3];
        int ground;
        texel_rmumusnu(a, x, y, z, &r, &mu, &mu_s, &nu, &ground);
        compute_single_scattering(a, T, r, mu, mu_s, nu, ground, ray, mie);
        real v0[4] = {ray[0], ray[1], ray[2], 0.0};
        real v1[4] = {mie[0], mie[1], mie[2], 0.0};
        real v2[4] = {ray[0], ray[1], ray[2], mie[0]};
        store4(a, dR + 4 * idx, v0, 1);
        store4(a, dM + 4 * idx, v1, 1);
        store4(a, S + 4 * idx, v2, 1);
    }
}

/* ------------------------------------------------------------------ */
/* Pass 4: scattering density — scattering_density.comp:10-155         */
/* ------------------------------------------------------------------ */
static void compute_scattering_density(const atmo_t *a, const real *T, const real *dR,
                                       const real *dM, const real *dMS, const real *dE,
                                       real r, real mu, real mu_s, real nu, int order,
                                       real out[3]) {
    real zenith[3] = {0.0, 0.0, 1.0};
    real omega[3] = {sqrt(1.0 - mu * mu), 0.0, mu};
    real sun_dir_x = omega[0] == 0.0 ? 0.0 : (nu - mu * mu_s) / omega[0];
    real sun_dir_y = sqrt(fmax(1.0 - sun_dir_x * sun_dir_x - mu_s * mu_s, 0.0));
    real omega_s[3] = {sun_dir_x, sun_dir_y, mu_s};
    const int SAMPLE_COUNT = 16;
    const real dphi = PI / (real)SAMPLE_COUNT;
    const real dtheta = PI / (real)SAMPLE_COUNT;
    real rm[3] = {0, 0, 0};
    /* Loop-invariant in the reference too (evaluated per sample there). */
    real rayleigh_density = profile_density(&a->rayleigh, r - a->bottom);
    real mie_density = profile_density(&a->mie, r - a->bottom);
    for (int l = 0; l < SAMPLE_COUNT; ++l) {
        real theta = ((real)l + 0.5) * dtheta;
        real cos_theta = cos(theta);
        real sin_theta = sin(theta);
        int ground_l = ray_intersects_ground(a, r, cos_theta);
        real distance_to_ground = 0.0;
        real tg[3] = {0, 0, 0};
        real albedo[3] = {0, 0, 0};
        if (ground_l) {
            distance_to_ground = distance_to_bottom(a, r, cos_theta);
            get_transmittance(a, T, r, cos_theta, distance_to_ground, 1, tg);
            for (int c = 0; c < 3; ++c) albedo[c] = a->albedo[c];
        }
        for (int m = 0; m < 2 * SAMPLE_COUNT; ++m) {
            real phi = ((real)m + 0.5) * dphi;
            real omega_i[3] = {cos(phi) * sin_theta, sin(phi) * sin_theta, cos_theta};
            real domega_i = dtheta * dphi * sin(theta);
            real nu1 = omega_s[0] * omega_i[0] + omega_s[1] * omega_i[1] +
                       omega_s[2] * omega_i[2];
            real inc[3];
            get_scattering_order(a, dR, dM, dMS, r, omega_i[2], mu_s, nu1, ground_l,
                                 order - 1, inc);
            real gn[3];
            for (int c = 0; c < 3; ++c) gn[c] = zenith[c] * r + omega_i[c] * distance_to_ground;
            real len = sqrt(gn[0] * gn[0] + gn[1] * gn[1] + gn[2] * gn[2]);
            for (int c = 0; c < 3; ++c) gn[c] = gn[c] / len;
            real gmu = gn[0] * omega_s[0] + gn[1] * omega_s[1] + gn[2] * omega_s[2];
            real ge[3];
            get_irradiance(a, dE, a->bottom, gmu, ge);
            for (int c = 0; c < 3; ++c) inc[c] += tg[c] * albedo[c] * (1.0 / PI) * ge[c];
            real nu2 = omega[0] * omega_i[0] + omega[1] * omega_i[1] + omega[2] * omega_i[2];
            real pr = phase_rayleigh(nu2), pm = phase_mie(a->g, nu2);
            for (int c = 0; c < 3; ++c)
                rm[c] += inc[c] *
                         (a->rayleigh_scattering[c] * rayleigh_density * pr +
                          a->mie_scattering[c] * mie_density * pm) *
                         domega_i;
        }
    }
    for (int c = 0; c < 3; ++c) out[c] = rm[c];
}
static void density_texel(const atmo_t *a, const real *T, const real *dR, const real *dM,
                          const real *dMS, const real *dE, int order, long idx, real out[4]) {
    int w = a->s_nu * a->s_mus;
    unsigned x = (unsigned)(idx % w), y = (unsigned)((idx / w) % a->s_mu),
             z = (unsigned)(idx / ((long)w * a->s_mu));
    real r, mu, mu_s, nu, j[3];
    int ground;
    texel_rmumusnu(a, x, y, z, &r, &mu, &mu_s, &nu, &ground);
    compute_scattering_density(a, T, dR, dM, dMS, dE, r, mu, mu_s, nu, order, j);
    out[0] = j[0];
    out[1] = j[1];
    out[2] = j[2];
    out[3] = 0.0;
}
static void pass_density(const atmo_t *a, const real *T, const real *dR, const real *dM,
                         const real *dMS, const real *dE, int order, real *J) {
    long n = (long)a->s_nu * a->s_mus * a->s_mu * a->s_r;
#pragma omp parallel for schedule(dynamic, 16)
    for (long idx = 0; idx < n; ++idx) {
        real v[4];
        density_texel(a, T, dR, dM, dMS, dE, order, idx, v);
        store4(a, J + 4 * idx, v, 1);
    }
}

/* ------------------------------------------------------------------ */
/* Pass 5: indirect irradiance — indirect_irradiance.comp:10-74         */
/* ------------------------------------------------------------------ */
static void compute_indirect_irradiance(const atmo_t *a, const real *dR, const real *dM,
                                        const real *dMS, real r, real mu_s, int order,
                                        real out[3]) {
    const int SAMPLE_COUNT = 32;
    const real dphi = PI / (real)SAMPLE_COUNT;
    const real dtheta = PI / (real)SAMPLE_COUNT;
    real result[3] = {0, 0, 0};
    real omega_s[3] = {sqrt(1.0 - mu_s * mu_s), 0.0, mu_s};
    for (int j = 0; j < SAMPLE_COUNT / 2; ++j) {
        real theta = ((real)j + 0.5) * dtheta;
        for (int i = 0; i < 2 * SAMPLE_COUNT; ++i) {
            real phi = ((real)i + 0.5) * dphi;
            real omega[3] = {cos(phi) * sin(theta), sin(phi) * sin(theta), cos(theta)};
            real domega = dtheta * dphi * sin(theta);
            real nu = omega[0] * omega_s[0] + omega[1] * omega_s[1] + omega[2] * omega_s[2];
            real s[3];
            get_scattering_order(a, dR, dM, dMS, r, omega[2], mu_s, nu, 0, order, s);
            for (int c = 0; c < 3; ++c) result[c] += s[c] * omega[2] * domega;
        }
    }
    for (int c = 0; c < 3; ++c) out[c] = result[c];
}
static void pass_indirect(const atmo_t *a, const real *dR, const real *dM, const real *dMS,
                          int order, real *dE, real *E) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < a->i_r; ++j)
        for (int i = 0; i < a->i_mus; ++i) {
            real x_mu_s = (real)i / (real)(a->i_mus - 1);
            real x_r = (real)j / (real)(a->i_r - 1);
            real r, mu_s, e[3];
            rmus_from_irradiance_unit(a, x_mu_s, x_r, &r, &mu_s);
            compute_indirect_irradiance(a, dR, dM, dMS, r, mu_s, order, e);
            size_t o = 4 * ((size_t)j * a->i_mus + i);
            real v[4] = {e[0], e[1], e[2], 0.0};
            store4(a, dE + o, v, 0);
            real acc[4] = {e[0] + E[o], e[1] + E[o + 1], e[2] + E[o + 2], 0.0 + E[o + 3]};
            store4(a, E + o, acc, 0);
        }
}

/* ------------------------------------------------------------------ */
/* Pass 6: multiple scattering — multiple_scattering.comp:9-93          */
/* ------------------------------------------------------------------ */
static void compute_multiple_scattering(const atmo_t *a, const real *T, const real *J,
                                        real r, real mu, real mu_s, real nu, int ground,
                                        real out[3]) {
    const int SAMPLE_COUNT = 50;
    real dx = distance_to_nearest(a, r, mu, ground) / (real)SAMPLE_COUNT;
    real sum[3] = {0, 0, 0};
    for (int i = 0; i <= SAMPLE_COUNT; ++i) {
        real d_i = (real)i * dx;
        real r_i = clamp_radius(a, sqrt(d_i * d_i + 2.0 * r * mu * d_i + r * r));
        real mu_i = clamp_cosine((r * mu + d_i) / r_i);
        real mu_s_i = clamp_cosine((r * mu_s + d_i * nu) / r_i);
        real s[4], t[3];
        get_scattering_tex(a, J, r_i, mu_i, mu_s_i, nu, ground, s);
        get_transmittance(a, T, r, mu, d_i, ground, t);
        real w = (i == 0 || i == SAMPLE_COUNT) ? 0.5 : 1.0;
        for (int c = 0; c < 3; ++c) sum[c] += s[c] * t[c] * dx * w;
    }
    for (int c = 0; c < 3; ++c) out[c] = sum[c];
}
static void multiple_texel(const atmo_t *a, const real *T, const real *J, long idx,
                           real ms_out[4], real *nu_out) {
    int w = a->s_nu * a->s_mus;
    unsigned x = (unsigned)(idx % w), y = (unsigned)((idx / w) % a->s_mu),
             z = (unsigned)(idx / ((long)w * a->s_mu));
    real r, mu, mu_s, nu, ms[3];
    int ground;
    texel_rmumusnu(a, x, y, z, &r, &mu, &mu_s, &nu, &ground);
    compute_multiple_scattering(a, T, J, r, mu, mu_s, nu, ground, ms);
    ms_out[0] = ms[0];
    ms_out[1] = ms[1];
    ms_out[2] = ms[2];
    ms_out[3] = 0.0;
    *nu_out = nu;
}
static void pass_multiple(const atmo_t *a, const real *T, const real *J, real *dMS,
                          real *S) {
    long n = (long)a->s_nu * a->s_mus * a->s_mu * a->s_r;
#pragma omp parallel for schedule(dynamic, 64)
    for (long idx = 0; idx < n; ++idx) {
        real v[4], nu;
        multiple_texel(a, T, J, idx, v, &nu);
        store4(a, dMS + 4 * idx, v, 1);
        real pr = phase_rayleigh(nu);
        real acc[4] = {v[0] / pr + S[4 * idx], v[1] / pr + S[4 * idx + 1],
                       v[2] / pr + S[4 * idx + 2], 0.0 + S[4 * idx + 3]};
        store4(a, S + 4 * idx, acc, 1);
    }
}

/* ------------------------------------------------------------------ */
/* Sky and lighting lookups — shaders/render_sky.h, render_lighting.h   */
/* ------------------------------------------------------------------ */
static real dot3(const real a[3], const real b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
/* render_sky.h:9-19 */
static void extrapolated_single_mie(const atmo_t *a, const real s[4], real out[3]) {
    if (s[0] <= 0.0) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    for (int c = 0; c < 3; ++c)
        out[c] = s[c] * s[3] / s[0] * (a->rayleigh_scattering[0] / a->mie_scattering[0]) *
                 (a->mie_scattering[c] / a->rayleigh_scattering[c]);
}
/* render_sky.h:21-43 */
static void combined_scattering(const atmo_t *a, const real *S, real r, real mu, real mu_s,
                                real nu, int ground, real rgb[3], real mie[3]) {
    real uvwz[4];
    scattering_uvwz(a, r, mu, mu_s, nu, ground, uvwz);
    real tex_coord_x = uvwz[0] * (real)(a->s_nu - 1);
    real tex_x = floor(tex_coord_x);
    real lerp = tex_coord_x - tex_x;
    real x0 = (tex_x + uvwz[1]) / (real)a->s_nu;
    real x1 = (tex_x + 1.0 + uvwz[1]) / (real)a->s_nu;
    real s0[4], s1[4], c4[4];
    int w = a->s_nu * a->s_mus;
    sample3d(S, w, a->s_mu, a->s_r, x0, uvwz[2], uvwz[3], s0);
    sample3d(S, w, a->s_mu, a->s_r, x1, uvwz[2], uvwz[3], s1);
    for (int c = 0; c < 4; ++c) c4[c] = s0[c] * (1.0 - lerp) + s1[c] * lerp;
    extrapolated_single_mie(a, c4, mie);
    rgb[0] = c4[0];
    rgb[1] = c4[1];
    rgb[2] = c4[2];
}
/* render_sky.h:45-109; shadow_length > 0 restates the light-shaft branch the
 * reference keeps commented out (render_sky.h:81-106). */
static void sky_radiance(const atmo_t *a, const real *T, const real *S, const real cam_in[3],
                         const real view[3], const real sun[3], real shadow_length, real rad[3],
                         real trans[3]) {
    real camera[3] = {cam_in[0], cam_in[1], cam_in[2]};
    real r = sqrt(dot3(camera, camera));
    real rmu = dot3(camera, view);
    real dtop = -rmu - sqrt(rmu * rmu - r * r + a->top * a->top);
    if (dtop > 0.0) {
        for (int c = 0; c < 3; ++c) camera[c] = camera[c] + view[c] * dtop;
        r = a->top;
        rmu += dtop;
    } else if (r > a->top) {
        trans[0] = trans[1] = trans[2] = 1.0;
        rad[0] = rad[1] = rad[2] = 0.0;
        return;
    }
    real mu = rmu / r;
    real mu_s = dot3(camera, sun) / r;
    real nu = dot3(view, sun);
    int ground = ray_intersects_ground(a, r, mu);
    if (ground) {
        trans[0] = trans[1] = trans[2] = 0.0;
    } else {
        transmittance_to_top(a, T, r, mu, trans);
    }
    real sc[3], mie[3];
    if (shadow_length == 0.0) {
        combined_scattering(a, S, r, mu, mu_s, nu, ground, sc, mie);
    } else {
        real d = shadow_length;
        real r_p = clamp_radius(a, sqrt(d * d + 2.0 * r * mu * d + r * r));
        real mu_p = (r * mu + d) / r_p;
        real mu_s_p = (r * mu_s + d * nu) / r_p;
        combined_scattering(a, S, r_p, mu_p, mu_s_p, nu, ground, sc, mie);
        real st[3];
        get_transmittance(a, T, r, mu, shadow_length, ground, st);
        for (int c = 0; c < 3; ++c) {
            sc[c] = sc[c] * st[c];
            mie[c] = mie[c] * st[c];
        }
    }
    real pr = phase_rayleigh(nu), pm = phase_mie(a->g, nu);
    for (int c = 0; c < 3; ++c) rad[c] = sc[c] * pr + mie[c] * pm;
}
/* render_sky.h:111-191; shadow_length > 0 restates the commented-out
 * light-shaft lines (render_sky.h:158-160, :172-177). */
static void sky_radiance_to_point(const atmo_t *a, const real *T, const real *S,
                                  const real cam_in[3], const real view[3],
                                  const real point[3], const real sun[3], real shadow_length,
                                  real rad[3], real trans[3]) {
    real camera[3] = {cam_in[0], cam_in[1], cam_in[2]};
    real r = sqrt(dot3(camera, camera));
    real rmu = dot3(camera, view);
    real dtop = -rmu - sqrt(rmu * rmu - r * r + a->top * a->top);
    if (dtop > 0.0) {
        for (int c = 0; c < 3; ++c) camera[c] = camera[c] + view[c] * dtop;
        r = a->top;
        rmu += dtop;
    } else if (r > a->top) {
        trans[0] = trans[1] = trans[2] = 1.0;
        rad[0] = rad[1] = rad[2] = 0.0;
        return;
    }
    real mu = rmu / r;
    real mu_s = dot3(camera, sun) / r;
    real nu = dot3(view, sun);
    real pc[3] = {point[0] - camera[0], point[1] - camera[1], point[2] - camera[2]};
    real d = sqrt(dot3(pc, pc));
    int ground = ray_intersects_ground(a, r, mu);
    get_transmittance(a, T, r, mu, d, ground, trans);
    real sc[3], mie[3];
    combined_scattering(a, S, r, mu, mu_s, nu, ground, sc, mie);
    if (!isinf(d)) {
        real st[3] = {trans[0], trans[1], trans[2]};
        if (shadow_length > 0.0) {
            d = fmax(d - shadow_length, 0.0);
            get_transmittance(a, T, r, mu, d, ground, st);
        }
        real r_p = clamp_radius(a, sqrt(d * d + 2.0 * r * mu * d + r * r));
        real mu_p = (r * mu + d) / r_p;
        real mu_s_p = (r * mu_s + d * nu) / r_p;
        real sc_p[3], mie_p[3];
        combined_scattering(a, S, r_p, mu_p, mu_s_p, nu, ground, sc_p, mie_p);
        for (int c = 0; c < 3; ++c) {
            sc[c] = sc[c] - st[c] * sc_p[c];
            mie[c] = mie[c] - st[c] * mie_p[c];
        }
        real s4[4] = {sc[0], sc[1], sc[2], mie[0]};
        extrapolated_single_mie(a, s4, mie);
        real h = smoothstep_r(0.0, 0.01, mu_s);
        for (int c = 0; c < 3; ++c) mie[c] = mie[c] * h;
    }
    real pr = phase_rayleigh(nu), pm = phase_mie(a->g, nu);
    for (int c = 0; c < 3; ++c) rad[c] = sc[c] * pr + mie[c] * pm;
}
/* render_lighting.h:10-28 */
static void sun_and_sky_irradiance(const atmo_t *a, const real *T, const real *E,
                                   const real point[3], const real normal[3],
                                   const real sun[3], real direct[3], real sky[3]) {
    real r = sqrt(dot3(point, point));
    real mu_s = dot3(point, sun) / r;
    real e[3];
    get_irradiance(a, E, r, mu_s, e);
    real f = (1.0 + dot3(normal, point) / r) * 0.5;
    for (int c = 0; c < 3; ++c) sky[c] = e[c] * f;
    real ts[3];
    transmittance_to_sun(a, T, r, mu_s, ts);
    real ns = fmax(dot3(normal, sun), 0.0);
    for (int c = 0; c < 3; ++c) direct[c] = a->solar[c] * ts[c] * ns;
}

/* ================================================================== */
/* Exported entry points (ctypes)                                       */
/* ================================================================== */
EXPORT int FN(oracle_real_bytes)(void) { return (int)sizeof(real); }

EXPORT real FN(oracle_phase_rayleigh)(real nu) { return phase_rayleigh(nu); }
EXPORT real FN(oracle_phase_mie)(real g, real nu) { return phase_mie(g, nu); }
EXPORT real FN(oracle_tex_from_unit)(real x, int n) { return tex_from_unit(x, n); }
EXPORT real FN(oracle_unit_from_tex)(real u, int n) { return unit_from_tex(u, n); }
EXPORT real FN(oracle_frag_from_texel)(unsigned x, unsigned n) { return frag_from_texel(x, n); }
EXPORT real FN(oracle_round_half)(real v) { return round_half(v); }

EXPORT real FN(oracle_profile_density)(const fb_parameters *p, int which, real h) {
    atmo_t a;
    load_atmo(&a, p, 0);
    const profile_t *pr = which == 0 ? &a.rayleigh : (which == 1 ? &a.mie : &a.absorption);
    return profile_density(pr, h);
}
EXPORT real FN(oracle_distance_to_top)(const fb_parameters *p, real r, real mu) {
    atmo_t a;
    load_atmo(&a, p, 0);
    return distance_to_top(&a, r, mu);
}
EXPORT real FN(oracle_distance_to_bottom)(const fb_parameters *p, real r, real mu) {
    atmo_t a;
    load_atmo(&a, p, 0);
    return distance_to_bottom(&a, r, mu);
}
EXPORT int FN(oracle_ray_intersects_ground)(const fb_parameters *p, real r, real mu) {
    atmo_t a;
    load_atmo(&a, p, 0);
    return ray_intersects_ground(&a, r, mu);
}
EXPORT void FN(oracle_transmittance_uv)(const fb_parameters *p, real r, real mu, real *uv) {
    atmo_t a;
    load_atmo(&a, p, 0);
    transmittance_uv_from_rmu(&a, r, mu, uv);
}
EXPORT void FN(oracle_transmittance_rmu)(const fb_parameters *p, real x_mu, real x_r,
                                         real *rmu) {
    atmo_t a;
    load_atmo(&a, p, 0);
    rmu_from_unit_ranges(&a, x_mu, x_r, &rmu[0], &rmu[1]);
}
EXPORT real FN(oracle_optical_length)(const fb_parameters *p, int which, real r, real mu) {
    atmo_t a;
    load_atmo(&a, p, 0);
    const profile_t *pr = which == 0 ? &a.rayleigh : (which == 1 ? &a.mie : &a.absorption);
    return optical_length_to_top(&a, pr, r, mu);
}
EXPORT void FN(oracle_irradiance_uv)(const fb_parameters *p, real r, real mu_s, real *uv) {
    atmo_t a;
    load_atmo(&a, p, 0);
    irradiance_uv_from_rmus(&a, r, mu_s, uv);
}
EXPORT void FN(oracle_scattering_uvwz)(const fb_parameters *p, real r, real mu, real mu_s,
                                       real nu, int ground, real *uvwz) {
    atmo_t a;
    load_atmo(&a, p, 0);
    scattering_uvwz(&a, r, mu, mu_s, nu, ground, uvwz);
}
EXPORT int FN(oracle_scattering_from_uvwz)(const fb_parameters *p, const real *uvwz,
                                           real *rmumusnu) {
    atmo_t a;
    int ground;
    load_atmo(&a, p, 0);
    rmumusnu_from_uvwz(&a, uvwz, &rmumusnu[0], &rmumusnu[1], &rmumusnu[2], &rmumusnu[3],
                       &ground);
    return ground;
}
EXPORT int FN(oracle_scattering_texel)(const fb_parameters *p, unsigned x, unsigned y,
                                       unsigned z, real *rmumusnu) {
    atmo_t a;
    int ground;
    load_atmo(&a, p, 0);
    texel_rmumusnu(&a, x, y, z, &rmumusnu[0], &rmumusnu[1], &rmumusnu[2], &rmumusnu[3],
                   &ground);
    return ground;
}
EXPORT void FN(oracle_sample2d)(const real *tex, int w, int h, real u, real v, real *out) {
    sample2d(tex, w, h, u, v, out);
}
EXPORT void FN(oracle_sample3d)(const real *tex, int w, int h, int d, real u, real v,
                                real s, real *out) {
    sample3d(tex, w, h, d, u, v, s, out);
}
EXPORT void FN(oracle_get_transmittance)(const fb_parameters *p, const real *T, real r,
                                         real mu, real d, int ground, real *out) {
    atmo_t a;
    load_atmo(&a, p, 0);
    get_transmittance(&a, T, r, mu, d, ground, out);
}
EXPORT void FN(oracle_get_scattering)(const fb_parameters *p, const real *tex, real r,
                                      real mu, real mu_s, real nu, int ground, real *out) {
    atmo_t a;
    load_atmo(&a, p, 0);
    get_scattering_tex(&a, tex, r, mu, mu_s, nu, ground, out);
}
/* The scattering density of scattering_density.comp:10-107 at an arbitrary
 * (r, mu, mu_s, nu), not only at texel centres. */
EXPORT void FN(oracle_density_at)(const fb_parameters *p, int order, const real *T,
                                  const real *dR, const real *dM, const real *dMS,
                                  const real *dE, real r, real mu, real mu_s, real nu,
                                  real *out) {
    atmo_t a;
    load_atmo(&a, p, 0);
    compute_scattering_density(&a, T, dR, dM, dMS, dE, r, mu, mu_s, nu, order, out);
}
/* render_sky.h:9-19 on one RGBA scattering value. */
EXPORT void FN(oracle_extrapolated_single_mie)(const fb_parameters *p, const real *s,
                                               real *out) {
    atmo_t a;
    load_atmo(&a, p, 0);
    extrapolated_single_mie(&a, s, out);
}

/* Full passes over whole LUTs, in the reference's order
 * (src/precompute.rs:1726-1999). */
EXPORT void FN(oracle_pass_transmittance)(const fb_parameters *p, real *T) {
    atmo_t a;
    load_atmo(&a, p, 0);
    pass_transmittance(&a, T);
}
EXPORT void FN(oracle_pass_direct_irradiance)(const fb_parameters *p, const real *T,
                                              real *dE) {
    atmo_t a;
    load_atmo(&a, p, 0);
    pass_direct_irradiance(&a, T, dE);
}
EXPORT void FN(oracle_pass_single_scattering)(const fb_parameters *p, int fp16, const real *T,
                                              real *dR, real *dM, real *S) {
    atmo_t a;
    load_atmo(&a, p, fp16);
    pass_single_scattering(&a, T, dR, dM, S);
}
EXPORT void FN(oracle_pass_density)(const fb_parameters *p, int fp16, int order,
                                    const real *T, const real *dR, const real *dM,
                                    const real *dMS, const real *dE, real *J) {
    atmo_t a;
    load_atmo(&a, p, fp16);
    pass_density(&a, T, dR, dM, dMS, dE, order, J);
}
EXPORT void FN(oracle_pass_indirect)(const fb_parameters *p, int order, const real *dR,
                                     const real *dM, const real *dMS, real *dE, real *E) {
    atmo_t a;
    load_atmo(&a, p, 0);
    pass_indirect(&a, dR, dM, dMS, order, dE, E);
}
EXPORT void FN(oracle_pass_multiple)(const fb_parameters *p, int fp16, const real *T,
                                     const real *J, real *dMS, real *S) {
    atmo_t a;
    load_atmo(&a, p, fp16);
    pass_multiple(&a, T, J, dMS, S);
}

/* Single-texel evaluations over caller-given inputs (for large LUTs where
 * a full CPU pass is too slow): density / multiple scattering of texels
 * idx[0..n) of the packed 3D extent. */
EXPORT void FN(oracle_density_texels)(const fb_parameters *p, int order, const real *T,
                                      const real *dR, const real *dM, const real *dMS,
                                      const real *dE, const long *idx, long n, real *out) {
    atmo_t a;
    load_atmo(&a, p, 0);
#pragma omp parallel for schedule(dynamic, 4)
    for (long k = 0; k < n; ++k) density_texel(&a, T, dR, dM, dMS, dE, order, idx[k], out + 4 * k);
}
EXPORT void FN(oracle_multiple_texels)(const fb_parameters *p, const real *T, const real *J,
                                       const long *idx, long n, real *out, real *nu_out) {
    atmo_t a;
    load_atmo(&a, p, 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (long k = 0; k < n; ++k) multiple_texel(&a, T, J, idx[k], out + 4 * k, nu_out + k);
}
EXPORT void FN(oracle_single_texels)(const fb_parameters *p, const real *T, const long *idx,
                                     long n, real *rayleigh, real *mie) {
    atmo_t a;
    load_atmo(&a, p, 0);
    int w = a.s_nu * a.s_mus;
#pragma omp parallel for schedule(dynamic, 16)
    for (long k = 0; k < n; ++k) {
        long i = idx[k];
        unsigned x = (unsigned)(i % w), y = (unsigned)((i / w) % a.s_mu),
                 z = (unsigned)(i / ((long)w * a.s_mu));
        real r, mu, mu_s, nu;
        int ground;
        texel_rmumusnu(&a, x, y, z, &r, &mu, &mu_s, &nu, &ground);
        compute_single_scattering(&a, T, r, mu, mu_s, nu, ground, rayleigh + 3 * k, mie + 3 * k);
    }
}

/* The whole precompute (src/precompute.rs:1726-1999): transmittance,
 * direct irradiance, single scattering, clear irradiance, then for
 * order n = 2..=order: density(n), indirect(n-1), multiple.
 * Scratch is allocated here.  When per_order_scattering / per_order_irradiance
 * are non-NULL they receive, for each order k = 1..order, the order-k
 * contribution: scattering slot k-1 = delta (single: rayleigh rgb with
 * phase-free layout; k>=2: delta_multiple_scattering), irradiance slot k-1 =
 * delta_irradiance after the indirect pass of order k (slot 0 = direct). */
EXPORT int FN(oracle_precompute)(const fb_parameters *p, int fp16, real *T, real *E, real *S,
                                 real *per_order_scattering, real *per_order_irradiance) {
    atmo_t a;
    load_atmo(&a, p, fp16);
    size_t nt = (size_t)a.t_mu * a.t_r, ni = (size_t)a.i_mus * a.i_r,
           ns = (size_t)a.s_nu * a.s_mus * a.s_mu * a.s_r;
    real *dE = calloc(4 * ni, sizeof(real));
    real *dR = calloc(4 * ns, sizeof(real));
    real *dM = calloc(4 * ns, sizeof(real));
    real *dMS = calloc(4 * ns, sizeof(real));
    real *J = calloc(4 * ns, sizeof(real));
    if (!dE || !dR || !dM || !dMS || !J) {
        free(dE);
        free(dR);
        free(dM);
        free(dMS);
        free(J);
        return -1;
    }
    (void)nt;
    pass_transmittance(&a, T);
    pass_direct_irradiance(&a, T, dE);
    pass_single_scattering(&a, T, dR, dM, S);
    memset(E, 0, 4 * ni * sizeof(real));
    if (per_order_irradiance) memcpy(per_order_irradiance, dE, 4 * ni * sizeof(real));
    if (per_order_scattering) memcpy(per_order_scattering, dR, 4 * ns * sizeof(real));
    for (int order = 2; order <= (int)p->order; ++order) {
        pass_density(&a, T, dR, dM, dMS, dE, order, J);
        pass_indirect(&a, dR, dM, dMS, order - 1, dE, E);
        pass_multiple(&a, T, J, dMS, S);
        if (per_order_irradiance)
            memcpy(per_order_irradiance + 4 * ni * (order - 1), dE, 4 * ni * sizeof(real));
        if (per_order_scattering)
            memcpy(per_order_scattering + 4 * ns * (order - 1), dMS, 4 * ns * sizeof(real));
    }
    free(dE);
    free(dR);
    free(dM);
    free(dMS);
    free(J);
    return 0;
}

/* Per-ray sky lookups (render_sky.h:45-191).  point == NULL selects
 * GetSkyRadiance, else GetSkyRadianceToPoint.  Arrays are packed 3-vectors;
 * shadow (one length per ray, km) may be NULL (no light shafts). */
EXPORT void FN(oracle_sky_rays)(const fb_parameters *p, const real *T, const real *S,
                                const real *cameras, int camera_stride, const real *views,
                                const real *points, const real *suns, int sun_stride,
                                const real *shadow, long n, real *rad, real *trans) {
    atmo_t a;
    load_atmo(&a, p, 0);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        const real *cam = cameras + 3 * (camera_stride ? i : 0);
        const real *sun = suns + 3 * (sun_stride ? i : 0);
        const real l = shadow ? shadow[i] : 0.0;
        if (points)
            sky_radiance_to_point(&a, T, S, cam, views + 3 * i, points + 3 * i, sun, l,
                                  rad + 3 * i, trans + 3 * i);
        else
            sky_radiance(&a, T, S, cam, views + 3 * i, sun, l, rad + 3 * i, trans + 3 * i);
    }
}

/* render_sky.frag:24-35 over a width x height frame; fullscreen.vert:5-8
 * gives screen_coords = ((px+0.5)/w, (py+0.5)/h) at pixel centres.
 * inverse_viewproj is column-major [col][row]; depth NULL = 0 (infinite). */
EXPORT void FN(oracle_render_frame)(const fb_parameters *p, const real *T, const real *S,
                                    const float *inverse_viewproj, const float *camera,
                                    const float *sun, const float *depth, int w, int h,
                                    real *color, real *trans) {
    atmo_t a;
    load_atmo(&a, p, 0);
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < (long)w * h; ++idx) {
        int px = (int)(idx % w), py = (int)(idx / w);
        real sx = ((real)px + 0.5) / (real)w, sy = ((real)py + 0.5) / (real)h;
        real ndc[4] = {2.0 * sx - 1.0, 2.0 * sy - 1.0, 0.0, 1.0};
        real v4[4], wp[4];
        for (int row = 0; row < 4; ++row) {
            v4[row] = 0.0;
            for (int col = 0; col < 4; ++col) v4[row] += (real)inverse_viewproj[col * 4 + row] * ndc[col];
        }
        real len = sqrt(v4[0] * v4[0] + v4[1] * v4[1] + v4[2] * v4[2]);
        real view[3] = {v4[0] / len, v4[1] / len, v4[2] / len};
        ndc[2] = depth ? (real)depth[idx] : 0.0;
        for (int row = 0; row < 4; ++row) {
            wp[row] = 0.0;
            for (int col = 0; col < 4; ++col) wp[row] += (real)inverse_viewproj[col * 4 + row] * ndc[col];
        }
        real world[3] = {wp[0] / wp[3] * 1e-3, wp[1] / wp[3] * 1e-3, wp[2] / wp[3] * 1e-3};
        real cam[3] = {camera[0], camera[1], camera[2]};
        real sn[3] = {sun[0], sun[1], sun[2]};
        real rad[3], t[3];
        sky_radiance_to_point(&a, T, S, cam, view, world, sn, 0.0, rad, t);
        for (int c = 0; c < 3; ++c) {
            color[4 * idx + c] = rad[c];
            trans[4 * idx + c] = t[c];
        }
        color[4 * idx + 3] = 0.0;
        trans[4 * idx + 3] = 1.0;
    }
}

EXPORT void FN(oracle_sun_and_sky_irradiance)(const fb_parameters *p, const real *T,
                                              const real *E, const real *points,
                                              const real *normals, const real *suns,
                                              int sun_stride, long n, real *direct,
                                              real *sky) {
    atmo_t a;
    load_atmo(&a, p, 0);
    for (long i = 0; i < n; ++i)
        sun_and_sky_irradiance(&a, T, E, points + 3 * i, normals + 3 * i,
                               suns + 3 * (sun_stride ? i : 0), direct + 3 * i, sky + 3 * i);
}
