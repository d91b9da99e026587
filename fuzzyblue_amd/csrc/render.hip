// render.hip — per-pixel sky and lighting lookups from a resident LUT set:
// shaders/render_sky.h (GetSkyRadiance, GetSkyRadianceToPoint),
// render_sky.frag (unprojection, dual-source output) and render_lighting.h
// (GetSunAndSkyIrradiance), as compute kernels that index pixels directly
// (the full-screen triangle of shaders/fullscreen.vert is not needed).
#include <string.h>

#include "atmo_device.hpp"
#include "runtime.hpp"

namespace fbk {

__device__ __forceinline__ float dot3(float3 a, float3 b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

// render_sky.h:9-19
__device__ __forceinline__ float3 extrapolated_single_mie(const DevParams& P, float4 s) {
    if (s.x <= 0.0f) return make_float3(0.0f, 0.0f, 0.0f);
    float k = P.rayleigh_scattering[0] / P.mie_scattering[0];
    return make_float3(s.x * s.w / s.x * k * (P.mie_scattering[0] / P.rayleigh_scattering[0]),
                       s.y * s.w / s.x * k * (P.mie_scattering[1] / P.rayleigh_scattering[1]),
                       s.z * s.w / s.x * k * (P.mie_scattering[2] / P.rayleigh_scattering[2]));
}

// render_sky.h:21-43 (GLSL mix of the two nu slices).  H: the scattering LUT
// is binary16 (FB_FLAG_F16_STORAGE); every function and kernel below that
// reads it takes the same parameter.
template <bool H>
__device__ __forceinline__ float4 combined_scattering(const DevParams& P,
                                                      const float4* __restrict__ S, float r,
                                                      float mu, float mu_s, float nu, bool ground,
                                                      float3& mie) {
    float4 c = get_scattering_tex<float, H>(P, S, r, mu, mu_s, nu, ground);
    mie = extrapolated_single_mie(P, c);
    return c;
}

// render_sky.h:45-109.  SHADOW: take the light-shaft term the reference
// keeps commented out (render_sky.h:81-106) for shadow_length > 0; a
// shadow_length of 0 is the reference's path.
template <bool H, bool SHADOW = false>
__device__ void sky_radiance(const DevParams& P, const float4* __restrict__ T,
                             const float4* __restrict__ S, float3 camera, float3 view,
                             float3 sun, float3& rad, float3& trans,
                             float shadow_length = 0.0f) {
    float r = sqrtf(dot3(camera, camera));
    float rmu = dot3(camera, view);
    float dtop = -rmu - sqrtf(rmu * rmu - r * r + P.top * P.top);
    if (dtop > 0.0f) {
        camera = make_float3(camera.x + view.x * dtop, camera.y + view.y * dtop,
                             camera.z + view.z * dtop);
        r = P.top;
        rmu += dtop;
    } else if (r > P.top) {
        trans = make_float3(1.0f, 1.0f, 1.0f);
        rad = make_float3(0.0f, 0.0f, 0.0f);
        return;
    }
    float mu = rmu / r;
    float mu_s = dot3(camera, sun) / r;
    float nu = dot3(view, sun);
    bool ground = ray_intersects_ground<float>(P, r, mu);
    trans = ground ? make_float3(0.0f, 0.0f, 0.0f) : transmittance_to_top<float>(P, T, r, mu);
    float3 mie;
    float4 sc;
    if (!SHADOW || shadow_length == 0.0f) {
        sc = combined_scattering<H>(P, S, r, mu, mu_s, nu, ground, mie);
    } else {
        // Eq. (18): the scattering seen from x + l v, attenuated by
        // T(x, x + l v) (render_sky.h:86-105).
        const float d = shadow_length;
        const float r_p = clamp_radius<float>(P, sqrtf(d * d + 2.0f * r * mu * d + r * r));
        const float mu_p = (r * mu + d) / r_p;
        const float mu_s_p = (r * mu_s + d * nu) / r_p;
        sc = combined_scattering<H>(P, S, r_p, mu_p, mu_s_p, nu, ground, mie);
        const float3 st = get_transmittance<float>(P, T, r, mu, shadow_length, ground);
        sc = make_float4(sc.x * st.x, sc.y * st.y, sc.z * st.z, sc.w);
        mie = make_float3(mie.x * st.x, mie.y * st.y, mie.z * st.z);
    }
    float pr = phase_rayleigh(nu), pm = phase_mie(P.g, nu);
    rad = make_float3(sc.x * pr + mie.x * pm, sc.y * pr + mie.y * pm, sc.z * pr + mie.z * pm);
}

// render_sky.h:111-191, with the commented-out light-shaft term
// (:158-160, :172-177) when SHADOW and shadow_length > 0.
template <bool H, bool SHADOW = false>
__device__ void sky_radiance_to_point(const DevParams& P, const float4* __restrict__ T,
                                      const float4* __restrict__ S, float3 camera, float3 view,
                                      float3 point, float3 sun, float3& rad, float3& trans,
                                      float shadow_length = 0.0f) {
    float r = sqrtf(dot3(camera, camera));
    float rmu = dot3(camera, view);
    float dtop = -rmu - sqrtf(rmu * rmu - r * r + P.top * P.top);
    if (dtop > 0.0f) {
        camera = make_float3(camera.x + view.x * dtop, camera.y + view.y * dtop,
                             camera.z + view.z * dtop);
        r = P.top;
        rmu += dtop;
    } else if (r > P.top) {
        trans = make_float3(1.0f, 1.0f, 1.0f);
        rad = make_float3(0.0f, 0.0f, 0.0f);
        return;
    }
    float mu = rmu / r;
    float mu_s = dot3(camera, sun) / r;
    float nu = dot3(view, sun);
    float3 pc = make_float3(point.x - camera.x, point.y - camera.y, point.z - camera.z);
    float d = sqrtf(dot3(pc, pc));
    bool ground = ray_intersects_ground<float>(P, r, mu);
    trans = get_transmittance<float>(P, T, r, mu, d, ground);
    float3 mie;
    float4 sc = combined_scattering<H>(P, S, r, mu, mu_s, nu, ground, mie);
    if (!isinf(d)) {
        // Eq. (17): the second lookup at d - l, attenuated by T(x, x_{d-l}).
        float3 st = trans;
        if (SHADOW && shadow_length > 0.0f) {
            d = fmaxf(d - shadow_length, 0.0f);
            st = get_transmittance<float>(P, T, r, mu, d, ground);
        }
        float r_p = clamp_radius<float>(P, sqrtf(d * d + 2.0f * r * mu * d + r * r));
        float mu_p = (r * mu + d) / r_p;
        float mu_s_p = (r * mu_s + d * nu) / r_p;
        float3 mie_p;
        float4 sc_p = combined_scattering<H>(P, S, r_p, mu_p, mu_s_p, nu, ground, mie_p);
        sc.x = sc.x - st.x * sc_p.x;
        sc.y = sc.y - st.y * sc_p.y;
        sc.z = sc.z - st.z * sc_p.z;
        float mr = mie.x - st.x * mie_p.x;
        mie = extrapolated_single_mie(P, make_float4(sc.x, sc.y, sc.z, mr));
        float h = smoothstep_g<float>(0.0f, 0.01f, mu_s);
        mie = make_float3(mie.x * h, mie.y * h, mie.z * h);
    }
    float pr = phase_rayleigh(nu), pm = phase_mie(P.g, nu);
    rad = make_float3(sc.x * pr + mie.x * pm, sc.y * pr + mie.y * pm, sc.z * pr + mie.z * pm);
}

// The LUTs at one (r, mu_s).  For a camera inside the atmosphere every ray
// of a frame shares the camera's r and, for one sun, mu_s: the r and mu_s
// interpolation of the scattering lookup (scattering.h:139-155) and the r
// interpolation of the transmittance lookup (transmittance.h:26-33) are the
// same for all of its pixels.  A block evaluates them once into LDS --
// Q[k][y] for nu slice k = 0..NU (slice NU is the clamped edge the lookup
// of the last slice's upper neighbour reads) and mu texel y, then Trow[m]
// for each transmittance mu texel -- and each pixel interpolates in mu and
// nu only: 4 + 2 LDS reads instead of 16 + 4 texel fetches, the same value
// up to the order of rounding.
__host__ __device__ inline size_t sky_slice_bytes(const DevParams& P) {
    return sizeof(float4) * (size_t(P.s_nu + 1) * P.s_mu + P.t_mu);
}
constexpr size_t kSkySliceMaxBytes = 65536;

// Entries first, first + step, ... of the slice at (r, mu_s) into `lds`
// (the block's LDS, or a frame's slice in global memory).
template <bool H>
__device__ void build_sky_slice(const DevParams& P, const float4* __restrict__ T,
                                const float4* __restrict__ S, float r, float mu_s,
                                float4* __restrict__ lds, int first, int step) {
    const int NU = P.s_nu, MU = P.s_mu, W = NU * P.s_mus, TMU = P.t_mu;
    const Taps tz = linear_taps<float>(scattering_u_r<float>(P, r), P.s_r);
    const float u_mu_s = scattering_u_mu_s<float>(P, mu_s);
    constexpr bool h = H;
    for (int e = first; e < (NU + 1) * MU; e += step) {
        const int k = e / MU, y = e - k * MU;
        const Taps tx = linear_taps<float>((float(k) + u_mu_s) / float(NU), W);
        const size_t p0 = (size_t(tz.i0) * MU + y) * W;
        const size_t p1 = (size_t(tz.i1) * MU + y) * W;
        lds[e] = lerp4(lerp4(ld_tex(S, p0 + tx.i0, h), ld_tex(S, p0 + tx.i1, h), tx.w),
                       lerp4(ld_tex(S, p1 + tx.i0, h), ld_tex(S, p1 + tx.i1, h), tx.w), tz.w);
    }
    float u, v;
    transmittance_uv<float>(P, r, 1.0f, u, v);  // v depends on r only
    const Taps tv = linear_taps<float>(v, P.t_r);
    float4* trow = lds + (NU + 1) * MU;
    for (int m = first; m < TMU; m += step)
        trow[m] = lerp4(ldg4(T + tv.i0 * TMU + m), ldg4(T + tv.i1 * TMU + m), tv.w);
}

// a / b correctly rounded, given y = 1/b correctly rounded and b normal: q =
// a*y, then Markstein's correction q + (a - b*q)*y with the residual exact
// in an FMA (P. Markstein, "Computation of elementary functions on the IBM
// RISC System/6000 processor", IBM J. Res. Dev. 34(1), 1990) -- three
// instructions in place of an IEEE division's ten, the same bits.  The
// residual is a multiple of about 2^(e_a - 46): for numerators below 2^-80
// (other than 0) it can be subnormal or not representable, and a wavefront
// with such a lane takes the division (a wave-uniform branch; the empty asm
// statement keeps the compiler from evaluating the division unconditionally
// and selecting).
__device__ __forceinline__ float div_rn(float a, float b, float y) {
    const float q = a * y;
    float d = fmaf(fmaf(-q, b, a), y, q);
    if (__builtin_amdgcn_ballot_w64(a != 0.0f && fabsf(a) < 0x1p-80f)) {
        asm volatile("");
        d = a / b;
    }
    return d;
}

// The per-frame terms of sky_radiance_slice: everything that depends on
// the camera's r and the atmosphere only, computed once per block.
struct SliceConsts {
    float r, r_inv, rr, b2, t2;   // r, 1/r, r*r, bottom^2, top^2
    float t_dmin, t_den, t_inv;   // transmittance x_mu = (d - t_dmin) / t_den
    float t_u0, t_u1;             // tex_from_unit(x, t_mu) = t_u0 + x * t_u1
    float g_dmin, g_den, g_inv;   // scattering u_mu, rays to the ground
    float s_dmin, s_den, s_inv;   // ... and to the top
    float h2;                     // H^2
    float m_u0, m_u1;             // tex_from_unit(x, MU/2)
    float mie[3];                 // (rayleigh.r / mie.r) * (mie / rayleigh)
    float pm_k, pm_g1, pm_g2;     // phase_mie: k, 1 + g^2, 2g
};
__device__ inline SliceConsts slice_consts(const DevParams& P, float r) {
    SliceConsts c;
    const float b = P.bottom, t = P.top;
    c.r = r;
    c.r_inv = 1.0f / r;
    c.rr = r * r;
    c.b2 = b * b;
    c.t2 = t * t;
    const float H = sqrtf(t * t - b * b);
    const float rho = sqrtf(fmaxf(r * r - b * b, 0.0f));
    // transmittance.h:7-24
    c.t_dmin = t - r;
    c.t_den = rho + H - c.t_dmin;
    c.t_inv = 1.0f / c.t_den;
    c.t_u0 = 0.5f / float(P.t_mu);
    c.t_u1 = 1.0f - 1.0f / float(P.t_mu);
    // scattering.h:23-46 (d_max == d_min for the ground: x = 0)
    c.g_dmin = r - b;
    c.g_den = rho - c.g_dmin;
    c.g_inv = 1.0f / c.g_den;
    c.s_dmin = t - r;
    c.s_den = rho + H - c.s_dmin;
    c.s_inv = 1.0f / c.s_den;
    c.h2 = H * H;
    const int hm = P.s_mu / 2;
    c.m_u0 = 0.5f / float(hm);
    c.m_u1 = 1.0f - 1.0f / float(hm);
    // render_sky.h:9-19
    const float k = P.rayleigh_scattering[0] / P.mie_scattering[0];
    for (int i = 0; i < 3; ++i) c.mie[i] = k * (P.mie_scattering[i] / P.rayleigh_scattering[i]);
    // util.h:31-34
    const float g = P.g;
    c.pm_k = 3.0f / (8.0f * kPi) * (1.0f - g * g) / (2.0f + g * g);
    c.pm_g1 = 1.0f + g * g;
    c.pm_g2 = 2.0f * g;
    return c;
}

// sky_radiance for a camera inside the atmosphere (so no move to the top
// boundary), from the block's slice at the camera's r and mu_s.  Every
// lookup coordinate is the reference's expression in the same binary32
// operations, rounded the same (IEEE square roots; quotients by a per-frame
// denominator through div_rn, bit for bit the IEEE quotient): the ground
// test switches between the two halves of the scattering texture, a ray
// distance is a difference of two ~r-sized terms (down from 0.5 km: d = 0.5
// from terms near 6360), and near the horizon the LUTs change steeply
// between texels, so an ulp in a coordinate can move a value by 1e-3.
// After the lookups the arithmetic is well conditioned: extrapolated Mie's
// c.w / c.x and phase_mie's x^-1.5 use the hardware's reciprocal and rsq^3,
// a few ulp from the IEEE forms.
__device__ void sky_radiance_slice(const DevParams& P, const SliceConsts& C,
                                   const float4* __restrict__ lds, float3 camera, float3 view,
                                   float3 sun, float3& rad, float3& trans) {
    const int NU = P.s_nu, MU = P.s_mu;
    const float r = C.r;
    const float mu = div_rn(dot3(camera, view), r, C.r_inv);
    const float nu = dot3(view, sun);
    const bool ground = ray_intersects_ground<float>(P, r, mu);
    if (ground) {
        trans = make_float3(0.0f, 0.0f, 0.0f);
    } else {
        // distance_to_top (params.h:105-110), then transmittance.h:7-24's u
        const float disc = C.rr * (mu * mu - 1.0f) + C.t2;
        const float d = fmaxf(-r * mu + sqrtf(fmaxf(disc, 0.0f)), 0.0f);
        const float u = C.t_u0 + div_rn(d - C.t_dmin, C.t_den, C.t_inv) * C.t_u1;
        const Taps tu = linear_taps<float>(u, P.t_mu);
        const float4* trow = lds + (NU + 1) * MU;
        const float4 t = lerp4(trow[tu.i0], trow[tu.i1], tu.w);
        trans = make_float3(t.x, t.y, t.z);
    }
    // scattering.h:23-46
    const float r_mu = r * mu;
    const float disc = r_mu * r_mu - C.rr + C.b2;
    float u_mu;
    if (ground) {
        const float d = -r_mu - sqrtf(fmaxf(disc, 0.0f));
        const float x = C.g_den == 0.0f ? 0.0f : div_rn(d - C.g_dmin, C.g_den, C.g_inv);
        u_mu = 0.5f - 0.5f * (C.m_u0 + x * C.m_u1);
    } else {
        const float d = -r_mu + sqrtf(fmaxf(disc + C.h2, 0.0f));
        u_mu = 0.5f + 0.5f * (C.m_u0 + div_rn(d - C.s_dmin, C.s_den, C.s_inv) * C.m_u1);
    }
    const Taps ty = linear_taps<float>(u_mu, MU);
    // nu of unit vectors can round just below -1 (anti-solar rays): clamped
    // so the slice index stays >= 0.  The per-ray lookup's CLAMP_TO_EDGE
    // then gives the same value up to a weight of the rounding's size.
    const float tex_coord_x = (fminf(fmaxf(nu, -1.0f), 1.0f) + 1.0f) / 2.0f * float(NU - 1);
    const float tex_x = floorf(tex_coord_x);
    const float l = tex_coord_x - tex_x;
    const int k0 = int(tex_x);
    const float4 a0 = lerp4(lds[k0 * MU + ty.i0], lds[k0 * MU + ty.i1], ty.w);
    const float4 a1 = lerp4(lds[(k0 + 1) * MU + ty.i0], lds[(k0 + 1) * MU + ty.i1], ty.w);
    const float4 c = make_float4(a0.x * (1.0f - l) + a1.x * l, a0.y * (1.0f - l) + a1.y * l,
                                 a0.z * (1.0f - l) + a1.z * l, a0.w * (1.0f - l) + a1.w * l);
    // render_sky.h:9-19
    float3 mie = make_float3(0.0f, 0.0f, 0.0f);
    if (c.x > 0.0f) {
        const float q = c.w * __builtin_amdgcn_rcpf(c.x);
        mie = make_float3(c.x * q * C.mie[0], c.y * q * C.mie[1], c.z * q * C.mie[2]);
    }
    const float pr = phase_rayleigh(nu);
    const float y = __builtin_amdgcn_rsqf(C.pm_g1 - C.pm_g2 * nu);
    const float pm = C.pm_k * (1.0f + nu * nu) * (y * y * y);
    rad = make_float3(c.x * pr + mie.x * pm, c.y * pr + mie.y * pm, c.z * pr + mie.z * pm);
}

struct DrawArgs {
    float ivp[4][4];  // [column][row]
    float cam[3];
    float sun[3];
};

// render_sky.frag:24-35 for one pixel; fullscreen.vert:5-8 interpolates
// screen_coords to ((px+0.5)/w, (py+0.5)/h) at pixel centres.
template <bool H>
__global__ __launch_bounds__(256) void k_render_sky(DevParams P, const float4* __restrict__ T,
                                                    const float4* __restrict__ S, DrawArgs D,
                                                    const float* __restrict__ depth, int w, int h,
                                                    float4* __restrict__ color,
                                                    float4* __restrict__ trans_out,
                                                    int composite) {
    int px = blockIdx.x * 16 + (threadIdx.x % 16);
    int py = blockIdx.y * 16 + (threadIdx.x / 16);
    if (px >= w || py >= h) return;
    size_t idx = size_t(py) * w + px;
    float sx = (float(px) + 0.5f) / float(w), sy = (float(py) + 0.5f) / float(h);
    float n0 = 2.0f * sx - 1.0f, n1 = 2.0f * sy - 1.0f;
    float v[4];
    for (int row = 0; row < 4; ++row)
        v[row] = D.ivp[0][row] * n0 + D.ivp[1][row] * n1 + D.ivp[2][row] * 0.0f + D.ivp[3][row];
    float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    float3 view = make_float3(v[0] / len, v[1] / len, v[2] / len);
    float dz = depth ? depth[idx] : 0.0f;
    float wp[4];
    for (int row = 0; row < 4; ++row)
        wp[row] = D.ivp[0][row] * n0 + D.ivp[1][row] * n1 + D.ivp[2][row] * dz + D.ivp[3][row];
    float3 world = make_float3(wp[0] / wp[3] * 1e-3f, wp[1] / wp[3] * 1e-3f, wp[2] / wp[3] * 1e-3f);
    float3 rad, tr;
    sky_radiance_to_point<H>(P, T, S, make_float3(D.cam[0], D.cam[1], D.cam[2]), view, world,
                          make_float3(D.sun[0], D.sun[1], D.sun[2]), rad, tr);
    if (composite) {
        // Dual-source blend dst = src0 + dst * src1 (src/render.rs:124-136);
        // alpha: dst.a kept (src factor ZERO, dst factor ONE).
        float4 dst = color[idx];
        color[idx] = make_float4(rad.x + dst.x * tr.x, rad.y + dst.y * tr.y, rad.z + dst.z * tr.z,
                                 dst.w);
    } else {
        color[idx] = make_float4(rad.x, rad.y, rad.z, 0.0f);
    }
    if (trans_out) trans_out[idx] = make_float4(tr.x, tr.y, tr.z, 1.0f);
}

template <bool H>
__global__ __launch_bounds__(256) void k_render_rays(DevParams P, const float4* __restrict__ T,
                                                     const float4* __restrict__ S,
                                                     const float* __restrict__ cams, int cstride,
                                                     const float* __restrict__ views,
                                                     const float* __restrict__ points,
                                                     const float* __restrict__ suns, int sstride,
                                                     const float* __restrict__ shadow,
                                                     size_t n, float4* __restrict__ rad_out,
                                                     float4* __restrict__ trans_out) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float l = shadow ? shadow[i] : 0.0f;
    size_t ci = cstride ? i : 0, si = sstride ? i : 0;
    float3 cam = make_float3(cams[3 * ci], cams[3 * ci + 1], cams[3 * ci + 2]);
    float3 sun = make_float3(suns[3 * si], suns[3 * si + 1], suns[3 * si + 2]);
    float3 view = make_float3(views[3 * i], views[3 * i + 1], views[3 * i + 2]);
    float3 rad, tr;
    if (points) {
        float3 pt = make_float3(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
        sky_radiance_to_point<H, true>(P, T, S, cam, view, pt, sun, rad, tr, l);
    } else {
        sky_radiance<H, true>(P, T, S, cam, view, sun, rad, tr, l);
    }
    rad_out[i] = make_float4(rad.x, rad.y, rad.z, 0.0f);
    if (trans_out) trans_out[i] = make_float4(tr.x, tr.y, tr.z, 1.0f);
}

// Counter-based direction generator: a 32-bit integer hash (the
// xorshift-multiply "lowbias32" of C. Wellons' hash search, 2018) of (seed,
// frame, pixel, attempt), and Marsaglia's (1972) rejection method for a
// uniform point on the sphere.  32-bit integer multiplies only (a 64-bit
// hash costs several 32-bit multiplies per product), and only +, *, sqrt
// (IEEE) in floating point, so the host copy in fb_ray_direction()
// produces identical bits.
__host__ __device__ inline uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__host__ __device__ inline void ray_direction(uint64_t seed, uint32_t frame, uint32_t pixel,
                                              float out[3]) {
    // Per frame (block-uniform), then per pixel.
    const uint32_t fk = hash32(frame ^ hash32(uint32_t(seed) ^ hash32(uint32_t(seed >> 32))));
    const uint32_t key = hash32((pixel * 0x9e3779b1u) ^ fk);
    for (uint32_t attempt = 0; attempt < 64; ++attempt) {
        const uint32_t h1 = hash32(key + 2u * attempt), h2 = hash32(key + 2u * attempt + 1u);
        float v1 = float(h1 >> 8) * (2.0f / 16777216.0f) - 1.0f;
        float v2 = float(h2 >> 8) * (2.0f / 16777216.0f) - 1.0f;
        float s = v1 * v1 + v2 * v2;
        if (s < 1.0f && s > 0.0f) {
            float q = 2.0f * sqrtf(1.0f - s);
            out[0] = v1 * q;
            out[1] = v2 * q;
            out[2] = 1.0f - 2.0f * s;
            return;
        }
    }
    out[0] = 0.0f;
    out[1] = 0.0f;
    out[2] = 1.0f;
}

// Pixels per thread of the frame kernels (a block shares one slice).
constexpr int kFramePixelsPerThread = 16;
// Blocks of 256 threads that cover a frame's pixels.
__host__ __device__ inline unsigned frame_blocks(uint32_t pixels) {
    const uint32_t per = 256u * kFramePixelsPerThread;
    return (pixels + per - 1) / per;
}
// Pixel blocks per block of the per-ray frame kernel.
constexpr unsigned kPerRayStride = 8;

// The camera of a synthetic frame at altitude alt: on the z axis, at
// radius r.  A frame takes the slice path when the LUTs' slice fits in LDS
// (`slice`) and the camera is inside the atmosphere; the host's choice of
// kernel and the kernels' own checks use this one test.
__host__ __device__ inline float frame_camera_z(const DevParams& P, float alt) {
    return P.bottom + alt;
}
__host__ __device__ inline bool frame_uses_slice(const DevParams& P, float alt, bool slice) {
    const float z = frame_camera_z(P, alt);
    return slice && sqrtf(0.0f * 0.0f + 0.0f * 0.0f + z * z) <= P.top;
}

// The frame's camera position, sun direction and camera radius.
struct FrameView {
    float3 cam, sun;
    float r;
};
__device__ inline FrameView frame_view(const DevParams& P, float alt, float zs) {
    FrameView v;
    v.cam = make_float3(0.0f, 0.0f, frame_camera_z(P, alt));
    v.sun = make_float3(sinf(zs), 0.0f, cosf(zs));
    v.r = sqrtf(dot3(v.cam, v.cam));
    return v;
}
__host__ __device__ inline size_t sky_slice_float4s(const DevParams& P) {
    return sky_slice_bytes(P) / sizeof(float4);
}

// The LUT slice of every frame on the slice path, once per frame, into
// `out` (frame f at out + f * sky_slice_float4s): blockIdx.y = frame, the
// gridDim.x blocks of a frame share its entries.  A frame kernel's blocks
// then copy their frame's slice into LDS with coalesced loads instead of
// each gathering it from the scattering LUT (two texel rows per entry).
template <bool H>
__global__ __launch_bounds__(256) void k_frame_slices(DevParams P, const float4* __restrict__ T,
                                                      const float4* __restrict__ S,
                                                      const float* __restrict__ altitudes,
                                                      const float* __restrict__ zeniths,
                                                      float4* __restrict__ out) {
    const uint32_t f = blockIdx.y;
    const float alt = altitudes[f];
    if (!frame_uses_slice(P, alt, true)) return;
    const FrameView v = frame_view(P, alt, zeniths[f]);
    build_sky_slice<H>(P, T, S, v.r, dot3(v.cam, v.sun) / v.r, out + f * sky_slice_float4s(P),
                       int(blockIdx.x * blockDim.x + threadIdx.x),
                       int(gridDim.x * blockDim.x));
}
constexpr unsigned kFrameSliceBlocks = 8;  // blocks per frame of k_frame_slices

// Pixels [pix0, pix0 + 256*kFramePixelsPerThread) of one synthetic frame
// (camera at altitude alt on the z axis, sun at zenith angle zs in the x-z
// plane).  `slice`: the block's LDS holds room for the frame's LUT slice.
// SLICE: the kernel for frames on the slice path, else the one for the
// others (per-ray lookups); a block of the other kind returns at once.  Two
// kernels rather than one branch: the per-ray lookup's registers would
// halve the slice kernel's occupancy.  `frame_slice`: the frame's slice
// from k_frame_slices, or null to build it in the block.
template <bool H, bool SLICE>
__device__ void render_frame_block(const DevParams& P, const float4* __restrict__ T,
                                   const float4* __restrict__ S, float alt, float zs,
                                   uint32_t frame, uint32_t pixels, uint64_t seed, uint32_t pix0,
                                   bool slice, const float4* __restrict__ frame_slice,
                                   float4* __restrict__ rad_out,
                                   float4* __restrict__ trans_out) {
    extern __shared__ float4 s_slice[];
    if (frame_uses_slice(P, alt, slice) != SLICE) return;  // block-uniform
    const FrameView v = frame_view(P, alt, zs);
    const float3 cam = v.cam, sun = v.sun;
    const float r = v.r;
    SliceConsts sc;
    if constexpr (SLICE) {
        if (frame_slice) {
            const int n = int(sky_slice_float4s(P));
            for (int i = threadIdx.x; i < n; i += blockDim.x) s_slice[i] = frame_slice[i];
        } else {
            build_sky_slice<H>(P, T, S, r, dot3(cam, sun) / r, s_slice, threadIdx.x, blockDim.x);
        }
        sc = slice_consts(P, r);
        __syncthreads();
    }
    for (int j = 0; j < kFramePixelsPerThread; ++j) {
        const uint32_t p = pix0 + uint32_t(j) * blockDim.x + threadIdx.x;
        if (p >= pixels) break;
        float d[3];
        ray_direction(seed, frame, p, d);
        const float3 view = make_float3(d[0], d[1], d[2]);
        float3 rad, tr;
        if constexpr (SLICE)
            sky_radiance_slice(P, sc, s_slice, cam, view, sun, rad, tr);
        else
            sky_radiance<H>(P, T, S, cam, view, sun, rad, tr);
        const size_t o = size_t(frame) * pixels + p;  // frame-major output
        // Written once and not read back by the kernel: non-temporal stores.
        typedef float v4f __attribute__((ext_vector_type(4)));
        __builtin_nontemporal_store(v4f{rad.x, rad.y, rad.z, 0.0f},
                                    reinterpret_cast<v4f*>(rad_out + o));
        if (trans_out)
            __builtin_nontemporal_store(v4f{tr.x, tr.y, tr.z, 1.0f},
                                        reinterpret_cast<v4f*>(trans_out + o));
    }
}

template <bool H, bool SLICE>
__global__ __launch_bounds__(256) void k_render_frames(DevParams P, const float4* __restrict__ T,
                                                       const float4* __restrict__ S,
                                                       const float* __restrict__ altitudes,
                                                       const float* __restrict__ zeniths,
                                                       uint32_t pixels, uint64_t seed,
                                                       int slice,
                                                       const float4* __restrict__ slices,
                                                       float4* __restrict__ rad_out,
                                                       float4* __restrict__ trans_out) {
    const uint32_t f = blockIdx.y;
    // The per-ray kernel walks its frame's pixel blocks with a stride (its
    // grid is kPerRayStride times smaller), so that a launch whose frames
    // all take the slice path costs a few hundred blocks, not one per
    // pixel block.
    const uint32_t nb = SLICE ? gridDim.x : frame_blocks(pixels);
    for (uint32_t bx = blockIdx.x; bx < nb; bx += gridDim.x) {  // no LDS on the per-ray path
        render_frame_block<H, SLICE>(P, T, S, altitudes[f], zeniths[f], f, pixels, seed,
                                     bx * blockDim.x * kFramePixelsPerThread, slice != 0,
                                     slices ? slices + f * sky_slice_float4s(P) : nullptr,
                                     rad_out, trans_out);
    }
}

// render_lighting.h:10-28
__global__ __launch_bounds__(256) void k_sun_sky_irradiance(
    DevParams P, const float4* __restrict__ T, const float4* __restrict__ E,
    const float* __restrict__ points, const float* __restrict__ normals,
    const float* __restrict__ suns, int sstride, size_t n, float4* __restrict__ direct_out,
    float4* __restrict__ sky_out) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    size_t si = sstride ? i : 0;
    float3 pt = make_float3(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
    float3 nr = make_float3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
    float3 sun = make_float3(suns[3 * si], suns[3 * si + 1], suns[3 * si + 2]);
    float r = sqrtf(dot3(pt, pt));
    float mu_s = dot3(pt, sun) / r;
    float3 e = get_irradiance<float>(P, E, r, mu_s);
    float f = (1.0f + dot3(nr, pt) / r) * 0.5f;
    sky_out[i] = make_float4(e.x * f, e.y * f, e.z * f, 0.0f);
    float3 ts = transmittance_to_sun<float>(P, T, r, mu_s);
    float ns = fmaxf(dot3(nr, sun), 0.0f);
    direct_out[i] = make_float4(P.solar[0] * ts.x * ns, P.solar[1] * ts.y * ns,
                                P.solar[2] * ts.z * ns, 0.0f);
}

// One synthetic frame with scalar camera altitude / sun zenith (the serving
// loop's render request; same rays as frame 0 of k_render_frames).
template <bool H, bool SLICE>
__global__ __launch_bounds__(256) void k_render_frame1(DevParams P, const float4* __restrict__ T,
                                                       const float4* __restrict__ S, float alt,
                                                       float zs, uint32_t pixels, uint64_t seed,
                                                       int slice, float4* __restrict__ rad_out) {
    render_frame_block<H, SLICE>(P, T, S, alt, zs, 0u, pixels, seed,
                                 blockIdx.x * blockDim.x * kFramePixelsPerThread, slice != 0,
                                 nullptr, rad_out, nullptr);
}

// Dynamic LDS of a frame launch.
static size_t frame_lds(const DevParams& P) {
    const size_t b = sky_slice_bytes(P);
    return b <= kSkySliceMaxBytes ? b : 0;
}

hipError_t launch_render_frame(const DevParams& P, const float4* T, const float4* S, float alt,
                               float zs, uint32_t pixels, uint64_t seed, float4* rad,
                               hipStream_t s) {
    if (pixels == 0) return hipSuccess;
    const size_t lds = frame_lds(P);
    const int slice = lds ? 1 : 0;
    const unsigned g = frame_blocks(pixels);
#define FB_FRAME1(H, SL) \
    k_render_frame1<H, SL><<<g, 256, SL ? lds : 0, s>>>(P, T, S, alt, zs, pixels, seed, slice, rad)
    const bool h = P.flags & kFlagF16Storage;
    if (frame_uses_slice(P, alt, slice != 0))
        h ? FB_FRAME1(true, true) : FB_FRAME1(false, true);
    else
        h ? FB_FRAME1(true, false) : FB_FRAME1(false, false);
#undef FB_FRAME1
    return hipGetLastError();
}

}  // namespace fbk

using namespace fbrt;

static fb_status check_atm(fb_builder* b, const fb_atmosphere* a) {
    if (!b || !a) {
        set_error("NULL builder or atmosphere");
        return FB_EINVAL;
    }
    if (a->builder != b) {
        set_error("atmosphere belongs to another builder");
        return FB_EINVAL;
    }
    return FB_OK;
}

// Order the render after the precompute, and record it among the
// atmosphere's uses (destroy waits for every recorded use, on any stream).
static fb_status begin_use(fb_atmosphere* a, hipStream_t s) {
    FB_HIP_TRY(hipSetDevice(a->builder->device));
    FB_HIP_TRY(hipStreamWaitEvent(s, a->ready, 0));
    return FB_OK;
}
// The use is recorded even when a launch failed: work queued before the
// failing launch may still read the LUTs.
static fb_status end_use(fb_atmosphere* a, hipStream_t s, hipError_t e = hipSuccess) {
    if (e == hipSuccess) e = hipGetLastError();
    const fb_status st = atmosphere_mark_use(a, s);
    if (e != hipSuccess) return hip_fail(e, "render launch");
    return st;
}

extern "C" {

fb_status fb_render_sky(fb_builder* b, fb_stream stream, const fb_atmosphere* ac,
                        const fb_draw_parameters* draw, const float* depth, uint32_t width,
                        uint32_t height, float* color, float* transmittance, uint32_t flags) {
    fb_status st = check_atm(b, ac);
    if (st != FB_OK) return st;
    if (!draw || !color || width == 0 || height == 0) {
        set_error("invalid render arguments");
        return FB_EINVAL;
    }
    fb_atmosphere* a = const_cast<fb_atmosphere*>(ac);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((st = begin_use(a, s)) != FB_OK) return st;
    fbk::DrawArgs D;
    memcpy(D.ivp, draw->inverse_viewproj, sizeof(D.ivp));
    memcpy(D.cam, draw->camera_position, sizeof(D.cam));
    memcpy(D.sun, draw->sun_direction, sizeof(D.sun));
    dim3 grid((width + 15) / 16, (height + 15) / 16);
    auto* c4 = reinterpret_cast<float4*>(color);
    auto* t4 = reinterpret_cast<float4*>(transmittance);
    const int comp = int(flags & FB_RENDER_COMPOSITE);
    if (a->dev.flags & fbk::kFlagF16Storage)
        fbk::k_render_sky<true><<<grid, 256, 0, s>>>(a->dev, a->T.p, a->S.p, D, depth, int(width),
                                                      int(height), c4, t4, comp);
    else
        fbk::k_render_sky<false><<<grid, 256, 0, s>>>(a->dev, a->T.p, a->S.p, D, depth,
                                                       int(width), int(height), c4, t4, comp);
    return end_use(a, s);
}

fb_status fb_render_sky_rays_shadowed(fb_builder* b, fb_stream stream, const fb_atmosphere* ac,
                                      const float* cameras, uint32_t camera_stride,
                                      const float* views, const float* points,
                                      const float* suns, uint32_t sun_stride,
                                      const float* shadow_lengths, size_t n, float* radiance,
                                      float* transmittance) {
    fb_status st = check_atm(b, ac);
    if (st != FB_OK) return st;
    if (!cameras || !views || !suns || !radiance || camera_stride > 1 || sun_stride > 1) {
        set_error("invalid ray arguments");
        return FB_EINVAL;
    }
    if (n == 0) return FB_OK;
    fb_atmosphere* a = const_cast<fb_atmosphere*>(ac);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((st = begin_use(a, s)) != FB_OK) return st;
    const unsigned blocks = unsigned((n + 255) / 256);
    auto* r4 = reinterpret_cast<float4*>(radiance);
    auto* t4 = reinterpret_cast<float4*>(transmittance);
    if (a->dev.flags & fbk::kFlagF16Storage)
        fbk::k_render_rays<true><<<blocks, 256, 0, s>>>(a->dev, a->T.p, a->S.p, cameras,
                                                         int(camera_stride), views, points, suns,
                                                         int(sun_stride), shadow_lengths, n, r4,
                                                         t4);
    else
        fbk::k_render_rays<false><<<blocks, 256, 0, s>>>(a->dev, a->T.p, a->S.p, cameras,
                                                          int(camera_stride), views, points, suns,
                                                          int(sun_stride), shadow_lengths, n, r4,
                                                          t4);
    return end_use(a, s);
}

fb_status fb_render_sky_rays(fb_builder* b, fb_stream stream, const fb_atmosphere* ac,
                             const float* cameras, uint32_t camera_stride, const float* views,
                             const float* points, const float* suns, uint32_t sun_stride,
                             size_t n, float* radiance, float* transmittance) {
    return fb_render_sky_rays_shadowed(b, stream, ac, cameras, camera_stride, views, points,
                                       suns, sun_stride, nullptr, n, radiance, transmittance);
}

fb_status fb_render_sky_frames(fb_builder* b, fb_stream stream, const fb_atmosphere* ac,
                               const float* altitudes_km, const float* sun_zenith_rad,
                               uint32_t n_frames, uint32_t pixels, uint64_t seed,
                               float* radiance, float* transmittance) {
    fb_status st = check_atm(b, ac);
    if (st != FB_OK) return st;
    if (!altitudes_km || !sun_zenith_rad || !radiance || n_frames > 65535) {
        set_error("invalid frame arguments");
        return FB_EINVAL;
    }
    if (n_frames == 0 || pixels == 0) return FB_OK;
    fb_atmosphere* a = const_cast<fb_atmosphere*>(ac);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((st = begin_use(a, s)) != FB_OK) return st;
    dim3 grid(fbk::frame_blocks(pixels), n_frames);
    const size_t lds = fbk::frame_lds(a->dev);
    auto* r4 = reinterpret_cast<float4*>(radiance);
    auto* t4 = reinterpret_cast<float4*>(transmittance);
    const int slice = lds ? 1 : 0;
    const bool h = a->dev.flags & fbk::kFlagF16Storage;
    // The altitudes are on the device: k_frame_slices and the slice kernel
    // take the frames on the slice path, the per-ray kernel the others (when
    // no slice fits, all of them); each returns at once from the other's
    // frames.  The frames' slices live in stream-ordered scratch
    // (n_frames * at most 64 KB), freed after the launches that read them.
    float4* slices = nullptr;
    if (slice) {
        hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&slices), size_t(n_frames) * lds, s);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("frame slice scratch: %s", hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? FB_ENOMEM : FB_EHIP;
        }
        dim3 sg(fbk::kFrameSliceBlocks, n_frames);
        if (h)
            fbk::k_frame_slices<true><<<sg, 256, 0, s>>>(a->dev, a->T.p, a->S.p, altitudes_km,
                                                          sun_zenith_rad, slices);
        else
            fbk::k_frame_slices<false><<<sg, 256, 0, s>>>(a->dev, a->T.p, a->S.p, altitudes_km,
                                                           sun_zenith_rad, slices);
    }
#define FB_FRAMES(H, SL)                                                                       \
    fbk::k_render_frames<H, SL><<<grid, 256, SL ? lds : 0, s>>>(a->dev, a->T.p, a->S.p,        \
                                                               altitudes_km, sun_zenith_rad, \
                                                               pixels, seed, slice, slices, r4, t4)
    if (slice) h ? FB_FRAMES(true, true) : FB_FRAMES(false, true);
    grid.x = (grid.x + fbk::kPerRayStride - 1) / fbk::kPerRayStride;
    h ? FB_FRAMES(true, false) : FB_FRAMES(false, false);
#undef FB_FRAMES
    hipError_t e = hipGetLastError();
    if (slices) {
        const hipError_t fe = hipFreeAsync(slices, s);
        if (e == hipSuccess) e = fe;
    }
    return end_use(a, s, e);
}

void fb_ray_direction(uint64_t seed, uint32_t frame, uint32_t pixel, float out[3]) {
    fbk::ray_direction(seed, frame, pixel, out);
}

fb_status fb_sun_and_sky_irradiance(fb_builder* b, fb_stream stream, const fb_atmosphere* ac,
                                    const float* points, const float* normals, const float* suns,
                                    uint32_t sun_stride, size_t n, float* sun_irradiance,
                                    float* sky_irradiance) {
    fb_status st = check_atm(b, ac);
    if (st != FB_OK) return st;
    if (!points || !normals || !suns || !sun_irradiance || !sky_irradiance || sun_stride > 1) {
        set_error("invalid lighting arguments");
        return FB_EINVAL;
    }
    if (n == 0) return FB_OK;
    fb_atmosphere* a = const_cast<fb_atmosphere*>(ac);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((st = begin_use(a, s)) != FB_OK) return st;
    fbk::k_sun_sky_irradiance<<<unsigned((n + 255) / 256), 256, 0, s>>>(
        a->dev, a->T.p, a->E.p, points, normals, suns, int(sun_stride), n,
        reinterpret_cast<float4*>(sun_irradiance), reinterpret_cast<float4*>(sky_irradiance));
    return end_use(a, s);
}

}  // extern "C"
