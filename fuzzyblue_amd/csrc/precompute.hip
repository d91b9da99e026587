// precompute.hip — the six precompute passes of the reference as CDNA4
// kernels (shaders/transmittance.comp, direct_irradiance.comp,
// single_scattering.comp, scattering_density.comp, indirect_irradiance.comp,
// multiple_scattering.comp), plus the irradiance clear.
//
// Every LUT is a row-major float4 array in HBM (x fastest).  Launch
// wrappers at the bottom take a DevParams by value and raw device pointers;
// the host runtime (runtime.hip) owns allocation and ordering.
#include <math.h>
#include <type_traits>

#include "atmo_device.hpp"
#include "kernels.hpp"

// The constant address space (read-only for the whole kernel): loads through
// it are scalar when uniform, and addresses loaded from it are known to
// point to global memory.  The host pass only parses device code, without
// address spaces.
#if defined(__HIP_DEVICE_COMPILE__)
#define FB_CONSTANT_AS __attribute__((address_space(4)))
#else
#define FB_CONSTANT_AS
#endif

namespace fbk {

// ------------------------------------------------------------------ tables
// Direction tables of the two spherical quadratures, computed once on the
// host in binary32 exactly as the reference evaluates them
// (scattering_density.comp:34-66, indirect_irradiance.comp:20-33), and in
// binary64 for the precision mode's geometry.
//   density:  l = 0..15 (theta), m = 0..31 (phi), dtheta = dphi = pi/16
//   indirect: j = 0..15 (theta), i = 0..63 (phi), dtheta = dphi = pi/32
__constant__ float4 c_den_dir[kDenTheta * kDenPhi];  // (cos phi sin t, sin phi sin t, cos t, domega)
__constant__ float c_den_cos[kDenTheta];
__constant__ double c_den_cos_d[kDenTheta];
__constant__ float4 c_ind_dir[kIndTheta * kIndPhi];
__constant__ float c_ind_cos[kIndTheta];
__constant__ double c_ind_cos_d[kIndTheta];

hipError_t upload_tables() {
    static float4 den[kDenTheta * kDenPhi], ind[kIndTheta * kIndPhi];
    static float dcos[kDenTheta], icos[kIndTheta];
    static double dcos_d[kDenTheta], icos_d[kIndTheta];
    const float PI = 3.14159265358979323846f;
    {
        const float dphi = PI / float(16), dtheta = PI / float(16);
        const double dphi_d = M_PI / 16.0, dtheta_d = M_PI / 16.0;
        for (int l = 0; l < kDenTheta; ++l) {
            float theta = (float(l) + 0.5f) * dtheta;
            float ct = cosf(theta), st = sinf(theta);
            dcos[l] = ct;
            dcos_d[l] = cos((double(l) + 0.5) * dtheta_d);
            for (int m = 0; m < kDenPhi; ++m) {
                float phi = (float(m) + 0.5f) * dphi;
                float domega = dtheta * dphi * sinf(theta);
                den[l * kDenPhi + m] = make_float4(cosf(phi) * st, sinf(phi) * st, ct, domega);
            }
            (void)dphi_d;
        }
    }
    {
        const float dphi = PI / float(32), dtheta = PI / float(32);
        const double dtheta_d = M_PI / 32.0;
        for (int j = 0; j < kIndTheta; ++j) {
            float theta = (float(j) + 0.5f) * dtheta;
            icos[j] = cosf(theta);
            icos_d[j] = cos((double(j) + 0.5) * dtheta_d);
            for (int i = 0; i < kIndPhi; ++i) {
                float phi = (float(i) + 0.5f) * dphi;
                float domega = dtheta * dphi * sinf(theta);
                ind[j * kIndPhi + i] =
                    make_float4(cosf(phi) * sinf(theta), sinf(phi) * sinf(theta), cosf(theta), domega);
            }
        }
    }
    hipError_t e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_den_dir), den, sizeof(den))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_den_cos), dcos, sizeof(dcos))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_den_cos_d), dcos_d, sizeof(dcos_d))) != hipSuccess)
        return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_ind_dir), ind, sizeof(ind))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_ind_cos), icos, sizeof(icos))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_ind_cos_d), icos_d, sizeof(icos_d))) != hipSuccess)
        return e;
    return hipSuccess;
}

template <typename G>
__device__ __forceinline__ G den_cos(int l);
template <>
__device__ __forceinline__ float den_cos<float>(int l) { return c_den_cos[l]; }
template <>
__device__ __forceinline__ double den_cos<double>(int l) { return c_den_cos_d[l]; }
template <typename G>
__device__ __forceinline__ G ind_cos(int j);
template <>
__device__ __forceinline__ float ind_cos<float>(int j) { return c_ind_cos[j]; }
template <>
__device__ __forceinline__ double ind_cos<double>(int j) { return c_ind_cos_d[j]; }

// Running sum with optional Neumaier compensation (precision mode).
template <bool COMP>
struct Sum {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float x) {
        if (COMP) {
            float t = s + x;
            if (fabsf(s) >= fabsf(x))
                c += (s - t) + x;
            else
                c += (x - t) + s;
            s = t;
        } else {
            s = s + x;
        }
    }
    __device__ __forceinline__ float value() const { return COMP ? s + c : s; }
};

// ============================================================ 1. transmittance
// transmittance.comp:8-46,71-80.  LANES lanes per texel: lane s sums the
// trapezoid samples k = s, s+LANES, ... of all three optical lengths (the
// reference recomputes r_i per profile with identical inputs), then the
// partial sums are combined with lane shuffles: the precision mode, 16
// lanes with compensated partial sums.  The FP32 mode keeps the reference's
// sequential summation order, so its rounding matches the reference's
// (transmittance_block_ordered below).
// T = exp(-(beta_R L_R + beta_M,ext L_M + beta_abs L_O)) (transmittance.comp:
// 34-46) for texel (i, j); with OD, also the optical depth itself, for
// transmittances below binary32's range (see log_transmittance_to_top).
__device__ __forceinline__ void store_transmittance(const DevParams& P, int i, int j, float LR,
                                                    float LM, float LO, float4* __restrict__ T,
                                                    float4* __restrict__ OD) {
    const float t0 = P.rayleigh_scattering[0] * LR + P.mie_extinction[0] * LM +
                     P.absorption_extinction[0] * LO;
    const float t1 = P.rayleigh_scattering[1] * LR + P.mie_extinction[1] * LM +
                     P.absorption_extinction[1] * LO;
    const float t2 = P.rayleigh_scattering[2] * LR + P.mie_extinction[2] * LM +
                     P.absorption_extinction[2] * LO;
    T[j * P.t_mu + i] = make_float4(expf(-t0), expf(-t1), expf(-t2), 1.0f);
    if (OD) OD[j * P.t_mu + i] = make_float4(t0, t1, t2, 0.0f);
}

template <typename G, bool COMP, int LANES>
__device__ __forceinline__ void transmittance_block(const DevParams& P, const int bx, const int by, const int bz, float4* __restrict__ T, float4* __restrict__ OD) {
    const int sub = threadIdx.x % LANES;
    const int texel = (bx * blockDim.x + threadIdx.x) / LANES;
    const int n = P.t_mu * P.t_r;
    const bool active = texel < n;
    const int i = active ? texel % P.t_mu : 0, j = active ? texel / P.t_mu : 0;
    G x_mu = G(i) / G(P.t_mu - 1);
    G x_r = G(j) / G(P.t_r - 1);
    G r, mu;
    rmu_from_transmittance_unit<G>(P, x_mu, x_r, r, mu);
    const int SAMPLE_COUNT = 500;
    G dx = distance_to_top<G>(P, r, mu) / G(SAMPLE_COUNT);
    float dxf = float(dx);
    Sum<COMP> lr, lm, lo;
    // Two samples per iteration: their square-root and exp chains overlap
    // (35-36 -> 32 us for 256x64 texels; the sums still add in k order).
#pragma unroll 2
    for (int k = sub; k <= SAMPLE_COUNT; k += LANES) {
        G d_i = G(k) * dx;
        // (r_i - bottom exactly as the oracle forms it: the density profile
        // switches layer at altitude 0, so a sample at the horizon ray's
        // tangent point must round the same way.)
        G r_i = sqrt(d_i * d_i + G(2) * r * mu * d_i + r * r);
        float alt = float(r_i - G(P.bottom));
        float w = (k == 0 || k == SAMPLE_COUNT) ? 0.5f : 1.0f;
        lr.add(profile_density(P.prof[0], alt) * w * dxf);
        lm.add(profile_density(P.prof[1], alt) * w * dxf);
        lo.add(profile_density(P.prof[2], alt) * w * dxf);
    }
    float LR = lr.value(), LM = lm.value(), LO = lo.value();
    for (int off = 1; off < LANES; off <<= 1) {
        LR += __shfl_xor(LR, off, 64);
        LM += __shfl_xor(LM, off, 64);
        LO += __shfl_xor(LO, off, 64);
    }
    if (!active || sub != 0) return;
    store_transmittance(P, i, j, LR, LM, LO, T, OD);
}

// Lanes per texel of the FP32-mode transmittance (a wavefront; the kernels
// pick transmittance_block_ordered for this lane count).
constexpr int kOrderedLanes = 64;

// FP32 mode: the reference's summation order — per profile, s = s + term
// for k = 0..500 in sequence (transmittance.comp:16-31) — with the terms
// computed by a wavefront's 64 lanes in parallel (each term by the same
// binary32 operations as in the sequential loop) and summed in order, one
// lane per profile.  Bit-identical to one lane doing all of it.
template <typename G>
__device__ __forceinline__ void transmittance_block_ordered(const DevParams& P, const int bx,
                                                            float4* __restrict__ T) {
    constexpr int SAMPLE_COUNT = 500;
    // Per wavefront of a 256-thread block; rows padded to 504 floats so that
    // each starts 16-byte aligned (read as float4 below).
    __shared__ __align__(16) float s_terms[4][3][504];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int texel = bx * int(blockDim.x / 64) + wave;
    const int n = P.t_mu * P.t_r;
    const bool active = texel < n;
    const int i = active ? texel % P.t_mu : 0, j = active ? texel / P.t_mu : 0;
    G x_mu = G(i) / G(P.t_mu - 1);
    G x_r = G(j) / G(P.t_r - 1);
    G r, mu;
    rmu_from_transmittance_unit<G>(P, x_mu, x_r, r, mu);
    G dx = distance_to_top<G>(P, r, mu) / G(SAMPLE_COUNT);
    float dxf = float(dx);
    for (int k = lane; k <= SAMPLE_COUNT; k += 64) {
        G d_i = G(k) * dx;
        G r_i = sqrt(d_i * d_i + G(2) * r * mu * d_i + r * r);
        float alt = float(r_i - G(P.bottom));
        float w = (k == 0 || k == SAMPLE_COUNT) ? 0.5f : 1.0f;
        s_terms[wave][0][k] = profile_density(P.prof[0], alt) * w * dxf;
        s_terms[wave][1][k] = profile_density(P.prof[1], alt) * w * dxf;
        s_terms[wave][2][k] = profile_density(P.prof[2], alt) * w * dxf;
    }
    __syncthreads();
    float L = 0.0f;
    if (lane < 3) {
        // Terms 0..499 four at a time (still added one by one, in order),
        // then term 500.
        const float* tp = s_terms[wave][lane];
        const float4* tp4 = reinterpret_cast<const float4*>(tp);
#pragma unroll 5
        for (int q = 0; q < SAMPLE_COUNT / 4; ++q) {
            const float4 v = tp4[q];
            L = L + v.x;
            L = L + v.y;
            L = L + v.z;
            L = L + v.w;
        }
        L = L + tp[SAMPLE_COUNT];
    }
    const float LR = __shfl(L, 0, 64), LM = __shfl(L, 1, 64), LO = __shfl(L, 2, 64);
    if (!active || lane != 0) return;
    store_transmittance(P, i, j, LR, LM, LO, T, nullptr);
}

// ====================================================== 2. direct irradiance
// direct_irradiance.comp:10-46
template <typename G>
__device__ __forceinline__ void direct_irradiance_block(const DevParams& P, const int bx, const int by, const int bz,
                                                           const float4* __restrict__ T,
                                                           float4* __restrict__ dE) {
    int i = bx * blockDim.x + threadIdx.x;
    int j = by;
    if (i >= P.i_mus || j >= P.i_r) return;
    G x_mu_s = G(i) / G(P.i_mus - 1);
    G x_r = G(j) / G(P.i_r - 1);
    G r, mu_s;
    rmus_from_irradiance_unit<G>(P, x_mu_s, x_r, r, mu_s);
    G alpha_s = G(P.sun_angular_radius);
    G avg = mu_s < -alpha_s
                ? G(0)
                : (mu_s > alpha_s ? mu_s : (mu_s + alpha_s) * (mu_s + alpha_s) / (G(4) * alpha_s));
    float a = float(avg);
    float3 t = transmittance_to_top<G>(P, T, r, mu_s);
    dE[j * P.i_mus + i] =
        make_float4(P.solar[0] * t.x * a, P.solar[1] * t.y * a, P.solar[2] * t.z * a, 0.0f);
}


// ===================================================== 3. single scattering
// single_scattering.comp:10-101.
//
// As for multiple scattering, the texels of one (r, mu) row share the view
// ray: for each of the 51 samples, d_i, r_d, the view transmittance
// T(r, mu, d_i), the densities at r_d and the transmittance-LUT row taps
// and horizon terms of r_d are computed once per row (row_geometry_block,
// in the per-set geometry pass) and read from the row table.  Each texel
// then evaluates only its sun transmittance T(r_d, mu_s_d) (one bilinear
// lookup) and accumulates.

// Multiple scattering interpolates each sample's value from four density
// rows, (r0, mu0), (r0, mu1), (r1, mu0), (r1, mu1).  Consecutive samples of
// a ray often share rows (the same cell, or the next one in mu or r: at
// default sizes 79 distinct rows serve the 204 taps of an (r, mu) row), so
// the kernel keeps four row "slots" in registers from sample to sample.
// Row (r index z, mu index y) always lives in slot 2*(z & 1) + (y & 1):
// the rows of one sample differ by at most one step in z and in y, so they
// never collide (clamped duplicates are the same row), and a row shared by
// consecutive samples stays where it is.  Per sample, the table gives the
// slots to load (a used slot is loaded unless the previous sample used it
// for the same row), the row offsets to load, and each slot's weight: the
// summed bilinear weights of its taps, 0 for an unused slot.
struct MsSlots {
    int mask;   // bit s: load slot s
    int4 load;  // row offset loaded into slot s (0 when not loaded)
    float4 w;   // slot weights
};
// i: this lane's sample; z[k], y[k], off[k], wt[k]: its taps (A, B, C, D).
// Every lane of the wavefront calls this (lane i holds sample i; lanes past
// the last sample repeat it).
__device__ __forceinline__ MsSlots ms_row_slots(int i, const int z[4], const int y[4],
                                                const int off[4], const float wt[4]) {
    int row[4] = {-1, -1, -1, -1};  // row offset held by each slot, -1 if unused
    float w[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int s = ((z[k] & 1) << 1) | (y[k] & 1);
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (t == s) {
                row[t] = off[k];
                w[t] += wt[k];
            }
    }
    MsSlots r{0, make_int4(0, 0, 0, 0), make_float4(w[0], w[1], w[2], w[3])};
    int ld[4] = {0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int prev = __shfl_up(row[t], 1, 64);  // sample i - 1's slot t
        if (row[t] >= 0 && (i == 0 || prev != row[t])) {
            r.mask |= 1 << t;
            ld[t] = row[t];
        }
    }
    r.load = make_int4(ld[0], ld[1], ld[2], ld[3]);
    return r;
}

// The single- and multiple-scattering row tables of row (y, z): thread i
// computes sample i.  The multiple-scattering passes march the same 51
// samples (multiple_scattering.comp:16-47, the same dx and d_i).
template <typename G>
__device__ __forceinline__ void row_geometry_block(const DevParams& P, const int y, const int z,
                                                   const float4* __restrict__ T,
                                                   const float4* __restrict__ OD,
                                                   float4* __restrict__ SR,
                                                   float4* __restrict__ MR) {
    const int tid = threadIdx.x;
    const int MU = P.s_mu, W = P.s_nu * P.s_mus;
    const TexelCoord<G> rc = scattering_texel_coord<G>(P, 0u, unsigned(y), unsigned(z));
    const G r = rc.r, mu = rc.mu;
    const bool ground = rc.ground;
    const G dx = distance_to_nearest<G>(P, r, mu, ground) / G(kMsSamples - 1);
    // Lanes 51..63 evaluate sample 50 again and write nothing: every lane
    // takes part in the cross-lane reads of ms_row_slots.
    const int i = min(tid, kMsSamples - 1);
    const bool writer = tid < kMsSamples;
    const G d_i = G(i) * dx;
    const G r_d = clamp_radius<G>(P, sqrt(d_i * d_i + G(2) * r * mu * d_i + r * r));
    float3 tv;
    if constexpr (sizeof(G) == sizeof(double))
        tv = get_transmittance_precise(P, T, OD, r, mu, d_i, ground);
    else
        tv = get_transmittance<G>(P, T, r, mu, d_i, ground);
    const float w = (i == 0 || i == kMsSamples - 1) ? 0.5f : 1.0f;
    const float alt = float(r_d - G(P.bottom));
    const G sin_h = G(P.bottom) / r_d;
    const G cos_h = -sqrt(fmax(G(1) - sin_h * sin_h, G(0)));
    // transmittance.h:7-24 terms that depend on r_d only.
    const G H = sqrt(G(P.top) * G(P.top) - G(P.bottom) * G(P.bottom));
    const G rho = safe_sqrt<G>(r_d * r_d - G(P.bottom) * G(P.bottom));
    const G d_min = G(P.top) - r_d, d_max = rho + H;
    const Taps tr = linear_taps<G>(tex_from_unit<G>(rho / H, P.t_r), P.t_r);
    if (writer) {
        float4* o = SR + (size_t(z) * MU + y) * kSsRowFloat4s;
        o[i] = make_float4(tv.x * w, tv.y * w, tv.z * w, float(d_i));
        o[kMsSamples + i] = make_float4(float(r_d), profile_density(P.prof[0], alt),
                                        profile_density(P.prof[1], alt), float(sin_h));
        o[2 * kMsSamples + i] =
            make_float4(float(cos_h), float(d_min), float(d_max - d_min), tr.w);
        o[3 * kMsSamples + i] =
            make_float4(float(G(1) / r_d), float(G(1) / (d_max - d_min)),
                        float(fmax(G(P.top) * G(P.top) - r_d * r_d, G(0))),
                        float(G(1) / (G(2) * sin_h * G(P.sun_angular_radius))));
        o[4 * kMsSamples + i] = make_float4(__int_as_float(tr.i0 * P.t_mu),
                                            __int_as_float(tr.i1 * P.t_mu), float(dx), 0.0f);
    }
    // Multiple scattering: T(r, mu, d_i) x dx x weight, then the (mu, r)
    // taps of the sample point (r_i, mu_i).
    const G mu_i = clamp_cosine<G>((r * mu + d_i) / r_d);
    const Taps ty = linear_taps<G>(scattering_u_mu<G>(P, r_d, mu_i, ground), MU);
    const Taps tz = linear_taps<G>(scattering_u_r<G>(P, r_d), P.s_r);
    const float f = float(dx) * w;
    float4* m = MR + (size_t(z) * MU + y) * kMsRowFloat4s;
    // The four density rows of the sample, (r0, mu0), (r0, mu1), (r1, mu0),
    // (r1, mu1), as offsets into J, and their bilinear weights.
    const int tz4[4] = {tz.i0, tz.i0, tz.i1, tz.i1}, ty4[4] = {ty.i0, ty.i1, ty.i0, ty.i1};
    int off[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) off[k] = (tz4[k] * MU + ty4[k]) * W;
    const G wy = G(ty.w), wz = G(tz.w);
    const float wt[4] = {float((G(1) - wy) * (G(1) - wz)), float(wy * (G(1) - wz)),
                         float((G(1) - wy) * wz), float(wy * wz)};
    const MsSlots sl = ms_row_slots(i, tz4, ty4, off, wt);
    if (!writer) return;
    m[i] = make_float4(tv.x * f, tv.y * f, tv.z * f, float(d_i));
    m[kMsSamples + i] = make_float4(float(G(1) / r_d), __int_as_float(sl.mask), 0.0f, 0.0f);
    reinterpret_cast<int4*>(m + 2 * kMsSamples)[i] = sl.load;
    m[3 * kMsSamples + i] = sl.w;
}

// A raw buffer resource over `bytes` bytes at p (descriptor word 3 for
// untyped 32-bit loads on gfx9-family parts); loads past the end return 0.
// Offsets are 32-bit, so the tap addresses need no 64-bit arithmetic.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), short(0), bytes, 0x00020000);
}
__device__ __forceinline__ float3 buffer_rgb(__amdgpu_buffer_rsrc_t r, int byte_offset) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, byte_offset, 0, 0);
    return make_float3(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]));
}
__device__ __forceinline__ float3 lerp3(float3 a, float3 b, float w) {
    return make_float3(fmaf(w, b.x - a.x, a.x), fmaf(w, b.y - a.y, a.y), fmaf(w, b.z - a.z, a.z));
}

// The sun-direction coordinate of sample (tv, gq, lq, iv) of texel
// (mu_s, nu): mu_s_d, the texel coordinate t of T(r_d, mu_s_d) along the
// LUT's mu axis (clamped to [0, TMU-1]) and the smoothstep sun-disc factor
// (transmittance.h:26-33, 63-74).
template <bool PRECISE>
__device__ __forceinline__ void ss_sun_coord(float rf, float mu_s, float nu, float4 tv, float4 gq,
                                             float4 lq, float4 iv, float top2, int TMU,
                                             float& t, float& sm) {
    const float r_d = gq.x, d_i = tv.w;
    const float tmax = float(TMU - 1);
    float mu_s_d;
    if constexpr (PRECISE) {
        // Sun-ray distance to the top boundary without cancellation: for
        // mu_s_d > 0 the root is (top^2 - r_d^2) / (r_d mu_s_d + sqrt(disc)),
        // with top^2 - r_d^2 from the FP64 set-up.  Reciprocals of the
        // per-sample divisors come from FP64 too.
        mu_s_d = __builtin_amdgcn_fmed3f(fmaf(rf, mu_s, d_i * nu) * iv.x, -1.0f, 1.0f);
        const float rm = r_d * mu_s_d;
        const float sq = __builtin_amdgcn_sqrtf(fmaf(rm, rm, iv.z));
        const float d = mu_s_d > 0.0f ? iv.z * __builtin_amdgcn_rcpf(rm + sq) : sq - rm;
        // tex_from_unit then linear_taps: texel coordinate x_mu*(N-1).
        t = __builtin_amdgcn_fmed3f((d - lq.y) * iv.y * tmax, 0.0f, tmax);
    } else {
        mu_s_d = clamp_cosine<float>((rf * mu_s + d_i * nu) / r_d);
        float disc = r_d * r_d * (mu_s_d * mu_s_d - 1.0f) + top2;
        float d = clamp_distance<float>(-r_d * mu_s_d + safe_sqrt<float>(disc));
        float x_mu = (d - lq.y) / lq.z;
        // linear_taps of tex_from_unit(x_mu): u*N - 0.5 = x_mu*(N-1), clamped.
        const float u = tex_from_unit<float>(x_mu, TMU);
        t = fminf(fmaxf(u * float(TMU) - 0.5f, 0.0f), tmax);
    }
    // smoothstep(-sin_h alpha, sin_h alpha, mu_s_d - cos_h) with the
    // divisor's reciprocal from the table.
    const float u = __builtin_amdgcn_fmed3f(fmaf(mu_s_d - lq.x, iv.w, 0.5f), 0.0f, 1.0f);
    sm = u * u * (3.0f - 2.0f * u);
}


// STAGED: per sample, the block first interpolates the sample's two
// transmittance rows (its r taps) into one row in LDS (double buffered, one
// barrier per sample) and each texel interpolates that row in mu: all
// texels of the (r, mu) row look up the same two LUT rows, so this replaces
// four scattered loads per texel and sample by coalesced loads of the rows.
// Otherwise (transmittance LUTs wider than kSsMaxStagedMu) each texel reads
// its four taps itself.
template <typename G, bool COMP, bool H, bool STAGED>
__device__ __forceinline__ void single_scattering_block(const DevParams& P, const int bx, const int by, const int bz,
                                                           const float4* __restrict__ T,
                                                           const float4* __restrict__ SR,
                                                           float4* __restrict__ dR,
                                                           float4* __restrict__ dM,
                                                           float4* __restrict__ S) {
    // The row table (read-only here, every read block-uniform: scalar loads
    // through the constant address space).
    typedef FB_CONSTANT_AS const float4 CF4;
    CF4* row = (CF4*)(SR + (size_t(by) * P.s_mu + bx) * kSsRowFloat4s);
    CF4* s_tv = row;                    // T(r,mu,d_i) * weight (rgb), d_i
    CF4* s_geo = row + kMsSamples;      // r_d, rho_rayleigh, rho_mie, sin_theta_h
    CF4* s_lut = row + 2 * kMsSamples;  // cos_theta_h, d_min, d_max - d_min, T-row weight
    CF4* s_inv = row + 3 * kMsSamples;  // 1/r_d, 1/(d_max - d_min), top^2 - r_d^2, 1/(2 sin_theta_h alpha)
    CF4* s_trw = row + 4 * kMsSamples;  // T-row offsets (int bits), dx
    constexpr bool kPrecise = sizeof(G) == sizeof(double);
    const int NU = P.s_nu, MUS = P.s_mus, MU = P.s_mu;
    const int W = NU * MUS;
    const int y = bx, z = by;
    const int tid = threadIdx.x, nt = blockDim.x;
    const G r = scattering_texel_coord<G>(P, 0u, unsigned(y), unsigned(z)).r;
    const float rf = float(r);
    const float top2 = P.top * P.top;
    const int TMU = P.t_mu;
    const float dxf = s_trw[0].z;
    auto store = [&](int x, const Sum<COMP>* rs, const Sum<COMP>* ms) {
        float3 ray = make_float3(rs[0].value() * dxf * P.solar[0] * P.rayleigh_scattering[0],
                                 rs[1].value() * dxf * P.solar[1] * P.rayleigh_scattering[1],
                                 rs[2].value() * dxf * P.solar[2] * P.rayleigh_scattering[2]);
        float3 mie = make_float3(ms[0].value() * dxf * P.solar[0] * P.mie_scattering[0],
                                 ms[1].value() * dxf * P.solar[1] * P.mie_scattering[1],
                                 ms[2].value() * dxf * P.solar[2] * P.mie_scattering[2]);
        const size_t idx = (size_t(z) * MU + y) * W + x;
        st_tex(dR, idx, make_float4(ray.x, ray.y, ray.z, 0.0f), H);
        st_tex(dM, idx, make_float4(mie.x, mie.y, mie.z, 0.0f), H);
        st_tex(S, idx, make_float4(ray.x, ray.y, ray.z, mie.x), H);
    };
    // One sample's contribution: transmittances and densities.
    auto accumulate = [&](Sum<COMP>* rs, Sum<COMP>* ms, float4 tv, float4 gq, float3 ts,
                          float sm) {
        float t0 = tv.x * (ts.x * sm), t1 = tv.y * (ts.y * sm), t2 = tv.z * (ts.z * sm);
        rs[0].add(t0 * gq.y);
        rs[1].add(t1 * gq.y);
        rs[2].add(t2 * gq.y);
        ms[0].add(t0 * gq.z);
        ms[1].add(t1 * gq.z);
        ms[2].add(t2 * gq.z);
    };
    if constexpr (STAGED) {
        extern __shared__ float4 s_trow[];  // [2][TMU]
        float tmu_s[kSsTexelsPerThread], tnu[kSsTexelsPerThread];
        Sum<COMP> rs[kSsTexelsPerThread][3], ms[kSsTexelsPerThread][3];
#pragma unroll
        for (int q = 0; q < kSsTexelsPerThread; ++q) {
            const int x = tid + q * nt;
            tmu_s[q] = 0.0f;
            tnu[q] = 0.0f;
            if (x < W) {
                TexelCoord<G> c =
                    scattering_texel_coord<G>(P, unsigned(x), unsigned(y), unsigned(z));
                tmu_s[q] = float(c.mu_s);
                tnu[q] = float(c.nu);
            }
        }
        // Sample i's two T rows, lerped in r with the row weight, into dst.
        auto stage = [&](int i, float4* dst) {
            const float4 tr4 = s_trw[i];
            const float w = s_lut[i].w;
            const float4* t0 = T + __float_as_int(tr4.x);
            const float4* t1 = T + __float_as_int(tr4.y);
            for (int j = tid; j < TMU; j += nt) {
                const float4 a = ldg4(t0 + j), b = ldg4(t1 + j);
                dst[j] = make_float4(fmaf(w, b.x - a.x, a.x), fmaf(w, b.y - a.y, a.y),
                                     fmaf(w, b.z - a.z, a.z), 0.0f);
            }
        };
        stage(0, s_trow);
        __syncthreads();
        for (int i = 0; i < kMsSamples; ++i) {
            const float4* cur = s_trow + (i & 1) * TMU;
            if (i + 1 < kMsSamples) stage(i + 1, s_trow + ((i + 1) & 1) * TMU);
            const float4 tv = s_tv[i], gq = s_geo[i], lq = s_lut[i], iv = s_inv[i];
#pragma unroll
            for (int q = 0; q < kSsTexelsPerThread; ++q) {
                if (tid + q * nt >= W) continue;
                float t, sm;
                ss_sun_coord<kPrecise>(rf, tmu_s[q], tnu[q], tv, gq, lq, iv, top2, TMU, t, sm);
                const int i0 = int(t);  // t >= 0: truncation is floor
                const int i1 = min(i0 + 1, TMU - 1);
                const float f = t - float(i0);
                const float4 a = cur[i0], b = cur[i1];
                const float3 ts = make_float3(fmaf(f, b.x - a.x, a.x), fmaf(f, b.y - a.y, a.y),
                                              fmaf(f, b.z - a.z, a.z));
                accumulate(rs[q], ms[q], tv, gq, ts, sm);
            }
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < kSsTexelsPerThread; ++q)
            if (tid + q * nt < W) store(tid + q * nt, rs[q], ms[q]);
    } else {
        for (int x = tid; x < W; x += nt) {
            TexelCoord<G> c = scattering_texel_coord<G>(P, unsigned(x), unsigned(y), unsigned(z));
            const float mu_s = float(c.mu_s), nu = float(c.nu);
            Sum<COMP> rs[3], ms[3];
            // The row table's entries of sample i+1 are loaded while sample i
            // is evaluated (scalar loads, waited for at their first use).
            float4 tv_n = s_tv[0], gq_n = s_geo[0], lq_n = s_lut[0], iv_n = s_inv[0];
            float4 tr4_n = s_trw[0];
            for (int i = 0; i < kMsSamples; ++i) {
                const float4 tv = tv_n, gq = gq_n, lq = lq_n, iv = iv_n, tr4 = tr4_n;
                const int in = i + 1 < kMsSamples ? i + 1 : i;
                tv_n = s_tv[in];
                gq_n = s_geo[in];
                lq_n = s_lut[in];
                iv_n = s_inv[in];
                tr4_n = s_trw[in];
                // The sample's two T rows (r taps) as one buffer: row 0 at byte
                // 0, row 1 at `drow` (0 when both taps are the same row).
                const int row0 = __float_as_int(tr4.x),
                          drow = (__float_as_int(tr4.y) - row0) * 16;
                const __amdgpu_buffer_rsrc_t tb = buffer_rsrc(T + row0, drow + TMU * 16);
                float t, sm;
                ss_sun_coord<kPrecise>(rf, mu_s, nu, tv, gq, lq, iv, top2, TMU, t, sm);
                const int i0 = int(t);
                const int o0 = i0 * 16, o1 = min(i0 + 1, TMU - 1) * 16;
                const float f = t - float(i0);
                const float3 ts =
                    lerp3(lerp3(buffer_rgb(tb, o0), buffer_rgb(tb, o1), f),
                          lerp3(buffer_rgb(tb, drow + o0), buffer_rgb(tb, drow + o1), f), lq.w);
                accumulate(rs, ms, tv, gq, ts, sm);
            }
            store(x, rs, ms);
        }
    }
}

// Rows of at most kSsChunkMaxW texels on the direct-lookup path
// (ss_chunked): the block's ss_chunks(P) groups of W threads each integrate
// a contiguous chunk of the row's samples -- per sample, the arithmetic of
// single_scattering_block's direct path -- and the chunk sums are added in
// chunk order.  The row table's reads depend on the group's sample: vector
// loads.
template <typename G, bool H>
__device__ __forceinline__ void single_scattering_block_chunked(
    const DevParams& P, const int y, const int z, const float4* __restrict__ T,
    const float4* __restrict__ SR, float4* __restrict__ dR, float4* __restrict__ dM,
    float4* __restrict__ S) {
    extern __shared__ float s_ssp[];  // [chunk][W][6]: Rayleigh rgb, Mie rgb
    const float4* row = SR + (size_t(z) * P.s_mu + y) * kSsRowFloat4s;
    const float4* s_tv = row;
    const float4* s_geo = row + kMsSamples;
    const float4* s_lut = row + 2 * kMsSamples;
    const float4* s_inv = row + 3 * kMsSamples;
    const float4* s_trw = row + 4 * kMsSamples;
    constexpr bool kPrecise = sizeof(G) == sizeof(double);
    const int NU = P.s_nu, MUS = P.s_mus, MU = P.s_mu;
    const int W = NU * MUS;
    const int CH = ss_chunks(P);
    const int tid = threadIdx.x;
    const int g = tid / W, x = tid - g * W;
    const bool in_chunk = g < CH;
    const G r = scattering_texel_coord<G>(P, 0u, unsigned(y), unsigned(z)).r;
    const float rf = float(r);
    const float top2 = P.top * P.top;
    const int TMU = P.t_mu;
    if (in_chunk) {
        const int i0 = g * kMsSamples / CH, i1 = (g + 1) * kMsSamples / CH;
        TexelCoord<G> c = scattering_texel_coord<G>(P, unsigned(x), unsigned(y), unsigned(z));
        const float mu_s = float(c.mu_s), nu = float(c.nu);
        float rs0 = 0.0f, rs1 = 0.0f, rs2 = 0.0f, ms0 = 0.0f, ms1 = 0.0f, ms2 = 0.0f;
        for (int i = i0; i < i1; ++i) {
            const float4 tv = s_tv[i], gq = s_geo[i], lq = s_lut[i], iv = s_inv[i],
                         tr4 = s_trw[i];
            const int row0 = __float_as_int(tr4.x), drow = (__float_as_int(tr4.y) - row0) * 16;
            const __amdgpu_buffer_rsrc_t tb = buffer_rsrc(T + row0, drow + TMU * 16);
            float t, sm;
            ss_sun_coord<kPrecise>(rf, mu_s, nu, tv, gq, lq, iv, top2, TMU, t, sm);
            const int k0 = int(t);
            const int o0 = k0 * 16, o1 = min(k0 + 1, TMU - 1) * 16;
            const float f = t - float(k0);
            const float3 ts =
                lerp3(lerp3(buffer_rgb(tb, o0), buffer_rgb(tb, o1), f),
                      lerp3(buffer_rgb(tb, drow + o0), buffer_rgb(tb, drow + o1), f), lq.w);
            const float t0 = tv.x * (ts.x * sm), t1 = tv.y * (ts.y * sm), t2 = tv.z * (ts.z * sm);
            rs0 = rs0 + t0 * gq.y;
            rs1 = rs1 + t1 * gq.y;
            rs2 = rs2 + t2 * gq.y;
            ms0 = ms0 + t0 * gq.z;
            ms1 = ms1 + t1 * gq.z;
            ms2 = ms2 + t2 * gq.z;
        }
        float* p = s_ssp + size_t(g * W + x) * 6;
        p[0] = rs0;
        p[1] = rs1;
        p[2] = rs2;
        p[3] = ms0;
        p[4] = ms1;
        p[5] = ms2;
    }
    __syncthreads();
    if (g == 0) {
        float v[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] = s_ssp[size_t(x) * 6 + k];
        for (int cg = 1; cg < CH; ++cg)
#pragma unroll
            for (int k = 0; k < 6; ++k) v[k] = v[k] + s_ssp[size_t(cg * W + x) * 6 + k];
        const float dxf = s_trw[0].z;
        const float3 ray = make_float3(v[0] * dxf * P.solar[0] * P.rayleigh_scattering[0],
                                       v[1] * dxf * P.solar[1] * P.rayleigh_scattering[1],
                                       v[2] * dxf * P.solar[2] * P.rayleigh_scattering[2]);
        const float3 mie = make_float3(v[3] * dxf * P.solar[0] * P.mie_scattering[0],
                                       v[4] * dxf * P.solar[1] * P.mie_scattering[1],
                                       v[5] * dxf * P.solar[2] * P.mie_scattering[2]);
        const size_t idx = (size_t(z) * MU + y) * W + x;
        st_tex(dR, idx, make_float4(ray.x, ray.y, ray.z, 0.0f), H);
        st_tex(dM, idx, make_float4(mie.x, mie.y, mie.z, 0.0f), H);
        st_tex(S, idx, make_float4(ray.x, ray.y, ray.z, mie.x), H);
    }
}

// ===================================================== 4. scattering density
// scattering_density.comp:10-155.
//
// For a texel (r, mu, mu_s, nu) every incident-radiance lookup of the
// 16x32-direction quadrature is GetScattering(r, cos(theta_l), mu_s, nu1),
// so its (u_r, u_mu, u_mu_s) coordinates depend only on the texel's r and
// mu_s and on theta_l; only u_nu varies with phi.  All MU*NU texels that
// share (r, mu_s) therefore read the same 16 "nu profiles": for each theta
// row, the trilinear (mu_s, mu, r) sample of every nu slice k = -1..NU
// (slices -1 and NU reproduce CLAMP_TO_EDGE at the LUT's x edges).  A block
// builds those profiles in LDS once; each sample is then a 1D lerp between
// two LDS entries — the same value the reference's two trilinear texture
// fetches produce, with 8x-64x fewer gathers.
//
// Grid: (ceil(MU*NU/256), MUS, R); each block serves one (r, mu_s) tile.
template <typename G>
__device__ __forceinline__ Taps slice_taps(const DevParams& P, int slice, G u_mu_s) {
    G x = (G(slice) + u_mu_s) / G(P.s_nu);
    return linear_taps<G>(x, P.s_nu * P.s_mus);
}

// H: the texture is binary16 (FB_FLAG_F16_STORAGE).
template <bool H>
__device__ __forceinline__ float4 trilinear_taps_t(const DevParams& P,
                                                   const float4* __restrict__ tex, const Taps& tx,
                                                   const Taps& ty, const Taps& tz) {
    const int W = P.s_nu * P.s_mus;
    const size_t p00 = (size_t(tz.i0) * P.s_mu + ty.i0) * W;
    const size_t p01 = (size_t(tz.i0) * P.s_mu + ty.i1) * W;
    const size_t p10 = (size_t(tz.i1) * P.s_mu + ty.i0) * W;
    const size_t p11 = (size_t(tz.i1) * P.s_mu + ty.i1) * W;
    auto at = [&](size_t row, int x) { return ld_tex(tex, row + x, H); };
    return lerp4(lerp4(lerp4(at(p00, tx.i0), at(p00, tx.i1), tx.w),
                       lerp4(at(p01, tx.i0), at(p01, tx.i1), tx.w), ty.w),
                 lerp4(lerp4(at(p10, tx.i0), at(p10, tx.i1), tx.w),
                       lerp4(at(p11, tx.i0), at(p11, tx.i1), tx.w), ty.w),
                 tz.w);
}

// A profile entry: the value at slot k (rgb) and the difference to slot
// k+1, so a lerp is one fma per channel.  32-byte stride; a lookup reads
// 24 bytes (one b128 and one b64 LDS load).
struct ProfEntry {
    float4 a;  // v.r, v.g, v.b, d.r
    float2 b;  // d.g, d.b
    float2 pad;
};

// Linear filtering at texel coordinate t over entries e = k+1, k in
// [-1, kmax]: entry 0 holds the clamped left edge and entry kmax+1 the
// clamped right edge, which reproduces CLAMP_TO_EDGE.  Callers pass a
// finite t (their inputs are clamped unit vectors); the med3 clamp keeps
// the entry index inside [0, kmax+1].
__device__ __forceinline__ float3 prof_lerp(const ProfEntry* __restrict__ prof, float t,
                                            float kmax) {
    float k = floorf(t);
    float f = t - k;
    k = __builtin_amdgcn_fmed3f(k, -1.0f, kmax);
    const ProfEntry* e = prof + (int(k) + 1);
    float4 a = e->a;
    float2 b = e->b;
    return make_float3(fmaf(a.w, f, a.x), fmaf(b.x, f, a.y), fmaf(b.y, f, a.z));
}

// 1/x^1.5 for x >= (1-|g|)^2 > 0: the raw v_rsq_f32 (no denormal rescaling
// is needed for this range).
__device__ __forceinline__ float rsqrt_cubed(float x) {
    float r = __builtin_amdgcn_rsqf(x);
    return r * r * r;
}

__device__ __forceinline__ ProfEntry make_entry(float4 v, float4 w) {
    ProfEntry e;
    e.a = make_float4(v.x, v.y, v.z, w.x - v.x);
    e.b = make_float2(w.y - v.y, w.z - v.z);
    e.pad = make_float2(0.0f, 0.0f);
    return e;
}

// Order 2 (single scattering as incident radiance): the Rayleigh and Mie
// profile entries of one slot side by side, so that one lookup reads both
// with three b128 LDS reads instead of two b128 + b64 pairs.
struct PairEntry {
    float4 a;  // R: v.r, v.g, v.b, d.r
    float4 b;  // R: d.g, d.b; M: v.r, v.g
    float4 c;  // M: v.b, d.r, d.g, d.b
};
__device__ __forceinline__ PairEntry make_pair_entry(float4 rv, float4 rw, float4 mv, float4 mw) {
    PairEntry e;
    e.a = make_float4(rv.x, rv.y, rv.z, rw.x - rv.x);
    e.b = make_float4(rw.y - rv.y, rw.z - rv.z, mv.x, mv.y);
    e.c = make_float4(mv.z, mw.x - mv.x, mw.y - mv.y, mw.z - mv.z);
    return e;
}

// Per-texel constants of the density quadrature.
struct DenTexel {
    float omx, omz;          // omega = (omx, 0, omz)
    float osx, osy, osz;     // omega_s
    float osxh, osyh;        // omega_s.x, .y x (NU-1)/2
    float g1, g2;            // 1 + g^2, -2g
    float hnu, hE;           // (NU-1)/2, (IRR_MUS-1)/2
    float hnu1;              // (NU-1)/2 + 1
    const ProfEntry* erow;   // delta_irradiance at r = bottom
};
// Per-(texel, theta) constants.
struct DenRow {
    float C, omzc;           // cos(theta)*omega_s.z, cos(theta)*omega.z
    float Ch1;               // (C + 1) x (NU-1)/2 + 1
    const ProfEntry* pr;     // nu profile of multiple scattering (order n-1 >= 2)
    const PairEntry* pp;     // nu profiles of single Rayleigh and Mie (order 1)
    const float4* dir;       // c_den_dir row: (cos phi sin t, sin phi sin t, cos t, dw)
    float3 gv;               // T_g * albedo / pi (ground rows)
    float Ag, Bg, Cg1;       // entry coordinate of the ground normal . omega_s:
                             // cx*Ag + sy*Bg + Cg1 = (n.omega_s + 1) x
                             // (IRR_MUS-1)/2 + 1
};

// Linear filtering at texel coordinate t, given as e = t + 1: entry k + 1
// holds slot k, so floor(e) is the entry and fract(e) the weight, as
// prof_lerp computes for t >= -1.  Coordinates come from unit vectors, so
// rounding puts t at most a few ulp outside [0, kmax]: for t in (-1, 0), e
// is in (0, 1) and lerps slots -1 and 0, as the reference's two fetches do
// for a nu coordinate just below its range; just above kmax, int(e) is still
// the last entry and the weight of its delta a few ulp.  With e > 0,
// truncation is floor: the entry and weight take one conversion and one
// fract, and no clamp.  The +1 is folded into the per-texel and per-row
// terms the callers add anyway.
__device__ __forceinline__ float3 prof_lerp_i(const ProfEntry* __restrict__ prof, float e1) {
    const int k = int(e1);
    const float f = __builtin_amdgcn_fractf(e1);
    const ProfEntry* e = prof + k;
    float4 a = e->a;
    float2 b = e->b;
    return make_float3(fmaf(a.w, f, a.x), fmaf(b.x, f, a.y), fmaf(b.y, f, a.z));
}
// The same for a Rayleigh / Mie pair of profiles.
__device__ __forceinline__ void pair_lerp_i(const PairEntry* __restrict__ prof, float e1,
                                            float3& R, float3& M) {
    const int k = int(e1);
    const float f = __builtin_amdgcn_fractf(e1);
    const PairEntry* e = prof + k;
    const float4 a = e->a, b = e->b, c = e->c;
    R = make_float3(fmaf(a.w, f, a.x), fmaf(b.x, f, a.y), fmaf(b.y, f, a.z));
    M = make_float3(fmaf(c.y, f, b.z), fmaf(c.z, f, b.w), fmaf(c.w, f, c.x));
}

// The 16 direction pairs (phi_m, phi_{31-m}) of one theta row.  Both members
// share cos(phi), hence nu2 = omega.omega_i and the phase functions; their
// sin(phi) terms have opposite signs.  The row's Rayleigh and Mie sums are
// kept apart (aR = sum of q L, aM = sum of q w L, with q = 1 + nu2^2 and w
// the Cornette-Shanks power): their scattering coefficients are constant
// over the row and are applied once by the caller.
// One direction pair of density_row: the lookups of phi_m and phi_{31-m}
// for one texel, added to its row sums with the pair's q and q w.
template <bool SINGLE, bool GROUND>
__device__ __forceinline__ void density_pair(const DenTexel& dt, const DenRow& row, float cx,
                                             float sy, float q, float qw, float3& aR,
                                             float3& aM) {
    float3 La, Lb;
    if (SINGLE) {
        const float base = fmaf(cx, dt.osx, row.C), sb = sy * dt.osy;
        const float nua = base + sb, nub = base - sb;
        // entry coordinates (nu + 1) (NU-1)/2 + 1
        const float ta = fmaf(nua, dt.hnu, dt.hnu1), tb = fmaf(nub, dt.hnu, dt.hnu1);
        // Profiles are pre-scaled by kR (rayleigh) and kM (mie).
        float3 Ma, Mb;
        pair_lerp_i(row.pp, ta, La, Ma);
        pair_lerp_i(row.pp, tb, Lb, Mb);
        const float qa = fmaf(nua, nua, 1.0f), qb = fmaf(nub, nub, 1.0f);
        const float wa = rsqrt_cubed(fmaf(dt.g2, nua, dt.g1));
        const float wb = rsqrt_cubed(fmaf(dt.g2, nub, dt.g1));
        La = make_float3(qa * fmaf(Ma.x, wa, La.x), qa * fmaf(Ma.y, wa, La.y),
                         qa * fmaf(Ma.z, wa, La.z));
        Lb = make_float3(qb * fmaf(Mb.x, wb, Lb.x), qb * fmaf(Mb.y, wb, Lb.y),
                         qb * fmaf(Mb.z, wb, Lb.z));
    } else {
        // entry coordinate (nu + 1) (NU-1)/2 + 1 with the scale and
        // offset folded into the per-texel and per-row terms
        const float base = fmaf(cx, dt.osxh, row.Ch1), sb = sy * dt.osyh;
        La = prof_lerp_i(row.pr, base + sb);
        Lb = prof_lerp_i(row.pr, base - sb);
    }
    if (GROUND) {
        const float gb = fmaf(cx, row.Ag, row.Cg1), gs = sy * row.Bg;
        const float3 Ea = prof_lerp_i(dt.erow, gb + gs);
        const float3 Eb = prof_lerp_i(dt.erow, gb - gs);
        La = make_float3(fmaf(row.gv.x, Ea.x, La.x), fmaf(row.gv.y, Ea.y, La.y),
                         fmaf(row.gv.z, Ea.z, La.z));
        Lb = make_float3(fmaf(row.gv.x, Eb.x, Lb.x), fmaf(row.gv.y, Eb.y, Lb.y),
                         fmaf(row.gv.z, Eb.z, Lb.z));
    }
    const float Sx = La.x + Lb.x, Sy = La.y + Lb.y, Sz = La.z + Lb.z;
    aR.x = fmaf(q, Sx, aR.x);
    aR.y = fmaf(q, Sy, aR.y);
    aR.z = fmaf(q, Sz, aR.z);
    aM.x = fmaf(qw, Sx, aM.x);
    aM.y = fmaf(qw, Sy, aM.y);
    aM.z = fmaf(qw, Sz, aM.z);
}

// The 16 direction pairs (phi_m, phi_{31-m}) of one theta row.  Both members
// share cos(phi), hence nu2 = omega.omega_i and the phase functions; their
// sin(phi) terms have opposite signs.  The row's Rayleigh and Mie sums are
// kept apart (aR = sum of q L, aM = sum of q w L, with q = 1 + nu2^2 and w
// the Cornette-Shanks power): their scattering coefficients are constant
// over the row and are applied once by the caller.
template <bool SINGLE, bool GROUND>
__device__ __forceinline__ void density_row(const DenTexel& dt, const DenRow& row, float3& aR,
                                            float3& aM) {
#pragma unroll
    for (int m = 0; m < kDenPhi / 2; ++m) {
        const float4 o = row.dir[m];
        const float cx = o.x, sy = o.y;
        const float nu2 = fmaf(cx, dt.omx, row.omzc);
        const float q = fmaf(nu2, nu2, 1.0f);
        const float qw = q * rsqrt_cubed(fmaf(dt.g2, nu2, dt.g1));
        density_pair<SINGLE, GROUND>(dt, row, cx, sy, q, qw, aR, aM);
    }
}
// density_row for N texels of one mu (one thread): nu2, q and q w depend
// on mu and the direction only, so the texels share them; each texel's own
// operations are density_row's, in the same order (the same bits).
template <bool SINGLE, bool GROUND, int N>
__device__ __forceinline__ void density_row_n(const DenTexel* dt, const DenRow* row, float3* aR,
                                              float3* aM) {
#pragma unroll
    for (int m = 0; m < kDenPhi / 2; ++m) {
        const float4 o = row[0].dir[m];
        const float cx = o.x, sy = o.y;
        const float nu2 = fmaf(cx, dt[0].omx, row[0].omzc);
        const float q = fmaf(nu2, nu2, 1.0f);
        const float qw = q * rsqrt_cubed(fmaf(dt[0].g2, nu2, dt[0].g1));
#pragma unroll
        for (int u = 0; u < N; ++u)
            density_pair<SINGLE, GROUND>(dt[u], row[u], cx, sy, q, qw, aR[u], aM[u]);
    }
}

// Dynamic LDS of a density block: the raw profiles (float4 per texture,
// theta row and nu slot), whose space the irradiance entries reuse once the
// profile entries are built; the entries (PairEntry at order 2, ProfEntry
// after); the x taps of the NU + 2 slots.  Kept small: at order 2 it sets
// how many blocks share a CU.
inline size_t density_lds_bytes(int nu, int i_mus, bool single) {
    const int ntex = single ? 2 : 1;
    const size_t raw = sizeof(float4) * size_t(ntex) * kDenTheta * (nu + 2);
    const size_t erow = sizeof(ProfEntry) * size_t(i_mus + 1);
    const size_t ent = (single ? sizeof(PairEntry) : sizeof(ProfEntry)) * kDenTheta * (nu + 1);
    return std::max(raw, erow) + ent + sizeof(Taps) * size_t(nu + 2);
}

// The density geometry table (dg_row / dg_x) of one LUT set; block b < R
// fills the rows of r index b, block R + j the x taps of mu_s index j.
// scattering_density.comp:34-66 for the directions, :45-60 for the ground
// term; the taps are GetScatteringTextureUvwzFromRMuMuSNu's (scattering.h
// :7-60) for r = r(z), mu = cos(theta_l), mu_s = mu_s(j).
template <typename G>
__device__ __forceinline__ void density_geometry_block(const DevParams& P, const int b,
                                                       const float4* __restrict__ T,
                                                       const float4* __restrict__ OD,
                                                       float4* __restrict__ DG) {
    const int tid = threadIdx.x;
    const int R = P.s_r, NU = P.s_nu;
    if (b < R) {
        const int z = b;
        const G r = scattering_texel_coord<G>(P, 0u, 0u, unsigned(z)).r;
        float4* row = DG + dg_row(z);
        if (tid < kDenTheta) {
            const int l = tid;
            const G ct = den_cos<G>(l);
            const bool g_l = ray_intersects_ground<G>(P, r, ct);
            const Taps ty = linear_taps<G>(scattering_u_mu<G>(P, r, ct, g_l), P.s_mu);
            float4 gv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float ginv = 0.0f;
            if (g_l) {
                const G dg = distance_to_bottom<G>(P, r, ct);
                float3 tg;
                if constexpr (sizeof(G) == sizeof(double))
                    tg = get_transmittance_precise(P, T, OD, r, ct, dg, true);
                else
                    tg = get_transmittance<G>(P, T, r, ct, dg, true);
                gv = make_float4(tg.x * P.albedo[0] * (1.0f / kPi),
                                 tg.y * P.albedo[1] * (1.0f / kPi),
                                 tg.z * P.albedo[2] * (1.0f / kPi), float(dg));
                // The ground normal's length does not depend on phi.
                const float4 o = c_den_dir[l * kDenPhi];
                const G st2 = G(o.x) * G(o.x) + G(o.y) * G(o.y);
                const G gz = r + ct * dg;
                ginv = float(G(1) / sqrt(st2 * dg * dg + gz * gz));
            }
            row[l] = gv;
            row[kDenTheta + l] =
                make_float4(__int_as_float(ty.i0), __int_as_float(ty.i1), ty.w, ginv);
        } else if (tid == kDenTheta) {
            const Taps tz = linear_taps<G>(scattering_u_r<G>(P, r), P.s_r);
            row[2 * kDenTheta] =
                make_float4(__int_as_float(tz.i0), __int_as_float(tz.i1), tz.w, 0.0f);
        }
    } else {
        const int j = b - R;
        if (tid < NU + 2) {
            const G mu_s = scattering_texel_coord<G>(P, unsigned(j), 0u, 0u).mu_s;
            const Taps tx = slice_taps<G>(P, tid - 1, scattering_u_mu_s<G>(P, mu_s));
            DG[dg_x(R, j, NU) + tid] =
                make_float4(__int_as_float(tx.i0), __int_as_float(tx.i1), tx.w, 0.0f);
        }
    }
}

// Per-sample body of the density quadrature for one pair of directions
// (phi_m, phi_{31-m}): both share cos(phi), hence nu2 and the two phase
// functions; their sin(phi) terms have opposite signs.
template <typename G, bool SINGLE, bool H, int TPX = 1>
__device__ __forceinline__ void scattering_density_block(const DevParams& P, const int bx, const int by, const int bz, const float4* __restrict__ DG, const float4* __restrict__ dR,
    const float4* __restrict__ dM, const float4* __restrict__ dMS, const float4* __restrict__ dE,
    float4* __restrict__ J) {
    extern __shared__ float4 s_dyn[];
    __shared__ float4 s_gv[kDenTheta];    // (T_g*albedo/pi).rgb, distance_to_ground
    __shared__ float s_ginv[kDenTheta];   // 1/|r*zenith + omega_i*d_g|
    __shared__ Taps s_ty[kDenTheta];      // mu taps of each theta row
    __shared__ Taps s_tz;                 // r taps of the tile
    __shared__ int s_ground_mask;

    constexpr int NTEX = SINGLE ? 2 : 1;
    const int NU = P.s_nu, MUS = P.s_mus, MU = P.s_mu, NE = P.i_mus;
    const int W = NU * MUS;
    const int nslots = NU + 2, nent = NU + 1;
    const int j = by;  // mu_s index within a nu slice
    const int z = bz;  // r index
    const int tid = threadIdx.x;
    const int raw_f4 = NTEX * kDenTheta * nslots;
    float4* s_raw = s_dyn;                                   // [NTEX][16][nslots]
    ProfEntry* s_erow = reinterpret_cast<ProfEntry*>(s_dyn);  // [NE+1], after step 2
    float4* s_ent = s_dyn + max(raw_f4, 2 * (NE + 1));
    ProfEntry* s_prof = reinterpret_cast<ProfEntry*>(s_ent);  // [16][nent] (order >= 3)
    PairEntry* s_pair = reinterpret_cast<PairEntry*>(s_ent);  // [16][nent] (order 2)
    Taps* s_tx = reinterpret_cast<Taps*>(SINGLE ? static_cast<void*>(s_pair + kDenTheta * nent)
                                                : static_cast<void*>(s_prof + kDenTheta * nent));

    // 0. Tile geometry from the density geometry table: the mu taps and
    // ground terms of each theta row (they depend on r only), the r taps,
    // and the x taps of each nu slot (they depend on mu_s only).
    const int slot0 = int(blockDim.x) >= 64 + nslots ? 64 : kDenTheta;
    if (tid < kDenTheta) {
        const float4* dgz = DG + dg_row(z);
        const float4 t = dgz[kDenTheta + tid];
        s_gv[tid] = dgz[tid];
        s_ginv[tid] = t.w;
        s_ty[tid] = Taps{__float_as_int(t.x), __float_as_int(t.y), t.z};
        const unsigned long long b = __ballot(t.w != 0.0f);
        if (tid == 0) s_ground_mask = int(b & 0xffffull);
    } else if (tid >= slot0 && tid < slot0 + nslots) {
        const float4 x = DG[dg_x(P.s_r, j, NU) + (tid - slot0)];
        s_tx[tid - slot0] = Taps{__float_as_int(x.x), __float_as_int(x.y), x.z};
        if (tid == slot0) {
            const float4 c = DG[dg_row(z) + 2 * kDenTheta];
            s_tz = Taps{__float_as_int(c.x), __float_as_int(c.y), c.z};
        }
    }
    __syncthreads();
    // 1. Raw profiles: 16 theta rows x (NU+2) nu slots (-1..NU) per texture.
    {
        const Taps tz = s_tz;
        for (int e = tid; e < NTEX * kDenTheta * nslots; e += blockDim.x) {
            int tex_sel = e / (kDenTheta * nslots);
            int rem = e % (kDenTheta * nslots);
            int l = rem / nslots, slot = rem % nslots;
            const float4* tex = SINGLE ? (tex_sel == 0 ? dR : dM) : dMS;
            s_raw[e] = trilinear_taps_t<H>(P, tex, s_tx[slot], s_ty[l], tz);
        }
    }
    __syncthreads();
    // 2. (value, delta) entries.
    // Order 1 (single scattering) profiles carry the phase constants:
    // rayleigh x 3/(16 pi), mie x 3/(8 pi)(1-g^2)/(2+g^2) (util.h:26-34).
    const float kRs = SINGLE ? 3.0f / (16.0f * kPi) : 1.0f;
    const float kMs =
        SINGLE ? 3.0f / (8.0f * kPi) * (1.0f - P.g * P.g) / (2.0f + P.g * P.g) : 1.0f;
    for (int e = tid; e < kDenTheta * nent; e += blockDim.x) {
        int row = e / nent, k = e % nent;  // entry k holds slot k-1
        auto scaled = [&](int rr, int kk, float sc) {
            const float4 v = s_raw[rr * nslots + kk];
            return make_float4(v.x * sc, v.y * sc, v.z * sc, 0.0f);
        };
        if (SINGLE)
            s_pair[e] = make_pair_entry(scaled(row, k, kRs), scaled(row, k + 1, kRs),
                                        scaled(kDenTheta + row, k, kMs),
                                        scaled(kDenTheta + row, k + 1, kMs));
        else
            s_prof[e] = make_entry(scaled(row, k, kRs), scaled(row, k + 1, kRs));
    }
    __syncthreads();
    // 3. delta_irradiance at r = bottom (irradiance.h:20-30 with x_r = 0)
    // as (value, delta) entries, in the raw profiles' space.
    {
        const Taps ty = linear_taps<G>(tex_from_unit<G>(G(0), P.i_r), P.i_r);
        auto er = [&](int x) {
            return lerp4(ldg4(dE + ty.i0 * NE + x), ldg4(dE + ty.i1 * NE + x), ty.w);
        };
        for (int e = tid; e <= NE; e += blockDim.x)
            s_erow[e] = make_entry(er(min(max(e - 1, 0), NE - 1)), er(min(e, NE - 1)));
    }
    __syncthreads();

    const float g = P.g;
    const float kR = 3.0f / (16.0f * kPi);
    const float kM = 3.0f / (8.0f * kPi) * (1.0f - g * g) / (2.0f + g * g);
    const int gmask = s_ground_mask;
    // omega, omega_s of texel (nu index k, mu index y) (scattering_density.
    // comp:28-32), built in G.
    auto texel_setup = [&](int k, int y, DenTexel& dt) {
        const unsigned x = unsigned(k * MUS + j);
        TexelCoord<G> c = scattering_texel_coord<G>(P, x, unsigned(y), unsigned(z));
        G om_x = sqrt(G(1) - c.mu * c.mu);
        G sdx = om_x == G(0) ? G(0) : (c.nu - c.mu * c.mu_s) / om_x;
        G sdy = sqrt(fmax(G(1) - sdx * sdx - c.mu_s * c.mu_s, G(0)));
        dt.omx = float(om_x);
        dt.omz = float(c.mu);
        // Unit-vector components, clamped so every lookup coordinate is finite.
        dt.osx = fminf(fmaxf(float(sdx), -1.0f), 1.0f);
        dt.osy = fminf(fmaxf(float(sdy), -1.0f), 1.0f);
        dt.osz = float(c.mu_s);
        dt.osxh = dt.osx * (0.5f * float(NU - 1));
        dt.osyh = dt.osy * (0.5f * float(NU - 1));
        dt.g1 = 1.0f + g * g;
        dt.g2 = -2.0f * g;
        dt.hnu = 0.5f * float(NU - 1);
        dt.hnu1 = dt.hnu + 1.0f;
        dt.hE = 0.5f * float(NE - 1);
        dt.erow = s_erow;
        return c;
    };
    // Theta row l of a texel (ct = cos(theta)).
    auto make_row = [&](const DenTexel& dt, int l, float ct, float rf) {
        DenRow row;
        row.C = ct * dt.osz;
        row.Ch1 = fmaf(row.C, dt.hnu, dt.hnu1);
        row.omzc = ct * dt.omz;
        row.pr = s_prof + l * nent;
        row.pp = s_pair + l * nent;
        row.dir = c_den_dir + l * kDenPhi;
        if ((gmask >> l) & 1) {
            const float4 gv = s_gv[l];
            const float inv = s_ginv[l];
            const float dgi = gv.w * inv;
            row.gv = make_float3(gv.x, gv.y, gv.z);
            row.Ag = dgi * dt.osx * dt.hE;
            row.Bg = dgi * dt.osy * dt.hE;
            row.Cg1 = fmaf((rf + ct * gv.w) * inv * dt.osz, dt.hE, dt.hE + 1.0f);
        }
        return row;
    };

    if constexpr (TPX > 1) {
        // Thread p of the tile: texels (mu index y, nu index k + u NU/TPX),
        // u < TPX, which share mu (NU a multiple of TPX; one thread per
        // group of TPX texels).
        const int part = NU / TPX;
        const int p = bx * int(blockDim.x) + tid;
        if (p >= MU * part) return;
        const int k0 = p % part, y = p / part;
        DenTexel dts[TPX];
        TexelCoord<G> c0 = texel_setup(k0, y, dts[0]);
#pragma unroll
        for (int u = 1; u < TPX; ++u) texel_setup(k0 + u * part, y, dts[u]);
        const float rf = float(c0.r);
        const float alt = float(c0.r - G(P.bottom));
        const float rho_r = profile_density(P.prof[0], alt);
        const float rho_m = profile_density(P.prof[1], alt);
        const float cr0 = kR * P.rayleigh_scattering[0] * rho_r,
                    cr1 = kR * P.rayleigh_scattering[1] * rho_r,
                    cr2 = kR * P.rayleigh_scattering[2] * rho_r;
        const float cm0 = kM * P.mie_scattering[0] * rho_m,
                    cm1 = kM * P.mie_scattering[1] * rho_m,
                    cm2 = kM * P.mie_scattering[2] * rho_m;
        float3 acc[TPX];
#pragma unroll
        for (int u = 0; u < TPX; ++u) acc[u] = make_float3(0.0f, 0.0f, 0.0f);
        for (int l = 0; l < kDenTheta; ++l) {
            float3 aR[TPX], aM[TPX];
            DenRow rows[TPX];
            const float4 d0 = c_den_dir[l * kDenPhi];
            const float ct = d0.z, dom = d0.w;
#pragma unroll
            for (int u = 0; u < TPX; ++u) {
                aR[u] = make_float3(0.0f, 0.0f, 0.0f);
                aM[u] = make_float3(0.0f, 0.0f, 0.0f);
                rows[u] = make_row(dts[u], l, ct, rf);
            }
            if ((gmask >> l) & 1)
                density_row_n<SINGLE, true, TPX>(dts, rows, aR, aM);
            else
                density_row_n<SINGLE, false, TPX>(dts, rows, aR, aM);
#pragma unroll
            for (int u = 0; u < TPX; ++u) {
                acc[u].x = fmaf(cm0 * dom, aM[u].x, fmaf(cr0 * dom, aR[u].x, acc[u].x));
                acc[u].y = fmaf(cm1 * dom, aM[u].y, fmaf(cr1 * dom, aR[u].y, acc[u].y));
                acc[u].z = fmaf(cm2 * dom, aM[u].z, fmaf(cr2 * dom, aR[u].z, acc[u].z));
            }
        }
        const size_t rowbase = (size_t(z) * MU + y) * W;
#pragma unroll
        for (int u = 0; u < TPX; ++u)
            st_tex(J, rowbase + size_t((k0 + u * part) * MUS + j),
                   make_float4(acc[u].x, acc[u].y, acc[u].z, 0.0f), H);
        return;
    } else {
        // S consecutive lanes per texel, lane l0 of a texel taking theta rows
        // l0, l0 + S, ... (S = 1: one thread walks all 16 rows).
        const int S = P.den_split;
        const int t = bx * (int(blockDim.x) / S) + tid / S;  // texel within the tile
        if (t >= MU * NU) return;  // a whole group of S lanes
        const int l0 = tid % S;
        const int k = t % NU, y = t / NU;
        const unsigned x = unsigned(k * MUS + j);
        DenTexel dt;
        const TexelCoord<G> c = texel_setup(k, y, dt);
        const float rf = float(c.r);
        const float alt = float(c.r - G(P.bottom));
        const float rho_r = profile_density(P.prof[0], alt);
        const float rho_m = profile_density(P.prof[1], alt);
        const float cr0 = kR * P.rayleigh_scattering[0] * rho_r,
                    cr1 = kR * P.rayleigh_scattering[1] * rho_r,
                    cr2 = kR * P.rayleigh_scattering[2] * rho_r;
        const float cm0 = kM * P.mie_scattering[0] * rho_m,
                    cm1 = kM * P.mie_scattering[1] * rho_m,
                    cm2 = kM * P.mie_scattering[2] * rho_m;
        float3 acc = make_float3(0.0f, 0.0f, 0.0f);
        for (int l = l0; l < kDenTheta; l += S) {
            float3 aR = make_float3(0.0f, 0.0f, 0.0f), aM = make_float3(0.0f, 0.0f, 0.0f);
            const float4 d0 = c_den_dir[l * kDenPhi];
            const float ct = d0.z, dom = d0.w;
            const DenRow row = make_row(dt, l, ct, rf);
            if ((gmask >> l) & 1)
                density_row<SINGLE, true>(dt, row, aR, aM);
            else
                density_row<SINGLE, false>(dt, row, aR, aM);
            // Scattering coefficients x phase constants x solid angle.
            acc.x = fmaf(cm0 * dom, aM.x, fmaf(cr0 * dom, aR.x, acc.x));
            acc.y = fmaf(cm1 * dom, aM.y, fmaf(cr1 * dom, aR.y, acc.y));
            acc.z = fmaf(cm2 * dom, aM.z, fmaf(cr2 * dom, aR.z, acc.z));
        }
        if (S > 1) {
            // The S rows' sums of one texel, in S consecutive lanes of one
            // wavefront (S divides 64): a tree reduction.
            for (int o = S / 2; o > 0; o >>= 1) {
                acc.x += __shfl_xor(acc.x, o);
                acc.y += __shfl_xor(acc.y, o);
                acc.z += __shfl_xor(acc.z, o);
            }
            if (l0 != 0) return;
        }
        st_tex(J, (size_t(z) * MU + y) * W + x, make_float4(acc.x, acc.y, acc.z, 0.0f), H);
    }
}


// =================================================== 5. indirect irradiance
// indirect_irradiance.comp:10-74.  One wavefront per irradiance texel: lane
// i takes phi_i (64 lanes = the 64 phi samples), loops the 16 theta rows,
// and the wave reduces.  As in the density pass, the lookups of one theta
// row share (u_r, u_mu, u_mu_s), so each wave first builds its 16 nu
// profiles in LDS.
constexpr int kIndWaves = 4;  // waves per block (fewer when LDS would exceed 64 KiB)

__host__ __device__ inline size_t indirect_lds_bytes_per_wave(int nu, bool single) {
    const int ntex = single ? 2 : 1;
    return sizeof(float4) * size_t(ntex) * kIndTheta * (nu + 2) +
           sizeof(ProfEntry) * size_t(ntex) * kIndTheta * (nu + 1);
}

template <typename G, bool SINGLE, bool H>
__device__ __forceinline__ void indirect_irradiance_block(const DevParams& P, const int bx, const int by, const int bz, const float4* __restrict__ dR, const float4* __restrict__ dM,
    const float4* __restrict__ dMS, float4* __restrict__ dE, float4* __restrict__ E) {
    extern __shared__ float4 s_dyn[];
    constexpr int NTEX = SINGLE ? 2 : 1;
    const int waves = blockDim.x / 64;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int texel = bx * waves + wave;
    const int n = P.i_mus * P.i_r;
    const bool active = texel < n;
    const int ti = active ? texel % P.i_mus : 0, tj = active ? texel / P.i_mus : 0;
    const int NU = P.s_nu, nslots = NU + 2, nent = NU + 1;
    const size_t per_wave = indirect_lds_bytes_per_wave(NU, SINGLE);
    float4* s_raw = reinterpret_cast<float4*>(reinterpret_cast<char*>(s_dyn) + wave * per_wave);
    ProfEntry* s_prof = reinterpret_cast<ProfEntry*>(s_raw + NTEX * kIndTheta * nslots);

    G x_mu_s = G(ti) / G(P.i_mus - 1);
    G x_r = G(tj) / G(P.i_r - 1);
    G r, mu_s;
    rmus_from_irradiance_unit<G>(P, x_mu_s, x_r, r, mu_s);
    const G u_mu_s = scattering_u_mu_s<G>(P, mu_s);
    const Taps tz = linear_taps<G>(scattering_u_r<G>(P, r), P.s_r);
    for (int e = lane; e < NTEX * kIndTheta * nslots; e += 64) {
        int tex_sel = e / (kIndTheta * nslots);
        int rem = e % (kIndTheta * nslots);
        int jj = rem / nslots, slot = rem % nslots;
        G ct = ind_cos<G>(jj);
        Taps ty = linear_taps<G>(scattering_u_mu<G>(P, r, ct, false), P.s_mu);
        const float4* tex = SINGLE ? (tex_sel == 0 ? dR : dM) : dMS;
        s_raw[e] = trilinear_taps_t<H>(P, tex, slice_taps<G>(P, slot - 1, u_mu_s), ty, tz);
    }
    __syncthreads();
    for (int e = lane; e < NTEX * kIndTheta * nent; e += 64) {
        int row = e / nent, k = e % nent;
        s_prof[e] = make_entry(s_raw[row * nslots + k], s_raw[row * nslots + k + 1]);
    }
    __syncthreads();
    if (!active) return;

    const float osx = float(sqrt(G(1) - mu_s * mu_s)), osz = float(mu_s);
    const float g = P.g;
    const float kR = 3.0f / (16.0f * kPi);
    const float kM = 3.0f / (8.0f * kPi) * (1.0f - g * g) / (2.0f + g * g);
    const float g1 = 1.0f + g * g, g2 = -2.0f * g;
    const float hnu = 0.5f * float(NU - 1);
    const int kmax = NU - 1;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for (int jj = 0; jj < kIndTheta; ++jj) {
        const float4 om = c_ind_dir[jj * kIndPhi + lane];
        const float nu = fmaf(om.x, osx, om.z * osz);
        const float t = fmaf(nu, hnu, hnu);
        float3 v = prof_lerp(s_prof + jj * nent, t, kmax);
        if (SINGLE) {
            float3 m = prof_lerp(s_prof + (kIndTheta + jj) * nent, t, kmax);
            float q = fmaf(nu, nu, 1.0f);
            float rs = rsqrtf(fmaf(g2, nu, g1));
            float pR = kR * q, pM = kM * q * (rs * rs * rs);
            v = make_float3(fmaf(m.x, pM, v.x * pR), fmaf(m.y, pM, v.y * pR),
                            fmaf(m.z, pM, v.z * pR));
        }
        const float w = om.z * om.w;
        a0 = fmaf(v.x, w, a0);
        a1 = fmaf(v.y, w, a1);
        a2 = fmaf(v.z, w, a2);
    }
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off, 64);
        a1 += __shfl_xor(a1, off, 64);
        a2 += __shfl_xor(a2, off, 64);
    }
    if (lane == 0) {
        int o = tj * P.i_mus + ti;
        dE[o] = make_float4(a0, a1, a2, 0.0f);
        float4 prev = E[o];
        E[o] = make_float4(a0 + prev.x, a1 + prev.y, a2 + prev.z, 0.0f + prev.w);
    }
}

// =================================================== 6. multiple scattering
// multiple_scattering.comp:9-93.
//
// All texels of one (r, mu) row (x = nu*MUS + mu_s varies) march the same
// ray: d_i, r_i, mu_i, the view transmittance T(r, mu, d_i) and the
// (u_mu, u_r) taps of every sample are shared, only mu_s_i (and the fixed
// nu slice) differ per texel.  A block owns one row: it computes the 51
// sample descriptors once, and for each sample stages the (mu, r)-bilinear
// row of the density LUT in LDS (double buffered, one barrier per sample);
// each texel then lerps 2 nu slices x 2 mu_s taps from LDS.
constexpr int kMsTexelsPerThread = 4;
// Threads per block of the multiple-scattering kernels: up to one per
// texel of the widest row (NU * MUS <= 1024).
constexpr int kMsMaxThreads = 1024;
// Staged density rows per block (sample i is evaluated from one while
// sample i+1 is staged into the other).
constexpr int kMsStagedRows = 2;

// A staged density value: rgb only (12 bytes, 4-byte aligned), so the
// four taps of a sample cost 48 bytes of LDS reads, not 64.
struct Rgb {
    float x, y, z;
};

__device__ __forceinline__ Rgb to_rgb(const Rgb& v) { return v; }
__device__ __forceinline__ Rgb to_rgb(const Half4& v) {
    return Rgb{float(v.x), float(v.y), float(v.z)};
}

// H: the density LUT J, the delta and the scattering LUT are binary16
// (FB_FLAG_F16_STORAGE); the density rows then stay binary16 in the slots
// until put() widens them.
template <typename G, bool COMP, int TPT, bool H>
__device__ __forceinline__ void multiple_scattering_block(const DevParams& P, const int bx, const int by, const int bz,
                                                             const float4* __restrict__ MR,
                                                             const float4* __restrict__ J,
                                                             float4* __restrict__ dMS,
                                                             float4* __restrict__ S) {
    extern __shared__ Rgb s_row[];              // [2][(NU + 1) * (MUS + 1)]: staged density rows
    const int NU = P.s_nu, MUS = P.s_mus, MU = P.s_mu;
    const int W = NU * MUS;
    const int y = bx, z = by;
    const int tid = threadIdx.x, nt = blockDim.x;

    // Row geometry: the 51 samples' terms, written by the geometry pass (the
    // same samples, single_scattering.comp:30-65).  Every read of it is
    // block-uniform, and the table is read-only here: read through the
    // constant address space, the reads are scalar loads.
    typedef FB_CONSTANT_AS const float4 CF4;
    typedef FB_CONSTANT_AS const int4 CI4;
    CF4* desc = (CF4*)(MR + (size_t(z) * MU + y) * kMsRowFloat4s);
    auto d_tw = [&](int i) { return desc[i]; };  // T(r,mu,d_i)*dx*weight (rgb), d_i
    auto d_geo = [&](int i) { return desc[kMsSamples + i]; };  // 1/r_i, slot load mask
    auto d_off = [&](int i) { return ((CI4*)(desc + 2 * kMsSamples))[i]; };  // slot rows
    auto d_wt = [&](int i) { return desc[3 * kMsSamples + i]; };  // slot weights
    const G r = scattering_texel_coord<G>(P, 0u, unsigned(y), unsigned(z)).r;

    // Per-texel constants: mu_s, nu, nu slice and its lerp weight.
    float tmu_s[TPT], tnu[TPT], tslice[TPT], tlerp[TPT];
    int spos[TPT];  // the texel's entry in a staged row
    Sum<COMP> acc[TPT][3];
    const float rf = float(r);
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        const int x = tid + q * nt;
        tmu_s[q] = 0.0f;
        tnu[q] = 0.0f;
        tslice[q] = 0.0f;
        tlerp[q] = 0.0f;
        spos[q] = x + x / MUS;
        if (x < W) {
            TexelCoord<G> c = scattering_texel_coord<G>(P, unsigned(x), unsigned(y), unsigned(z));
            tmu_s[q] = float(c.mu_s);
            tnu[q] = float(c.nu);
            float u_nu = (tnu[q] + 1.0f) / 2.0f;
            float tcx = u_nu * float(NU - 1);
            tslice[q] = floorf(tcx);
            tlerp[q] = tcx - tslice[q];
        }
    }
    __syncthreads();

    // GetScatteringTextureUvwzFromRMuMuSNu's mu_s coordinate (scattering.h:
    // 48-56) folded into the texel coordinate of slice `tslice`:
    // x_tex = x_mu_s*(MUS-1) + tslice*MUS, with reciprocals hoisted; in a
    // staged row, slice k starts at entry k*(MUS + 1).
    const float b = P.bottom, b2 = P.bottom * P.bottom;
    const float H2 = P.top * P.top - P.bottom * P.bottom;
    const float d_min = P.top - P.bottom;
    const float inv_range = 1.0f / (sqrtf(H2) - d_min);
    const float invA = 1.0f / (-2.0f * P.mu_s_min * P.bottom * inv_range);
    const int SS = MUS + 1;
    const float musm1 = float(MUS - 1), fSS = float(SS);

    // The four density-row slots of this thread's texels x = tid + q*nt,
    // carried from sample to sample (ms_row_slots): each sample loads only
    // the rows it does not share with the previous one.
    typedef typename std::conditional<H, Half4, Rgb>::type Slot;
    Slot slot[TPT][4];
#pragma unroll
    for (int q = 0; q < TPT; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k) slot[q][k] = Slot{};
    // Row loads as buffer loads: the row start (block-uniform) goes in the
    // scalar offset and the texel's byte offset in a loop-invariant vector
    // offset, so no address arithmetic writes a register between the loads
    // and each load lands directly in its slot.  Loads past the LUT's end
    // (texels x >= W of the last row) return 0 and are never used.
    constexpr int TB = H ? 8 : 16;  // bytes per texel of J
    const __amdgpu_buffer_rsrc_t jb = buffer_rsrc(J, int(size_t(P.s_r) * MU * W * TB));
    int xoff[TPT];
#pragma unroll
    for (int q = 0; q < TPT; ++q) xoff[q] = (tid + q * nt) * TB;
    auto ld = [&](int o, int q) {
        if constexpr (H) {
            return __builtin_bit_cast(Half4,
                                      __builtin_amdgcn_raw_buffer_load_b64(jb, xoff[q], o * TB, 0));
        } else {
            const auto v = __builtin_amdgcn_raw_buffer_load_b96(jb, xoff[q], o * TB, 0);
            return Rgb{__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2])};
        }
    };
    // Issue the loads of the sample whose slot mask and row offsets are
    // (mask, o); they stay in flight until put() reads them.
    auto advance = [&](int mask, int4 o) {
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            if (mask & 1) slot[q][0] = ld(o.x, q);
            if (mask & 2) slot[q][1] = ld(o.y, q);
            if (mask & 4) slot[q][2] = ld(o.z, q);
            if (mask & 8) slot[q][3] = ld(o.w, q);
        }
    };
    // Row layout: nu slice k (k = 0..NU-1) at entries k*SS .. k*SS+MUS-1,
    // then one pad entry; then a zero slice NU.  Pads and the zero slice are
    // read only with weight 0 (the mu_s tap after a slice's last entry has
    // weight 0, and so has the nu slice above the last one), so every tap
    // index goes unclamped.  The pad entry shifts consecutive slices by one
    // LDS bank (taps of neighbouring slices at the same mu_s would otherwise
    // share banks).
    const int WP = (NU + 1) * SS;
    for (int e = tid; e < NU + 1 + MUS; e += nt) {
        const int k = e < NU ? e * SS + MUS : NU * SS + (e - NU);  // pads, then the zero slice
#pragma unroll
        for (int b = 0; b < kMsStagedRows; ++b) s_row[b * WP + k] = Rgb{0.0f, 0.0f, 0.0f};
    }
    // The (mu, r)-bilinear density value of sample i at texel x (the slot
    // weights are the bilinear weights), into the staged row dst.
    auto put = [&](int i, const Slot* s, Rgb* dst, int x) {
        const float4 w = d_wt(i);
        const Rgb t[4] = {to_rgb(s[0]), to_rgb(s[1]), to_rgb(s[2]), to_rgb(s[3])};
        const Rgb v{fmaf(w.w, t[3].x, fmaf(w.z, t[2].x, fmaf(w.y, t[1].x, w.x * t[0].x))),
                    fmaf(w.w, t[3].y, fmaf(w.z, t[2].y, fmaf(w.y, t[1].y, w.x * t[0].y))),
                    fmaf(w.w, t[3].z, fmaf(w.z, t[2].z, fmaf(w.y, t[1].z, w.x * t[0].z)))};
        dst[x] = v;
    };
    auto put_all = [&](int i, Rgb* dst) {
#pragma unroll
        for (int q = 0; q < TPT; ++q)
            if (tid + q * nt < W) put(i, slot[q], dst, spos[q]);
    };
    // Sample i of every texel of this thread, from the staged row `cur`.
    auto eval = [&](int i, const Rgb* __restrict__ cur) {
        const float4 tw = d_tw(i);
        const float inv_r_i = d_geo(i).x, d_i = tw.w;
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            const int x = tid + q * nt;
            if (x < W) {
                // mu_s at the sample (multiple_scattering.comp:38), in [-1, 1].
                float mu_s_i = fmaf(rf, tmu_s[q], d_i * tnu[q]) * inv_r_i;
                mu_s_i = __builtin_amdgcn_fmed3f(mu_s_i, -1.0f, 1.0f);
                float disc = fmaf(b2, mu_s_i * mu_s_i, H2);
                float d = fmaxf(fmaf(-b, mu_s_i, __builtin_amdgcn_sqrtf(disc)), 0.0f);
                float a = (d - d_min) * inv_range;
                // max(1 - a/A, 0) / (1 + a), kept <= 1 (a >= 0 up to rounding).
                float xm = __builtin_amdgcn_fmed3f(
                    fmaf(-a, invA, 1.0f) * __builtin_amdgcn_rcpf(1.0f + a), 0.0f, 1.0f);
                float xt = fmaf(xm, musm1, tslice[q] * fSS);
                float fl = floorf(xt);
                float w = xt - fl;
                int i0 = int(fl);
                // 0 <= i0 - tslice*SS <= MUS - 1 (0 <= xm <= 1, 0 <= tslice
                // <= NU - 1): the taps i0, i0+1, i0+SS, i0+SS+1 stay inside
                // the padded row (see the layout above).
                const Rgb p0 = cur[i0], p1 = cur[i0 + 1], p2 = cur[i0 + SS],
                          p3 = cur[i0 + SS + 1];
                float l = tlerp[q];
                float v0 = fmaf(w, p1.x - p0.x, p0.x), v1 = fmaf(w, p1.y - p0.y, p0.y),
                      v2 = fmaf(w, p1.z - p0.z, p0.z);
                float u0 = fmaf(w, p3.x - p2.x, p2.x), u1 = fmaf(w, p3.y - p2.y, p2.y),
                      u2 = fmaf(w, p3.z - p2.z, p2.z);
                acc[q][0].add(fmaf(l, u0 - v0, v0) * tw.x);
                acc[q][1].add(fmaf(l, u1 - v1, v1) * tw.y);
                acc[q][2].add(fmaf(l, u2 - v2, v2) * tw.z);
            }
        }
    };
    // (Each thread stages exactly its own texels x = tid + q*nt: W <= TPT*nt.)
    advance(__float_as_int(d_geo(0).y), d_off(0));
    put_all(0, s_row);
    __syncthreads();
    // The slot mask and row offsets of the sample after next are read (scalar
    // loads) one iteration ahead of their use.
    int mask_n = __float_as_int(d_geo(1).y);
    int4 o_n = d_off(1);
    for (int i = 0; i < kMsSamples; ++i) {
        const Rgb* cur = s_row + (i & 1) * WP;
        Rgb* nxt = s_row + ((i + 1) & 1) * WP;
        const bool next = i + 1 < kMsSamples;
        // The loads of sample i+1 are issued first, so that their latency
        // overlaps the evaluation of sample i.
        if (next) advance(mask_n, o_n);
        const int i2 = i + 2 < kMsSamples ? i + 2 : kMsSamples - 1;
        mask_n = __float_as_int(d_geo(i2).y);
        o_n = d_off(i2);
        eval(i, cur);
        if (next) put_all(i + 1, nxt);
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        const int x = tid + q * nt;
        if (x < W) {
            const size_t idx = (size_t(z) * MU + y) * W + x;
            float m0 = acc[q][0].value(), m1 = acc[q][1].value(), m2 = acc[q][2].value();
            st_tex(dMS, idx, make_float4(m0, m1, m2, 0.0f), H);
            float pr = phase_rayleigh(tnu[q]);
            float4 prev = ld_tex(S, idx, H);
            st_tex(S, idx,
                   make_float4(m0 / pr + prev.x, m1 / pr + prev.y, m2 / pr + prev.z,
                               0.0f + prev.w),
                   H);
        }
    }
}


// Rows of at most kMsChunkMaxW texels (ms_chunked): the block's
// ms_chunks(P) groups of W threads each integrate a contiguous chunk of the
// row's samples -- per sample, the arithmetic of multiple_scattering_block
// with TPT = 1, on the group's own staged rows and row slots -- and the
// chunk sums are added in chunk order.  The row's table is staged in LDS
// (its reads depend on the group's sample).
template <typename G, bool H>
__device__ __forceinline__ void multiple_scattering_block_chunked(
    const DevParams& P, const int y, const int z, const float4* __restrict__ MR,
    const float4* __restrict__ J, float4* __restrict__ dMS, float4* __restrict__ S) {
    // The row table [kMsRowFloat4s] and the slots' rows at each sample
    // [kMsSamples], then [chunk][2][(NU + 1) * (MUS + 1)] staged rows, then
    // [chunk][W] sums.
    extern __shared__ float4 s_ms[];
    float4* s_desc = s_ms;
    int4* s_held = reinterpret_cast<int4*>(s_ms + kMsRowFloat4s);
    Rgb* s_row = reinterpret_cast<Rgb*>(s_ms + kMsRowFloat4s + kMsSamples);
    const int NU = P.s_nu, MUS = P.s_mus, MU = P.s_mu;
    const int W = NU * MUS;
    const int CH = ms_chunks(P);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int g = tid / W, x = tid - g * W;
    const bool in_chunk = g < CH;  // threads past CH * W only take the barriers
    const int i0 = in_chunk ? g * kMsSamples / CH : 0;
    const int i1 = in_chunk ? (g + 1) * kMsSamples / CH : 0;
    const int steps = (kMsSamples + CH - 1) / CH;  // the longest chunk

    const float4* desc = MR + (size_t(z) * MU + y) * kMsRowFloat4s;
    for (int e = tid; e < kMsRowFloat4s; e += nt) s_desc[e] = desc[e];
    auto d_tw = [&](int i) { return s_desc[i]; };
    auto d_geo = [&](int i) { return s_desc[kMsSamples + i]; };
    auto d_off = [&](int i) { return reinterpret_cast<const int4*>(s_desc + 2 * kMsSamples)[i]; };
    auto d_wt = [&](int i) { return s_desc[3 * kMsSamples + i]; };
    const G r = scattering_texel_coord<G>(P, 0u, unsigned(y), unsigned(z)).r;
    const float rf = float(r);

    float tmu_s = 0.0f, tnu = 0.0f, tslice = 0.0f, tlerp = 0.0f;
    if (in_chunk) {
        TexelCoord<G> c = scattering_texel_coord<G>(P, unsigned(x), unsigned(y), unsigned(z));
        tmu_s = float(c.mu_s);
        tnu = float(c.nu);
        float u_nu = (tnu + 1.0f) / 2.0f;
        float tcx = u_nu * float(NU - 1);
        tslice = floorf(tcx);
        tlerp = tcx - tslice;
    }
    const int spos = x + x / MUS;

    const float b = P.bottom, b2 = P.bottom * P.bottom;
    const float H2 = P.top * P.top - P.bottom * P.bottom;
    const float d_min = P.top - P.bottom;
    const float inv_range = 1.0f / (sqrtf(H2) - d_min);
    const float invA = 1.0f / (-2.0f * P.mu_s_min * P.bottom * inv_range);
    const int SS = MUS + 1;
    const float musm1 = float(MUS - 1), fSS = float(SS);

    typedef typename std::conditional<H, Half4, Rgb>::type Slot;
    Slot slot[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) slot[k] = Slot{};
    constexpr int TB = H ? 8 : 16;
    const __amdgpu_buffer_rsrc_t jb = buffer_rsrc(J, int(size_t(P.s_r) * MU * W * TB));
    const int xoff = x * TB;
    auto ld = [&](int o) {
        if constexpr (H) {
            return __builtin_bit_cast(Half4, __builtin_amdgcn_raw_buffer_load_b64(jb, xoff, o * TB, 0));
        } else {
            const auto v = __builtin_amdgcn_raw_buffer_load_b96(jb, xoff, o * TB, 0);
            return Rgb{__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2])};
        }
    };
    auto advance = [&](int mask, int4 o) {
        if (mask & 1) slot[0] = ld(o.x);
        if (mask & 2) slot[1] = ld(o.y);
        if (mask & 4) slot[2] = ld(o.z);
        if (mask & 8) slot[3] = ld(o.w);
    };
    // Each chunk's two staged rows, laid out as in multiple_scattering_block
    // (pad entry per nu slice, zero slice NU).
    const int WP = (NU + 1) * SS;
    const int PADS = NU + 1 + MUS;
    for (int e = tid; e < CH * PADS; e += nt) {
        const int cg = e / PADS, e0 = e - cg * PADS;
        const int k = e0 < NU ? e0 * SS + MUS : NU * SS + (e0 - NU);
        s_row[(2 * cg) * WP + k] = Rgb{0.0f, 0.0f, 0.0f};
        s_row[(2 * cg + 1) * WP + k] = Rgb{0.0f, 0.0f, 0.0f};
    }
    Rgb* rows = s_row + 2 * (in_chunk ? g : 0) * WP;
    auto put = [&](int i, Rgb* dst) {
        const float4 w = d_wt(i);
        const Rgb t[4] = {to_rgb(slot[0]), to_rgb(slot[1]), to_rgb(slot[2]), to_rgb(slot[3])};
        dst[spos] = Rgb{fmaf(w.w, t[3].x, fmaf(w.z, t[2].x, fmaf(w.y, t[1].x, w.x * t[0].x))),
                        fmaf(w.w, t[3].y, fmaf(w.z, t[2].y, fmaf(w.y, t[1].y, w.x * t[0].y))),
                        fmaf(w.w, t[3].z, fmaf(w.z, t[2].z, fmaf(w.y, t[1].z, w.x * t[0].z)))};
    };
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    auto eval = [&](int i, const Rgb* __restrict__ cur) {
        const float4 tw = d_tw(i);
        const float inv_r_i = d_geo(i).x, d_i = tw.w;
        float mu_s_i = fmaf(rf, tmu_s, d_i * tnu) * inv_r_i;
        mu_s_i = __builtin_amdgcn_fmed3f(mu_s_i, -1.0f, 1.0f);
        float disc = fmaf(b2, mu_s_i * mu_s_i, H2);
        float d = fmaxf(fmaf(-b, mu_s_i, __builtin_amdgcn_sqrtf(disc)), 0.0f);
        float a = (d - d_min) * inv_range;
        float xm = __builtin_amdgcn_fmed3f(
            fmaf(-a, invA, 1.0f) * __builtin_amdgcn_rcpf(1.0f + a), 0.0f, 1.0f);
        float xt = fmaf(xm, musm1, tslice * fSS);
        float fl = floorf(xt);
        float w = xt - fl;
        int k0 = int(fl);
        const Rgb p0 = cur[k0], p1 = cur[k0 + 1], p2 = cur[k0 + SS], p3 = cur[k0 + SS + 1];
        float l = tlerp;
        float v0 = fmaf(w, p1.x - p0.x, p0.x), v1 = fmaf(w, p1.y - p0.y, p0.y),
              v2 = fmaf(w, p1.z - p0.z, p0.z);
        float u0 = fmaf(w, p3.x - p2.x, p2.x), u1 = fmaf(w, p3.y - p2.y, p2.y),
              u2 = fmaf(w, p3.z - p2.z, p2.z);
        a0 = a0 + fmaf(l, u0 - v0, v0) * tw.x;
        a1 = a1 + fmaf(l, u1 - v1, v1) * tw.y;
        a2 = a2 + fmaf(l, u2 - v2, v2) * tw.z;
    };
    __syncthreads();  // the table and the rows' pads
    // The rows the four slots hold at each sample: a slot that a sample uses
    // without loading it holds the row it was last loaded with (the table's
    // offset is 0 for slots a sample does not load).  Lane j of the first
    // wavefront takes sample j: per slot, the last sample <= j that loads
    // it, by a prefix maximum over the lanes.
    if (tid < 64) {
        const int j = tid;
        const int m = j < kMsSamples ? __float_as_int(d_geo(j).y) : 0;
        int last[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            int v = (m >> t) & 1 ? j : -1;
#pragma unroll
            for (int k = 1; k < 64; k <<= 1) {
                const int u = __shfl_up(v, k, 64);
                if (j >= k) v = max(v, u);
            }
            last[t] = v;
        }
        if (j < kMsSamples)
            s_held[j] = make_int4(last[0] >= 0 ? d_off(last[0]).x : 0,
                                  last[1] >= 0 ? d_off(last[1]).y : 0,
                                  last[2] >= 0 ? d_off(last[2]).z : 0,
                                  last[3] >= 0 ? d_off(last[3]).w : 0);
    }
    __syncthreads();
    if (in_chunk) {
        advance(15, s_held[i0]);
        put(i0, rows);
    }
    __syncthreads();
    for (int s = 0; s < steps; ++s) {
        const int i = i0 + s;
        const bool has = in_chunk && i < i1;
        const bool next = in_chunk && i + 1 < i1;
        if (next) advance(__float_as_int(d_geo(i + 1).y), d_off(i + 1));
        if (has) eval(i, rows + (s & 1) * WP);
        if (next) put(i + 1, rows + ((s + 1) & 1) * WP);
        __syncthreads();
    }
    Rgb* part = s_row + 2 * CH * WP;
    if (in_chunk) part[g * W + x] = Rgb{a0, a1, a2};
    __syncthreads();
    if (g == 0) {
        float m0 = part[x].x, m1 = part[x].y, m2 = part[x].z;
        for (int c = 1; c < CH; ++c) {
            const Rgb q = part[c * W + x];
            m0 = m0 + q.x;
            m1 = m1 + q.y;
            m2 = m2 + q.z;
        }
        const size_t idx = (size_t(z) * MU + y) * W + x;
        st_tex(dMS, idx, make_float4(m0, m1, m2, 0.0f), H);
        float pr = phase_rayleigh(tnu);
        float4 prev = ld_tex(S, idx, H);
        st_tex(S, idx,
               make_float4(m0 / pr + prev.x, m1 / pr + prev.y, m2 / pr + prev.z, 0.0f + prev.w),
               H);
    }
}

// ================================================== kernel entry points
// Plain kernels: one atmosphere, block coordinates from blockIdx, the
// parameters as a kernel argument.  Batched kernels (kb_*): one launch over
// many atmospheres; each block reads a 32-bit tile word (atmosphere slot in
// the top 8 bits, block index within that atmosphere's grid in the low 24)
// and the atmosphere's parameters and LUT pointers from HBM arrays.
__device__ __forceinline__ void decode_tile(uint32_t t, int& atm, int& local) {
    atm = int(t >> 24);
    local = int(t & 0xffffffu);
}
// Batched kernels read the LUT addresses from memory.  Read through the
// constant address space (the arrays are written by the host before the
// launch and only read here), the loaded addresses are known to point to
// global memory, so the compiler emits global rather than flat accesses
// (flat ones also count against the LDS wait counter, serialising LDS and
// memory waits).
typedef FB_CONSTANT_AS const LutPtrs ConstLutPtrs;
__device__ __forceinline__ LutPtrs lut_ptrs(const LutPtrs* L, int atm) {
    return ((ConstLutPtrs*)L)[atm];
}
template <typename G, bool COMP, int LANES>
__global__ __launch_bounds__(256) void k_transmittance(DevParams P, float4* __restrict__ T,
                                                       float4* __restrict__ OD) {
    if constexpr (LANES == kOrderedLanes)
        transmittance_block_ordered<G>(P, blockIdx.x, T);
    else
        transmittance_block<G, COMP, LANES>(P, blockIdx.x, 0, 0, T, OD);
}
template <typename G, bool COMP, int LANES>
__global__ __launch_bounds__(256) void kb_transmittance(const DevParams* __restrict__ Ps,
                                                        const LutPtrs* __restrict__ L,
                                                        const uint32_t* __restrict__ tiles, unsigned ntiles) {
    // Persistent when the grid is smaller than the tile list: the block
    // takes tiles blockIdx.x, +gridDim.x, ...; the barrier keeps one
    // tile's LDS reads apart from the next tile's writes.
    for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (t != blockIdx.x) __syncthreads();
        int atm, local;
        decode_tile(tiles[t], atm, local);
        const LutPtrs l = lut_ptrs(L, atm);
        if constexpr (LANES == kOrderedLanes)
            transmittance_block_ordered<G>(Ps[atm], local, l.T);
        else
            transmittance_block<G, COMP, LANES>(Ps[atm], local, 0, 0, l.T, l.OD);
    }
}

template <typename G>
__global__ __launch_bounds__(256) void k_direct_irradiance(DevParams P, const float4* __restrict__ T,
                                                           float4* __restrict__ dE) {
    direct_irradiance_block<G>(P, blockIdx.x, blockIdx.y, 0, T, dE);
}
template <typename G>
__global__ __launch_bounds__(256) void kb_direct_irradiance(const DevParams* __restrict__ Ps,
                                                            const LutPtrs* __restrict__ L,
                                                            const uint32_t* __restrict__ tiles, unsigned ntiles) {
    // Persistent when the grid is smaller than the tile list: the block
    // takes tiles blockIdx.x, +gridDim.x, ...; the barrier keeps one
    // tile's LDS reads apart from the next tile's writes.
    for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (t != blockIdx.x) __syncthreads();
        int atm, local;
        decode_tile(tiles[t], atm, local);
        const DevParams& P = Ps[atm];
        const int gx = (P.i_mus + 63) / 64;
        const LutPtrs l = lut_ptrs(L, atm);
        direct_irradiance_block<G>(P, local % gx, local / gx, 0, l.T, l.dE);
    }
}

// Workgroups go to the 8 XCDs round-robin by linear id, and each XCD has
// its own L2.  With `remap` set, XCD k instead runs the k-th contiguous
// share of the logical blocks, so blocks of neighbouring rows / tiles, which
// read overlapping LUT rows, share an L2.  A bijection on [0, N).
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_remap(unsigned L, unsigned N) {
    const unsigned xcd = L % kXcds, local = L / kXcds;
    const unsigned per = N / kXcds, rem = N % kXcds;
    return xcd * per + min(xcd, rem) + local;
}
__device__ __forceinline__ unsigned logical_block(bool remap) {
    const unsigned L = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    return remap ? xcd_remap(L, gridDim.x * gridDim.y * gridDim.z) : L;
}

template <typename G, bool COMP, bool H, bool STAGED>
__global__ __launch_bounds__(256) void k_single_scattering(DevParams P, const float4* __restrict__ T,
                                                           const float4* __restrict__ SR,
                                                           float4* __restrict__ dR,
                                                           float4* __restrict__ dM,
                                                           float4* __restrict__ S,
                                                           unsigned remap) {
    const unsigned b = logical_block(remap);
    single_scattering_block<G, COMP, H, STAGED>(P, b % P.s_mu, b / P.s_mu, 0, T, SR, dR, dM, S);
}
template <typename G, bool COMP, bool H, bool STAGED>
__global__ __launch_bounds__(256) void kb_single_scattering(const DevParams* __restrict__ Ps,
                                                            const LutPtrs* __restrict__ L,
                                                            const uint32_t* __restrict__ tiles, unsigned ntiles) {
    // Persistent when the grid is smaller than the tile list: the block
    // takes tiles blockIdx.x, +gridDim.x, ...; the barrier keeps one
    // tile's LDS reads apart from the next tile's writes.
    for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (t != blockIdx.x) __syncthreads();
        int atm, local;
        decode_tile(tiles[t], atm, local);
        const DevParams& P = Ps[atm];
        const LutPtrs l = lut_ptrs(L, atm);
        single_scattering_block<G, COMP, H, STAGED>(P, local % P.s_mu, local / P.s_mu, 0, l.T,
                                                    l.SR, l.dR, l.dM, l.S);
    }
}

template <typename G, bool H>
__global__ __launch_bounds__(kSsChunkThreads) void k_single_scattering_chunked(
    DevParams P, const float4* __restrict__ T, const float4* __restrict__ SR,
    float4* __restrict__ dR, float4* __restrict__ dM, float4* __restrict__ S, unsigned remap) {
    const unsigned b = logical_block(remap);
    single_scattering_block_chunked<G, H>(P, b % P.s_mu, b / P.s_mu, T, SR, dR, dM, S);
}
template <typename G, bool H>
__global__ __launch_bounds__(kSsChunkThreads) void kb_single_scattering_chunked(
    const DevParams* __restrict__ Ps, const LutPtrs* __restrict__ L,
    const uint32_t* __restrict__ tiles, unsigned ntiles) {
    for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (t != blockIdx.x) __syncthreads();
        int atm, local;
        decode_tile(tiles[t], atm, local);
        const DevParams& P = Ps[atm];
        const LutPtrs l = lut_ptrs(L, atm);
        single_scattering_block_chunked<G, H>(P, local % P.s_mu, local / P.s_mu, l.T, l.SR, l.dR,
                                              l.dM, l.S);
    }
}

template <typename G, bool SINGLE, bool H, int TPX = 1>
__global__ __launch_bounds__(256, den_occupancy(TPX, SINGLE)) void k_scattering_density(
    DevParams P, const float4* __restrict__ DG, const float4* __restrict__ dR,
    const float4* __restrict__ dM, const float4* __restrict__ dMS, const float4* __restrict__ dE,
    float4* __restrict__ J, unsigned remap) {
    const unsigned b = logical_block(remap), gx = gridDim.x;
    scattering_density_block<G, SINGLE, H, TPX>(P, b % gx, (b / gx) % P.s_mus,
                                                b / (gx * P.s_mus), DG, dR, dM, dMS, dE, J);
}

// Per-set geometry: blocks [0, R + MUS) fill the density geometry table,
// the next MU * R blocks the single- and multiple-scattering row tables.
template <typename G>
__device__ __forceinline__ void geometry_block(const DevParams& P, const int b,
                                               const float4* __restrict__ T,
                                               const float4* __restrict__ OD,
                                               float4* __restrict__ DG, float4* __restrict__ SR,
                                               float4* __restrict__ MR) {
    const int nd = P.s_r + P.s_mus;
    if (b < nd) {
        density_geometry_block<G>(P, b, T, OD, DG);
    } else {
        const int rb = b - nd;
        row_geometry_block<G>(P, rb % P.s_mu, rb / P.s_mu, T, OD, SR, MR);
    }
}
template <typename G>
__global__ __launch_bounds__(64) void k_geometry(DevParams P, const float4* __restrict__ T,
                                                 const float4* __restrict__ OD,
                                                 float4* __restrict__ DG, float4* __restrict__ SR,
                                                 float4* __restrict__ MR) {
    geometry_block<G>(P, blockIdx.x, T, OD, DG, SR, MR);
}
template <typename G>
__global__ __launch_bounds__(64) void kb_geometry(const DevParams* __restrict__ Ps,
                                                  const LutPtrs* __restrict__ L,
                                                  const uint32_t* __restrict__ tiles,
                                                  unsigned ntiles) {
    for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
        int atm, local;
        decode_tile(tiles[t], atm, local);
        const LutPtrs l = lut_ptrs(L, atm);
        geometry_block<G>(Ps[atm], local, l.T, l.OD, l.DG, l.SR, l.MR);
    }
}
template <typename G, bool SINGLE, bool H, int TPX = 1>
__global__ __launch_bounds__(256, den_occupancy(TPX, SINGLE)) void kb_scattering_density(
    const DevParams* __restrict__ Ps, const LutPtrs* __restrict__ L,
    const uint32_t* __restrict__ tiles, unsigned ntiles) {
    // Persistent when the grid is smaller than the tile list: the block
    // takes tiles blockIdx.x, +gridDim.x, ...; the barrier keeps one
    // tile's LDS reads apart from the next tile's writes.
    for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (t != blockIdx.x) __syncthreads();
        int atm, local;
        decode_tile(tiles[t], atm, local);
        const DevParams& P = Ps[atm];
        const LutPtrs l = lut_ptrs(L, atm);
        const int per = int(blockDim.x) / P.den_split;  // texels (pairs at TPX 2) per block
        const int gx = (P.s_mu * P.s_nu / TPX + per - 1) / per;
        scattering_density_block<G, SINGLE, H, TPX>(P, local % gx, (local / gx) % P.s_mus,
                                                    local / (gx * P.s_mus), l.DG, l.dR, l.dM,
                                                    l.dMS, l.dE, l.J);
    }
}

template <typename G, bool SINGLE, bool H>
__global__ __launch_bounds__(64 * kIndWaves) void k_indirect_irradiance(
    DevParams P, const float4* __restrict__ dR, const float4* __restrict__ dM,
    const float4* __restrict__ dMS, float4* __restrict__ dE, float4* __restrict__ E) {
    indirect_irradiance_block<G, SINGLE, H>(P, blockIdx.x, 0, 0, dR, dM, dMS, dE, E);
}
template <typename G, bool SINGLE, bool H>
__global__ __launch_bounds__(64 * kIndWaves) void kb_indirect_irradiance(
    const DevParams* __restrict__ Ps, const LutPtrs* __restrict__ L,
    const uint32_t* __restrict__ tiles, unsigned ntiles) {
    // Persistent when the grid is smaller than the tile list: the block
    // takes tiles blockIdx.x, +gridDim.x, ...; the barrier keeps one
    // tile's LDS reads apart from the next tile's writes.
    for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (t != blockIdx.x) __syncthreads();
        int atm, local;
        decode_tile(tiles[t], atm, local);
        const LutPtrs l = lut_ptrs(L, atm);
        indirect_irradiance_block<G, SINGLE, H>(Ps[atm], local, 0, 0, l.dR, l.dM, l.dMS, l.dE,
                                                l.E);
    }
}

template <typename G, bool COMP, int TPT, bool H>
__global__ __launch_bounds__(kMsMaxThreads) void k_multiple_scattering(DevParams P,
                                                             const float4* __restrict__ MR,
                                                             const float4* __restrict__ J,
                                                             float4* __restrict__ dMS,
                                                             float4* __restrict__ S,
                                                             unsigned remap) {
    const unsigned b = logical_block(remap);
    multiple_scattering_block<G, COMP, TPT, H>(P, b % P.s_mu, b / P.s_mu, 0, MR, J, dMS, S);
}
template <typename G, bool COMP, int TPT, bool H>
__global__ __launch_bounds__(kMsMaxThreads) void kb_multiple_scattering(const DevParams* __restrict__ Ps,
                                                              const LutPtrs* __restrict__ L,
                                                              const uint32_t* __restrict__ tiles, unsigned ntiles) {
    // Persistent when the grid is smaller than the tile list: the block
    // takes tiles blockIdx.x, +gridDim.x, ...; the barrier keeps one
    // tile's LDS reads apart from the next tile's writes.
    for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (t != blockIdx.x) __syncthreads();
        int atm, local;
        decode_tile(tiles[t], atm, local);
        const DevParams& P = Ps[atm];
        const LutPtrs l = lut_ptrs(L, atm);
        multiple_scattering_block<G, COMP, TPT, H>(P, local % P.s_mu, local / P.s_mu, 0, l.MR,
                                                   l.J, l.dMS, l.S);
    }
}

template <typename G, bool H>
__global__ __launch_bounds__(kMsChunkMaxThreads) void k_multiple_scattering_chunked(
    DevParams P, const float4* __restrict__ MR, const float4* __restrict__ J,
    float4* __restrict__ dMS, float4* __restrict__ S, unsigned remap) {
    const unsigned b = logical_block(remap);
    multiple_scattering_block_chunked<G, H>(P, b % P.s_mu, b / P.s_mu, MR, J, dMS, S);
}
template <typename G, bool H>
__global__ __launch_bounds__(kMsChunkMaxThreads) void kb_multiple_scattering_chunked(
    const DevParams* __restrict__ Ps, const LutPtrs* __restrict__ L,
    const uint32_t* __restrict__ tiles, unsigned ntiles) {
    for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (t != blockIdx.x) __syncthreads();
        int atm, local;
        decode_tile(tiles[t], atm, local);
        const DevParams& P = Ps[atm];
        const LutPtrs l = lut_ptrs(L, atm);
        multiple_scattering_block_chunked<G, H>(P, local % P.s_mu, local / P.s_mu, l.MR, l.J,
                                                l.dMS, l.S);
    }
}

// ================================================================== clear
__global__ void k_fill_zero(float4* __restrict__ p, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
// Clear every atmosphere's irradiance LUT (src/precompute.rs:1802-1816).
__global__ void kb_clear_irradiance(const DevParams* __restrict__ Ps,
                                    const LutPtrs* __restrict__ L,
                                    const uint32_t* __restrict__ tiles, unsigned ntiles) {
    // Persistent when the grid is smaller than the tile list: the block
    // takes tiles blockIdx.x, +gridDim.x, ...; the barrier keeps one
    // tile's LDS reads apart from the next tile's writes.
    for (unsigned t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (t != blockIdx.x) __syncthreads();
        int atm, local;
        decode_tile(tiles[t], atm, local);
        const size_t n = size_t(Ps[atm].i_mus) * Ps[atm].i_r;
        const size_t i = size_t(local) * blockDim.x + threadIdx.x;
        if (i < n) lut_ptrs(L, atm).E[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// ========================================================== launch wrappers
static inline unsigned cdiv(size_t a, size_t b) { return unsigned((a + b - 1) / b); }

hipError_t launch_transmittance(const DevParams& P, float4* T, float4* OD, hipStream_t s) {
    size_t texels = size_t(P.t_mu) * P.t_r;
    if (P.flags & 1)
        k_transmittance<double, true, 16><<<cdiv(texels * 16, 256), 256, 0, s>>>(P, T, OD);
    else
        k_transmittance<float, false, kOrderedLanes>
            <<<cdiv(texels * kOrderedLanes, 256), 256, 0, s>>>(P, T, nullptr);
    return hipGetLastError();
}

hipError_t launch_direct_irradiance(const DevParams& P, const float4* T, float4* dE,
                                    hipStream_t s) {
    dim3 grid(cdiv(P.i_mus, 64), P.i_r);
    if (P.flags & 1)
        k_direct_irradiance<double><<<grid, 64, 0, s>>>(P, T, dE);
    else
        k_direct_irradiance<float><<<grid, 64, 0, s>>>(P, T, dE);
    return hipGetLastError();
}

hipError_t launch_geometry(const DevParams& P, const float4* T, const float4* OD, float4* DG,
                           float4* SR, float4* MR, hipStream_t s) {
    const unsigned blocks = unsigned(P.s_r + P.s_mus + P.s_mu * P.s_r);
    if (P.flags & 1)
        k_geometry<double><<<blocks, 64, 0, s>>>(P, T, OD, DG, SR, MR);
    else
        k_geometry<float><<<blocks, 64, 0, s>>>(P, T, OD, DG, SR, MR);
    return hipGetLastError();
}

template <bool H>
static hipError_t launch_single_scattering_h(const DevParams& P, const float4* T,
                                             const float4* SR, float4* dR, float4* dM, float4* S,
                                             const Tuning& tu, hipStream_t s) {
    const unsigned threads = tuned_threads(P, kPassSingle, 2, tu);
    const unsigned remap = (tu.remap >> kPassSingle) & 1u;
    dim3 grid(P.s_mu, P.s_r);
    if (ss_chunked(P)) {
        const size_t lds = ss_chunked_lds_bytes(P);
        if (P.flags & 1)
            k_single_scattering_chunked<double, H>
                <<<grid, kSsChunkThreads, lds, s>>>(P, T, SR, dR, dM, S, remap);
        else
            k_single_scattering_chunked<float, H>
                <<<grid, kSsChunkThreads, lds, s>>>(P, T, SR, dR, dM, S, remap);
    } else if (ss_staged(P)) {
        const size_t lds = ss_staged_lds_bytes(P.t_mu);
        if (P.flags & 1)
            k_single_scattering<double, false, H, true>
                <<<grid, threads, lds, s>>>(P, T, SR, dR, dM, S, remap);
        else
            k_single_scattering<float, false, H, true>
                <<<grid, threads, lds, s>>>(P, T, SR, dR, dM, S, remap);
    } else if (P.flags & 1) {
        k_single_scattering<double, false, H, false><<<grid, threads, 0, s>>>(P, T, SR, dR, dM, S, remap);
    } else {
        k_single_scattering<float, false, H, false><<<grid, threads, 0, s>>>(P, T, SR, dR, dM, S, remap);
    }
    return hipGetLastError();
}
hipError_t launch_single_scattering(const DevParams& P, const float4* T, const float4* SR,
                                    float4* dR, float4* dM, float4* S, const Tuning& tu,
                                    hipStream_t s) {
    return (P.flags & kFlagF16Storage)
               ? launch_single_scattering_h<true>(P, T, SR, dR, dM, S, tu, s)
               : launch_single_scattering_h<false>(P, T, SR, dR, dM, S, tu, s);
}

template <bool H>
static hipError_t launch_scattering_density_h(const DevParams& P, int order, const float4* DG,
                                              const float4* dR, const float4* dM,
                                              const float4* dMS, const float4* dE, float4* J,
                                              const Tuning& tu, hipStream_t s) {
    const unsigned threads = tuned_threads(P, kPassDensity, order, tu);
    const unsigned remap = (tu.remap >> kPassDensity) & 1u;
    dim3 grid(cdiv(size_t(P.s_mu) * P.s_nu, threads / unsigned(P.den_split)), P.s_mus, P.s_r);
    // The density of order n reads the scattering of order n-1; order 1 is
    // single Rayleigh + Mie (scattering.h:157-179).
    const bool single = (order - 1) == 1;
    size_t lds = density_lds_bytes(P.s_nu, P.i_mus, single);
    const bool precise = P.flags & 1;
    const int tpx = den_tpx(P);
    if (tpx > 1) {
        grid.x = cdiv(size_t(P.s_mu) * (P.s_nu / tpx), threads);
#define FB_KD(G, SG, X) \
    k_scattering_density<G, SG, H, X><<<grid, threads, lds, s>>>(P, DG, dR, dM, dMS, dE, J, remap)
        if (precise && single) FB_KD(double, true, 2);
        else if (precise) FB_KD(double, false, 2);
        else if (single) FB_KD(float, true, 2);
        else FB_KD(float, false, 2);
#undef FB_KD
        return hipGetLastError();
    }
    if (precise && single)
        k_scattering_density<double, true, H><<<grid, threads, lds, s>>>(P, DG, dR, dM, dMS, dE, J, remap);
    else if (precise)
        k_scattering_density<double, false, H><<<grid, threads, lds, s>>>(P, DG, dR, dM, dMS, dE, J, remap);
    else if (single)
        k_scattering_density<float, true, H><<<grid, threads, lds, s>>>(P, DG, dR, dM, dMS, dE, J, remap);
    else
        k_scattering_density<float, false, H><<<grid, threads, lds, s>>>(P, DG, dR, dM, dMS, dE, J, remap);
    return hipGetLastError();
}
hipError_t launch_scattering_density(const DevParams& P, int order, const float4* DG,
                                     const float4* dR, const float4* dM, const float4* dMS,
                                     const float4* dE, float4* J, const Tuning& tu,
                                     hipStream_t s) {
    return (P.flags & kFlagF16Storage)
               ? launch_scattering_density_h<true>(P, order, DG, dR, dM, dMS, dE, J, tu, s)
               : launch_scattering_density_h<false>(P, order, DG, dR, dM, dMS, dE, J, tu, s);
}

template <bool H>
static hipError_t launch_indirect_irradiance_h(const DevParams& P, int order, const float4* dR,
                                               const float4* dM, const float4* dMS, float4* dE,
                                               float4* E, const Tuning& tu, hipStream_t s) {
    size_t n = size_t(P.i_mus) * P.i_r;
    const bool single = order == 1;
    size_t per_wave = indirect_lds_bytes_per_wave(P.s_nu, single);
    const int waves = int(tuned_threads(P, kPassIndirect, order + 1, tu) / 64);
    size_t lds = per_wave * waves;
    dim3 grid(cdiv(n, waves)), block(64 * waves);
    const bool precise = P.flags & 1;
    if (precise && single)
        k_indirect_irradiance<double, true, H><<<grid, block, lds, s>>>(P, dR, dM, dMS, dE, E);
    else if (precise)
        k_indirect_irradiance<double, false, H><<<grid, block, lds, s>>>(P, dR, dM, dMS, dE, E);
    else if (single)
        k_indirect_irradiance<float, true, H><<<grid, block, lds, s>>>(P, dR, dM, dMS, dE, E);
    else
        k_indirect_irradiance<float, false, H><<<grid, block, lds, s>>>(P, dR, dM, dMS, dE, E);
    return hipGetLastError();
}
hipError_t launch_indirect_irradiance(const DevParams& P, int order, const float4* dR,
                                      const float4* dM, const float4* dMS, float4* dE, float4* E,
                                      const Tuning& tu, hipStream_t s) {
    return (P.flags & kFlagF16Storage)
               ? launch_indirect_irradiance_h<true>(P, order, dR, dM, dMS, dE, E, tu, s)
               : launch_indirect_irradiance_h<false>(P, order, dR, dM, dMS, dE, E, tu, s);
}

template <bool H>
static hipError_t launch_multiple_scattering_h(const DevParams& P, const float4* MR,
                                               const float4* J, float4* dMS, float4* S,
                                               const Tuning& tu, hipStream_t s) {
    const int W = P.s_nu * P.s_mus;
    const int threads = int(tuned_threads(P, kPassMultiple, 2, tu));
    const unsigned remap = (tu.remap >> kPassMultiple) & 1u;
    dim3 grid(P.s_mu, P.s_r);
    if (ms_chunked(P)) {
        const size_t lds = ms_chunked_lds_bytes(P);
        const unsigned ct = ms_chunk_threads(P);
        if (P.flags & 1)
            k_multiple_scattering_chunked<double, H><<<grid, ct, lds, s>>>(P, MR, J, dMS, S,
                                                                          remap);
        else
            k_multiple_scattering_chunked<float, H><<<grid, ct, lds, s>>>(P, MR, J, dMS, S,
                                                                         remap);
        return hipGetLastError();
    }
    if (W > threads * kMsTexelsPerThread) return hipErrorInvalidValue;
    size_t lds = sizeof(Rgb) * kMsStagedRows * size_t(P.s_nu + 1) * (P.s_mus + 1);
    // Texels per thread: 1 up to 256 texels per row, else 2 or 4.  The
    // arithmetic of a texel does not depend on it.
    const int tpt = W <= threads ? 1 : (W <= 2 * threads ? 2 : 4);
    const bool precise = P.flags & 1;
#define FB_MS(G, TPT) \
    k_multiple_scattering<G, false, TPT, H><<<grid, threads, lds, s>>>(P, MR, J, dMS, S, remap)
    if (precise)
        tpt == 1 ? FB_MS(double, 1) : tpt == 2 ? FB_MS(double, 2) : FB_MS(double, 4);
    else
        tpt == 1 ? FB_MS(float, 1) : tpt == 2 ? FB_MS(float, 2) : FB_MS(float, 4);
#undef FB_MS
    return hipGetLastError();
}
hipError_t launch_multiple_scattering(const DevParams& P, const float4* MR, const float4* J,
                                      float4* dMS, float4* S, const Tuning& tu, hipStream_t s) {
    return (P.flags & kFlagF16Storage)
               ? launch_multiple_scattering_h<true>(P, MR, J, dMS, S, tu, s)
               : launch_multiple_scattering_h<false>(P, MR, J, dMS, S, tu, s);
}

// ------------------------------------------------------------ tuning
// Default and tuned block sizes.  `order` is the launch's scattering order
// (density n, the indirect pass of the order-n loop iteration, multiple n).
unsigned default_threads(const DevParams& P, int pass, int order) {
    const unsigned W = unsigned(P.s_nu * P.s_mus);
    const unsigned row = (W + 63) / 64 * 64;
    switch (pass) {
        case kPassSingle:
            return std::min(256u, row);
        case kPassMultiple:  // one texel per thread
            return std::min(unsigned(kMsMaxThreads), row);
        case kPassDensity:
            return 256;
        case kPassIndirect: {
            const size_t per_wave = indirect_lds_bytes_per_wave(P.s_nu, (order - 1) == 1);
            size_t waves = 65536 / per_wave;
            waves = waves < 1 ? 1 : (waves > size_t(kIndWaves) ? size_t(kIndWaves) : waves);
            return unsigned(64 * waves);
        }
    }
    return 256;
}

unsigned tuned_threads(const DevParams& P, int pass, int order, const Tuning& tu) {
    unsigned t = 0;
    switch (pass) {
        case kPassSingle: t = tu.single; break;
        case kPassDensity: t = order == 2 ? tu.density2 : tu.density; break;
        case kPassIndirect: t = 64 * (order == 2 ? tu.indirect1 : tu.indirect); break;
        case kPassMultiple: t = tu.multiple; break;
    }
    return (t != 0 && tuning_valid(P, pass, order, t)) ? t : default_threads(P, pass, order);
}

bool tuning_valid(const DevParams& P, int pass, int order, unsigned threads) {
    if (threads == 0 || threads % 64 != 0) return false;
    const unsigned W = unsigned(P.s_nu * P.s_mus);
    if (pass == kPassMultiple)  // above 256 threads, at most one wavefront more than the row
        return threads <= unsigned(kMsMaxThreads) && (threads <= 256 || threads < W + 64) &&
               W <= threads * unsigned(kMsTexelsPerThread);
    if (threads > 256) return false;
    switch (pass) {
        case kPassSingle:  // the staged kernel holds at most 4 texels per thread
            return !ss_staged(P) || W <= threads * unsigned(kSsTexelsPerThread);
        case kPassDensity:
            return true;
        case kPassIndirect:
            return indirect_lds_bytes_per_wave(P.s_nu, (order - 1) == 1) * (threads / 64) <=
                   65536;
    }
    return false;
}

// ------------------------------------------------------------ batched
PassShape batch_pass_shape(const DevParams& P, int pass, int order, unsigned threads) {
    const size_t ni = size_t(P.i_mus) * P.i_r;
    const bool precise = P.flags & 1;
    switch (pass) {
        case kPassTransmittance: {
            const size_t lanes = precise ? 16 : kOrderedLanes;
            return {cdiv(size_t(P.t_mu) * P.t_r * lanes, 256), 256, 0};
        }
        case kPassDirect:
            return {cdiv(P.i_mus, 64) * unsigned(P.i_r), 64, 0};
        case kPassSingle:
            if (ss_chunked(P))
                return {unsigned(P.s_mu) * unsigned(P.s_r), kSsChunkThreads,
                        ss_chunked_lds_bytes(P)};
            return {unsigned(P.s_mu) * unsigned(P.s_r), threads,
                    ss_staged(P) ? ss_staged_lds_bytes(P.t_mu) : 0};
        case kPassClear:
            return {cdiv(ni, 256), 256, 0};
        case kPassGeometry:
            return {unsigned(P.s_r + P.s_mus + P.s_mu * P.s_r), 64, 0};
        case kPassDensity:
            return {cdiv(size_t(P.s_mu) * P.s_nu / unsigned(den_tpx(P)),
                         threads / unsigned(P.den_split)) *
                        unsigned(P.s_mus) * unsigned(P.s_r),
                    threads, density_lds_bytes(P.s_nu, P.i_mus, (order - 1) == 1)};
        case kPassIndirect: {  // one wave per texel
            const unsigned waves = threads / 64;
            return {cdiv(ni, waves), threads,
                    indirect_lds_bytes_per_wave(P.s_nu, (order - 1) == 1) * waves};
        }
        case kPassMultiple:
            if (ms_chunked(P))
                return {unsigned(P.s_mu) * unsigned(P.s_r), ms_chunk_threads(P),
                        ms_chunked_lds_bytes(P)};
            return {unsigned(P.s_mu) * unsigned(P.s_r), threads,
                    sizeof(Rgb) * kMsStagedRows * size_t(P.s_nu + 1) * (P.s_mus + 1)};
    }
    return {0, 0, 0};
}

template <bool H>
static hipError_t launch_batch_h(int pass, int order, bool precise, const DevParams* Ps,
                                 const LutPtrs* L, const uint32_t* tiles, unsigned ntiles,
                                 unsigned threads, size_t lds, unsigned grid_blocks, unsigned tpt,
                                 hipStream_t s) {
    if (ntiles == 0) return hipSuccess;
    const dim3 grid(grid_blocks && grid_blocks < ntiles ? grid_blocks : ntiles), block(threads);
    const bool single = (order - 1) == 1;
    switch (pass) {
        case kPassTransmittance:
            if (precise)
                kb_transmittance<double, true, 16><<<grid, block, 0, s>>>(Ps, L, tiles, ntiles);
            else
                kb_transmittance<float, false, kOrderedLanes>
                    <<<grid, block, 0, s>>>(Ps, L, tiles, ntiles);
            break;
        case kPassDirect:
            if (precise)
                kb_direct_irradiance<double><<<grid, block, 0, s>>>(Ps, L, tiles, ntiles);
            else
                kb_direct_irradiance<float><<<grid, block, 0, s>>>(Ps, L, tiles, ntiles);
            break;
        case kPassSingle:  // tpt: kSsStaged, kSsDirect or kSsChunked
            if (tpt == kSsChunked) {
                if (precise)
                    kb_single_scattering_chunked<double, H><<<grid, block, lds, s>>>(Ps, L, tiles, ntiles);
                else
                    kb_single_scattering_chunked<float, H><<<grid, block, lds, s>>>(Ps, L, tiles, ntiles);
            } else if (precise && tpt)
                kb_single_scattering<double, false, H, true><<<grid, block, lds, s>>>(Ps, L, tiles, ntiles);
            else if (precise)
                kb_single_scattering<double, false, H, false><<<grid, block, 0, s>>>(Ps, L, tiles, ntiles);
            else if (tpt)
                kb_single_scattering<float, false, H, true><<<grid, block, lds, s>>>(Ps, L, tiles, ntiles);
            else
                kb_single_scattering<float, false, H, false><<<grid, block, 0, s>>>(Ps, L, tiles, ntiles);
            break;
        case kPassClear:
            kb_clear_irradiance<<<grid, block, 0, s>>>(Ps, L, tiles, ntiles);
            break;
        case kPassGeometry:
            if (precise)
                kb_geometry<double><<<grid, block, 0, s>>>(Ps, L, tiles, ntiles);
            else
                kb_geometry<float><<<grid, block, 0, s>>>(Ps, L, tiles, ntiles);
            break;
        case kPassDensity:  // tpt: texels per thread (den_tpx)
#define FB_KBD(G, SG, X) kb_scattering_density<G, SG, H, X><<<grid, block, lds, s>>>(Ps, L, tiles, ntiles)
            if (tpt == 2) {
                if (precise && single) FB_KBD(double, true, 2);
                else if (precise) FB_KBD(double, false, 2);
                else if (single) FB_KBD(float, true, 2);
                else FB_KBD(float, false, 2);
            } else {
                if (precise && single) FB_KBD(double, true, 1);
                else if (precise) FB_KBD(double, false, 1);
                else if (single) FB_KBD(float, true, 1);
                else FB_KBD(float, false, 1);
            }
#undef FB_KBD
            break;
        case kPassIndirect:
            if (precise && single)
                kb_indirect_irradiance<double, true, H><<<grid, block, lds, s>>>(Ps, L, tiles, ntiles);
            else if (precise)
                kb_indirect_irradiance<double, false, H><<<grid, block, lds, s>>>(Ps, L, tiles, ntiles);
            else if (single)
                kb_indirect_irradiance<float, true, H><<<grid, block, lds, s>>>(Ps, L, tiles, ntiles);
            else
                kb_indirect_irradiance<float, false, H><<<grid, block, lds, s>>>(Ps, L, tiles, ntiles);
            break;
        case kPassMultiple:
#define FB_KBMS(G, T) kb_multiple_scattering<G, false, T, H><<<grid, block, lds, s>>>(Ps, L, tiles, ntiles)
            if (tpt == kMsChunkedTpt) {
                if (precise)
                    kb_multiple_scattering_chunked<double, H><<<grid, block, lds, s>>>(Ps, L, tiles, ntiles);
                else
                    kb_multiple_scattering_chunked<float, H><<<grid, block, lds, s>>>(Ps, L, tiles, ntiles);
            } else if (precise)
                tpt == 1 ? FB_KBMS(double, 1) : tpt == 2 ? FB_KBMS(double, 2) : FB_KBMS(double, 4);
            else
                tpt == 1 ? FB_KBMS(float, 1) : tpt == 2 ? FB_KBMS(float, 2) : FB_KBMS(float, 4);
#undef FB_KBMS
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
hipError_t launch_batch(int pass, int order, bool precise, bool f16, const DevParams* Ps,
                        const LutPtrs* L, const uint32_t* tiles, unsigned ntiles,
                        unsigned threads, size_t lds, unsigned grid_blocks, unsigned tpt,
                        hipStream_t s) {
    return f16 ? launch_batch_h<true>(pass, order, precise, Ps, L, tiles, ntiles, threads, lds,
                                      grid_blocks, tpt, s)
               : launch_batch_h<false>(pass, order, precise, Ps, L, tiles, ntiles, threads, lds,
                                       grid_blocks, tpt, s);
}

// Blocks of one batched launch that fit on the device at once (per-CU
// occupancy of the launch's kernel x CUs); 0 if it cannot be determined.
template <bool H>
static unsigned batch_resident_blocks_h(int pass, int order, bool precise, unsigned threads,
                                        size_t lds, unsigned tpt) {
    const bool single = (order - 1) == 1;
    const void* fn = nullptr;
    switch (pass) {
        case kPassTransmittance:
            fn = precise ? (const void*)kb_transmittance<double, true, 16>
                         : (const void*)kb_transmittance<float, false, kOrderedLanes>;
            break;
        case kPassDirect:
            fn = precise ? (const void*)kb_direct_irradiance<double>
                         : (const void*)kb_direct_irradiance<float>;
            break;
        case kPassSingle:
            if (tpt == kSsChunked)
                fn = precise ? (const void*)kb_single_scattering_chunked<double, H>
                             : (const void*)kb_single_scattering_chunked<float, H>;
            else if (tpt)
                fn = precise ? (const void*)kb_single_scattering<double, false, H, true>
                             : (const void*)kb_single_scattering<float, false, H, true>;
            else
                fn = precise ? (const void*)kb_single_scattering<double, false, H, false>
                             : (const void*)kb_single_scattering<float, false, H, false>;
            break;
        case kPassClear:
            fn = (const void*)kb_clear_irradiance;
            break;
        case kPassGeometry:
            fn = precise ? (const void*)kb_geometry<double> : (const void*)kb_geometry<float>;
            break;
        case kPassDensity:
            if (tpt == 2)
                fn = precise ? (single ? (const void*)kb_scattering_density<double, true, H, 2>
                                       : (const void*)kb_scattering_density<double, false, H, 2>)
                             : (single ? (const void*)kb_scattering_density<float, true, H, 2>
                                       : (const void*)kb_scattering_density<float, false, H, 2>);
            else
                fn = precise ? (single ? (const void*)kb_scattering_density<double, true, H, 1>
                                       : (const void*)kb_scattering_density<double, false, H, 1>)
                             : (single ? (const void*)kb_scattering_density<float, true, H, 1>
                                       : (const void*)kb_scattering_density<float, false, H, 1>);
            break;
        case kPassIndirect:
            fn = precise ? (single ? (const void*)kb_indirect_irradiance<double, true, H>
                                   : (const void*)kb_indirect_irradiance<double, false, H>)
                         : (single ? (const void*)kb_indirect_irradiance<float, true, H>
                                   : (const void*)kb_indirect_irradiance<float, false, H>);
            break;
        case kPassMultiple:
            if (tpt == kMsChunkedTpt)
                fn = precise ? (const void*)kb_multiple_scattering_chunked<double, H>
                             : (const void*)kb_multiple_scattering_chunked<float, H>;
            else if (tpt == 1)
                fn = precise ? (const void*)kb_multiple_scattering<double, false, 1, H>
                             : (const void*)kb_multiple_scattering<float, false, 1, H>;
            else if (tpt == 2)
                fn = precise ? (const void*)kb_multiple_scattering<double, false, 2, H>
                             : (const void*)kb_multiple_scattering<float, false, 2, H>;
            else
                fn = precise ? (const void*)kb_multiple_scattering<double, false, 4, H>
                             : (const void*)kb_multiple_scattering<float, false, 4, H>;
            break;
    }
    int per_cu = 0, dev = 0, cus = 0;
    if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, int(threads), lds) !=
                   hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return unsigned(per_cu) * unsigned(cus);
}
unsigned batch_resident_blocks(int pass, int order, bool precise, bool f16, unsigned threads,
                               size_t lds, unsigned tpt) {
    return f16 ? batch_resident_blocks_h<true>(pass, order, precise, threads, lds, tpt)
               : batch_resident_blocks_h<false>(pass, order, precise, threads, lds, tpt);
}

hipError_t launch_fill_zero(float4* p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_fill_zero<<<cdiv(n, 256), 256, 0, s>>>(p, n);
    return hipGetLastError();
}

}  // namespace fbk
