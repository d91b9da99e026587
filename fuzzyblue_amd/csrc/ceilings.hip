// ceilings.hip — micro-kernels that measure two gfx950 ceilings the
// precompute kernels are held against (libfb_ceilings.so, loaded by the
// benchmark only; the product library does not contain them):
//
//  * VALU issue rates of single operations in the operand forms the
//    kernels use (VGPR, SGPR or inline-constant sources): one wave64 FP32
//    instruction per two cycles per SIMD is 1.229e12 instructions/s over
//    256 CUs at 2.4 GHz (the published 157.3 TFLOP/s in FMAs), and some
//    forms issue at half or a quarter of that;
//  * the LDS read rate of ds_read_b128 / ds_read_b64 in the access pattern
//    of the scattering-density quadrature: each lane reads a 24-byte
//    (value, delta) entry — one b128 and one b64 — of a 32-byte-strided
//    profile, a wavefront touching a few distinct entries at a time.
//
// Each kernel returns a per-thread checksum so no load or FMA is dead.
#include <hip/hip_runtime.h>

namespace {

constexpr int kUnroll = 16;

__global__ __launch_bounds__(256) void k_fma(float* __restrict__ out, int iters, float a,
                                             float b) {
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = float(threadIdx.x) * 1e-3f + float(k);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
            for (int k = 0; k < 8; ++k) x[k] = fmaf(x[k], a, b);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += x[k];
    out[size_t(blockIdx.x) * blockDim.x + threadIdx.x] = s;
}

// Chains of one VALU operation: OP 0 v_fmac_f32 (acc += a * b, all
// operands in VGPRs), 1 v_add_f32, 2 v_mul_f32, 3 v_fma_f32 with three
// VGPR operands (the VOP3 form: the result replaces a multiplicand),
// 4 v_med3_f32, 5 v_lshl_add_u32, 6 v_fma_f32 with a scalar (SGPR)
// multiplicand, 7 v_fma_f32 with the inline constant 1.0 as addend, 8
// v_mul_f32 with a scalar operand, 9 v_max_f32, 10 v_fract_f32, 11
// v_lshl_add_u32 with VGPR operands only, 12 v_mul_f32 with a scalar operand
// alternating with v_mul_f32 on VGPRs, 13 v_rsq_f32, 14 v_cvt_i32_f32.
template <int OP>
__global__ __launch_bounds__(256) void k_valu(float* __restrict__ out, int iters, int shift,
                                              float sa) {
    float x[8], a[8], c[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        x[k] = float(threadIdx.x) * 1e-3f + float(k);
        a[k] = 0.999f - float(k) * 1e-4f - float(threadIdx.x) * 1e-7f;
        c[k] = 1e-3f + float(k) * 1e-5f + float(threadIdx.x) * 1e-9f;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (OP == 0) x[k] = fmaf(a[k], c[k], x[k]);
                if (OP == 1) x[k] = x[k] + a[k];
                if (OP == 2) x[k] = x[k] * a[k];
                if (OP == 3) x[k] = fmaf(x[k], a[k], c[k]);
                if (OP == 4) x[k] = __builtin_amdgcn_fmed3f(x[k], a[k], c[k]);
                if (OP == 5)
                    x[k] = __int_as_float((__float_as_int(x[k]) << shift) + __float_as_int(a[k]));
                if (OP == 6) x[k] = fmaf(x[k], sa, c[k]);
                if (OP == 7) x[k] = fmaf(x[k], a[k], 1.0f);
                if (OP == 8) x[k] = x[k] * sa;
                if (OP == 9) x[k] = fmaxf(x[k] * 0.5f, a[k]);
                if (OP == 10) x[k] = __builtin_amdgcn_fractf(x[k] + a[k]);
                if (OP == 11)
                    x[k] = __int_as_float((__float_as_int(x[k]) << (__float_as_int(c[k]) & 7)) +
                                          __float_as_int(a[k]));
                if (OP == 12) x[k] = (k & 1) ? x[k] * sa : x[k] * a[k];
                if (OP == 13) x[k] = __builtin_amdgcn_rsqf(x[k]);
                if (OP == 14) x[k] = float(int(x[k] * a[k]));
            }
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += x[k];
    out[size_t(blockIdx.x) * blockDim.x + threadIdx.x] = s;
}

// Split layout: the 16-byte (value, delta.r) part of entry e at float4
// e, its 8-byte (delta.g, delta.b) part in a second array at float2 e.
__global__ __launch_bounds__(256) void k_lds_split(float* __restrict__ out, int iters,
                                                   int distinct) {
    __shared__ float4 a4[64];
    __shared__ float2 b2[64];
    for (int e = threadIdx.x; e < 64; e += blockDim.x) {
        a4[e] = make_float4(float(e), 1.0f, 2.0f, 3.0f);
        b2[e] = make_float2(float(e), 1.0f);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int base = lane % distinct;
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float2 acc2 = make_float2(0.0f, 0.0f);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int e = (base + i * kUnroll + u) & 63;
            const float4 v = a4[e];
            const float2 w = b2[e];
            acc.x += v.x;
            acc.y += v.y;
            acc.z += v.z;
            acc.w += v.w;
            acc2.x += w.x;
            acc2.y += w.y;
        }
    }
    out[size_t(blockIdx.x) * blockDim.x + threadIdx.x] =
        acc.x + acc.y + acc.z + acc.w + acc2.x + acc2.y;
}

// mode 0: b128 + b64 of one entry per read (the density lookup);
// mode 1: b128 only; mode 2: b64 only.
// distinct: entries touched by one wavefront per read (1..64): lane l reads
// entry (l % distinct + i) & 63 at step i.
template <int MODE>
__global__ __launch_bounds__(256) void k_lds(float* __restrict__ out, int iters, int distinct) {
    __shared__ float4 lds[128];  // 64 entries x 32 bytes
    for (int e = threadIdx.x; e < 128; e += blockDim.x)
        lds[e] = make_float4(float(e), 1.0f, 2.0f, 3.0f);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int base = lane % distinct;
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float2 acc2 = make_float2(0.0f, 0.0f);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int e = (base + i * kUnroll + u) & 63;
            if (MODE != 2) {
                const float4 v = lds[2 * e];
                acc.x += v.x;
                acc.y += v.y;
                acc.z += v.z;
                acc.w += v.w;
            }
            if (MODE != 1) {
                const float4* p = &lds[2 * e + 1];
                const float2 w = *reinterpret_cast<const float2*>(p);
                acc2.x += w.x;
                acc2.y += w.y;
            }
        }
    }
    out[size_t(blockIdx.x) * blockDim.x + threadIdx.x] =
        acc.x + acc.y + acc.z + acc.w + acc2.x + acc2.y;
}

}  // namespace

extern "C" {

// Runs `kind` (0: v_fma_f32 with a scalar operand; 1..3: LDS modes 0..2;
// 4..6: v_fmac_f32, v_add_f32, v_mul_f32 chains; 7: split-layout b128 +
// b64 entry reads; 8..19: k_valu operations 3..14) with `blocks` x 256 threads
// and `iters` outer iterations, `reps` times back to back on the default
// stream, and returns the mean milliseconds per launch in *ms.  `out` holds
// blocks * 256 floats.  Returns a hipError_t value (0 on success).
int fb_ceiling_run(int kind, int blocks, int iters, int distinct, int reps, float* out,
                   float* ms) {
    if (blocks <= 0 || iters <= 0 || reps <= 0 || !out || !ms || distinct < 1 ||
        distinct > 64 || kind < 0 || kind > 19)
        return int(hipErrorInvalidValue);
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e != hipSuccess) return int(e);
    auto launch = [&]() {
        switch (kind) {
            case 0: k_fma<<<blocks, 256>>>(out, iters, 0.999f, 1e-3f); break;
            case 1: k_lds<0><<<blocks, 256>>>(out, iters, distinct); break;
            case 2: k_lds<1><<<blocks, 256>>>(out, iters, distinct); break;
            case 3: k_lds<2><<<blocks, 256>>>(out, iters, distinct); break;
            case 4: k_valu<0><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 5: k_valu<1><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 6: k_valu<2><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 8: k_valu<3><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 9: k_valu<4><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 10: k_valu<5><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 11: k_valu<6><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 12: k_valu<7><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 13: k_valu<8><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 14: k_valu<9><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 15: k_valu<10><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 16: k_valu<11><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 17: k_valu<12><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 18: k_valu<13><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 19: k_valu<14><<<blocks, 256>>>(out, iters, 1, 0.999f); break;
            case 7: k_lds_split<<<blocks, 256>>>(out, iters, distinct); break;
        }
    };
    launch();  // warm-up
    e = hipEventRecord(a, nullptr);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        launch();
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(b, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
    if (e == hipSuccess) *ms /= float(reps);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return int(e);
}

// Per launch: VALU operations (lanes x instructions; kinds 0, 4..6) or LDS
// bytes read (kinds 1..3, 7).
double fb_ceiling_work(int kind, int blocks, int iters) {
    const double lanes = double(blocks) * 256.0 * double(iters) * kUnroll;
    switch (kind) {
        case 0: case 4: case 5: case 6: case 8: case 9: case 10: case 11: case 12: case 13:
        case 14: case 15: case 16: case 17: case 18: case 19: return lanes * 8.0;
        case 1: case 7: return lanes * 24.0;
        case 2: return lanes * 16.0;
        default: return lanes * 8.0;
    }
}

}  // extern "C"
