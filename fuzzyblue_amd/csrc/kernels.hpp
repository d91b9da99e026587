// kernels.hpp — host-visible launch interface of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fbk {

struct DevParams;

// Quadrature sizes of the reference (scattering_density.comp:34,
// indirect_irradiance.comp:20).
constexpr int kDenTheta = 16;
constexpr int kDenPhi = 32;
constexpr int kIndTheta = 16;
constexpr int kIndPhi = 64;

// Size limits the kernels' LDS layouts accept (checked by the host).
constexpr int kMaxNu = 32;
constexpr int kMaxIrrMuS = 256;
constexpr int kMaxRowTexels = 1024;  // NU * MUS: one scattering row per block
// Trapezoid samples of the single and multiple scattering integrals:
// SAMPLE_COUNT + 1 (single_scattering.comp:42, multiple_scattering.comp:21).
constexpr int kMsSamples = 51;
// Multiple-scattering row table: per (r, mu) row and sample, the view
// transmittance x dx x trapezoid weight (rgb) and d_i; (1/r_i, mask of the
// density-row slots to load); the row offsets loaded into those slots; the
// slots' bilinear weights.
constexpr int kMsRowFloat4s = 4 * kMsSamples;
// Single-scattering row table: per (r, mu) row and sample, the view
// transmittance x trapezoid weight (rgb) and d_i; (r_d, Rayleigh density,
// Mie density, sin of the horizon angle); (cos of the horizon angle,
// d_min, d_max - d_min, T-row weight); (1/r_d, 1/(d_max - d_min),
// top^2 - r_d^2); (T-row offsets as int bits, dx).
constexpr int kSsRowFloat4s = 5 * kMsSamples;

hipError_t upload_tables();

// Device addresses of one atmosphere's LUTs (batched launches read an HBM
// array of these, indexed by the tile's atmosphere slot).
struct LutPtrs {
    float4 *T, *E, *S, *dE, *dR, *dM, *dMS, *J, *MR;
    float4* OD;  // optical depths beside T (precision mode; else null)
    float4* DG;  // density geometry table (dg_* below)
    float4* SR;  // single-scattering row table (kSsRowFloat4s per (r, mu) row)
};

// Passes of a batched schedule (values match fb_pass).
enum BatchPass {
    kPassTransmittance = 0,
    kPassDirect = 1,
    kPassSingle = 2,
    kPassClear = 3,
    kPassDensity = 4,
    kPassIndirect = 5,
    kPassMultiple = 6,
    kPassGeometry = 7
};

// Density geometry table, built once per LUT set after the transmittance
// pass (it depends on r, theta and mu_s only, never on the order): for each
// r index z, 16 ground entries (T_g * albedo / pi rgb, d_g) and 16 row
// entries (mu taps i0, i1 as int bits, weight, 1/|ground normal| or 0 when
// the row misses the ground), then the r taps; after all z, for each mu_s
// index j the x taps of nu slots -1..NU.
constexpr int kDgRow = 2 * 16 + 1;
__host__ __device__ inline size_t dg_row(int z) { return size_t(z) * kDgRow; }
__host__ __device__ inline size_t dg_x(int R, int j, int NU) {
    return size_t(R) * kDgRow + size_t(j) * (NU + 2);
}
__host__ __device__ inline size_t dg_count(int R, int MUS, int NU) { return dg_x(R, MUS, NU); }

// Block sizes an autotuner may choose per LUT-size class; 0 = default.
// None of them changes the arithmetic of a texel (each texel's sums run in
// the same order whatever the block shape), so every choice gives
// bit-identical LUTs.
struct Tuning {
    unsigned single = 0;     // threads per block, single scattering
    unsigned density2 = 0;   // threads per block, density of order 2
    unsigned density = 0;    // threads per block, density of order >= 3
    unsigned indirect1 = 0;  // wavefronts per block, indirect irradiance of order 1
    unsigned indirect = 0;   // wavefronts per block, indirect irradiance of order >= 2
    unsigned multiple = 0;   // threads per block, multiple scattering
    unsigned remap = 0;      // bit p (a BatchPass): XCD-contiguous block order
};
// `order` is the scattering order of the launch: density n, the indirect
// pass of the order-n iteration (n - 1 bounces), multiple n.
unsigned default_threads(const DevParams& P, int pass, int order);
// Single scattering stages the r-interpolated transmittance row of each
// sample in LDS (two rows, double buffered), with at most
// kSsTexelsPerThread texels per thread, when the row of texels is wide
// enough to pay for staging a whole transmittance row per sample — at
// least two texels per three transmittance mu columns (measured: staged
// 1.2x faster at 256 texels per 256 columns, 3.5x at 1024 per 512;
// unstaged 1.5x faster at 64 per 128, 2x at 16 per 256) — and the LUT is
// at most kSsMaxStagedMu wide (64 KiB).  Otherwise each texel reads its
// transmittance taps directly.
constexpr int kSsTexelsPerThread = 4;
// Scattering density with fewer texels than this runs kDenTheta threads per
// texel, one theta row each, summed by a lane reduction: one thread per
// texel would leave most of the device idle, each thread walking all 512
// directions in sequence.
constexpr size_t kDenSplitTexels = 65536;
// Density texels per thread: 2 when the LUT set is not theta-split, NU is
// even and an (r, mu_s) tile holds at least kDenPairMinTexels texels — the
// thread then takes texels (mu, nu) and (mu, nu + NU/2), which share
// nu2 = omega.omega_i, q and the Cornette-Shanks power of every direction
// pair (computed once for both; each texel's own operations and their order
// are unchanged, so the values are the same bits), and a block's profile
// build serves twice the texels.
constexpr int kDenPairMinTexels = 512;
// Blocks per CU the density kernels are compiled for: 8 at one texel per
// thread (64 VGPRs), 5 at two (91-96).  Four texels per thread (3 blocks,
// 138-143 VGPRs) measured 6% slower at double and 30% at default sizes; at
// two, 4 blocks for order 2 (no spills) was 1% slower than 5.
__host__ __device__ constexpr int den_occupancy(int tpx, bool) { return tpx == 1 ? 8 : 5; }
__host__ __device__ inline int den_tpx(const DevParams& P) {
    return (P.den_split == 1 && P.s_nu % 2 == 0 && P.s_mu * P.s_nu >= kDenPairMinTexels) ? 2
                                                                                          : 1;
}
constexpr int kSsMaxStagedMu = 2048;
// Multiple scattering over rows of at most kMsChunkMaxW texels (small LUT
// sets, whose launches are a few hundred one-wavefront blocks, each walking
// 51 barrier-separated samples): a block of ms_chunk_threads() threads splits
// its row's samples into ms_chunks() contiguous chunks, one per group of W
// threads, each integrated with its own staged rows and row slots, and adds
// the chunk sums in chunk order.  Plain and batched builds choose it alike (by the row
// width), so a batch member stays bit-identical to a plain build.  Blocks
// of 256 threads (16 chunks of 3-4 samples at W = 16): one chunk per sample
// in up to 1024 threads measured no faster.  Rows of 64 texels (the
// half-resolution class) stay unchunked: chunked, their plain launches ran
// 1.25x faster, but the 256-atmosphere mix served 0.7% and 256 served
// half-resolution sets 2.5% slower (batched launches of many sets are not
// latency-bound, and chunks add row loads).
constexpr int kMsChunkMaxW = 32;
constexpr unsigned kMsChunkMaxThreads = 256;
__host__ __device__ inline bool ms_chunked(const DevParams& P) { return P.s_nu * P.s_mus <= kMsChunkMaxW; }
__host__ __device__ inline unsigned ms_chunk_threads(const DevParams& P) {
    const unsigned t = (unsigned(kMsSamples * P.s_nu * P.s_mus) + 63u) & ~63u;
    return t < kMsChunkMaxThreads ? t : kMsChunkMaxThreads;
}
__host__ __device__ inline int ms_chunks(const DevParams& P) {
    const int c = int(ms_chunk_threads(P)) / (P.s_nu * P.s_mus);
    return c < kMsSamples ? c : kMsSamples;
}
// The texels-per-thread code of a batched multiple-scattering launch that
// runs the chunked kernel.
constexpr unsigned kMsChunkedTpt = 0;
// The row's table and the rows held by the four slots at each sample
// (float4 each), then staged rows (two per chunk) and the chunks' partial
// sums, 12-byte rgb.
__host__ __device__ inline size_t ms_chunked_lds_bytes(const DevParams& P) {
    const size_t ch = size_t(ms_chunks(P));
    const size_t wp = size_t(P.s_nu + 1) * size_t(P.s_mus + 1);
    return 16 * size_t(kMsRowFloat4s + kMsSamples) +
           12 * (ch * 2 * wp + ch * size_t(P.s_nu * P.s_mus));
}
inline bool ss_staged(const DevParams& P) {
    return P.t_mu <= kSsMaxStagedMu && 2 * P.t_mu <= 3 * P.s_nu * P.s_mus;
}
inline size_t ss_staged_lds_bytes(int t_mu) { return 2 * size_t(t_mu) * 16; }
// Unstaged single scattering over rows of at most kSsChunkMaxW texels (small
// LUT sets): each thread of a one-wavefront block would walk 51 samples,
// each waiting for its transmittance gathers.  A block of kSsChunkThreads
// threads instead splits the samples into ss_chunks() contiguous chunks, one
// per group of W threads, and adds the chunk sums in chunk order (chosen by
// the sizes alike in plain and batched builds).
constexpr int kSsChunkMaxW = 32;
constexpr unsigned kSsChunkThreads = 256;
__host__ __device__ inline int ss_chunks(const DevParams& P) {
    const int c = int(kSsChunkThreads) / (P.s_nu * P.s_mus);
    return c < kMsSamples ? c : kMsSamples;
}
inline bool ss_chunked(const DevParams& P) {
    return !ss_staged(P) && P.s_nu * P.s_mus <= kSsChunkMaxW;
}
// Per chunk and texel, the Rayleigh and Mie partial sums (6 floats).
inline size_t ss_chunked_lds_bytes(const DevParams& P) {
    return 24 * size_t(ss_chunks(P)) * size_t(P.s_nu * P.s_mus);
}
// The kernel code of a batched single-scattering launch: staged, direct
// lookups, or chunked direct lookups.
constexpr unsigned kSsDirect = 0, kSsStaged = 1, kSsChunked = 2;

unsigned tuned_threads(const DevParams& P, int pass, int order, const Tuning& tu);
bool tuning_valid(const DevParams& P, int pass, int order, unsigned threads);

// Launch shape of one atmosphere's share of a batched pass: the number of
// blocks (= tiles), threads per block and dynamic LDS bytes, for a block
// size `threads` (used by the tunable passes).
struct PassShape {
    unsigned blocks;
    unsigned threads;
    size_t lds;
};
PassShape batch_pass_shape(const DevParams& P, int pass, int order, unsigned threads);

// One launch of `pass` over a tile list (atmosphere slot << 24 | block).
// `order` is the scattering order of the launch (density n, indirect n-1's
// caller order n, multiple n); threads / lds are the maxima over the tiles'
// atmospheres.
// With grid_blocks in (0, ntiles) the launch is persistent: grid_blocks
// blocks walk the tile list (0 = one block per tile).
// tpt: multiple scattering's texels per thread (1, 2 or 4; every row of the
// launch must fit threads x tpt texels).
// f16: the set's scattering-type LUTs are binary16 (FB_FLAG_F16_STORAGE).
hipError_t launch_batch(int pass, int order, bool precise, bool f16, const DevParams* d_params,
                        const LutPtrs* d_luts, const uint32_t* d_tiles, unsigned ntiles,
                        unsigned threads, size_t lds, unsigned grid_blocks, unsigned tpt,
                        hipStream_t s);
// Blocks of a batched launch of `pass` resident on the device at once.
unsigned batch_resident_blocks(int pass, int order, bool precise, bool f16, unsigned threads,
                               size_t lds, unsigned tpt);

// OD: optical-depth LUT written beside T (precision mode), or null.
hipError_t launch_transmittance(const DevParams& P, float4* T, float4* OD, hipStream_t s);
hipError_t launch_direct_irradiance(const DevParams& P, const float4* T, float4* dE,
                                    hipStream_t s);
// The per-set geometry tables (kPassGeometry): DG, SR and MR.
hipError_t launch_geometry(const DevParams& P, const float4* T, const float4* OD, float4* DG,
                           float4* SR, float4* MR, hipStream_t s);
hipError_t launch_single_scattering(const DevParams& P, const float4* T, const float4* SR,
                                    float4* dR, float4* dM, float4* S, const Tuning& tu,
                                    hipStream_t s);
hipError_t launch_scattering_density(const DevParams& P, int order, const float4* DG,
                                     const float4* dR, const float4* dM, const float4* dMS,
                                     const float4* dE, float4* J, const Tuning& tu,
                                     hipStream_t s);
// `order` here is the number of bounces summed (the loop's n - 1).
hipError_t launch_indirect_irradiance(const DevParams& P, int order, const float4* dR,
                                      const float4* dM, const float4* dMS, float4* dE, float4* E,
                                      const Tuning& tu, hipStream_t s);
hipError_t launch_multiple_scattering(const DevParams& P, const float4* MR, const float4* J,
                                      float4* dMS, float4* S, const Tuning& tu, hipStream_t s);
hipError_t launch_fill_zero(float4* p, size_t n, hipStream_t s);

}  // namespace fbk
