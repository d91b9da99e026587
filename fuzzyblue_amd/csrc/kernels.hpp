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

hipError_t upload_tables();

hipError_t launch_transmittance(const DevParams& P, float4* T, hipStream_t s);
hipError_t launch_direct_irradiance(const DevParams& P, const float4* T, float4* dE,
                                    hipStream_t s);
hipError_t launch_single_scattering(const DevParams& P, const float4* T, float4* dR,
                                    float4* dM, float4* S, hipStream_t s);
hipError_t launch_scattering_density(const DevParams& P, int order, const float4* T,
                                     const float4* dR, const float4* dM, const float4* dMS,
                                     const float4* dE, float4* J, hipStream_t s);
hipError_t launch_indirect_irradiance(const DevParams& P, int order, const float4* dR,
                                      const float4* dM, const float4* dMS, float4* dE, float4* E,
                                      hipStream_t s);
hipError_t launch_multiple_scattering(const DevParams& P, const float4* T, const float4* J,
                                      float4* dMS, float4* S, hipStream_t s);
hipError_t launch_fill_zero(float4* p, size_t n, hipStream_t s);

}  // namespace fbk
