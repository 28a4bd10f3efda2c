// Device helpers shared by the gfx950 nbnxm kernels (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "nbnxm_amd.h"

namespace nbnxm_amd
{
namespace dev
{

constexpr int c_cl = NBNXM_CLUSTER_SIZE;
constexpr int c_ncps = NBNXM_NUM_CLUSTER_PER_SUPERCLUSTER;
constexpr int c_jg = NBNXM_JGROUP_SIZE;
constexpr float c_fixedScale = 4294967296.0f;        // 2^NBNXM_FIXED_FRAC_BITS
constexpr double c_fixedScaleD = 4294967296.0;
constexpr float c_fixedMaxF = 1073741824.0f;         // 2^30: keeps both 32-bit halves defined
constexpr double c_oneSqrtPi = 0.56418958354775628695;

__device__ __forceinline__ double reduceWaveD(double v)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1)
    {
        v += __shfl_xor(v, m);
    }
    return v;
}

// FP32 force -> int64 fixed point (2^-32 units) in two exact 32-bit halves:
// hi = floor(v), lo = (v - floor(v)) * 2^32 in [0, 2^32), truncated.  Six
// VALU operations instead of a generic float-to-int64 conversion; a
// deterministic function of v, which is all that order-independent
// accumulation needs.  |v| is clamped to 2^30 kJ/mol/nm.
__device__ __forceinline__ unsigned long long toFixed(float v)
{
    v = __builtin_amdgcn_fmed3f(v, -c_fixedMaxF, c_fixedMaxF);
    const float hi = __builtin_floorf(v);
    const float lo = (v - hi) * c_fixedScale;
    const unsigned int loBits = static_cast<unsigned int>(lo);
    const unsigned int hiBits = static_cast<unsigned int>(static_cast<int>(hi));
    return (static_cast<unsigned long long>(hiBits) << 32) | loBits;
}

__device__ __forceinline__ unsigned long long toFixedD(double v)
{
    double s = v * c_fixedScaleD;
    s = fmin(fmax(s, -9.0e18), 9.0e18);
    return static_cast<unsigned long long>(__double2ll_rn(s));
}

// Analytical Ewald force correction: the FP32 rational of pmecorrF
// (src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel_utils.cuh:407-443), same operation
// order, hardware reciprocal as the reference's fast-math build.
__device__ __forceinline__ float pmecorrF(float z2)
{
    const float FN6 = -1.7357322914161492954e-8F;
    const float FN5 = 1.4703624142580877519e-6F;
    const float FN4 = -0.000053401640219807709149F;
    const float FN3 = 0.0010054721316683106153F;
    const float FN2 = -0.019278317264888380590F;
    const float FN1 = 0.069670166153766424023F;
    const float FN0 = -0.75225204789749321333F;
    const float FD4 = 0.0011193462567257629232F;
    const float FD3 = 0.014866955030185295499F;
    const float FD2 = 0.11583842382862377919F;
    const float FD1 = 0.50736591960530292870F;
    const float FD0 = 1.0F;

    const float z4 = z2 * z2;
    float polyFD0 = __builtin_fmaf(FD4, z4, FD2);
    float polyFD1 = __builtin_fmaf(FD3, z4, FD1);
    polyFD0 = __builtin_fmaf(polyFD0, z4, FD0);
    polyFD0 = __builtin_fmaf(polyFD1, z2, polyFD0);
    polyFD0 = __builtin_amdgcn_rcpf(polyFD0);

    float polyFN0 = __builtin_fmaf(FN6, z4, FN4);
    float polyFN1 = __builtin_fmaf(FN5, z4, FN3);
    polyFN0 = __builtin_fmaf(polyFN0, z4, FN2);
    polyFN1 = __builtin_fmaf(polyFN1, z4, FN1);
    polyFN0 = __builtin_fmaf(polyFN0, z4, FN0);
    polyFN0 = __builtin_fmaf(polyFN1, z2, polyFN0);
    return polyFN0 * polyFD0;
}

// r^2 in the fixed evaluation order shared with the CPU prune restatement.
__device__ __forceinline__ float norm2Fixed(float dx, float dy, float dz)
{
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

} // namespace dev
} // namespace nbnxm_amd
