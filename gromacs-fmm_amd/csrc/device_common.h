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

// Lane exchanges of gfx950 that stay in the VALU (no LDS): with both operands
// equal, v_permlane16_swap / v_permlane32_swap leave each lane its own value
// in one result and the value of lane ^ 16 (lane ^ 32) in the other, so the
// sum of the two is one butterfly step.
__device__ __forceinline__ float addXor16(float v)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

__device__ __forceinline__ float addXor32(float v)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// One DPP move of a double (both halves with the same control).
template<int ctrl>
__device__ __forceinline__ double dppD(double v)
{
    const unsigned long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b & 0xFFFFFFFFull), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), ctrl, 0xF, 0xF, false);
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(static_cast<unsigned int>(hi)) << 32)
                                                       | static_cast<unsigned int>(lo)));
}

template<bool k32>
__device__ __forceinline__ double addXorRowsD(double v)
{
    const unsigned long long b = __double_as_longlong(v);
    const unsigned int lo = static_cast<unsigned int>(b & 0xFFFFFFFFull);
    const unsigned int hi = static_cast<unsigned int>(b >> 32);
    const auto l = k32 ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false) : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto h = k32 ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false) : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const double a = __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(h[0]) << 32) | l[0]));
    const double c = __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(h[1]) << 32) | l[1]));
    return a + c;
}

// Sum over the wave, the same bits in every lane: quad xor 1, xor 2, then
// the quads of 8 (row_half_mirror) and the halves of 16 (row_mirror), whose
// lanes already hold equal values, then rows xor 16 and xor 32.
__device__ __forceinline__ double reduceWaveD(double v)
{
    v += dppD<0xB1>(v);  // quad_perm [1,0,3,2]
    v += dppD<0x4E>(v);  // quad_perm [2,3,0,1]
    v += dppD<0x141>(v); // row_half_mirror
    v += dppD<0x140>(v); // row_mirror
    v = addXorRowsD<false>(v);
    return addXorRowsD<true>(v);
}

// A wave-uniform float computed in the vector unit, moved to a scalar register
// (a VOP2 or VOP3 instruction takes one scalar operand), so that it holds no
// vector register through the list walk.
__device__ __forceinline__ float uniformF(float v)
{
    // readfirstlane folds away on a value the compiler knows to be uniform,
    // which would leave it in a vector register; the empty asm hides that.
    // The move itself stays the builtin, so the compiler's hazard checks
    // (wait states after a VALU write of a scalar register) apply to it.
    asm("" : "+v"(v));
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// FP32 force -> int64 fixed point (2^-32 units) in two exact 32-bit halves:
// hi = floor(v), lo = (v - floor(v)) * 2^32 in [0, 2^32), truncated.  Six
// VALU operations instead of a generic float-to-int64 conversion; a
// deterministic function of v, which is all that order-independent
// accumulation needs.  |v| is clamped to 2^30 kJ/mol/nm: bound is c_fixedMaxF,
// which a caller in a loop passes from a scalar register (uniformF, read once
// before the loop), so that the clamp is one med3 whose two bounds are that
// register, once negated (a source modifier), and no vector register holds a
// constant.
__device__ __forceinline__ unsigned long long toFixed(float v, float bound = c_fixedMaxF)
{
    v = __builtin_amdgcn_fmed3f(v, -bound, bound);
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
