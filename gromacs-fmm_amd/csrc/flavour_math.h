// Pair interaction of every kernel flavour, shared by the tuned list walk's
// flavour-class instantiations and the standalone flavour kernel (not part of
// the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "device_common.h"
#include "nbnxm_kernels.h"

namespace nbnxm_amd
{
namespace dev
{

// erfc(z) with relative accuracy for z in [0, 4.5]: erfc(z) = exp(-z^2) R(t),
// t = 1/(1 + 0.4 z), R a degree-6 polynomial fitted to erfc(z) exp(z^2) in FP64
// by relative least squares reweighted towards the minimax fit (max relative
// error 2.4e-7 in FP64, 5.4e-7 evaluated in FP32 with fused multiply-adds).
// Used for pairs without exclusions, where 1 - erf(z) would lose the small
// erfc(z) near the cut-off.  Takes t; the caller forms it with beta folded in.
constexpr float c_erfcxPolyScale = 0.4f;
__device__ __forceinline__ float erfcxPolyT(float t)
{
    float p = -0.216648832f;
    p = __builtin_fmaf(p, t, 0.7811663151f);
    p = __builtin_fmaf(p, t, -0.5001354218f);
    p = __builtin_fmaf(p, t, 0.5890582204f);
    p = __builtin_fmaf(p, t, 0.1001023054f);
    p = __builtin_fmaf(p, t, 0.2481535822f);
    p = __builtin_fmaf(p, t, -0.001696089399f);
    return p;
}

// Flavour constants in the kernel's arithmetic type, derived the way the
// reference kernel derives them from its FP32 parameter block
// (nbnxm_gpu_data_mgmt.cpp:93-159, nbnxm_cuda_kernel.cuh:311-376).
template<typename Real>
struct FlavourConsts
{
    Real rcoulomb_sq, rvdw_sq, epsfac, beta, beta2, beta3, sh_ewald, two_k_rf, c_rf;
    Real cpot_rep, cpot_disp, rvdw_switch;
    Real disp_c2, disp_c3, rep_c2, rep_c3, sw_c3, sw_c4, sw_c5;
    Real lje_coeff2, lje_coeff6_6, sh_lj_ewald;
    Real tableScale;
    // Exclusion-free analytical Ewald: exp(-b^2 r^2) = exp2(r^2 negBeta2Log2e),
    // t = 1/(1 + betaT r), and 2 beta/sqrt(pi).
    Real negBeta2Log2e, betaT, twoBetaSqrtPi;

    __device__ explicit FlavourConsts(const NbFlavourParams& p)
    {
        const Real rc = static_cast<Real>(p.rcoulomb);
        const Real rvdw = static_cast<Real>(p.rvdw);
        rcoulomb_sq = rc * rc;
        rvdw_sq = rvdw * rvdw;
        epsfac = static_cast<Real>(p.epsfac);
        beta = static_cast<Real>(p.beta);
        beta2 = beta * beta;
        beta3 = beta * beta * beta;
        sh_ewald = static_cast<Real>(p.sh_ewald);
        two_k_rf = static_cast<Real>(2.0 * p.k_rf);
        c_rf = static_cast<Real>(p.c_rf);
        cpot_rep = static_cast<Real>(p.cpot_rep);
        cpot_disp = static_cast<Real>(p.cpot_disp);
        rvdw_switch = static_cast<Real>(p.rvdw_switch);
        disp_c2 = static_cast<Real>(p.disp_c2);
        disp_c3 = static_cast<Real>(p.disp_c3);
        rep_c2 = static_cast<Real>(p.rep_c2);
        rep_c3 = static_cast<Real>(p.rep_c3);
        sw_c3 = static_cast<Real>(p.sw_c3);
        sw_c4 = static_cast<Real>(p.sw_c4);
        sw_c5 = static_cast<Real>(p.sw_c5);
        const Real ljc = static_cast<Real>(p.ewaldcoeff_lj);
        lje_coeff2 = ljc * ljc;
        lje_coeff6_6 = lje_coeff2 * lje_coeff2 * lje_coeff2 * (Real(1) / Real(6));
        sh_lj_ewald = static_cast<Real>(p.sh_lj_ewald);
        tableScale = static_cast<Real>(p.ewaldTableScale);
        negBeta2Log2e = -beta2 * static_cast<Real>(1.44269504088896340736);
        betaT = static_cast<Real>(c_erfcxPolyScale) * beta;
        twoBetaSqrtPi = beta * static_cast<Real>(2.0 * c_oneSqrtPi);
    }
};

// What a flavour class fixes at compile time; -1 = read from the parameters.
template<int kFlav>
struct FlavourTraits
{
    static constexpr bool fixed = kFlav >= c_flavRfPotShift;
    static constexpr int elec = kFlav == c_flavRfPotShift ? NBNXM_ELEC_RF : (fixed ? NBNXM_ELEC_EWALD : -1);
    static constexpr int modifier = kFlav == c_flavEwaldForceSwitch
                                            ? NBNXM_VDW_MOD_FORCESWITCH
                                            : (kFlav == c_flavEwaldPotSwitch ? NBNXM_VDW_MOD_POTSWITCH : (fixed ? NBNXM_VDW_MOD_POTSHIFT : -1));
    static constexpr bool table = kFlav == c_flavEwaldTabPotShift;
    static constexpr bool twin = kFlav == c_flavEwaldTwin;
    static constexpr bool ljPme = kFlav == c_flavEwaldLjPme;
};

template<int kFlav>
__device__ __forceinline__ bool flavRf(const NbFlavourParams& f)
{
    if constexpr (FlavourTraits<kFlav>::fixed)
    {
        return FlavourTraits<kFlav>::elec == NBNXM_ELEC_RF;
    }
    else
    {
        return f.eeltype == NBNXM_ELEC_RF;
    }
}

template<int kFlav>
__device__ __forceinline__ bool flavLjPme(const NbFlavourParams& f)
{
    if constexpr (FlavourTraits<kFlav>::fixed)
    {
        return FlavourTraits<kFlav>::ljPme;
    }
    else
    {
        return f.ljPme != 0;
    }
}

template<int kFlav>
__device__ __forceinline__ bool flavModifier(const NbFlavourParams& f, int modifier)
{
    if constexpr (FlavourTraits<kFlav>::fixed)
    {
        return FlavourTraits<kFlav>::modifier == modifier;
    }
    else
    {
        return f.vdwModifier == modifier;
    }
}

// The LJ cut-off check of twin cut-offs and the tabulated correction: fixed by
// their own classes, read from the parameters in run-time flavours.
template<int kFlav>
__device__ __forceinline__ bool flavTwin(const NbFlavourParams& f)
{
    if constexpr (FlavourTraits<kFlav>::fixed)
    {
        return FlavourTraits<kFlav>::twin;
    }
    else
    {
        return f.vdwCutoffCheck != 0;
    }
}

template<int kFlav>
__device__ __forceinline__ bool flavTable(const NbFlavourParams& f)
{
    if constexpr (FlavourTraits<kFlav>::fixed)
    {
        return FlavourTraits<kFlav>::table;
    }
    else
    {
        return f.tabulatedEwald != 0;
    }
}

template<typename Real>
__device__ __forceinline__ Real rsqrtR(Real x)
{
    if constexpr (std::is_same_v<Real, float>)
    {
        return __builtin_amdgcn_rsqf(x);
    }
    else
    {
        return 1.0 / sqrt(x);
    }
}

template<typename Real>
__device__ __forceinline__ Real expR(Real x)
{
    if constexpr (std::is_same_v<Real, float>)
    {
        return __expf(x);
    }
    else
    {
        return exp(x);
    }
}

// One pair, in the order of the reference kernel body (nbnxm_cuda_kernel.cuh:514-665):
// clamp r^2, LJ with the PotShift constants, force switch (unmasked c6/c12 as
// the reference GPU kernels use them), LJ-PME grid part, potential switch, LJ
// cut-off check of twin cut-offs, Coulomb.  Returns F/r; adds energies.
// kFlav fixes the flavour at compile time where its class allows.
template<typename Real, bool computeEnergy, int kFlav = c_flavRuntime>
__device__ __forceinline__ Real pairForce(const NbFlavourParams& flav,
                                          const FlavourConsts<Real>& k,
                                          Real r2,
                                          Real int_bit,
                                          Real qq,
                                          Real c6,
                                          Real c12,
                                          Real c6grid,
                                          Real& eLJ,
                                          Real& eEl,
                                          const float* tableF = nullptr,
                                          bool exclFree = false)
{
    constexpr Real oneSixth = Real(1) / Real(6);
    constexpr Real oneTwelveth = Real(1) / Real(12);
    r2 = r2 > Real(NBNXM_MIN_DISTANCE_SQUARED) ? r2 : Real(NBNXM_MIN_DISTANCE_SQUARED);
    const Real inv_r = rsqrtR(r2);
    const Real inv_r2 = inv_r * inv_r;
    const Real inv_r6_nm = inv_r2 * inv_r2 * inv_r2;
    const Real inv_r6 = inv_r6_nm * int_bit;
    Real F_invr = inv_r6 * (c12 * inv_r6 - c6) * inv_r2;
    Real E_lj_p = int_bit * (c12 * (inv_r6 * inv_r6 + k.cpot_rep) * oneTwelveth - c6 * (inv_r6 + k.cpot_disp) * oneSixth);
    if (flavModifier<kFlav>(flav, NBNXM_VDW_MOD_FORCESWITCH))
    {
        const Real r = r2 * inv_r;
        Real rs = r - k.rvdw_switch;
        rs = rs >= Real(0) ? rs : Real(0);
        F_invr += (-c6 * (k.disp_c2 + k.disp_c3 * rs) + c12 * (k.rep_c2 + k.rep_c3 * rs)) * rs * rs * inv_r;
        if constexpr (computeEnergy)
        {
            E_lj_p += (c6 * (k.disp_c2 / Real(3) + k.disp_c3 / Real(4) * rs) - c12 * (k.rep_c2 / Real(3) + k.rep_c3 / Real(4) * rs))
                      * rs * rs * rs;
        }
    }
    if (flavLjPme<kFlav>(flav))
    {
        const Real cr2 = k.lje_coeff2 * r2;
        const Real expmcr2 = expR(-cr2);
        const Real poly = Real(1) + cr2 + Real(0.5) * cr2 * cr2;
        F_invr += c6grid * (inv_r6_nm - expmcr2 * (inv_r6_nm * poly + k.lje_coeff6_6)) * inv_r2;
        if constexpr (computeEnergy)
        {
            E_lj_p += oneSixth * c6grid * (inv_r6_nm * (Real(1) - expmcr2 * poly) + k.sh_lj_ewald * int_bit);
        }
    }
    if (flavModifier<kFlav>(flav, NBNXM_VDW_MOD_POTSWITCH))
    {
        const Real r = r2 * inv_r;
        Real rs = r - k.rvdw_switch;
        rs = rs >= Real(0) ? rs : Real(0);
        const Real sw = Real(1) + (k.sw_c3 + (k.sw_c4 + k.sw_c5 * rs) * rs) * rs * rs * rs;
        const Real dsw = (Real(3) * k.sw_c3 + (Real(4) * k.sw_c4 + Real(5) * k.sw_c5 * rs) * rs) * rs * rs;
        F_invr = F_invr * sw - inv_r * E_lj_p * dsw;
        E_lj_p *= sw;
    }
    if (flavTwin<kFlav>(flav) && !(r2 < k.rvdw_sq))
    {
        F_invr = Real(0);
        E_lj_p = Real(0);
    }
    if constexpr (computeEnergy)
    {
        eLJ += E_lj_p;
    }
    if (flavRf<kFlav>(flav))
    {
        F_invr += qq * (int_bit * inv_r2 * inv_r - k.two_k_rf);
        if constexpr (computeEnergy)
        {
            eEl += qq * (int_bit * inv_r + Real(0.5) * k.two_k_rf * r2 - k.c_rf);
        }
    }
    else if (flavTable<kFlav>(flav))
    {
        // interpolate_coulomb_force_r (nbnxm_cuda_kernel_utils.cuh:345-385): linear
        // interpolation of the F(r) table, lerp as two fused multiply-adds.
        const Real normalized = k.tableScale * (r2 * inv_r);
        const int index = static_cast<int>(normalized);
        const Real fraction = normalized - static_cast<Real>(index);
        Real corr;
        if constexpr (std::is_same_v<Real, float>)
        {
            // The FP32 table: the caller's LDS copy in the tabulated class
            // (a compile-time choice, so its reads stay LDS reads).
            const float* t = FlavourTraits<kFlav>::table ? tableF : flav.ewaldTable;
            const float d0 = t[index];
            const float d1 = t[index + 1];
            corr = __builtin_fmaf(fraction, d1, __builtin_fmaf(-fraction, d0, d0));
        }
        else
        {
            const double d0 = flav.ewaldTableD[index];
            const double d1 = flav.ewaldTableD[index + 1];
            corr = fma(fraction, d1, fma(-fraction, d0, d0));
        }
        F_invr += qq * (int_bit * inv_r2 - corr) * inv_r;
        if constexpr (computeEnergy)
        {
            if constexpr (std::is_same_v<Real, float>)
            {
                // Both paths give V/qq; one accumulation after them keeps the
                // accumulator in place (no copies at the join).
                float elArg;
                if (exclFree)
                {
                    // erfc(beta r) = exp(-beta^2 r^2) R(t) instead of 1 - erff.
                    const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(k.betaT, r2 * inv_r, 1.0f));
                    const float e = __builtin_amdgcn_exp2f(r2 * k.negBeta2Log2e);
                    elArg = __builtin_fmaf(e * erfcxPolyT(t), inv_r, -k.sh_ewald);
                }
                else
                {
                    elArg = inv_r * (int_bit - erff(r2 * inv_r * k.beta)) - int_bit * k.sh_ewald;
                }
                eEl = __builtin_fmaf(qq, elArg, eEl);
            }
            else
            {
                eEl += qq * (inv_r * (int_bit - erf(r2 * inv_r * k.beta)) - int_bit * k.sh_ewald);
            }
        }
    }
    else if constexpr (std::is_same_v<Real, float>)
    {
        [[maybe_unused]] float elArg = 0.0f;
        if (exclFree)
        {
            // No excluded lane (int_bit = 1, wave-uniform; the F+E kernels of the
            // tuned list walk pass it, where it replaces erff): erfc(beta r) =
            // e R(t) with e = exp(-beta^2 r^2) serves force and energy,
            // F/(qq r) = e (R/r + 2 beta/sqrt(pi)) / r^2, V/qq = e R / r - sh_ewald.
            const float r = r2 * inv_r;
            const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(k.betaT, r, 1.0f));
            const float R = erfcxPolyT(t);
            const float e = __builtin_amdgcn_exp2f(r2 * k.negBeta2Log2e);
            F_invr = __builtin_fmaf(qq * e, __builtin_fmaf(R, inv_r, k.twoBetaSqrtPi) * inv_r2, F_invr);
            if constexpr (computeEnergy)
            {
                elArg = __builtin_fmaf(e * R, inv_r, -k.sh_ewald);
            }
        }
        else
        {
            F_invr += qq * (int_bit * inv_r2 * inv_r + pmecorrF(k.beta2 * r2) * k.beta3);
            if constexpr (computeEnergy)
            {
                elArg = inv_r * (int_bit - erff(r2 * inv_r * k.beta)) - int_bit * k.sh_ewald;
            }
        }
        if constexpr (computeEnergy)
        {
            eEl = __builtin_fmaf(qq, elArg, eEl);
        }
    }
    else
    {
        const Real br = k.beta * r2 * inv_r;
        const Real erfbr = erf(br);
        F_invr += qq * (int_bit * inv_r2 * inv_r + Real(2) * k.beta * Real(c_oneSqrtPi) * exp(-br * br) * inv_r2 - erfbr * inv_r2 * inv_r);
        if constexpr (computeEnergy)
        {
            eEl += qq * (inv_r * (int_bit - erfbr) - int_bit * k.sh_ewald);
        }
    }
    return F_invr;
}

// LJ-PME grid C6 of a type pair (index numTypes * type_i + type_j).
template<typename Real>
__device__ __forceinline__ Real ljPmeGridC6(const NbFlavourParams& flav, int pairIndex)
{
    if constexpr (std::is_same_v<Real, float>)
    {
        return flav.ljPmeGrid[pairIndex];
    }
    else
    {
        return flav.ljPmeGridD[pairIndex];
    }
}

} // namespace dev
} // namespace nbnxm_amd
