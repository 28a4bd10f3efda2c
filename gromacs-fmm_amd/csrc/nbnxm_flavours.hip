// gfx950 kernel for every interaction flavour other than the tuned default
// (Ewald analytical + LJ cut-off with potential shift, rvdw = rcoulomb):
// tabulated Ewald, reaction field / plain cut-off Coulomb, the twin cut-off LJ
// check, LJ force switch, LJ potential switch and the LJ-PME grid correction
// with the geometric or Lorentz-Berthelot grid rule.  These are the kernels the reference generates from one
// body with preprocessor flavours (src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel.cuh,
// selected in nbnxm_cuda.cu:171-254); here one body takes the flavour as
// wave-uniform kernel arguments, so a flavour costs scalar branches, not a
// separate code object.
//
// Layout: one wave per pair-list entry, lane l = i-atom (l & 7) x j-atom (l >> 3)
// of each 8x8 cluster pair, i-cluster data in registers for the whole entry,
// j-forces reduced over the 8 i-lanes, i-forces over the 8 j-lanes at the end.
// Real = float is the FP32 kernel (the reference's arithmetic); Real = double
// is the FP64 kernel with exact erf/exp.  Forces go to int64 fixed point or,
// for FP32 with float accumulation, to float atomics.
#include "device_common.h"
#include "flavour_math.h"
#include "nbnxm_kernels.h"

#include <cmath>
#include <type_traits>
#include <vector>

namespace nbnxm_amd
{

namespace
{

using namespace dev;

template<typename Real>
__device__ __forceinline__ Real sumOverTi(Real v)
{
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    return v;
}

template<typename Real>
__device__ __forceinline__ Real sumOverTj(Real v)
{
    v += __shfl_xor(v, 8);
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

template<typename Real, int accumulation>
__device__ __forceinline__ void addForce(const NbKernelArgs& a, int index, Real v)
{
    if constexpr (accumulation == NBNXM_ACCUM_FLOAT)
    {
        atomicAdd(a.fFloat + index, static_cast<float>(v));
    }
    else if constexpr (std::is_same_v<Real, float>)
    {
        atomicAdd(a.fFixed + index, toFixed(v));
    }
    else
    {
        atomicAdd(a.fFixed + index, toFixedD(v));
    }
}

template<typename Real, bool computeEnergy, int accumulation>
__global__ __launch_bounds__(256) void nbnxmFlavourKernel(const NbKernelArgs a, const NbFlavourParams flav, int vdwType, int wavesPerBlock)
{
    const int lane = threadIdx.x & 63;
    const int ti = lane & 7;
    const int tj = lane >> 3;
    const int work = blockIdx.x * wavesPerBlock + (threadIdx.x >> 6);
    if (work >= a.nsci)
    {
        return;
    }
    const FlavourConsts<Real> k(flav);
    const int entry = __builtin_amdgcn_readfirstlane(a.sciOrder ? a.sciOrder[work] : work);
    const nbnxm_sci_t s = a.sci[entry];
    const bool centralShift = (s.shift % NBNXM_NUM_SHIFTS) == NBNXM_CENTRAL_SHIFT;
    const float4 sh = a.shiftVec[s.shift];
    const bool useTable = (vdwType == NBNXM_VDW_CUT_TYPE_TABLE);
    const bool useGeom = (vdwType == NBNXM_VDW_CUT_COMB_GEOM);
    const bool needTypes = useTable || flav.ljPme;

    // i-cluster atoms of this lane, shifted, charges scaled by epsfac.
    Real xi[c_ncps], yi[c_ncps], zi[c_ncps], qi[c_ncps];
    int typei[c_ncps];
    float2 combi[c_ncps];
    for (int c = 0; c < c_ncps; c++)
    {
        const int ai = (s.sci * c_ncps + c) * c_cl + ti;
        const float4 v = a.xq[ai];
        if constexpr (std::is_same_v<Real, float>)
        {
            xi[c] = v.x + sh.x;
            yi[c] = v.y + sh.y;
            zi[c] = v.z + sh.z;
        }
        else
        {
            xi[c] = static_cast<Real>(v.x) + sh.x;
            yi[c] = static_cast<Real>(v.y) + sh.y;
            zi[c] = static_cast<Real>(v.z) + sh.z;
        }
        qi[c] = static_cast<Real>(v.w) * k.epsfac;
        typei[c] = needTypes ? a.atomTypes[ai] : 0;
        combi[c] = useTable ? make_float2(0.f, 0.f) : a.ljComb[ai];
    }

    Real eLJ = Real(0), eEl = Real(0);
    double eLJAcc = 0.0, eElAcc = 0.0;
    if constexpr (computeEnergy)
    {
        // Self terms on the diagonal entry (nbnxm_cuda_kernel.cuh:381-418), on the tj == 0 lanes.
        if (centralShift && s.cjPackedEnd > s.cjPackedBegin && a.cjPacked[s.cjPackedBegin].cj[0] == s.sci * c_ncps && tj == 0)
        {
            double q2 = 0.0, c6self = 0.0;
            for (int c = 0; c < c_ncps; c++)
            {
                const double q = a.xq[(s.sci * c_ncps + c) * c_cl + ti].w;
                q2 += q * q;
                if (flav.ljPme)
                {
                    c6self += a.nbfp[typei[c] * (a.numTypes + 1)].x;
                }
            }
            const double selfFactor = (flav.eeltype == NBNXM_ELEC_RF) ? 0.5 * flav.c_rf : flav.beta * c_oneSqrtPi;
            eElAcc -= flav.epsfac * q2 * selfFactor;
            if (flav.ljPme)
            {
                const double b2 = flav.ewaldcoeff_lj * flav.ewaldcoeff_lj;
                eLJAcc += c6self / 12.0 * (b2 * b2 * b2 / 6.0);
            }
        }
    }
    const bool nonSelfLane = !(centralShift && tj <= ti);
    Real fci[c_ncps][3];
    for (int c = 0; c < c_ncps; c++)
    {
        fci[c][0] = fci[c][1] = fci[c][2] = Real(0);
    }

    for (int jp = s.cjPackedBegin; jp < s.cjPackedEnd; jp++)
    {
        const unsigned int imask = a.cjPacked[jp].imei[0].imask;
        if (imask == 0)
        {
            continue;
        }
        const unsigned int wexcl = a.excl[a.cjPacked[jp].imei[0].excl_ind].pair[lane];
        for (int jm = 0; jm < c_jg; jm++)
        {
            if (!(imask & (0xFFu << (jm * c_ncps))))
            {
                continue;
            }
            const int cj = a.cjPacked[jp].cj[jm];
            const int aj = cj * c_cl + tj;
            const float4 xqj = a.xq[aj];
            const int typej = needTypes ? a.atomTypes[aj] : 0;
            const float2 combj = useTable ? make_float2(0.f, 0.f) : a.ljComb[aj];
            const Real qj = static_cast<Real>(xqj.w);
            Real fcj[3] = { Real(0), Real(0), Real(0) };
            for (int c = 0; c < c_ncps; c++)
            {
                const int bit = jm * c_ncps + c;
                if (!(imask & (1u << bit)))
                {
                    continue;
                }
                const Real dx = xi[c] - static_cast<Real>(xqj.x);
                const Real dy = yi[c] - static_cast<Real>(xqj.y);
                const Real dz = zi[c] - static_cast<Real>(xqj.z);
                Real r2;
                if constexpr (std::is_same_v<Real, float>)
                {
                    r2 = norm2Fixed(dx, dy, dz);
                }
                else
                {
                    r2 = dx * dx + dy * dy + dz * dz;
                }
                const bool notDiag = (s.sci * c_ncps + c) != cj;
                if (r2 < k.rcoulomb_sq && (nonSelfLane || notDiag))
                {
                    const Real int_bit = ((wexcl >> bit) & 1u) ? Real(1) : Real(0);
                    Real c6, c12;
                    if (useTable)
                    {
                        const float2 p = a.nbfp[a.numTypes * typei[c] + typej];
                        c6 = p.x;
                        c12 = p.y;
                    }
                    else if (useGeom)
                    {
                        c6 = static_cast<Real>(combi[c].x) * combj.x;
                        c12 = static_cast<Real>(combi[c].y) * combj.y;
                    }
                    else
                    {
                        // convert_sigma_epsilon_to_c6_c12 (nbnxm_cuda_kernel_utils.cuh:76-86).
                        const Real sigma = static_cast<Real>(combi[c].x) + combj.x;
                        const Real epsilon = static_cast<Real>(combi[c].y) * combj.y;
                        const Real sigma2 = sigma * sigma;
                        const Real sigma6 = sigma2 * sigma2 * sigma2;
                        c6 = epsilon * sigma6;
                        c12 = c6 * sigma6;
                    }
                    const Real c6grid = flav.ljPme ? ljPmeGridC6<Real>(flav, a.numTypes * typei[c] + typej) : Real(0);
                    const Real F_invr = pairForce<Real, computeEnergy>(flav, k, r2, int_bit, qi[c] * qj, c6, c12, c6grid, eLJ, eEl);
                    fcj[0] -= dx * F_invr;
                    fcj[1] -= dy * F_invr;
                    fcj[2] -= dz * F_invr;
                    fci[c][0] += dx * F_invr;
                    fci[c][1] += dy * F_invr;
                    fci[c][2] += dz * F_invr;
                }
            }
            const Real fj0 = sumOverTi(fcj[0]);
            const Real fj1 = sumOverTi(fcj[1]);
            const Real fj2 = sumOverTi(fcj[2]);
            if (ti < 3)
            {
                addForce<Real, accumulation>(a, 3 * aj + ti, ti == 0 ? fj0 : (ti == 1 ? fj1 : fj2));
            }
            if constexpr (computeEnergy)
            {
                // FP64 per j-cluster keeps the long energy sums accurate.
                eLJAcc += static_cast<double>(eLJ);
                eElAcc += static_cast<double>(eEl);
                eLJ = Real(0);
                eEl = Real(0);
            }
        }
    }
    double fsX = 0.0, fsY = 0.0, fsZ = 0.0;
    for (int c = 0; c < c_ncps; c++)
    {
        const Real f0 = sumOverTj(fci[c][0]);
        const Real f1 = sumOverTj(fci[c][1]);
        const Real f2 = sumOverTj(fci[c][2]);
        const int ai = (s.sci * c_ncps + c) * c_cl + ti;
        if (tj < 3)
        {
            addForce<Real, accumulation>(a, 3 * ai + tj, tj == 0 ? f0 : (tj == 1 ? f1 : f2));
        }
        fsX += f0;
        fsY += f1;
        fsZ += f2;
    }
    if (computeEnergy || a.computeFshift)
    {
        const double keep = (tj == 0) ? 1.0 : 0.0;
        const double v0 = reduceWaveD(eLJAcc);
        const double v1 = reduceWaveD(eElAcc);
        const bool doShift = a.computeFshift && !centralShift;
        const double s0 = doShift ? reduceWaveD(fsX * keep) : 0.0;
        const double s1 = doShift ? reduceWaveD(fsY * keep) : 0.0;
        const double s2 = doShift ? reduceWaveD(fsZ * keep) : 0.0;
        if (lane < 8)
        {
            const double out = (lane == 0) ? v0 : (lane == 1) ? v1 : (lane == 2) ? s0 : (lane == 3) ? s1 : (lane == 4) ? s2 : 0.0;
            a.entryPartials[8 * static_cast<size_t>(entry) + lane] = out;
        }
    }
}

template<typename Real, bool E, int A>
hipError_t launchFlavour(const NbKernelArgs& a, const NbFlavourParams& flav, int vdwType, hipStream_t stream)
{
    constexpr int wpb = 4;
    const int blocks = (a.nsci + wpb - 1) / wpb;
    if (blocks == 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL((nbnxmFlavourKernel<Real, E, A>), dim3(blocks), dim3(64 * wpb), 0, stream, a, flav, vdwType, wpb);
    return hipGetLastError();
}

template<bool E>
hipError_t launchFlavourAcc(const NbKernelArgs& a, const NbFlavourParams& flav, int accumulation, int vdwType, hipStream_t stream)
{
    switch (accumulation)
    {
        case NBNXM_ACCUM_FLOAT: return launchFlavour<float, E, NBNXM_ACCUM_FLOAT>(a, flav, vdwType, stream);
        case NBNXM_ACCUM_FIXED64: return launchFlavour<float, E, NBNXM_ACCUM_FIXED64>(a, flav, vdwType, stream);
        case NBNXM_ACCUM_FP64: return launchFlavour<double, E, NBNXM_ACCUM_FP64>(a, flav, vdwType, stream);
        default: return hipErrorInvalidValue;
    }
}

} // namespace

void ljPmeGridTables(int rule, int numTypes, const float* nbfp, float* grid, double* gridD)
{
    const int nt = numTypes;
    // nbfp_comb per type (atomdata.cpp:299-330): geometric sqrt(6C6_ii);
    // Lorentz-Berthelot 0.5 sixthroot(12C12/6C6) and sqrt(6C6^2/12C12), zero
    // for a type without LJ; in the reference's real (FP32) arithmetic.
    std::vector<float> a(nt), b(nt);
    std::vector<double> ad(nt), bd(nt);
    for (int t = 0; t < nt; t++)
    {
        const float c6 = nbfp[2 * (static_cast<size_t>(t) * nt + t)];
        const float c12 = nbfp[2 * (static_cast<size_t>(t) * nt + t) + 1];
        if (rule == NBNXM_LJPME_LB)
        {
            const bool lj = c6 > 0 && c12 > 0;
            a[t] = lj ? static_cast<float>(0.5 * std::sqrt(std::cbrt(c12 / c6))) : 0.0f;
            b[t] = lj ? std::sqrt(c6 * c6 / c12) : 0.0f;
            ad[t] = lj ? 0.5 * std::sqrt(std::cbrt(static_cast<double>(c12) / c6)) : 0.0;
            bd[t] = lj ? std::sqrt(static_cast<double>(c6) * c6 / c12) : 0.0;
        }
        else
        {
            a[t] = std::sqrt(c6);
            ad[t] = std::sqrt(static_cast<double>(c6));
        }
    }
    for (int i = 0; i < nt; i++)
    {
        for (int j = 0; j < nt; j++)
        {
            const size_t k = static_cast<size_t>(i) * nt + j;
            if (rule == NBNXM_LJPME_LB)
            {
                // sigma = s_i + s_j, epsilon = e_i e_j, c6grid = epsilon sigma2^3 (:317-324).
                const float sigma = a[i] + a[j];
                const float epsilon = b[i] * b[j];
                const float sigma2 = sigma * sigma;
                grid[k] = epsilon * sigma2 * sigma2 * sigma2;
                const double sd = ad[i] + ad[j];
                const double s2 = sd * sd;
                gridD[k] = bd[i] * bd[j] * s2 * s2 * s2;
            }
            else
            {
                grid[k] = a[i] * a[j];
                gridD[k] = ad[i] * ad[j];
            }
        }
    }
}

NbFlavourParams flavourParams(const nbnxm_interaction_const_t& ic)
{
    NbFlavourParams p;
    p.eeltype = ic.eeltype;
    p.vdwModifier = ic.vdwModifier;
    p.ljPme = ic.ljPme;
    p.vdwCutoffCheck = ic.rvdw < ic.rcoulomb ? 1 : 0;
    p.rcoulomb = ic.rcoulomb;
    p.rvdw = ic.rvdw;
    p.epsfac = ic.epsfac;
    p.beta = ic.ewaldcoeff_q;
    p.sh_ewald = ic.sh_ewald;
    p.k_rf = ic.k_rf;
    p.c_rf = ic.c_rf;
    p.cpot_rep = ic.repulsion_shift_cpot;
    p.cpot_disp = ic.dispersion_shift_cpot;
    p.rvdw_switch = ic.rvdw_switch;
    p.disp_c2 = ic.dispersion_shift_c2;
    p.disp_c3 = ic.dispersion_shift_c3;
    p.rep_c2 = ic.repulsion_shift_c2;
    p.rep_c3 = ic.repulsion_shift_c3;
    p.sw_c3 = ic.vdw_switch_c3;
    p.sw_c4 = ic.vdw_switch_c4;
    p.sw_c5 = ic.vdw_switch_c5;
    p.ewaldcoeff_lj = ic.ewaldcoeff_lj;
    p.sh_lj_ewald = ic.sh_lj_ewald;
    p.tabulatedEwald = ic.tabulatedEwald;
    // Flavour classes compiled with their flavour fixed (one instantiation each).
    const bool plain = !ic.ljPme && !ic.tabulatedEwald && ic.rvdw == ic.rcoulomb;
    const bool plainTabulated = !ic.ljPme && ic.tabulatedEwald && ic.rvdw == ic.rcoulomb;
    if (plain && ic.eeltype == NBNXM_ELEC_RF && ic.vdwModifier == NBNXM_VDW_MOD_POTSHIFT)
    {
        p.flavourClass = c_flavRfPotShift;
    }
    else if (plain && ic.eeltype == NBNXM_ELEC_EWALD && ic.vdwModifier == NBNXM_VDW_MOD_FORCESWITCH)
    {
        p.flavourClass = c_flavEwaldForceSwitch;
    }
    else if (plainTabulated && ic.eeltype == NBNXM_ELEC_EWALD && ic.vdwModifier == NBNXM_VDW_MOD_POTSHIFT)
    {
        p.flavourClass = c_flavEwaldTabPotShift;
    }
    else if (plain && ic.eeltype == NBNXM_ELEC_EWALD && ic.vdwModifier == NBNXM_VDW_MOD_POTSWITCH)
    {
        p.flavourClass = c_flavEwaldPotSwitch;
    }
    else if (!ic.ljPme && !ic.tabulatedEwald && ic.rvdw < ic.rcoulomb && ic.eeltype == NBNXM_ELEC_EWALD
             && ic.vdwModifier == NBNXM_VDW_MOD_POTSHIFT)
    {
        p.flavourClass = c_flavEwaldTwin;
    }
    else if (ic.ljPme && !ic.tabulatedEwald && ic.rvdw == ic.rcoulomb && ic.eeltype == NBNXM_ELEC_EWALD
             && ic.vdwModifier == NBNXM_VDW_MOD_POTSHIFT)
    {
        p.flavourClass = c_flavEwaldLjPme;
    }
    else
    {
        p.flavourClass = c_flavRuntime;
    }
    return p;
}

hipError_t launchNonbondedFlavour(const NbKernelArgs& a, const NbFlavourParams& flav, int accumulation, int vdwType, bool computeEnergy, hipStream_t stream)
{
    if (vdwType < NBNXM_VDW_CUT_TYPE_TABLE || vdwType > NBNXM_VDW_CUT_COMB_LB || (flav.ljPme && (!flav.ljPmeGrid || !flav.ljPmeGridD || !a.atomTypes || !a.nbfp))
        || (flav.tabulatedEwald && (!flav.ewaldTable || !flav.ewaldTableD || !(flav.ewaldTableScale > 0))))
    {
        return hipErrorInvalidValue;
    }
    return computeEnergy ? launchFlavourAcc<true>(a, flav, accumulation, vdwType, stream)
                         : launchFlavourAcc<false>(a, flav, accumulation, vdwType, stream);
}

} // namespace nbnxm_amd
