// The tuned nonbonded kernel of the nbnxm cluster-pair path (gfx950, CDNA4,
// wave64), shared by the translation units that instantiate its flavours.
//
// One wavefront owns one pair-list entry (an i-super-cluster of 8 clusters x 8
// atoms and a run of packed j-groups).  Lane l handles i-atom ti = l & 7 and
// j-atom tj = l >> 3 of every 8x8 cluster pair, the thread layout of the
// reference kernel (src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel.cuh:216-218) with
// the cluster-pair split removed so the whole 8x8 pair is one wave.
//  - i-cluster coordinates, charges and LJ parameters live in registers for the
//    whole entry (the reference stages them in shared memory, :330-351);
//  - j-forces are reduced over the 8 i-lanes with three DPP adds per component
//    (row_half_mirror, quad_perm xor1, quad_perm xor2) instead of warp shuffles
//    (:677, nbnxm_cuda_kernel_utils.cuh:467-496);
//  - forces are accumulated as int64 fixed point (2^-32 units), so every
//    evaluation is bit-reproducible whatever the atomic order;
//  - energies and shift forces are reduced in FP64 per entry and written to a
//    per-entry slot, then summed per system in a fixed order by a second kernel.
#pragma once

#include "device_common.h"
#include "flavour_math.h"
#include "nbnxm_kernels.h"

namespace nbnxm_amd
{

namespace
{

using namespace dev;

template<int ctrl>
__device__ __forceinline__ float dppF(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, false));
}

// Sum over the 8 consecutive lanes of a group; every lane of the group gets the
// same bits (the adds are pairwise commutative images of each other).
__device__ __forceinline__ float reduceOver8Lanes(float v)
{
    v += dppF<0x141>(v); // row_half_mirror: lane k <-> 7-k within 8
    v += dppF<0xB1>(v);  // quad_perm [1,0,3,2]
    v += dppF<0x4E>(v);  // quad_perm [2,3,0,1]
    return v;
}

// Sums of (x, y, z) over the 8 consecutive lanes of a group with 8 VALU
// instructions instead of 9 DPP adds and 2 selects: a reduce-scatter.  After
// the mirror step the lower 4 lanes carry x and y pair sums, the upper 4 the z
// pair sums; the xor-2 step splits the lower quad into x (lanes 0, 1) and
// y (lanes 2, 3).  Returns x in group lanes 0 and 1, y in lanes 2 and 3, z in
// lanes 4 to 7.  lower = (lane & 7) < 4, xLanes = (lane & 7) < 2,
// yLanes = (lane & 7) is 2 or 3.
__device__ __forceinline__ float reduceScatterOver8Lanes(float x, float y, float z, bool lower, bool xLanes, bool yLanes)
{
    // row_half_mirror pairs lane k with 7 - k: a lower lane reads its upper
    // partner's x, an upper lane its lower partner's z.
    const float r1 = (lower ? x : z) + dppF<0x141>(lower ? z : x);
    const float r2 = y + dppF<0x141>(y);
    // quad_perm [2,3,0,1] pairs k with k ^ 2: lanes 0, 1 read x from lanes
    // 2, 3, which read y from lanes 0, 1; upper lanes keep z.
    const float r3 = (yLanes ? r2 : r1) + dppF<0x4E>(xLanes ? r2 : r1);
    return r3 + dppF<0xB1>(r3); // quad_perm [1,0,3,2]
}

// OR of a 32-bit value over the whole wave, as a wave-uniform value.
__device__ __forceinline__ unsigned int waveOr(unsigned int v)
{
    v |= static_cast<unsigned int>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
    v |= static_cast<unsigned int>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x4E, 0xF, 0xF, false));  // quad_perm [2,3,0,1]
    v |= static_cast<unsigned int>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x141, 0xF, 0xF, false)); // row_half_mirror
    v |= static_cast<unsigned int>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x140, 0xF, 0xF, false)); // row_mirror
    return static_cast<unsigned int>(__builtin_amdgcn_readlane(static_cast<int>(v), 0) | __builtin_amdgcn_readlane(static_cast<int>(v), 16)
                                     | __builtin_amdgcn_readlane(static_cast<int>(v), 32) | __builtin_amdgcn_readlane(static_cast<int>(v), 48));
}

// Atomic add to element `index` (a 32-bit lane offset) of a global array at a
// wave-uniform base, issued with the scalar-base address form (base in scalar
// registers plus a 32-bit vector offset).  Base and offset pass through an
// empty asm: otherwise the compiler hoists the widened lane offset out of the
// caller's loop as a 64-bit vector address (two registers live through the
// list walk) and adds the base in the vector unit.  The asm costs one vector
// move (a copy of the offset) per call.
template<typename T>
using GlobalPtr = __attribute__((address_space(1))) T*;

template<typename T>
__device__ __forceinline__ void atomicAddAtLane(T* uniformBase, unsigned int index, T value)
{
    GlobalPtr<char> base = (GlobalPtr<char>)(uniformBase);
    unsigned int byteOffset = index * static_cast<unsigned int>(sizeof(T));
    asm("" : "+s"(base), "+v"(byteOffset));
    GlobalPtr<T> p = (GlobalPtr<T>)(base + static_cast<size_t>(byteOffset));
    __hip_atomic_fetch_add(p, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The lane index (0-63) from the lane counter, opaque to the compiler: a pass
// of the list walk or the final reduction that takes its lane-dependent
// indices and LDS addresses from its own read keeps none of them live in a
// register through the other passes (the compiler would keep one shared copy
// of every such value alive across the whole kernel).
__device__ __forceinline__ int laneIdOpaque()
{
    int l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
    return l;
}

// Sum over lanes with the same (lane & 7): xor 8 (row_ror:8), then rows
// xor 16 and xor 32 with the VALU lane swaps.
__device__ __forceinline__ float reduceOverStride8(float v)
{
    v += dppF<0x128>(v);
    v = addXor16(v);
    return addXor32(v);
}

// Read-only list and parameter data through the constant address space: with
// wave-uniform indices the loads become scalar (s_load, counted by lgkmcnt),
// so waiting for them never waits on the vector-memory counter that also
// tracks the in-flight force atomics.
template<typename T>
using ConstPtr = const __attribute__((address_space(4))) T*;

template<typename T>
__device__ __forceinline__ ConstPtr<T> asConst(const T* p)
{
    return (ConstPtr<T>)(p);
}

__device__ __forceinline__ nbnxm_sci_t loadSci(ConstPtr<nbnxm_sci_t> p, int i)
{
    const ConstPtr<int> q = (ConstPtr<int>)(p + i);
    nbnxm_sci_t s;
    s.sci = q[0];
    s.shift = q[1];
    s.cjPackedBegin = q[2];
    s.cjPackedEnd = q[3];
    return s;
}

__device__ __forceinline__ float4 loadF4(ConstPtr<float4> p, int i)
{
    const ConstPtr<float> q = (ConstPtr<float>)(p + i);
    return make_float4(q[0], q[1], q[2], q[3]);
}

__device__ __forceinline__ float2 loadF2(ConstPtr<float2> p, int i)
{
    const ConstPtr<float> q = (ConstPtr<float>)(p + i);
    return make_float2(q[0], q[1]);
}

// Occupancy floor: 3 waves per SIMD (at most 168 VGPRs) for the run-time
// flavour's force-only kernels with a combination rule, which the allocator
// otherwise leaves one register above it; no floor elsewhere (capping the
// tuned force-only kernel at 128 VGPRs, 4 waves per SIMD, spills and measured
// 25 % slower than its 3-wave allocation).
// kFlav = c_flavTuned is the tuned flavour (analytical Ewald, LJ potential
// shift, rvdw = rcoulomb); every other flavour class runs the same list walk
// with the pair interaction of flavour_math.h: fixed at compile time for the
// common classes, read from flav (wave-uniform, scalar branches) for
// c_flavRuntime.
// Bytes of the LJ type table in dynamic LDS, rounded up to 16 so that what
// follows it stays 16-byte aligned.
__host__ __device__ constexpr size_t typeTableLdsBytes(int numTypes)
{
    return (sizeof(float2) * numTypes * numTypes + 15) / 16 * 16;
}

// Per-wave LDS record of 16-byte slots, after the tables in dynamic LDS: the
// staged group's 32 j-atoms (xq, then the LJ data in .x/.y), the 64 i-atoms'
// LJ combination parameters (.x/.y) and, with the i-cluster data in LDS, their
// shifted coordinates and charges times epsfac.  One slot size for all four
// arrays lets a pass address each side with one lane register (the slot of
// j-atom tj, of i-atom ti, or of the lane) plus constant offsets.
constexpr int c_slotJXq = 0;
constexpr int c_slotJLj = 32;
constexpr int c_slotILj = 64;
constexpr int c_slotIXq = 128;
__host__ __device__ constexpr int waveStageSlots(bool iClusterInLds)
{
    return iClusterInLds ? 192 : 128;
}

// Occupancy floor (waves per SIMD) of the instantiations that have a variant
// with the i-cluster data in LDS (0: none; 1: the natural allocation).  Each
// pass of the list walk takes its lane indices from its own lane-counter read,
// uniform constants sit in scalar registers and the j-force atomics use the
// scalar-base address form, so the LDS variants need 61-85 VGPRs (F) and
// 74-112 (F+E) against 93-116 and 106-143 with the i-cluster data in
// registers: every compile-time class has an LDS variant, at its natural
// allocation.  The tuned F+E kernel with a combination rule allocates 87 VGPRs
// (5 waves: its exclusion-free pass needs 70, the exclusion pass and the final
// reductions more) and is held to 72 for a seventh wave: its spills are stored
// and reloaded in the exclusion pass only, and the batched ensemble launch
// measured 1.218 ms (5 waves) -> 1.155 (6) -> 1.139 ms (7).  The register
// variants need no floor (the tuned F+E kernel held to 4 waves, 128 VGPRs,
// spilled 8 registers and measured 2 % slower on a 96,000-atom box in an
// earlier build).
template<bool computeEnergy, int vdwType, int kFlav>
constexpr int iClusterLdsWaves()
{
    if (kFlav == c_flavRuntime)
    {
        return 0;
    }
    if (computeEnergy && kFlav == c_flavTuned && vdwType != NBNXM_VDW_CUT_TYPE_TABLE)
    {
        return 7;
    }
    return 1;
}

// Waves per SIMD the register allocation must allow.
template<bool computeEnergy, int vdwType, int kFlav, bool kILds>
constexpr int minWavesPerSimd()
{
    if (kILds)
    {
        return iClusterLdsWaves<computeEnergy, vdwType, kFlav>();
    }
    return (kFlav == c_flavRuntime && !computeEnergy && vdwType != NBNXM_VDW_CUT_TYPE_TABLE) ? 3 : 1;
}

// kILds: the i-cluster coordinates and charges are read from per-wave LDS for
// every cluster pair instead of being held in 32 registers.
// kPrune: the combined prune and interaction kernel of a fresh list without
// dynamic pruning (the _prune kernels, nbnxm_cuda_kernel.cuh:145-157,494-501,
// 680-685,722-741): each cluster pair of the list with no atom pair inside
// rlist leaves the group's mask, which is written back (and to the outer mask),
// and the entry's pruned cluster-pair count gives its bin, the key of the
// entry sort.
template<bool computeEnergy, int vdwType, int accumulation, int kFlav, bool kILds = false, bool kPrune = false>
__global__ __launch_bounds__(256, (minWavesPerSimd<computeEnergy, vdwType, kFlav, kILds>())) void nbnxmNonbondedKernel(const NbKernelArgs a, const NbFlavourParams flav, int wavesPerBlock, int stageOffset)
{
    static_assert(!(kPrune && kILds), "the combined prune kernel keeps the i-clusters in registers");
    constexpr bool generic = kFlav != c_flavTuned;
    const int lane = threadIdx.x & 63;
    const int ti = lane & 7;
    const int tj = lane >> 3;
    // Wave-uniform by construction; readfirstlane lets the compiler keep it scalar.
    const int work = __builtin_amdgcn_readfirstlane(blockIdx.x * wavesPerBlock + (threadIdx.x >> 6));
    // The LJ type table in LDS (dynamic size numTypes^2 float2, staged by the
    // whole block before any wave leaves): per-pair (6C6, 12C12) reads become
    // LDS reads, which do not wait behind the in-flight force atomics as a
    // vector load's vmcnt wait would.
    // The tabulated Ewald class stages its FP32 correction table after it, so
    // that the two table reads per pair are LDS reads too.
    // LJ-PME (flavour classes with the type table) stages the grid C6 of every
    // type pair after it, read per pair with the type-table index.
    constexpr bool ldsEwaldTable = FlavourTraits<kFlav>::table;
    HIP_DYNAMIC_SHARED(float4, sDyn);
    [[maybe_unused]] float2* sNbfp = reinterpret_cast<float2*>(sDyn);
    [[maybe_unused]] float* sEwald = reinterpret_cast<float*>(reinterpret_cast<char*>(sDyn) + typeTableLdsBytes(vdwType == NBNXM_VDW_CUT_TYPE_TABLE ? a.numTypes : 0));
    [[maybe_unused]] float* sGrid = sEwald;
    if constexpr (vdwType == NBNXM_VDW_CUT_TYPE_TABLE || ldsEwaldTable)
    {
        if constexpr (vdwType == NBNXM_VDW_CUT_TYPE_TABLE)
        {
            const int numEntries = a.numTypes * a.numTypes;
            for (int i = threadIdx.x; i < numEntries; i += blockDim.x)
            {
                sNbfp[i] = a.nbfp[i];
            }
            if constexpr (generic && !ldsEwaldTable)
            {
                if (flavLjPme<kFlav>(flav))
                {
                    for (int i = threadIdx.x; i < numEntries; i += blockDim.x)
                    {
                        sGrid[i] = flav.ljPmeGrid[i];
                    }
                }
            }
        }
        if constexpr (ldsEwaldTable)
        {
            for (int i = threadIdx.x; i < flav.ewaldTableSize; i += blockDim.x)
            {
                sEwald[i] = flav.ewaldTable[i];
            }
        }
        __syncthreads();
    }
    if (work >= a.nsci)
    {
        return;
    }
    const ConstPtr<nbnxm_cj_packed_t> cjList = asConst(a.cjPacked);
    const ConstPtr<float4> xqc = asConst(a.xq);
    const int entry = __builtin_amdgcn_readfirstlane(a.sciOrder ? asConst(a.sciOrder)[work] : work);
    const nbnxm_sci_t s = loadSci(asConst(a.sci), entry);
    const int sys = s.shift / NBNXM_NUM_SHIFTS;
    const int shiftInSys = s.shift - sys * NBNXM_NUM_SHIFTS;
    const bool centralShift = (shiftInSys == NBNXM_CENTRAL_SHIFT);

    const float4 pA = loadF4(asConst(a.sysParamsA), sys);
    const float2 pB = loadF2(asConst(a.sysParamsB), sys);
    const float rcoulomb_sq = pA.x;
    const float epsfac = pA.y;
    const float beta = pA.z;
    const float beta2 = uniformF(beta * beta);
    const float beta3 = uniformF(beta * beta * beta);
    const float ewaldShift = pA.w;
    const float cpotRep = pB.x;
    const float cpotDisp = pB.y;
    const float4 sh = loadF4(asConst(a.shiftVec), s.shift);
    [[maybe_unused]] const FlavourConsts<float> fk(flav);

    // i-atom coordinates (shifted) and charges (times epsfac) of all 8
    // clusters in registers, or with kILds in per-wave LDS; the type-table row
    // offsets in registers; combination parameters in per-wave LDS.
    static_assert(!kILds || iClusterLdsWaves<computeEnergy, vdwType, kFlav>() > 0, "no LDS variant for this instantiation");
    constexpr bool iInLds = kILds;
    // This wave's LDS record (stageOffset: bytes of the tables before the records).
    float4* const slots = reinterpret_cast<float4*>(reinterpret_cast<char*>(sDyn) + stageOffset)
                          + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * waveStageSlots(iInLds);
    float xi[iInLds ? 1 : c_ncps], yi[iInLds ? 1 : c_ncps], zi[iInLds ? 1 : c_ncps], qi[iInLds ? 1 : c_ncps];
    [[maybe_unused]] int tOffI[c_ncps]; // type table: row offset numTypes * type_i, in registers
#pragma unroll
    for (int c = 0; c < c_ncps; c++)
    {
        const int ai = (s.sci * c_ncps + c) * c_cl + ti;
        const float4 v = loadF4(xqc, ai);
        if constexpr (iInLds)
        {
            slots[c_slotIXq + c * c_cl + ti] = make_float4(v.x + sh.x, v.y + sh.y, v.z + sh.z, v.w * epsfac);
        }
        else
        {
            xi[c] = v.x + sh.x;
            yi[c] = v.y + sh.y;
            zi[c] = v.z + sh.z;
            qi[c] = v.w * epsfac;
        }
        if constexpr (vdwType == NBNXM_VDW_CUT_TYPE_TABLE)
        {
            tOffI[c] = a.numTypes * a.atomTypes[ai];
        }
    }
    if constexpr (vdwType != NBNXM_VDW_CUT_TYPE_TABLE)
    {
        const int ai = s.sci * c_ncps * c_cl + lane; // lane = c*8 + ti
        *reinterpret_cast<float2*>(slots + c_slotILj + lane) = loadF2(asConst(a.ljComb), ai);
    }

    float fciX[c_ncps], fciY[c_ncps], fciZ[c_ncps];
#pragma unroll
    for (int c = 0; c < c_ncps; c++)
    {
        fciX[c] = 0.0f;
        fciY[c] = 0.0f;
        fciZ[c] = 0.0f;
    }

    using EAcc = typename std::conditional<accumulation == NBNXM_ACCUM_FLOAT, float, double>::type;
    EAcc eLJAcc = 0;
    EAcc eElAcc = 0;
    if constexpr (computeEnergy)
    {
        // Self energy on the diagonal entry (:381-418): -epsfac*beta/sqrt(pi)*sum q^2,
        // added once per i-atom by the tj == 0 lanes.
        if (centralShift && s.cjPackedEnd > s.cjPackedBegin && cjList[s.cjPackedBegin].cj[0] == s.sci * c_ncps && tj == 0)
        {
            if constexpr (generic)
            {
                // Coulomb: -epsfac sum q^2 times beta/sqrt(pi) (Ewald) or c_rf/2
                // (reaction field); LJ-PME: + sum 6C6_ii/12 beta_lj^6/6.
                double q2 = 0.0, c6self = 0.0;
#pragma unroll
                for (int c = 0; c < c_ncps; c++)
                {
                    const int ai = (s.sci * c_ncps + c) * c_cl + ti;
                    const double q = loadF4(xqc, ai).w;
                    q2 += q * q;
                    if (flavLjPme<kFlav>(flav))
                    {
                        c6self += a.nbfp[a.atomTypes[ai] * (a.numTypes + 1)].x;
                    }
                }
                const double selfFactor = flavRf<kFlav>(flav) ? 0.5 * flav.c_rf : flav.beta * c_oneSqrtPi;
                eElAcc -= static_cast<EAcc>(flav.epsfac * q2 * selfFactor);
                if (flavLjPme<kFlav>(flav))
                {
                    const double b2 = flav.ewaldcoeff_lj * flav.ewaldcoeff_lj;
                    eLJAcc += static_cast<EAcc>(c6self / 12.0 * (b2 * b2 * b2 / 6.0));
                }
            }
            else if constexpr (accumulation == NBNXM_ACCUM_FLOAT)
            {
                float q2 = 0.0f;
#pragma unroll
                for (int c = 0; c < c_ncps; c++)
                {
                    const float qc = iInLds ? slots[c_slotIXq + c * c_cl + ti].w : qi[iInLds ? 0 : c];
                    q2 += qc * qc;
                }
                eElAcc -= q2 / epsfac * beta * static_cast<float>(c_oneSqrtPi);
            }
            else
            {
                // The self term is the largest energy contribution and cancels
                // most of the pair sum: evaluate it with FP64 charges and constants.
                const double* pD = a.sysParamsD + 8 * sys;
                double q2 = 0.0;
#pragma unroll
                for (int c = 0; c < c_ncps; c++)
                {
                    const double q = loadF4(xqc, (s.sci * c_ncps + c) * c_cl + ti).w;
                    q2 += q * q;
                }
                eElAcc -= pD[1] * q2 * pD[2] * c_oneSqrtPi;
            }
        }
    }

    const int firstCluster = __builtin_amdgcn_readfirstlane(s.sci * c_ncps);
    // Constants folded for the exclusion-free path.
    const float negBeta2Log2e = uniformF(-beta2 * 1.44269504088896340736f); // exp(-b^2 r^2) = exp2(r^2 * this)
    const float betaT = uniformF(c_erfcxPolyScale * beta);                   // t = 1/(1 + 0.4 beta r)
    const float twoBetaSqrtPi = uniformF(beta * static_cast<float>(2.0 * c_oneSqrtPi));
    const float twoCpotDisp = uniformF(2.0f * cpotDisp);
    const float fixedBound = uniformF(c_fixedMaxF);

    // The packed group's 32 j-atoms are staged in the wave's LDS record.  The
    // next group is fetched into registers (one coalesced load per lane)
    // before the current group's force atomics are issued, so waiting for
    // j-data never waits for those atomics (vmcnt counts both, in order).

    // Combined prune: the current group's mask less its cluster pairs with no
    // pair inside rlist, and the entry's count of kept cluster pairs.
    [[maybe_unused]] const float pruneRlist_sq = a.pruneRlist_sq;
    [[maybe_unused]] unsigned int prunedMask = 0u;
    [[maybe_unused]] int prunedPairCount = 0;

    {
        // Two passes over the entry's groups, each a loop with one path, so
        // the accumulators stay in their registers through it (a loop that
        // chose between the two group bodies per group kept the i-force
        // accumulators in two register sets and copied them at every
        // iteration): pass 0 takes the groups without an exclusion mask,
        // pass 1 the groups with one.  Pass 0 records where the latter are
        // (a scalar bit mask of the first 64 groups, a scan beyond), so
        // pass 1 finds them without reading the list again.  Every group is
        // evaluated once, in the same order in both placements of the
        // i-cluster data and in the combined prune kernel.
        const int begin = s.cjPackedBegin;
        unsigned long long exclGroups = 0ull; // bit g - begin: group g has an exclusion mask
        bool exclBeyond = false;              // such a group at begin + 64 or later
        auto nextPlain = [&](int g) {
            while (g < s.cjPackedEnd)
            {
                const unsigned int m = cjList[g].imei[0].imask;
                const int ex = cjList[g].imei[0].excl_ind;
                if (m != 0u && ex == 0)
                {
                    break;
                }
                if (m != 0u)
                {
                    if (g - begin < 64)
                    {
                        exclGroups |= 1ull << (g - begin);
                    }
                    else
                    {
                        exclBeyond = true;
                    }
                }
                g++;
            }
            return g;
        };
        auto nextExcl = [&](int g) {
            if (g - begin < 64)
            {
                const unsigned long long m = exclGroups & (~0ull << (g - begin));
                if (m != 0ull)
                {
                    return begin + static_cast<int>(__builtin_ctzll(m));
                }
                g = begin + 64;
            }
            if (!exclBeyond)
            {
                return s.cjPackedEnd;
            }
            while (g < s.cjPackedEnd && (cjList[g].imei[0].imask == 0u || cjList[g].imei[0].excl_ind == 0))
            {
                g++;
            }
            return g;
        };
        auto walk = [&](auto passTag) {
            constexpr bool exclPass = decltype(passTag)::value;
            // Lane indices of this pass from its own read of the lane counter:
            // no lane-dependent index or LDS address stays live in a register
            // through the other pass and the final reductions.
            const int lane = laneIdOpaque();
            const int ti = lane & 7;
            const int tj = lane >> 3;
            // Newton / self masking on the diagonal of the central shift (:423): on a
            // cluster pair with ci == cj only lanes with tj > ti interact; their lane
            // mask (a constant of the layout) as a wave-uniform value.
            const unsigned long long upperLanes = __builtin_amdgcn_ballot_w64(tj > ti);
            const bool lowerLane = ti < 4;
            const bool xLane = ti < 2;
            const bool yLane = (ti >> 1) == 1;
            const int stageAtom = lane & 31; // j-cluster stageAtom/8, atom stageAtom%8
            const float4* const jSlot = slots + tj; // j-atom tj's slots
            const float4* const iSlot = slots + ti; // i-atom ti's slots

            // Deferred j-force atomics: the reduced force component of lane ti < 3
            // for each j-cluster of the group, issued at the start of the next group.
            float pendV[c_jg];
            int pendCj[c_jg];
            unsigned int pendMask = 0u;

            // One packed group of 4 j-clusters against the 8 i-clusters, instantiated
            // with and without exclusion handling.  In a group with exclusions, only
            // cluster pairs that hold an excluded lane take the interaction-bit path;
            // the others, like every pair of an exclusion-free group, take the erfc
            // path, where one exp(-z^2) serves both Coulomb force and energy.
            // Energies are summed in FP32 over the group and added in FP64 per group;
            // the LJ energy is kept as 12*E without its constant shift, plus the
            // sums of C6 and C12 that carry the shift.
            auto group = [&](const int cj0, const int cj1, const int cj2, const int cj3, const unsigned int imask, const unsigned int wexcl,
                             const unsigned int groupExclBits, auto withExcl) {
                constexpr bool hasExcl = decltype(withExcl)::value;
                // Bit b set: some lane of cluster pair b has an excluded (or Newton
                // masked) pair (waveOr(~wexcl) of the group).
                const unsigned int exclBits = hasExcl ? groupExclBits : 0u;
                // Bit b set: cluster pair b lies on the diagonal of the central shift
                // (i-cluster firstCluster + c is j-cluster cj), at most one per j-cluster.
                unsigned int diagBits = 0u;
                if (centralShift)
                {
                    const int cjs[c_jg] = { cj0, cj1, cj2, cj3 };
#pragma unroll
                    for (int jm = 0; jm < c_jg; jm++)
                    {
                        const unsigned int d = static_cast<unsigned int>(cjs[jm] - firstCluster);
                        diagBits |= (d < static_cast<unsigned int>(c_ncps)) ? (1u << (jm * c_ncps + d)) : 0u;
                    }
                }
                float eLJ12 = 0.0f, eSumC6 = 0.0f, eSumC12 = 0.0f, eElG = 0.0f;
                [[maybe_unused]] float eLJg = 0.0f;
#pragma unroll
                for (int jm = 0; jm < c_jg; jm++)
                {
                    if (!(imask & (0xFFu << (jm * c_ncps))))
                    {
                        continue;
                    }
                    const int cj = (jm == 0) ? cj0 : (jm == 1) ? cj1 : (jm == 2) ? cj2 : cj3;
                    const float4 xqj = jSlot[c_slotJXq + jm * c_cl];
                    const float2 ljStaged = *reinterpret_cast<const float2*>(jSlot + c_slotJLj + jm * c_cl);
                    const int typej = __float_as_int(ljStaged.x);
                    const float2 ljcj = ljStaged;
                    float fcjX = 0.0f, fcjY = 0.0f, fcjZ = 0.0f;
                    unsigned long long anyInRange = 0ull;
#pragma unroll
                    for (int c = 0; c < c_ncps; c++)
                    {
                        const int bit = jm * c_ncps + c;
                        if (!(imask & (1u << bit)))
                        {
                            continue;
                        }
                        float4 iv;
                        if constexpr (iInLds)
                        {
                            iv = iSlot[c_slotIXq + c * c_cl];
                        }
                        else
                        {
                            iv = make_float4(xi[c], yi[c], zi[c], qi[c]);
                        }
                        const float dx = iv.x - xqj.x;
                        const float dy = iv.y - xqj.y;
                        const float dz = iv.z - xqj.z;
                        float r2 = norm2Fixed(dx, dy, dz);
                        if constexpr (kPrune)
                        {
                            // No pair of the 64 inside rlist: the cluster pair leaves the list (:494-501).
                            if (__builtin_amdgcn_ballot_w64(r2 < pruneRlist_sq) == 0ull)
                            {
                                prunedMask &= ~(1u << bit);
                            }
                        }
                        // On the diagonal of the central shift only pairs with j > i
                        // interact (nonSelfInteraction, :423,508), whatever the list's
                        // exclusion words say: the lanes with tj <= ti leave the
                        // in-range mask.  Wave-uniform: exclHere from the group's OR of
                        // its exclusion words, the lane mask from list indices, so the
                        // rule costs a scalar AND of masks rather than a select per lane.
                        const bool exclHere = hasExcl && ((exclBits >> bit) & 1u);
                        const unsigned long long keepLanes = ((diagBits >> bit) & 1u) ? upperLanes : ~0ull;
                        const unsigned long long inRangeLanes = __builtin_amdgcn_ballot_w64(r2 < rcoulomb_sq) & keepLanes;
                        anyInRange |= inRangeLanes;
                        // With the i-cluster data in LDS: a wave-uniform test skips
                        // the pair work when no lane is in range; inside it the lanes
                        // out of range keep F_invr = 0, so the force updates after the
                        // pair math run on the whole wave and accumulate in place (no
                        // copies of the accumulators for the lanes a divergent region
                        // would leave untouched).  With the data in registers the
                        // updates stay in the divergent region, which measured faster
                        // there (3-4 % per launch on a 96,000-atom box), while the
                        // form above took 4.5 % off the LDS variant.
                        constexpr bool forcesAfterPairMath = iInLds;
                        if (forcesAfterPairMath && inRangeLanes == 0ull)
                        {
                            continue;
                        }
                        float F_invr = 0.0f;
                        auto accumulateForces = [&]() {
                            fcjX = __builtin_fmaf(-dx, F_invr, fcjX);
                            fcjY = __builtin_fmaf(-dy, F_invr, fcjY);
                            fcjZ = __builtin_fmaf(-dz, F_invr, fcjZ);
                            fciX[c] = __builtin_fmaf(dx, F_invr, fciX[c]);
                            fciY[c] = __builtin_fmaf(dy, F_invr, fciY[c]);
                            fciZ[c] = __builtin_fmaf(dz, F_invr, fciZ[c]);
                        };
                        if (__builtin_amdgcn_inverse_ballot_w64(inRangeLanes))
                        {
                            float c6, c12;
                            float sigma = 0.0f, epsilon = 0.0f;
                            if constexpr (vdwType == NBNXM_VDW_CUT_TYPE_TABLE)
                            {
                                // One LDS read: the i row offset is a register.
                                const float2 p = sNbfp[tOffI[c] + typej];
                                c6 = p.x;
                                c12 = p.y;
                            }
                            else if constexpr (vdwType == NBNXM_VDW_CUT_COMB_GEOM)
                            {
                                const float2 ljci = *reinterpret_cast<const float2*>(iSlot + c_slotILj + c * c_cl);
                                c6 = ljci.x * ljcj.x;
                                c12 = ljci.y * ljcj.y;
                            }
                            else
                            {
                                const float2 ljci = *reinterpret_cast<const float2*>(iSlot + c_slotILj + c * c_cl);
                                sigma = ljci.x + ljcj.x;
                                epsilon = ljci.y * ljcj.y;
                                const float sigma2 = sigma * sigma;
                                const float sigma6 = sigma2 * sigma2 * sigma2;
                                c6 = epsilon * sigma6;
                                c12 = c6 * sigma6;
                            }
                            // Only lanes inside the cut-off get here: the lower clamp
                            // (:536) is all that applies.  r2 >= 0, so the float order
                            // is the order of the bit patterns: one integer max with a
                            // literal operand (a float max would first quiet NaNs).
                            r2 = __int_as_float(max(__float_as_int(r2), __float_as_int(NBNXM_MIN_DISTANCE_SQUARED)));
                            const float inv_r = __builtin_amdgcn_rsqf(r2);
                            const float inv_r2 = inv_r * inv_r;
                            const float qq = iv.w * xqj.w;
                            // Energy terms of the tuned paths, accumulated after them.
                            [[maybe_unused]] float eLj6 = 0.0f, eLjArg = 0.0f, eElArg = 0.0f;
                            if constexpr (generic)
                            {
                                const float int_bit = (hasExcl && exclHere) ? static_cast<float>((wexcl >> bit) & 1u) : 1.0f;
                                float c6grid = 0.0f;
                                if constexpr (vdwType == NBNXM_VDW_CUT_TYPE_TABLE && !ldsEwaldTable)
                                {
                                    c6grid = flavLjPme<kFlav>(flav) ? sGrid[tOffI[c] + typej] : 0.0f;
                                }
                                F_invr = pairForce<float, computeEnergy, kFlav>(flav, fk, r2, int_bit, qq, c6, c12, c6grid, eLJg, eElG,
                                                                                ldsEwaldTable ? sEwald : nullptr, computeEnergy && !(hasExcl && exclHere));
                            }
                            else if (hasExcl && exclHere)
                            {
                                // The reference form with the interaction bit (:548-659).
                                const float int_bit = static_cast<float>((wexcl >> bit) & 1u);
                                const float inv_r6 = int_bit * (inv_r2 * inv_r2 * inv_r2);
                                if constexpr (vdwType == NBNXM_VDW_CUT_COMB_LB && !computeEnergy)
                                {
                                    const float sig_r = sigma * inv_r;
                                    const float sig_r2 = sig_r * sig_r;
                                    const float sig_r6 = int_bit * (sig_r2 * sig_r2 * sig_r2);
                                    F_invr = epsilon * sig_r6 * (sig_r6 - 1.0f) * inv_r2;
                                }
                                else
                                {
                                    F_invr = inv_r6 * (c12 * inv_r6 - c6) * inv_r2;
                                }
                                const float z2 = beta2 * r2;
                                F_invr += qq * (int_bit * inv_r2 * inv_r + pmecorrF(z2) * beta3);
                                if constexpr (computeEnergy)
                                {
                                    eLj6 = inv_r6;
                                    eLjArg = __builtin_fmaf(c12, inv_r6, -2.0f * c6);
                                    eElArg = inv_r * (int_bit - erff(r2 * inv_r * beta)) - int_bit * ewaldShift;
                                    eSumC6 = __builtin_fmaf(int_bit, c6, eSumC6);
                                    eSumC12 = __builtin_fmaf(int_bit, c12, eSumC12);
                                }
                            }
                            else
                            {
                                const float inv_r6 = inv_r2 * inv_r2 * inv_r2;
                                if constexpr (vdwType == NBNXM_VDW_CUT_COMB_LB && !computeEnergy)
                                {
                                    const float sig_r = sigma * inv_r;
                                    const float sig_r2 = sig_r * sig_r;
                                    const float sig_r6 = sig_r2 * sig_r2 * sig_r2;
                                    F_invr = epsilon * sig_r6 * (sig_r6 - 1.0f) * inv_r2 + qq * (inv_r2 * inv_r + pmecorrF(beta2 * r2) * beta3);
                                }
                                else if constexpr (computeEnergy)
                                {
                                    // F/(qq r) = e (R/r + 2 beta/sqrt(pi)) / r^2, V/qq = e R / r - sh_ewald,
                                    // with e = exp(-beta^2 r^2) and erfc(beta r) = e R.
                                    const float r = r2 * inv_r;
                                    const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(betaT, r, 1.0f));
                                    const float R = erfcxPolyT(t);
                                    const float e = __builtin_amdgcn_exp2f(r2 * negBeta2Log2e);
                                    const float w = qq * e;
                                    const float u = __builtin_fmaf(R, inv_r, twoBetaSqrtPi);
                                    const float tl = __builtin_fmaf(c12, inv_r6, -c6);
                                    F_invr = __builtin_fmaf(inv_r6, tl, w * u) * inv_r2;
                                    eLj6 = inv_r6;
                                    eLjArg = tl - c6;
                                    eElArg = __builtin_fmaf(e * R, inv_r, -ewaldShift);
                                    eSumC6 += c6;
                                    eSumC12 += c12;
                                }
                                else
                                {
                                    const float tl = __builtin_fmaf(c12, inv_r6, -c6);
                                    const float coul = qq * __builtin_fmaf(pmecorrF(beta2 * r2), beta3, inv_r2 * inv_r);
                                    F_invr = __builtin_fmaf(inv_r6 * tl, inv_r2, coul);
                                }
                            }
                            if constexpr (computeEnergy && !generic)
                            {
                                eLJ12 = __builtin_fmaf(eLj6, eLjArg, eLJ12);
                                eElG = __builtin_fmaf(qq, eElArg, eElG);
                            }
                            if constexpr (!forcesAfterPairMath)
                            {
                                accumulateForces();
                            }
                        }
                        if constexpr (forcesAfterPairMath)
                        {
                            accumulateForces();
                        }
                    }
                    // j-force: sum over the 8 i-lanes, skipped when no pair was inside
                    // the cut-off; lanes ti = 0, 1, 2 keep x, y, z for the deferred
                    // atomic (consecutive lanes on consecutive addresses coalesce).
                    if (anyInRange != 0ull)
                    {
                        pendV[jm] = reduceScatterOver8Lanes(fcjX, fcjY, fcjZ, lowerLane, xLane, yLane);
                        pendCj[jm] = cj;
                        pendMask |= 1u << jm;
                    }
                }
                if constexpr (computeEnergy && generic)
                {
                    eLJAcc += static_cast<EAcc>(eLJg);
                    eElAcc += static_cast<EAcc>(eElG);
                }
                else if constexpr (computeEnergy)
                {
                    const float eLJ = (eLJ12 + cpotRep * eSumC12 - twoCpotDisp * eSumC6) * (1.0f / 12.0f);
                    eLJAcc += static_cast<EAcc>(eLJ);
                    eElAcc += static_cast<EAcc>(eElG);
                }
            };

            // Lanes that issue the j-force atomics (ti = 0, 1, 2 for x, y, z) and the
            // lane part of the address: component ti of j-atom tj of a cluster.
            const bool jForceLane = (ti & 1) == 0 && ti < 6;
            const unsigned int jForceLaneOffset = 3 * tj + (ti >> 1);
            auto flushPending = [&]() {
#pragma unroll
                for (int jm = 0; jm < c_jg; jm++)
                {
                    if ((pendMask >> jm) & 1u)
                    {
                        if (jForceLane)
                        {
                            // Uniform cluster base (scalar registers) plus a 32-bit
                            // lane offset: the atomic's scalar-base address form.
                            const size_t base = static_cast<size_t>(pendCj[jm]) * (3 * c_cl);
                            if constexpr (accumulation == NBNXM_ACCUM_FLOAT)
                            {
                                atomicAddAtLane(a.fFloat + base, jForceLaneOffset, pendV[jm]);
                            }
                            else
                            {
                                atomicAddAtLane(a.fFixed + base, jForceLaneOffset, toFixed(pendV[jm], fixedBound));
                            }
                        }
                    }
                }
                pendMask = 0u;
            };

            // Lanes 0-31 fetch xq, lanes 32-63 the LJ data of j-atom stageAtom of group g.
            float4 pf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const bool stageK1 = (stageAtom >> 3) == 1;
            const bool stageK2 = (stageAtom >> 3) == 2;
            const bool stageK3 = (stageAtom >> 3) == 3;
            unsigned int pfExcl = 0xFFFFFFFFu;
            auto prefetch = [&](int g) {
                // Scalar loads first, then a per-lane select (a select of the
                // addresses would become a vector load whose wait also waits for the
                // in-flight force atomics).
                const int g0 = __builtin_amdgcn_readfirstlane(cjList[g].cj[0]);
                const int g1 = __builtin_amdgcn_readfirstlane(cjList[g].cj[1]);
                const int g2 = __builtin_amdgcn_readfirstlane(cjList[g].cj[2]);
                const int g3 = __builtin_amdgcn_readfirstlane(cjList[g].cj[3]);
                const int cjl = stageK1 ? g1 : (stageK2 ? g2 : (stageK3 ? g3 : g0));
                const int atom = cjl * c_cl + (stageAtom & 7);
                if (lane < 32)
                {
                    pf = loadF4(xqc, atom);
                }
                else if constexpr (vdwType == NBNXM_VDW_CUT_TYPE_TABLE)
                {
                    pf = make_float4(__int_as_float(asConst(a.atomTypes)[atom]), 0.0f, 0.0f, 0.0f);
                }
                else
                {
                    const float2 lj = loadF2(asConst(a.ljComb), atom);
                    pf = make_float4(lj.x, lj.y, 0.0f, 0.0f);
                }
                const int ex = cjList[g].imei[0].excl_ind;
                pfExcl = (ex != 0) ? asConst(a.excl)[ex].pair[lane] : 0xFFFFFFFFu;
            };
            unsigned int stagedExcl = 0xFFFFFFFFu;
            auto stage = [&]() {
                __builtin_amdgcn_wave_barrier();
                // Lanes 0-31 fill the xq slots, lanes 32-63 the LJ slots.
                slots[c_slotJXq + lane] = pf;
                stagedExcl = pfExcl;
                __builtin_amdgcn_wave_barrier();
            };


            auto nextInPass = [&](int g) { return exclPass ? nextExcl(g) : nextPlain(g); };
            int jp = nextInPass(begin);
            if (jp < s.cjPackedEnd)
            {
                prefetch(jp);
                stage();
            }
            while (jp < s.cjPackedEnd)
            {
                const unsigned int imask = cjList[jp].imei[0].imask;
                const int cj0 = cjList[jp].cj[0];
                const int cj1 = cjList[jp].cj[1];
                const int cj2 = cjList[jp].cj[2];
                const int cj3 = cjList[jp].cj[3];
                const unsigned int wexcl = stagedExcl;
                flushPending();
                const int jn = nextInPass(jp + 1);
                if (jn < s.cjPackedEnd)
                {
                    prefetch(jn);
                }
                if constexpr (kPrune)
                {
                    prunedMask = imask;
                }
                if constexpr (exclPass)
                {
                    group(cj0, cj1, cj2, cj3, imask, wexcl, waveOr(~wexcl), std::true_type{});
                }
                else
                {
                    group(cj0, cj1, cj2, cj3, imask, 0xFFFFFFFFu, 0u, std::false_type{});
                }
                if constexpr (kPrune)
                {
                    // The group's list words were read (scalar) before this
                    // store (its mask by the pass scan), and no other wave
                    // reads them.  Every lane stores the same word (one
                    // coalesced write): under a lane-0 branch the compiler
                    // would wait for these stores before the next group's
                    // staged data can be used.
                    a.pruneCjPacked[jp].imei[0].imask = prunedMask;
                    a.pruneImaskOuter[jp] = prunedMask;
                    prunedPairCount += __popc(prunedMask);
                }
                if (jn < s.cjPackedEnd)
                {
                    stage();
                }
                jp = jn;
            }
            flushPending();
        };
        walk(std::false_type{});
        walk(std::true_type{});
    }

    {
        // Lane indices of the final reductions from their own read of the lane
        // counter (see laneIdOpaque).
        const int lane = laneIdOpaque();
        const int ti = lane & 7;
        const int tj = lane >> 3;
        if constexpr (kPrune)
        {
            // The entry's bin, the key of the entry sort (:722-741); one wave per entry.
            if (lane == 0)
            {
                a.pruneSciCount[work] = max(NBNXM_SCI_HISTOGRAM_SIZE - prunedPairCount - 1, 0);
            }
        }

        // i-forces: sum over the 8 j-lanes, lanes tj = 0,1,2 store x,y,z.
        double fsX = 0.0, fsY = 0.0, fsZ = 0.0;
#pragma unroll
        for (int c = 0; c < c_ncps; c++)
        {
            const float fx = reduceOverStride8(fciX[c]);
            const float fy = reduceOverStride8(fciY[c]);
            const float fz = reduceOverStride8(fciZ[c]);
            const int ai = (s.sci * c_ncps + c) * c_cl + ti;
            if (tj < 3)
            {
                const float v = (tj == 0) ? fx : ((tj == 1) ? fy : fz);
                if (v != 0.0f)
                {
                    if constexpr (accumulation == NBNXM_ACCUM_FLOAT)
                    {
                        atomicAdd(a.fFloat + 3 * ai + tj, v);
                    }
                    else
                    {
                        atomicAdd(a.fFixed + 3 * ai + tj, toFixed(v, fixedBound));
                    }
                }
            }
            fsX += static_cast<double>(fx);
            fsY += static_cast<double>(fy);
            fsZ += static_cast<double>(fz);
        }

        if (computeEnergy || a.computeFshift)
        {
            // Every lane of an i-column holds the column's sums: take tj == 0 lanes.
            const double keep = (tj == 0) ? 1.0 : 0.0;
            double v0 = 0.0, v1 = 0.0;
            if constexpr (computeEnergy)
            {
                v0 = reduceWaveD(static_cast<double>(eLJAcc));
                v1 = reduceWaveD(static_cast<double>(eElAcc));
            }
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
}

// Largest LJ type table staged in LDS: 64 types (+ dummy), 33.8 KB.
constexpr size_t c_maxLdsTypeTableBytes = sizeof(float2) * c_maxLdsTableTypes * c_maxLdsTableTypes;
// Largest tabulated Ewald correction table staged in LDS (the tabulated class
// falls back to the run-time class, which reads it from global memory).
constexpr size_t c_maxLdsEwaldTableBytes = 32768;

template<bool E, int V, int A, int kFlav, bool kILds = false, bool kPrune = false>
hipError_t launchTunedKernel(const NbKernelArgs& a, const NbFlavourParams& flav, const NbLaunchConfig& cfg, hipStream_t stream)
{
    const int wpb = cfg.wavesPerBlock;
    const int blocks = (a.nsci + wpb - 1) / wpb;
    if (blocks == 0)
    {
        return hipSuccess;
    }
    // The type table kernel stages the whole table in LDS.
    const size_t typeBytes = (V == NBNXM_VDW_CUT_TYPE_TABLE) ? sizeof(float2) * a.numTypes * a.numTypes : 0;
    const size_t ewaldBytes = FlavourTraits<kFlav>::table ? sizeof(float) * flav.ewaldTableSize : 0;
    // The LJ-PME grid table follows the type table (never with the Ewald table).
    const bool gridInLds = kFlav != c_flavTuned && !FlavourTraits<kFlav>::table && V == NBNXM_VDW_CUT_TYPE_TABLE && flav.ljPme;
    const size_t gridBytes = gridInLds ? sizeof(float) * a.numTypes * a.numTypes : 0;
    if (typeBytes > c_maxLdsTypeTableBytes || ewaldBytes > c_maxLdsEwaldTableBytes || (gridInLds && !flav.ljPmeGrid))
    {
        return hipErrorInvalidValue;
    }
    // Tables first (16-byte aligned end), then one record per wave.
    const size_t tableBytes = (((V == NBNXM_VDW_CUT_TYPE_TABLE) ? typeTableLdsBytes(a.numTypes) : 0) + ewaldBytes + gridBytes + 15) / 16 * 16;
    const size_t ldsBytes = tableBytes + static_cast<size_t>(wpb) * waveStageSlots(kILds) * sizeof(float4);
    hipLaunchKernelGGL((nbnxmNonbondedKernel<E, V, A, kFlav, kILds, kPrune>), dim3(blocks), dim3(64 * wpb), ldsBytes, stream, a, flav, wpb,
                       static_cast<int>(tableBytes));
    return hipGetLastError();
}

// FP32 accumulation modes (float atomics or int64 fixed point) of one flavour class.
template<int kFlav>
hipError_t launchTunedFlavour(const NbKernelArgs& a, const NbFlavourParams& flav, int accumulation, int vdwType, bool computeEnergy, const NbLaunchConfig& cfg, hipStream_t stream)
{
    const bool iLds = iClusterDataInLds(cfg, a.nsci);
    // A fresh list without dynamic pruning: the combined prune kernel (its
    // outputs set, entries in list order).
    const bool prune = a.pruneCjPacked != nullptr;
    if (prune && (!a.pruneImaskOuter || !a.pruneSciCount || a.sciOrder))
    {
        return hipErrorInvalidValue;
    }
    auto acc = [&](auto e, auto v) -> hipError_t {
        constexpr bool E = decltype(e)::value;
        constexpr int V = decltype(v)::value;
        if (prune)
        {
            switch (accumulation)
            {
                case NBNXM_ACCUM_FLOAT: return launchTunedKernel<E, V, NBNXM_ACCUM_FLOAT, kFlav, false, true>(a, flav, cfg, stream);
                case NBNXM_ACCUM_FIXED64: return launchTunedKernel<E, V, NBNXM_ACCUM_FIXED64, kFlav, false, true>(a, flav, cfg, stream);
                default: return hipErrorInvalidValue;
            }
        }
        if constexpr (iClusterLdsWaves<E, V, kFlav>() > 0)
        {
            if (iLds)
            {
                switch (accumulation)
                {
                    case NBNXM_ACCUM_FLOAT: return launchTunedKernel<E, V, NBNXM_ACCUM_FLOAT, kFlav, true>(a, flav, cfg, stream);
                    case NBNXM_ACCUM_FIXED64: return launchTunedKernel<E, V, NBNXM_ACCUM_FIXED64, kFlav, true>(a, flav, cfg, stream);
                    default: return hipErrorInvalidValue;
                }
            }
        }
        switch (accumulation)
        {
            case NBNXM_ACCUM_FLOAT: return launchTunedKernel<E, V, NBNXM_ACCUM_FLOAT, kFlav>(a, flav, cfg, stream);
            case NBNXM_ACCUM_FIXED64: return launchTunedKernel<E, V, NBNXM_ACCUM_FIXED64, kFlav>(a, flav, cfg, stream);
            default: return hipErrorInvalidValue;
        }
    };
    auto vdw = [&](auto e) -> hipError_t {
        switch (vdwType)
        {
            case NBNXM_VDW_CUT_TYPE_TABLE: return acc(e, std::integral_constant<int, NBNXM_VDW_CUT_TYPE_TABLE>{});
            case NBNXM_VDW_CUT_COMB_GEOM: return acc(e, std::integral_constant<int, NBNXM_VDW_CUT_COMB_GEOM>{});
            case NBNXM_VDW_CUT_COMB_LB: return acc(e, std::integral_constant<int, NBNXM_VDW_CUT_COMB_LB>{});
            default: return hipErrorInvalidValue;
        }
    };
    return computeEnergy ? vdw(std::true_type{}) : vdw(std::false_type{});
}

} // namespace
} // namespace nbnxm_amd
