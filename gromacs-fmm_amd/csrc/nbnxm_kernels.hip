// gfx950 (CDNA4, wave64) kernels of the nbnxm cluster-pair path.
//
// One wavefront owns one pair-list entry (an i-super-cluster of 8 clusters x 8
// atoms and a run of packed j-groups).  Lane l handles i-atom ti = l & 7 and
// j-atom tj = l >> 3 of every 8x8 cluster pair, the thread layout of the
// reference kernel (src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel.cuh:216-218) with
// the cluster-pair split removed so the whole 8x8 pair is one wave.
//  - i-cluster coordinates, charges and LJ parameters live in registers for the
//    whole entry (the reference stages them in shared memory, :330-351);
//  - j-forces are reduced over the 8 i-lanes with a DPP reduce-scatter of the
//    three components (8 VALU instructions) instead of warp shuffles
//    (:677, nbnxm_cuda_kernel_utils.cuh:467-496);
//  - forces are accumulated as int64 fixed point (2^-32 units), so every
//    evaluation is bit-reproducible whatever the atomic order;
//  - energies and shift forces are reduced in FP64 per entry and written to a
//    per-entry slot, then summed per system in a fixed order by a second kernel.
#include "nbnxm_kernel_body.h"

#include <hipcub/hipcub.hpp>

namespace nbnxm_amd
{

namespace
{

using namespace dev;

// FP64 pair math with exact erf/exp (the GPU oracle mode), fixed-point forces.
template<bool computeEnergy, int vdwType>
__global__ __launch_bounds__(256) void nbnxmNonbondedKernelFp64(const NbKernelArgs a, int wavesPerBlock)
{
    const int lane = threadIdx.x & 63;
    const int ti = lane & 7;
    const int tj = lane >> 3;
    const int work = blockIdx.x * wavesPerBlock + (threadIdx.x >> 6);
    if (work >= a.nsci)
    {
        return;
    }
    const int entry = __builtin_amdgcn_readfirstlane(a.sciOrder ? a.sciOrder[work] : work);
    const nbnxm_sci_t s = a.sci[entry];
    const int sys = s.shift / NBNXM_NUM_SHIFTS;
    const int shiftInSys = s.shift - sys * NBNXM_NUM_SHIFTS;
    const bool centralShift = (shiftInSys == NBNXM_CENTRAL_SHIFT);
    const double* pD = a.sysParamsD + 8 * sys;
    const double rcoulomb_sq = pD[0];
    const double epsfac = pD[1];
    const double beta = pD[2];
    const double ewaldShift = pD[3];
    const double cpotRep = pD[4];
    const double cpotDisp = pD[5];
    const float4 sh = a.shiftVec[s.shift];

    double eLJ = 0.0, eEl = 0.0;
    if constexpr (computeEnergy)
    {
        if (centralShift && s.cjPackedEnd > s.cjPackedBegin && a.cjPacked[s.cjPackedBegin].cj[0] == s.sci * c_ncps && tj == 0)
        {
            double q2 = 0.0;
            for (int c = 0; c < c_ncps; c++)
            {
                const double q = a.xq[(s.sci * c_ncps + c) * c_cl + ti].w;
                q2 += q * q;
            }
            eEl -= epsfac * q2 * beta * c_oneSqrtPi;
        }
    }
    const bool nonSelfLane = !(centralShift && tj <= ti);
    double fsX = 0.0, fsY = 0.0, fsZ = 0.0;
    double fci[c_ncps][3];
    for (int c = 0; c < c_ncps; c++)
    {
        fci[c][0] = fci[c][1] = fci[c][2] = 0.0;
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
            double fcj[3] = { 0.0, 0.0, 0.0 };
            for (int c = 0; c < c_ncps; c++)
            {
                const int bit = jm * c_ncps + c;
                if (!(imask & (1u << bit)))
                {
                    continue;
                }
                const int ai = (s.sci * c_ncps + c) * c_cl + ti;
                const float4 xqi = a.xq[ai];
                const double dx = (static_cast<double>(xqi.x) + sh.x) - xqj.x;
                const double dy = (static_cast<double>(xqi.y) + sh.y) - xqj.y;
                const double dz = (static_cast<double>(xqi.z) + sh.z) - xqj.z;
                double r2 = dx * dx + dy * dy + dz * dz;
                const bool notDiag = (s.sci * c_ncps + c) != cj;
                if (r2 < rcoulomb_sq && (nonSelfLane || notDiag))
                {
                    const double int_bit = ((wexcl >> bit) & 1u) ? 1.0 : 0.0;
                    double c6, c12;
                    if constexpr (vdwType == NBNXM_VDW_CUT_TYPE_TABLE)
                    {
                        const float2 p = a.nbfp[a.numTypes * a.atomTypes[ai] + a.atomTypes[aj]];
                        c6 = p.x;
                        c12 = p.y;
                    }
                    else if constexpr (vdwType == NBNXM_VDW_CUT_COMB_GEOM)
                    {
                        const float2 li = a.ljComb[ai];
                        const float2 lj = a.ljComb[aj];
                        c6 = static_cast<double>(li.x) * lj.x;
                        c12 = static_cast<double>(li.y) * lj.y;
                    }
                    else
                    {
                        const float2 li = a.ljComb[ai];
                        const float2 lj = a.ljComb[aj];
                        const double sigma = static_cast<double>(li.x) + lj.x;
                        const double epsilon = static_cast<double>(li.y) * lj.y;
                        const double s2 = sigma * sigma;
                        const double s6 = s2 * s2 * s2;
                        c6 = epsilon * s6;
                        c12 = c6 * s6;
                    }
                    r2 = fmax(r2, static_cast<double>(NBNXM_MIN_DISTANCE_SQUARED));
                    const double r = sqrt(r2);
                    const double inv_r = 1.0 / r;
                    const double inv_r2 = inv_r * inv_r;
                    const double inv_r6 = int_bit * inv_r2 * inv_r2 * inv_r2;
                    double F_invr = inv_r6 * (c12 * inv_r6 - c6) * inv_r2;
                    const double qq = static_cast<double>(xqi.w) * epsfac * xqj.w;
                    const double br = beta * r;
                    const double erfbr = erf(br);
                    F_invr += qq * (int_bit * inv_r2 * inv_r + 2.0 * beta * c_oneSqrtPi * exp(-br * br) * inv_r2 - erfbr * inv_r2 * inv_r);
                    if constexpr (computeEnergy)
                    {
                        eLJ += int_bit * (c12 * (inv_r6 * inv_r6 + cpotRep) / 12.0 - c6 * (inv_r6 + cpotDisp) / 6.0);
                        eEl += qq * (inv_r * (int_bit - erfbr) - int_bit * ewaldShift);
                    }
                    fcj[0] -= dx * F_invr;
                    fcj[1] -= dy * F_invr;
                    fcj[2] -= dz * F_invr;
                    fci[c][0] += dx * F_invr;
                    fci[c][1] += dy * F_invr;
                    fci[c][2] += dz * F_invr;
                }
            }
            for (int d = 0; d < 3; d++)
            {
                double v = fcj[d];
                v += __shfl_xor(v, 1);
                v += __shfl_xor(v, 2);
                v += __shfl_xor(v, 4);
                fcj[d] = v;
            }
            if (ti < 3)
            {
                atomicAdd(a.fFixed + 3 * aj + ti, toFixedD(fcj[ti == 0 ? 0 : (ti == 1 ? 1 : 2)]));
            }
        }
    }
    for (int c = 0; c < c_ncps; c++)
    {
        double f3[3];
        for (int d = 0; d < 3; d++)
        {
            double v = fci[c][d];
            v += __shfl_xor(v, 8);
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            f3[d] = v;
        }
        const int ai = (s.sci * c_ncps + c) * c_cl + ti;
        if (tj < 3)
        {
            atomicAdd(a.fFixed + 3 * ai + tj, toFixedD(f3[tj == 0 ? 0 : (tj == 1 ? 1 : 2)]));
        }
        fsX += f3[0];
        fsY += f3[1];
        fsZ += f3[2];
    }
    if (computeEnergy || a.computeFshift)
    {
        const double keep = (tj == 0) ? 1.0 : 0.0;
        const double v0 = reduceWaveD(eLJ);
        const double v1 = reduceWaveD(eEl);
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

// Prune (src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel_pruneonly.cuh:104-343), split
// 1: one workgroup of 1, 2 or 4 waves per entry of the current rolling part
// (launchPrune chooses from the list's groups per entry), wave k of it
// checking packed groups k, k + waves, ... (an entry of a large box has
// hundreds of groups; one wave per entry would leave most SIMDs idle on lists
// of a few thousand entries and on rolling parts).  The
// list words of a wave's groups read 64 groups at a time with vector loads,
// groups with nothing to check skipped by a ballot, and the 32 j-atoms of the
// next group fetched into registers while the current one is checked, then
// staged in per-wave LDS, so no distance waits for a global load.  The
// distances and ballots are those of the reference.
constexpr int c_pruneWavesPerEntry = 4; // at most
// Waves per entry from the list's mean packed groups per entry (measured on
// fresh lists: 96,000-atom water, 26 groups per entry: 4 waves 32.9 us, 2
// waves 46.5; the 256-box ensemble, 14 groups per entry: 4 waves 500 us, 2
// waves 450, 1 wave 446).
constexpr double c_pruneGroupsFor4 = 20.0;
constexpr double c_pruneGroupsFor2 = 6.0;
constexpr int c_pruneMinWaves = 8192; // 8 per SIMD

template<bool haveFreshList>
__global__ __launch_bounds__(64 * c_pruneWavesPerEntry) void nbnxmPruneKernel(const PruneArgs a)
{
    const int lane = threadIdx.x & 63;
    const int ti = lane & 7;
    const int tj = lane >> 3;
    const int waveInEntry = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wavesPerEntry = __builtin_amdgcn_readfirstlane(blockDim.x >> 6);
    const int w = blockIdx.x;

    // The part index per block, as the reference keeps it (:119-127).  A
    // fresh list is pruned in one part: its index stays 0 (the counters are
    // reset after the prune), so the fresh kernel neither reads nor writes
    // it and its first dependent load is the entry's.
    int part = 0;
    if (!haveFreshList)
    {
        part = __builtin_amdgcn_readfirstlane(a.rollingPart[w]);
        __syncthreads();
        if (threadIdx.x == 0)
        {
            a.rollingPart[w] = (part + 1) % a.numParts;
        }
    }
    const int numInPart = (a.nsci - part + a.numParts - 1) / a.numParts;
    if (w >= numInPart)
    {
        return;
    }
    const int work = w * a.numParts + part;
    const int entry = __builtin_amdgcn_readfirstlane(a.sciOrder ? asConst(a.sciOrder)[work] : work);
    const nbnxm_sci_t s = loadSci(asConst(static_cast<const nbnxm_sci_t*>(a.sci)), entry);
    const float4 sh = loadF4(asConst(static_cast<const float4*>(a.shiftVec)), s.shift);
    float xi[c_ncps], yi[c_ncps], zi[c_ncps];
#pragma unroll
    for (int c = 0; c < c_ncps; c++)
    {
        const float4 v = a.xq[(s.sci * c_ncps + c) * c_cl + ti];
        xi[c] = v.x + sh.x;
        yi[c] = v.y + sh.y;
        zi[c] = v.z + sh.z;
    }
    // The wave's groups are g = begin + waveInEntry + wavesPerEntry k.  Their list words
    // are read 64 groups at a time, lane l holding group k0 + l (two vector
    // loads for the packed group's six words, one for its outer mask), and the
    // groups with bits to check are taken from a ballot: per group there is no
    // scalar load to wait for, only its j-atoms' xq, fetched into registers
    // while the group before it is checked and then staged in per-wave LDS.
    // The new masks collect in the lanes of their groups and are stored with
    // one vector store per chunk.  Each group is read and written by one wave
    // only.
    const int firstGroup = s.cjPackedBegin + waveInEntry;
    const int numGroups = (s.cjPackedEnd > firstGroup) ? (s.cjPackedEnd - firstGroup + wavesPerEntry - 1) / wavesPerEntry : 0;
    __shared__ float4 sXq[c_pruneWavesPerEntry][c_jg * c_cl];
    float4* stageXq = sXq[waveInEntry];
    auto stage = [&](const float4 v) {
        __builtin_amdgcn_wave_barrier();
        if (lane < 32)
        {
            stageXq[lane] = v;
        }
        __builtin_amdgcn_wave_barrier();
    };
    int prunedPairCount = 0;
    for (int k0 = 0; k0 < numGroups; k0 += 64)
    {
        const int myGroup = firstGroup + wavesPerEntry * (k0 + lane);
        const bool valid = k0 + lane < numGroups;
        int4 words = make_int4(0, 0, 0, 0);
        unsigned int inner = 0u, outer = 0u;
        if (valid)
        {
            // A packed group is 24 bytes: its j-cluster words are 8-byte aligned.
            const int2* const cj = reinterpret_cast<const int2*>(a.cjPacked[myGroup].cj);
            const int2 w01 = cj[0];
            const int2 w23 = cj[1];
            words = make_int4(w01.x, w01.y, w23.x, w23.y);
            inner = a.cjPacked[myGroup].imei[0].imask;
            if (!haveFreshList)
            {
                outer = a.imaskOuter[myGroup];
            }
        }
        // Bits still to check: every bit of a fresh list, else those of the
        // outer mask that the inner mask has dropped (:229-245).
        const unsigned int check = haveFreshList ? inner : (inner ^ outer);
        const unsigned long long todo = __builtin_amdgcn_ballot_w64(check != 0u);
        if (todo == 0ull)
        {
            continue;
        }
        // The j-atoms of the group in lane k: lanes 0-31 (and, repeating them,
        // 32-63) load atom lane & 7 of j-cluster (lane >> 3) & 3.
        auto prefetch = [&](int k) -> float4 {
            const int g0 = __builtin_amdgcn_readlane(words.x, k);
            const int g1 = __builtin_amdgcn_readlane(words.y, k);
            const int g2 = __builtin_amdgcn_readlane(words.z, k);
            const int g3 = __builtin_amdgcn_readlane(words.w, k);
            const int jc = (lane >> 3) & 3;
            const int cjl = (jc == 1) ? g1 : ((jc == 2) ? g2 : ((jc == 3) ? g3 : g0));
            return a.xq[cjl * c_cl + (lane & 7)];
        };
        unsigned int newInner = inner;
        unsigned int newOuter = outer;
        int k = __builtin_ctzll(todo);
        unsigned long long rest = todo & (todo - 1ull);
        stage(prefetch(k));
        while (true)
        {
            const int kn = rest ? static_cast<int>(__builtin_ctzll(rest)) : -1;
            float4 pf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (kn >= 0)
            {
                pf = prefetch(kn);
            }
            const unsigned int imaskCheck = static_cast<unsigned int>(__builtin_amdgcn_readlane(static_cast<int>(check), k));
            unsigned int imaskFull, imaskNew;
            if (haveFreshList)
            {
                imaskFull = imaskCheck;
                imaskNew = 0u;
            }
            else
            {
                imaskFull = static_cast<unsigned int>(__builtin_amdgcn_readlane(static_cast<int>(outer), k));
                imaskNew = static_cast<unsigned int>(__builtin_amdgcn_readlane(static_cast<int>(inner), k));
            }
#pragma unroll
            for (int jm = 0; jm < c_jg; jm++)
            {
                if (!(imaskCheck & (0xFFu << (jm * c_ncps))))
                {
                    continue;
                }
                const float4 xj = stageXq[jm * c_cl + tj];
#pragma unroll
                for (int c = 0; c < c_ncps; c++)
                {
                    const unsigned int mask_ji = 1u << (jm * c_ncps + c);
                    if (!(imaskCheck & mask_ji))
                    {
                        continue;
                    }
                    const float r2 = norm2Fixed(xi[c] - xj.x, yi[c] - xj.y, zi[c] - xj.z);
                    // rlistInner <= rlistOuter: a pair inside the inner radius
                    // keeps the cluster pair in both masks, so the outer test
                    // runs only for cluster pairs with none.
                    if (__builtin_amdgcn_ballot_w64(r2 < a.rlistInner_sq) != 0ull)
                    {
                        imaskNew |= mask_ji;
                    }
                    else if (haveFreshList && __builtin_amdgcn_ballot_w64(r2 < a.rlistOuter_sq) == 0ull)
                    {
                        imaskFull &= ~mask_ji;
                    }
                }
            }
            const bool mine = lane == k;
            newInner = mine ? imaskNew : newInner;
            if (haveFreshList)
            {
                newOuter = mine ? imaskFull : newOuter;
                prunedPairCount += __popc(imaskNew);
            }
            if (kn < 0)
            {
                break;
            }
            stage(pf);
            k = kn;
            rest &= rest - 1ull;
        }
        if (check != 0u)
        {
            a.cjPacked[myGroup].imei[0].imask = newInner;
            if (haveFreshList)
            {
                a.imaskOuter[myGroup] = newOuter;
            }
        }
    }
    if (haveFreshList)
    {
        // The entry's pruned cluster pairs over its waves, in wave order.
        __shared__ int sCount[c_pruneWavesPerEntry];
        if (lane == 0)
        {
            sCount[waveInEntry] = prunedPairCount;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            int total = 0;
            for (int k = 0; k < wavesPerEntry; k++)
            {
                total += sCount[k];
            }
            const int index = max(NBNXM_SCI_HISTOGRAM_SIZE - total - 1, 0);
            if (a.sciCount)
            {
                a.sciCount[work] = index;
            }
        }
    }
}

__device__ __forceinline__ void convertForce(int i, unsigned long long* fFixed, float* fOut, double* fOutDouble, int clear)
{
    const long long v = static_cast<long long>(fFixed[i]);
    const double d = static_cast<double>(v) * (1.0 / c_fixedScaleD);
    if (fOut)
    {
        fOut[i] = static_cast<float>(d);
    }
    if (fOutDouble)
    {
        fOutDouble[i] = d;
    }
    if (clear)
    {
        fFixed[i] = 0ull;
    }
}

__global__ void convertForcesKernel(unsigned long long* fFixed, int n3, float* fOut, double* fOutDouble, int clear)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3)
    {
        return;
    }
    convertForce(i, fFixed, fOut, fOutDouble, clear);
}

// One workgroup of c_reduceThreads per (system, slot): slot 0 = energies,
// slots 1..45 = shift vector s-1.  Thread t adds fields first..first+N-1 of
// the entry partials of list[k] for k = begin + t, begin + t + c_reduceThreads,
// ... in that order, four entries' loads issued before their adds; the
// workgroup then sums its waves' totals in wave order.  The order is fixed,
// so the sums are bitwise reproducible, and every launch path uses this
// function.
constexpr int c_reduceThreads = 256;

template<int N>
__device__ __forceinline__ void sumEntryPartials(const int* list, int begin, int end, const double* partials, int first, double (&acc)[N])
{
    constexpr int stride = c_reduceThreads;
    int k = begin + static_cast<int>(threadIdx.x);
    for (; k + 3 * stride < end; k += 4 * stride)
    {
        int e[4];
        double v[4][N];
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
            e[u] = list[k + u * stride];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
#pragma unroll
            for (int c = 0; c < N; c++)
            {
                v[u][c] = partials[8 * static_cast<size_t>(e[u]) + first + c];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
#pragma unroll
            for (int c = 0; c < N; c++)
            {
                acc[c] += v[u][c];
            }
        }
    }
    for (; k < end; k += stride)
    {
        const int e = list[k];
#pragma unroll
        for (int c = 0; c < N; c++)
        {
            acc[c] += partials[8 * static_cast<size_t>(e) + first + c];
        }
    }
    // Wave totals, then the workgroup's in wave order (thread 0 holds them).
    __shared__ double sWave[c_reduceThreads / 64][N];
#pragma unroll
    for (int c = 0; c < N; c++)
    {
        acc[c] = reduceWaveD(acc[c]);
    }
    if ((threadIdx.x & 63) == 0)
    {
#pragma unroll
        for (int c = 0; c < N; c++)
        {
            sWave[threadIdx.x >> 6][c] = acc[c];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
#pragma unroll
        for (int c = 0; c < N; c++)
        {
            double t = sWave[0][c];
            for (int w = 1; w < c_reduceThreads / 64; w++)
            {
                t += sWave[w][c];
            }
            acc[c] = t;
        }
    }
}

__device__ __forceinline__ void reduceEnergySlot(int sys,
                                                 int slot,
                                                 const double* partials,
                                                 const int* entryRanges,
                                                 const int* entryList,
                                                 const int* shiftRanges,
                                                 const int* shiftEntryList,
                                                 int computeFshift,
                                                 double* energiesOut,
                                                 double* fshiftOut)
{
    if (slot == 0)
    {
        double acc[2] = { 0.0, 0.0 };
        sumEntryPartials(entryList, entryRanges[sys], entryRanges[sys + 1], partials, 0, acc);
        if (threadIdx.x == 0)
        {
            energiesOut[2 * sys] = acc[0];
            energiesOut[2 * sys + 1] = acc[1];
        }
        return;
    }
    if (!computeFshift)
    {
        return;
    }
    const int sidx = sys * NBNXM_NUM_SHIFTS + slot - 1;
    double acc[3] = { 0.0, 0.0, 0.0 };
    sumEntryPartials(shiftEntryList, shiftRanges[sidx], shiftRanges[sidx + 1], partials, 2, acc);
    if (threadIdx.x == 0)
    {
        fshiftOut[3 * sidx + 0] = acc[0];
        fshiftOut[3 * sidx + 1] = acc[1];
        fshiftOut[3 * sidx + 2] = acc[2];
    }
}

__global__ __launch_bounds__(c_reduceThreads) void reduceEnergiesKernel(const double* partials,
                                                                        const int* entryRanges,
                                                                        const int* entryList,
                                                                        const int* shiftRanges,
                                                                        const int* shiftEntryList,
                                                                        int computeFshift,
                                                                        double* energiesOut,
                                                                        double* fshiftOut)
{
    reduceEnergySlot(blockIdx.x / (1 + NBNXM_NUM_SHIFTS), blockIdx.x % (1 + NBNXM_NUM_SHIFTS), partials, entryRanges, entryList, shiftRanges,
                     shiftEntryList, computeFshift, energiesOut, fshiftOut);
}

// The end of a batched step in one launch: blocks [0, convertBlocks) convert
// and clear the fixed-point forces (as convertForcesKernel), every wave of the
// blocks after them reduces one (system, slot) of the energies and shift
// forces (as reduceEnergiesKernel), and thread 0 of block 0 advances the step
// counter (as advanceStepKernel; the step's perturbation read it before the
// nonbonded launch).  One launch instead of three dependent ones.
__global__ __launch_bounds__(c_reduceThreads) void finishStepKernel(unsigned long long* fFixed,
                                                        int n3,
                                                        float* fOut,
                                                        double* fOutDouble,
                                                        int convertBlocks,
                                                        const double* partials,
                                                        const int* entryRanges,
                                                        const int* entryList,
                                                        const int* shiftRanges,
                                                        const int* shiftEntryList,
                                                        int numSystems,
                                                        int computeFshift,
                                                        double* energiesOut,
                                                        double* fshiftOut,
                                                        unsigned long long* step)
{
    if (blockIdx.x == 0 && threadIdx.x == 0 && step)
    {
        *step += 1ull;
    }
    if (static_cast<int>(blockIdx.x) < convertBlocks)
    {
        // Two components per thread: 16-byte loads and stores (the buffers
        // are allocation-aligned).
        const int i = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
        if (i + 1 < n3)
        {
            ulonglong2* const p = reinterpret_cast<ulonglong2*>(fFixed + i);
            const ulonglong2 v = *p;
            const double d0 = static_cast<double>(static_cast<long long>(v.x)) * (1.0 / c_fixedScaleD);
            const double d1 = static_cast<double>(static_cast<long long>(v.y)) * (1.0 / c_fixedScaleD);
            if (fOut)
            {
                *reinterpret_cast<float2*>(fOut + i) = make_float2(static_cast<float>(d0), static_cast<float>(d1));
            }
            if (fOutDouble)
            {
                *reinterpret_cast<double2*>(fOutDouble + i) = make_double2(d0, d1);
            }
            *p = make_ulonglong2(0ull, 0ull);
        }
        else if (i < n3)
        {
            convertForce(i, fFixed, fOut, fOutDouble, 1);
        }
        return;
    }
    // One workgroup per (system, slot) after the conversion blocks.
    const int pair = static_cast<int>(blockIdx.x) - convertBlocks;
    reduceEnergySlot(pair / (1 + NBNXM_NUM_SHIFTS), pair % (1 + NBNXM_NUM_SHIFTS), partials, entryRanges, entryList, shiftRanges, shiftEntryList,
                     computeFshift, energiesOut, fshiftOut);
}

// Single-sum virial, atom part: per block FP64 partial sums of x (x) f over the
// slots of one system (calc_x_times_f, src/gromacs/mdlib/calcvir.cpp:57-85).
// Block (b, sys) covers slots slotRanges[sys] + b*256 + t + k*blocksPerSystem*256;
// the wave and block reductions run in a fixed order, so the result is
// bitwise reproducible for fixed-point forces.  Filler slots are skipped.
__global__ __launch_bounds__(256) void virialPartialKernel(const float4* xq,
                                                           const unsigned long long* fFixed,
                                                           const float* fFloat,
                                                           const int* slotRanges,
                                                           int blocksPerSystem,
                                                           double* partials)
{
    const int sys = blockIdx.y;
    const int b = blockIdx.x;
    const int begin = slotRanges[sys];
    const int end = slotRanges[sys + 1];
    double acc[9];
#pragma unroll
    for (int c = 0; c < 9; c++)
    {
        acc[c] = 0.0;
    }
    for (int s = begin + b * 256 + static_cast<int>(threadIdx.x); s < end; s += blocksPerSystem * 256)
    {
        const float4 p = xq[s];
        if (p.x < -1.0e5f)
        {
            continue;
        }
        double f[3];
#pragma unroll
        for (int d = 0; d < 3; d++)
        {
            f[d] = fFixed ? static_cast<double>(static_cast<long long>(fFixed[3 * static_cast<size_t>(s) + d])) * (1.0 / c_fixedScaleD)
                          : static_cast<double>(fFloat[3 * static_cast<size_t>(s) + d]);
        }
        const double x[3] = { p.x, p.y, p.z };
#pragma unroll
        for (int d = 0; d < 3; d++)
        {
#pragma unroll
            for (int n = 0; n < 3; n++)
            {
                acc[3 * d + n] = fma(x[d], f[n], acc[3 * d + n]);
            }
        }
    }
    __shared__ double red[4][9];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < 9; c++)
    {
        const double v = reduceWaveD(acc[c]);
        if (lane == 0)
        {
            red[wave][c] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < 9)
    {
        const int c = threadIdx.x;
        partials[(static_cast<size_t>(sys) * blocksPerSystem + b) * 9 + c] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
    }
}

// Virial per system: -1/2 (sum_i x_i (x) f_i + sum_s shift_s (x) fshift_s)
// (calc_vir on shift forces and on forces, src/gromacs/mdlib/calcvir.cpp:89-133,
// as api/nblib/virials.cpp:55-84 combines them).  One block per system.
__global__ __launch_bounds__(64) void virialFinishKernel(const double* partials,
                                                         int blocksPerSystem,
                                                         const float4* shiftVec,
                                                         const double* fshift,
                                                         double* virial)
{
    const int sys = blockIdx.x;
    const int c = threadIdx.x;
    if (c >= 9)
    {
        return;
    }
    const int d = c / 3;
    const int n = c % 3;
    double xf = 0.0;
    for (int b = 0; b < blocksPerSystem; b++)
    {
        xf += partials[(static_cast<size_t>(sys) * blocksPerSystem + b) * 9 + c];
    }
    double sf = 0.0;
    for (int s = 0; s < NBNXM_NUM_SHIFTS; s++)
    {
        const float4 sv = shiftVec[sys * NBNXM_NUM_SHIFTS + s];
        const double svd = (d == 0) ? sv.x : ((d == 1) ? sv.y : sv.z);
        sf = fma(svd, fshift[3 * (sys * NBNXM_NUM_SHIFTS + s) + n], sf);
    }
    virial[9 * sys + c] = -0.5 * (sf + xf);
}

// Coordinate buffer op: xq.xyz of every real grid slot from the coordinates
// in atom order (nbnxn_gpu_x_to_nbat_x_kernel,
// src/gromacs/nbnxm/cuda/nbnxm_gpu_buffer_ops_internal.cu:66-105); the charge
// (w) and the filler slots are left as they are.
__global__ void xToXqKernel(const float* x, const int* atomIndices, int numSlots, float4* xq)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= numSlots)
    {
        return;
    }
    const int atom = atomIndices[slot];
    if (atom < 0)
    {
        return;
    }
    const float w = xq[slot].w;
    xq[slot] = make_float4(x[3 * static_cast<size_t>(atom)], x[3 * static_cast<size_t>(atom) + 1], x[3 * static_cast<size_t>(atom) + 2], w);
}

// Grid-order atom data of a list built on the device (the host's fillXq,
// fillTypes and fillLjComb): fillers at -1e6 with q = 0, the dummy type and
// zero combination parameters; also the atom -> slot map of the force
// reduction.  slotTypes / slotComb may be null.
__global__ void gatherAtomDataKernel(const int* atomIndices,
                                     int numSlots,
                                     const float4* xw,
                                     const float* q,
                                     const int* types,
                                     const float2* typeComb,
                                     int dummyType,
                                     float4* xq,
                                     int* slotTypes,
                                     float2* slotComb,
                                     int* cells)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= numSlots)
    {
        return;
    }
    const int atom = atomIndices[slot];
    if (atom < 0)
    {
        constexpr float farAway = -1000000.0f;
        xq[slot] = make_float4(farAway, farAway, farAway, 0.0f);
        if (slotTypes)
        {
            slotTypes[slot] = dummyType;
        }
        if (slotComb)
        {
            slotComb[slot] = make_float2(0.0f, 0.0f);
        }
        return;
    }
    const float4 p = xw[atom];
    xq[slot] = make_float4(p.x, p.y, p.z, q[atom]);
    if (slotTypes)
    {
        slotTypes[slot] = types[atom];
    }
    if (slotComb)
    {
        slotComb[slot] = typeComb[types[atom]];
    }
    cells[atom] = slot;
}

// Wrapped coordinates back into a 3-float-per-atom array.
__global__ void writeCoordinatesKernel(const float4* xw, int numAtoms, float* x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < numAtoms)
    {
        const float4 p = xw[i];
        x[3 * static_cast<size_t>(i)] = p.x;
        x[3 * static_cast<size_t>(i) + 1] = p.y;
        x[3 * static_cast<size_t>(i) + 2] = p.z;
    }
}

// Force reduction into atom order (reduceKernel,
// src/gromacs/mdlib/gpuforcereduction_impl_internal.cu:62-120): f[a] is set
// to (or, with accumulate, incremented by) the nonbonded force of slot
// cell[a], then fToAdd[a] is added when given.  Fixed-point forces are
// converted exactly and their accumulator slots cleared for the next launch.
template<bool fixedPoint, bool accumulate, bool addRvec>
__global__ void reduceForcesKernel(unsigned long long* fFixed, const float* fFloat, const int* cells, int numAtoms, const float* fToAdd, float* f)
{
    const int atom = blockIdx.x * blockDim.x + threadIdx.x;
    if (atom >= numAtoms)
    {
        return;
    }
    const size_t slot = static_cast<size_t>(cells[atom]);
#pragma unroll
    for (int d = 0; d < 3; d++)
    {
        float v;
        if constexpr (fixedPoint)
        {
            v = static_cast<float>(static_cast<double>(static_cast<long long>(fFixed[3 * slot + d])) * (1.0 / c_fixedScaleD));
            fFixed[3 * slot + d] = 0ull;
        }
        else
        {
            v = fFloat[3 * slot + d];
        }
        const size_t k = 3 * static_cast<size_t>(atom) + d;
        float out = accumulate ? f[k] + v : v;
        if constexpr (addRvec)
        {
            out += fToAdd[k];
        }
        f[k] = out;
    }
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), used as a counter-based
// generator: key = seed, counter = (atom, step).
__device__ __forceinline__ uint4 philox4x32_10(uint4 ctr, uint2 key)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++)
    {
        const uint32_t hi0 = __umulhi(M0, ctr.x);
        const uint32_t lo0 = M0 * ctr.x;
        const uint32_t hi1 = __umulhi(M1, ctr.z);
        const uint32_t lo1 = M1 * ctr.z;
        ctr = make_uint4(hi1 ^ ctr.y ^ key.x, lo1, hi0 ^ ctr.w ^ key.y, lo0);
        key.x += W0;
        key.y += W1;
    }
    return ctr;
}

__device__ __forceinline__ float uniformOpen(uint32_t u)
{
    return (static_cast<float>(u >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

__global__ void perturbKernel(const float4* x0, float4* xq, int first, int count, float sigma, uint2 key, const unsigned long long* step)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count)
    {
        return;
    }
    const int i = first + k;
    float4 v = x0[i];
    if (v.x > -1.0e5f && sigma > 0.0f)
    {
        const unsigned long long st = *step;
        const uint4 r = philox4x32_10(make_uint4(static_cast<uint32_t>(i), 0u, static_cast<uint32_t>(st), static_cast<uint32_t>(st >> 32)), key);
        const float u1 = uniformOpen(r.x), u2 = uniformOpen(r.y), u3 = uniformOpen(r.z), u4 = uniformOpen(r.w);
        const float m1 = sqrtf(-2.0f * logf(u1));
        const float m2 = sqrtf(-2.0f * logf(u3));
        const float a1 = 6.283185307179586f * u2;
        const float a2 = 6.283185307179586f * u4;
        v.x += sigma * m1 * cosf(a1);
        v.y += sigma * m1 * sinf(a1);
        v.z += sigma * m2 * cosf(a2);
    }
    xq[i] = v;
}

__global__ void advanceStepKernel(unsigned long long* step)
{
    if (threadIdx.x == 0)
    {
        *step += 1ull;
    }
}

__global__ void fillU64Kernel(unsigned long long* p, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
    {
        p[i] = 0ull;
    }
}

__global__ void fillI32Kernel(int* p, size_t n, int value)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
    {
        p[i] = value;
    }
}

// Work-order arrays of a list whose entries are on the device: 0..n-1 into
// three arrays, each entry's shift index as a sort key, and {0, n}.
__global__ void entryArraysKernel(const nbnxm_sci_t* sci, int n, int* entryList, int* sciOrder, int* iota, int* shiftKeys, int* entryRanges)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
    {
        entryList[i] = i;
        sciOrder[i] = i;
        iota[i] = i;
        shiftKeys[i] = sci[i].shift;
    }
    if (i == 0)
    {
        entryRanges[0] = 0;
        entryRanges[1] = n;
    }
}

// shiftRanges[k] = number of sorted keys below k, k = 0..numShifts.
__global__ void shiftRangesKernel(const int* sortedKeys, int n, int numShifts, int* shiftRanges)
{
    const int k = threadIdx.x;
    if (k > numShifts)
    {
        return;
    }
    int lo = 0, hi = n;
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if (sortedKeys[mid] < k)
        {
            lo = mid + 1;
        }
        else
        {
            hi = mid;
        }
    }
    shiftRanges[k] = lo;
}

__global__ void storeInt2Kernel(int* p, int a, int b)
{
    p[0] = a;
    p[1] = b;
}

template<bool E, int V>
hipError_t launchFp64(const NbKernelArgs& a, const NbLaunchConfig& cfg, hipStream_t stream)
{
    const int wpb = cfg.wavesPerBlock;
    const int blocks = (a.nsci + wpb - 1) / wpb;
    if (blocks == 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL((nbnxmNonbondedKernelFp64<E, V>), dim3(blocks), dim3(64 * wpb), 0, stream, a, wpb);
    return hipGetLastError();
}

hipError_t launchFp64Vdw(const NbKernelArgs& a, int vdwType, bool computeEnergy, const NbLaunchConfig& cfg, hipStream_t stream)
{
    auto vdw = [&](auto e) -> hipError_t {
        constexpr bool E = decltype(e)::value;
        switch (vdwType)
        {
            case NBNXM_VDW_CUT_TYPE_TABLE: return launchFp64<E, NBNXM_VDW_CUT_TYPE_TABLE>(a, cfg, stream);
            case NBNXM_VDW_CUT_COMB_GEOM: return launchFp64<E, NBNXM_VDW_CUT_COMB_GEOM>(a, cfg, stream);
            case NBNXM_VDW_CUT_COMB_LB: return launchFp64<E, NBNXM_VDW_CUT_COMB_LB>(a, cfg, stream);
            default: return hipErrorInvalidValue;
        }
    };
    return computeEnergy ? vdw(std::true_type{}) : vdw(std::false_type{});
}

} // namespace

hipError_t launchNonbonded(const NbKernelArgs& a, int accumulation, int vdwType, bool computeEnergy, const NbLaunchConfig& cfg, hipStream_t stream,
                           const NbFlavourParams* flav)
{
    if (cfg.wavesPerBlock < 1 || cfg.wavesPerBlock > 4)
    {
        return hipErrorInvalidValue;
    }
    if (flav)
    {
        // The flavour-class instantiations read the LJ-PME grid C6 by type pair: type table only.
        if (accumulation == NBNXM_ACCUM_FP64 || (flav->ljPme && (vdwType != NBNXM_VDW_CUT_TYPE_TABLE || !flav->ljPmeGrid))
            || (flav->tabulatedEwald && (!flav->ewaldTable || !(flav->ewaldTableScale > 0))))
        {
            return hipErrorInvalidValue;
        }
        return launchNonbondedFlavoured(a, *flav, accumulation, vdwType, computeEnergy, cfg, stream);
    }
    if (accumulation == NBNXM_ACCUM_FP64)
    {
        if (a.pruneCjPacked)
        {
            return hipErrorInvalidValue; // the FP64 mode prunes with the prune kernel
        }
        return launchFp64Vdw(a, vdwType, computeEnergy, cfg, stream);
    }
    return launchTunedFlavour<c_flavTuned>(a, NbFlavourParams{}, accumulation, vdwType, computeEnergy, cfg, stream);
}

hipError_t launchPrune(const PruneArgs& a, bool haveFreshList, hipStream_t stream)
{
    // One workgroup per entry of the largest part (a.rollingPart holds at
    // least that many part counters).
    const int blocks = (a.nsci + a.numParts - 1) / a.numParts;
    if (blocks == 0)
    {
        return hipSuccess;
    }
    // Waves per entry from the mean groups per entry: enough groups per wave
    // that its fixed costs (i-cluster loads, list words, the entry's count)
    // stay small, enough waves that a list of a few thousand entries fills
    // the GPU; and at least enough for c_pruneMinWaves waves in the launch (a
    // rolling part holds a fraction of the entries: 1,226 of a 6,127-entry
    // list in 5 parts took 8.6 us at 2 waves per entry, 6.8 us at 4).
    const double groupsPerEntry = static_cast<double>(a.numGroups) / a.nsci;
    const int byGroups = groupsPerEntry >= c_pruneGroupsFor4 ? 4 : (groupsPerEntry >= c_pruneGroupsFor2 ? 2 : 1);
    const int toFill = (c_pruneMinWaves + blocks - 1) / blocks;
    const int waves = std::min(c_pruneWavesPerEntry, std::max(byGroups, toFill));
    if (haveFreshList)
    {
        hipLaunchKernelGGL(nbnxmPruneKernel<true>, dim3(blocks), dim3(64 * waves), 0, stream, a);
    }
    else
    {
        hipLaunchKernelGGL(nbnxmPruneKernel<false>, dim3(blocks), dim3(64 * waves), 0, stream, a);
    }
    return hipGetLastError();
}

hipError_t launchConvertForces(unsigned long long* fFixed, int numAtoms, float* fOut, double* fOutDouble, hipStream_t stream, bool clear)
{
    const int n3 = 3 * numAtoms;
    if (n3 == 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL(convertForcesKernel, dim3((n3 + 255) / 256), dim3(256), 0, stream, fFixed, n3, fOut, fOutDouble, clear ? 1 : 0);
    return hipGetLastError();
}

hipError_t launchConvertForcesRange(unsigned long long* fFixed, int first, int count, float* fOut, double* fOutDouble, hipStream_t stream)
{
    if (count <= 0)
    {
        return hipSuccess;
    }
    const size_t off = 3 * static_cast<size_t>(first);
    return launchConvertForces(fFixed + off, count, fOut ? fOut + off : nullptr, fOutDouble ? fOutDouble + off : nullptr, stream);
}

hipError_t launchReduceEnergies(const double* entryPartials,
                                const int* entryRanges,
                                const int* entryList,
                                const int* shiftRanges,
                                const int* shiftEntryList,
                                int numSystems,
                                double* energiesOut,
                                double* fshiftOut,
                                int computeFshift,
                                hipStream_t stream)
{
    if (numSystems == 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL(reduceEnergiesKernel, dim3(numSystems * (1 + NBNXM_NUM_SHIFTS)), dim3(c_reduceThreads), 0, stream, entryPartials, entryRanges,
                       entryList, shiftRanges, shiftEntryList, computeFshift, energiesOut, fshiftOut);
    return hipGetLastError();
}

hipError_t launchFinishStep(unsigned long long* fFixed,
                            int numAtoms,
                            float* fOut,
                            double* fOutDouble,
                            const double* entryPartials,
                            const int* entryRanges,
                            const int* entryList,
                            const int* shiftRanges,
                            const int* shiftEntryList,
                            int numSystems,
                            double* energiesOut,
                            double* fshiftOut,
                            int computeFshift,
                            unsigned long long* step,
                            hipStream_t stream)
{
    const int n3 = fFixed ? 3 * numAtoms : 0;
    const int convertBlocks = (n3 + 2 * c_reduceThreads - 1) / (2 * c_reduceThreads); // two components per thread
    const int reduceBlocks = numSystems * (1 + NBNXM_NUM_SHIFTS);
    const int blocks = convertBlocks + reduceBlocks;
    if (blocks == 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL(finishStepKernel, dim3(blocks), dim3(c_reduceThreads), 0, stream, fFixed, n3, fOut, fOutDouble, convertBlocks, entryPartials,
                       entryRanges, entryList, shiftRanges, shiftEntryList, numSystems, computeFshift, energiesOut, fshiftOut, step);
    return hipGetLastError();
}

hipError_t launchVirialPartials(const float4* xq,
                                const unsigned long long* fFixed,
                                const float* fFloat,
                                const int* slotRanges,
                                int numSystems,
                                int blocksPerSystem,
                                double* partials,
                                hipStream_t stream)
{
    if (numSystems <= 0 || blocksPerSystem <= 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL(virialPartialKernel, dim3(blocksPerSystem, numSystems), dim3(256), 0, stream, xq, fFixed, fFloat, slotRanges, blocksPerSystem, partials);
    return hipGetLastError();
}

hipError_t launchVirialFinish(const double* partials, int numSystems, int blocksPerSystem, const float4* shiftVec, const double* fshift, double* virial, hipStream_t stream)
{
    if (numSystems <= 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL(virialFinishKernel, dim3(numSystems), dim3(64), 0, stream, partials, blocksPerSystem, shiftVec, fshift, virial);
    return hipGetLastError();
}

hipError_t launchGatherAtomData(const int* atomIndices,
                                int numSlots,
                                const float4* xw,
                                const float* q,
                                const int* types,
                                const float2* typeComb,
                                int dummyType,
                                float4* xq,
                                int* slotTypes,
                                float2* slotComb,
                                int* cells,
                                hipStream_t stream)
{
    if (numSlots <= 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL(gatherAtomDataKernel, dim3((numSlots + 255) / 256), dim3(256), 0, stream, atomIndices, numSlots, xw, q, types, typeComb, dummyType, xq,
                       slotTypes, slotComb, cells);
    return hipGetLastError();
}

hipError_t launchWriteCoordinates(const float4* xw, int numAtoms, float* x, hipStream_t stream)
{
    if (numAtoms <= 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL(writeCoordinatesKernel, dim3((numAtoms + 255) / 256), dim3(256), 0, stream, xw, numAtoms, x);
    return hipGetLastError();
}

hipError_t launchXToXq(const float* x, const int* atomIndices, int numSlots, float4* xq, hipStream_t stream)
{
    if (numSlots <= 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL(xToXqKernel, dim3((numSlots + 255) / 256), dim3(256), 0, stream, x, atomIndices, numSlots, xq);
    return hipGetLastError();
}

hipError_t launchReduceForces(unsigned long long* fFixed, const float* fFloat, const int* cells, int numAtoms, const float* fToAdd, float* f, bool accumulate, hipStream_t stream)
{
    if (numAtoms <= 0)
    {
        return hipSuccess;
    }
    const dim3 grid((numAtoms + 255) / 256), block(256);
    const bool fixedPoint = fFixed != nullptr;
    const bool addRvec = fToAdd != nullptr;
    auto launch = [&](auto kernel) { hipLaunchKernelGGL(kernel, grid, block, 0, stream, fFixed, fFloat, cells, numAtoms, fToAdd, f); };
    if (fixedPoint)
    {
        if (accumulate)
        {
            addRvec ? launch(reduceForcesKernel<true, true, true>) : launch(reduceForcesKernel<true, true, false>);
        }
        else
        {
            addRvec ? launch(reduceForcesKernel<true, false, true>) : launch(reduceForcesKernel<true, false, false>);
        }
    }
    else
    {
        if (accumulate)
        {
            addRvec ? launch(reduceForcesKernel<false, true, true>) : launch(reduceForcesKernel<false, true, false>);
        }
        else
        {
            addRvec ? launch(reduceForcesKernel<false, false, true>) : launch(reduceForcesKernel<false, false, false>);
        }
    }
    return hipGetLastError();
}

hipError_t launchPerturb(const float4* x0, float4* xq, int first, int count, float sigma, uint64_t seed, const unsigned long long* step, hipStream_t stream)
{
    if (count <= 0)
    {
        return hipSuccess;
    }
    const uint2 key = make_uint2(static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
    hipLaunchKernelGGL(perturbKernel, dim3((count + 255) / 256), dim3(256), 0, stream, x0, xq, first, count, sigma, key, step);
    return hipGetLastError();
}

hipError_t launchAdvanceStep(unsigned long long* step, hipStream_t stream)
{
    hipLaunchKernelGGL(advanceStepKernel, dim3(1), dim3(64), 0, stream, step);
    return hipGetLastError();
}

hipError_t launchFillU64(unsigned long long* p, size_t n, hipStream_t stream)
{
    if (n == 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL(fillU64Kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t launchEntryArrays(const nbnxm_sci_t* sci, int n, int* entryList, int* sciOrder, int* iota, int* shiftKeys, int* entryRanges, hipStream_t stream)
{
    const int blocks = std::max(1, (n + 255) / 256);
    hipLaunchKernelGGL(entryArraysKernel, dim3(blocks), dim3(256), 0, stream, sci, n, entryList, sciOrder, iota, shiftKeys, entryRanges);
    return hipGetLastError();
}

hipError_t launchShiftRanges(const int* sortedKeys, int n, int* shiftRanges, hipStream_t stream)
{
    static_assert(NBNXM_NUM_SHIFTS + 1 <= 64, "one wave covers the ranges");
    hipLaunchKernelGGL(shiftRangesKernel, dim3(1), dim3(64), 0, stream, sortedKeys, n, NBNXM_NUM_SHIFTS, shiftRanges);
    return hipGetLastError();
}

hipError_t launchStoreInt2(int* p, int a, int b, hipStream_t stream)
{
    hipLaunchKernelGGL(storeInt2Kernel, dim3(1), dim3(1), 0, stream, p, a, b);
    return hipGetLastError();
}

hipError_t launchFillI32(int* p, size_t n, int value, hipStream_t stream)
{
    if (n == 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL(fillI32Kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, p, n, value);
    return hipGetLastError();
}

} // namespace nbnxm_amd
