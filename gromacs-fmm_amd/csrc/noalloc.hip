// nbnxm_gpu_launch_kernel_noalloc: the nonbonded evaluation of
// nbnxm_gpu_launch_kernel and the energy / shift-force reduction of
// nbnxm_gpu_launch_cpyback on device buffers the caller owns, for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  Nothing here allocates: the launch's
// scratch (its parameter blocks and the FP64 energy / shift-force partials of
// every pair-list entry) lives in one file-scope device array of fixed size,
// whose device address is looked up once per device.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "capi_common.h"
#include "device_common.h"
#include "nbnxm_kernels.h"

namespace nbnxm_amd
{

namespace
{

using namespace dev;

// What one launch reads as its per-system parameter blocks
// (NbKernelArgs::sysParamsA/B/D of a single system).
struct ParamBlocks
{
    float4 a;
    float2 b;
    double d[8];
};

// Scratch layout: the parameter blocks, then 8 doubles per pair-list entry.
constexpr size_t c_scratchParamBytes = 256;
constexpr size_t c_scratchBytes = c_scratchParamBytes + 8 * sizeof(double) * static_cast<size_t>(NBNXM_NOALLOC_MAX_ENTRIES);
static_assert(sizeof(ParamBlocks) <= c_scratchParamBytes, "parameter blocks exceed their scratch slot");

// The scratch array.  Each GPU holds its own copy until the process ends; in
// host code this name stands for the array but is not its device address
// (that comes from hipGetSymbolAddress below, per device).
static __device__ __align__(256) unsigned char scratch[c_scratchBytes];

constexpr int c_maxDevices = 64;
std::mutex g_scratchMutex;
void* g_scratchAddress[c_maxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call on that device with it current and kept.
int scratchAddress(int device, unsigned char** out)
{
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchAddress[device])
    {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
        {
            return hipFail(e, "gpu_launch_kernel_noalloc: hipGetSymbolAddress(scratch)");
        }
        if (!p)
        {
            return fail(NBNXM_ERR_HIP, "gpu_launch_kernel_noalloc: no device address for the scratch array");
        }
        g_scratchAddress[device] = p;
    }
    *out = static_cast<unsigned char*>(g_scratchAddress[device]);
    return NBNXM_OK;
}

// Parameter blocks from kernel arguments: no host-to-device copy, so nothing
// pageable is read asynchronously and a capture holds kernels only.
__global__ void writeParamBlocksKernel(ParamBlocks* dst, ParamBlocks src)
{
    if (threadIdx.x == 0)
    {
        *dst = src;
    }
}

// The sums of reduceEnergiesKernel (nbnxm_kernels.hip) for one system whose
// entry lists the host never saw: slot 0 adds the energies of all entries in
// entry order, slot 1 + s the shift forces of the entries with shift s in
// entry order.  As there, one workgroup of 256 threads per slot: thread t
// adds positions t, t + 256, ... of the slot's list in that order, each wave
// sums its threads with reduceWaveD, and thread 0 adds the four wave totals
// in wave order, so the sums are the same bits.  The positions of shift s
// are found chunk by chunk (256 entries): a ballot per wave, the waves'
// match counts through LDS, and each match handed to the thread of its
// position.
constexpr int c_slotThreads = 256;

template<int N>
__device__ __forceinline__ void finishSlotSum(double (&acc)[N], double* out)
{
    __shared__ double sWave[c_slotThreads / 64][N];
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
            for (int w = 1; w < c_slotThreads / 64; w++)
            {
                t += sWave[w][c];
            }
            out[c] = t;
        }
    }
}

__global__ __launch_bounds__(c_slotThreads) void reduceEntriesInOrderKernel(const double* partials,
                                                                            const nbnxm_sci_t* sci,
                                                                            int nsci,
                                                                            int computeEnergy,
                                                                            int computeFshift,
                                                                            double* energies,
                                                                            double* fshift)
{
    const int slot = blockIdx.x;
    const int tid = threadIdx.x;
    if (slot == 0)
    {
        if (!computeEnergy)
        {
            return;
        }
        double acc[2] = { 0.0, 0.0 };
        for (int k = tid; k < nsci; k += c_slotThreads)
        {
            acc[0] += partials[8 * static_cast<size_t>(k)];
            acc[1] += partials[8 * static_cast<size_t>(k) + 1];
        }
        finishSlotSum(acc, energies);
        return;
    }
    if (!computeFshift)
    {
        return;
    }
    const int s = slot - 1;
    const int wave = tid >> 6;
    __shared__ double sv[3][c_slotThreads];
    __shared__ int sFilled[c_slotThreads];
    __shared__ int sWaveMatches[c_slotThreads / 64];
    sFilled[tid] = 0;
    __syncthreads();
    double acc[3] = { 0.0, 0.0, 0.0 };
    int count = 0; // positions of shift s before this chunk (uniform)
    for (int base = 0; base < nsci; base += c_slotThreads)
    {
        const int e = base + tid;
        const bool match = e < nsci && sci[e].shift == s;
        const unsigned long long m = __builtin_amdgcn_ballot_w64(match);
        if ((tid & 63) == 0)
        {
            sWaveMatches[wave] = __popcll(m);
        }
        __syncthreads();
        int before = 0, total = 0;
        for (int w = 0; w < c_slotThreads / 64; w++)
        {
            before += (w < wave) ? sWaveMatches[w] : 0;
            total += sWaveMatches[w];
        }
        if (match)
        {
            const int rank = static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<unsigned int>(m >> 32),
                                                                        __builtin_amdgcn_mbcnt_lo(static_cast<unsigned int>(m), 0u)));
            // At most c_slotThreads matches per chunk: distinct threads.
            const int t = (count + before + rank) & (c_slotThreads - 1);
            sv[0][t] = partials[8 * static_cast<size_t>(e) + 2];
            sv[1][t] = partials[8 * static_cast<size_t>(e) + 3];
            sv[2][t] = partials[8 * static_cast<size_t>(e) + 4];
            sFilled[t] = 1;
        }
        __syncthreads();
        if (sFilled[tid])
        {
            acc[0] += sv[0][tid];
            acc[1] += sv[1][tid];
            acc[2] += sv[2][tid];
            sFilled[tid] = 0;
        }
        count += total;
        __syncthreads();
    }
    finishSlotSum(acc, fshift + 3 * s);
}

} // namespace

} // namespace nbnxm_amd

using namespace nbnxm_amd;

extern "C" int nbnxm_gpu_launch_kernel_noalloc(const nbnxm_interaction_const_t* ic,
                                               const nbnxm_device_inputs_t* in,
                                               int accumulation,
                                               int computeEnergy,
                                               int computeFshift,
                                               unsigned long long* d_fFixed,
                                               float* d_fFloat,
                                               double* d_energies,
                                               double* d_fshift,
                                               void* stream)
{
    if (!ic || !in)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_launch_kernel_noalloc: null pointer");
    }
    if (!isDefaultFlavour(*ic))
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "gpu_launch_kernel_noalloc: analytical Ewald + LJ potential shift with rvdw = rcoulomb only");
    }
    if (accumulation < NBNXM_ACCUM_FLOAT || accumulation > NBNXM_ACCUM_FP64)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_launch_kernel_noalloc: unknown accumulation mode");
    }
    if (in->nsci < 0 || in->nsci > NBNXM_NOALLOC_MAX_ENTRIES)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_launch_kernel_noalloc: nsci must be in [0, NBNXM_NOALLOC_MAX_ENTRIES = "
                                               + std::to_string(NBNXM_NOALLOC_MAX_ENTRIES) + "], got " + std::to_string(in->nsci));
    }
    if (in->numAtoms < 0 || in->numAtoms % NBNXM_CLUSTER_SIZE != 0)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_launch_kernel_noalloc: numAtoms must be a non-negative multiple of 8");
    }
    const bool fixed = accumulation != NBNXM_ACCUM_FLOAT;
    if ((fixed && !d_fFixed) || (!fixed && !d_fFloat) || (in->numAtoms > 0 && (!in->xq || !in->shiftVec))
        || (in->nsci > 0 && (!in->sci || !in->cjPacked || !in->excl)) || (computeEnergy && !d_energies) || (computeFshift && !d_fshift))
    {
        return fail(NBNXM_ERR_INVALID, "gpu_launch_kernel_noalloc: missing device array");
    }
    if (in->vdwType == NBNXM_VDW_CUT_TYPE_TABLE)
    {
        // The FP32 kernels stage the type table in LDS.
        if (in->numTypes < 1 || !in->type || !in->nbfp || (accumulation != NBNXM_ACCUM_FP64 && in->numTypes > c_maxLdsTableTypes))
        {
            return fail(NBNXM_ERR_INVALID, "gpu_launch_kernel_noalloc: the type table needs type, nbfp and 1..65 types");
        }
    }
    else if (in->vdwType == NBNXM_VDW_CUT_COMB_GEOM || in->vdwType == NBNXM_VDW_CUT_COMB_LB)
    {
        if (in->numAtoms > 0 && !in->ljComb)
        {
            return fail(NBNXM_ERR_INVALID, "gpu_launch_kernel_noalloc: a combination rule needs ljComb");
        }
    }
    else
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "gpu_launch_kernel_noalloc: unknown vdwType");
    }
    int device = 0;
    NBNXM_HIP_CHECK(hipGetDevice(&device));
    if (device < 0 || device >= c_maxDevices)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "gpu_launch_kernel_noalloc: device index beyond 63");
    }
    unsigned char* base = nullptr;
    int rc = scratchAddress(device, &base);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (in->nsci == 0)
    {
        // No entry: no forces; energies and shift forces are zero.
        if (computeEnergy)
        {
            NBNXM_HIP_CHECK(launchFillI32(reinterpret_cast<int*>(d_energies), 4, 0, st));
        }
        if (computeFshift)
        {
            NBNXM_HIP_CHECK(launchFillI32(reinterpret_cast<int*>(d_fshift), 6 * NBNXM_NUM_SHIFTS, 0, st));
        }
        return NBNXM_OK;
    }

    // Every launch rewrites the parameter blocks and, for the entries it
    // processes, all 8 partials of each entry, so nothing in the scratch needs
    // clearing; the forces are added to the caller's accumulator.
    ParamBlocks pbk;
    systemParamBlocks(*ic, &pbk.a, &pbk.b, pbk.d);
    ParamBlocks* blocks = reinterpret_cast<ParamBlocks*>(base);
    hipLaunchKernelGGL(writeParamBlocksKernel, dim3(1), dim3(64), 0, st, blocks, pbk);
    NBNXM_HIP_CHECK(hipGetLastError());

    NbKernelArgs a;
    a.xq = reinterpret_cast<const float4*>(in->xq);
    a.atomTypes = in->type;
    a.ljComb = reinterpret_cast<const float2*>(in->ljComb);
    a.nbfp = reinterpret_cast<const float2*>(in->nbfp);
    a.numTypes = in->numTypes;
    a.shiftVec = reinterpret_cast<const float4*>(in->shiftVec);
    a.sci = in->sci;
    a.sciOrder = nullptr;
    a.nsci = in->nsci;
    a.cjPacked = in->cjPacked;
    a.excl = in->excl;
    a.sysParamsA = &blocks->a;
    a.sysParamsB = &blocks->b;
    a.sysParamsD = blocks->d;
    a.fFloat = d_fFloat;
    a.fFixed = d_fFixed;
    a.entryPartials = reinterpret_cast<double*>(base + c_scratchParamBytes);
    a.computeFshift = computeFshift ? 1 : 0;
    const NbLaunchConfig cfg;
    NBNXM_HIP_CHECK(launchNonbonded(a, accumulation, in->vdwType, computeEnergy != 0, cfg, st));
    if (computeEnergy || computeFshift)
    {
        hipLaunchKernelGGL(reduceEntriesInOrderKernel, dim3(1 + NBNXM_NUM_SHIFTS), dim3(c_slotThreads), 0, st, a.entryPartials, in->sci, in->nsci,
                           computeEnergy ? 1 : 0, computeFshift ? 1 : 0, d_energies, d_fshift);
        NBNXM_HIP_CHECK(hipGetLastError());
    }
    return NBNXM_OK;
}
