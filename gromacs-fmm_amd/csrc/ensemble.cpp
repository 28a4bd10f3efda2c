// Batched ensemble of independent systems: one concatenated list, one
// nonbonded launch per evaluation, hipGraph replay, plain per-system launches
// for comparison and an autotuner for the entry split per size bucket.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "capi_common.h"
#include "nbnxm_kernels.h"
#include "gpu_search.h"
#include "pairsearch.h"

using namespace nbnxm_amd;

namespace
{

constexpr int c_cl = NBNXM_CLUSTER_SIZE;
constexpr int c_ncps = NBNXM_NUM_CLUSTER_PER_SUPERCLUSTER;

struct SystemHost
{
    GridSearch grid;
    int numAtoms = 0;
    int slotOffset = 0; // first grid slot in the concatenated arrays
    int numSlots = 0;
    int baseBegin = 0;  // unsplit entries [baseBegin, baseEnd) of the global list
    int baseEnd = 0;
    int workBegin = 0;  // work items of this system in plainOrder
    int workEnd = 0;
    int buildStatus = NBNXM_OK;
};

struct HipEvent
{
    hipEvent_t e = nullptr;
    ~HipEvent()
    {
        if (e)
        {
            (void)hipEventDestroy(e);
        }
    }
};

} // namespace

struct nbnxm_ensemble
{
    nbnxm_ensemble_options_t opt{};
    nbnxm_interaction_const_t ic{};
    int vdwType = NBNXM_VDW_CUT_TYPE_TABLE;
    int numTypes = 0; // including the dummy type
    hipStream_t stream = nullptr;
    std::vector<SystemHost> sys;
    int totalSlots = 0;
    int64_t totalAtoms = 0;

    // Host copies of the unsplit list (global indices) and group masks after the fresh prune.
    std::vector<nbnxm_sci_t> baseSci;
    std::vector<int> baseSys;
    std::vector<nbnxm_cj_packed_t> cjPacked;
    std::vector<nbnxm_excl_t> excl;
    std::vector<unsigned int> innerMask;

    // Current work items.
    std::vector<nbnxm_sci_t> work;
    std::vector<int> workSys;
    NbLaunchConfig launchConfig;

    DeviceBuffer<float4> x0;
    DeviceBuffer<float4> xq;
    DeviceBuffer<int> types;
    DeviceBuffer<float2> ljComb;
    DeviceBuffer<float2> nbfp;
    DeviceBuffer<float4> shiftVec;
    DeviceBuffer<float4> paramsA;
    DeviceBuffer<float2> paramsB;
    DeviceBuffer<double> paramsD;
    DeviceBuffer<nbnxm_sci_t> dBaseSci;
    DeviceBuffer<nbnxm_sci_t> dWork;
    DeviceBuffer<int> dPlainOrder;
    DeviceBuffer<nbnxm_cj_packed_t> dCjPacked;
    DeviceBuffer<nbnxm_excl_t> dExcl;
    DeviceBuffer<unsigned int> dImaskOuter;
    DeviceBuffer<int> dRollingPart;
    DeviceBuffer<unsigned long long> fFixed;
    DeviceBuffer<float> fFloat;
    DeviceBuffer<double> fOutDouble;
    DeviceBuffer<double> partials;
    DeviceBuffer<double> energies;
    DeviceBuffer<double> fshift;
    DeviceBuffer<int> entryRanges;
    DeviceBuffer<int> entryList;
    DeviceBuffer<int> shiftRanges;
    DeviceBuffer<int> shiftEntryList;
    DeviceBuffer<unsigned long long> step;

    hipGraphExec_t graphExec = nullptr;
    int graphSteps = 0;
    int graphComputeEnergy = -1;
    float graphSigma = -1.0f;
    uint64_t graphSeed = 0;
    bool lastComputeEnergy = false;

    ~nbnxm_ensemble()
    {
        if (stream)
        {
            (void)hipStreamSynchronize(stream);
        }
        if (graphExec)
        {
            (void)hipGraphExecDestroy(graphExec);
        }
        if (stream)
        {
            (void)hipStreamDestroy(stream);
        }
    }
};

namespace
{

int dropGraph(nbnxm_ensemble* e)
{
    if (e->graphExec)
    {
        NBNXM_HIP_CHECK(hipGraphExecDestroy(e->graphExec));
        e->graphExec = nullptr;
    }
    e->graphSteps = 0;
    return NBNXM_OK;
}

// Splits every unsplit entry into work items of at most about maxPairs pruned
// cluster pairs (per size bucket), orders work items by descending pair count
// (stable) for the batched launch and per system for plain launches, and
// uploads items and the FP64 reduction maps.
int applySplit(nbnxm_ensemble* e, const std::vector<int>& bucketEdges, const std::vector<int>& maxPairs)
{
    const int nsys = static_cast<int>(e->sys.size());
    std::vector<nbnxm_sci_t> items;
    std::vector<int> itemSys;
    std::vector<int> itemCost;
    items.reserve(e->baseSci.size() * 2);
    for (int k = 0; k < nsys; k++)
    {
        const SystemHost& s = e->sys[k];
        const int bucket = static_cast<int>(std::upper_bound(bucketEdges.begin(), bucketEdges.end(), s.numAtoms) - bucketEdges.begin());
        const int limit = maxPairs.empty() ? 0 : maxPairs[std::min<size_t>(bucket, maxPairs.size() - 1)];
        for (int b = s.baseBegin; b < s.baseEnd; b++)
        {
            const nbnxm_sci_t base = e->baseSci[b];
            int total = 0;
            for (int g = base.cjPackedBegin; g < base.cjPackedEnd; g++)
            {
                total += __builtin_popcount(e->innerMask[g]);
            }
            if (limit <= 0 || total <= limit || base.cjPackedEnd - base.cjPackedBegin <= 1)
            {
                items.push_back(base);
                itemSys.push_back(k);
                itemCost.push_back(total);
                continue;
            }
            const int pieces = (total + limit - 1) / limit;
            const double target = static_cast<double>(total) / pieces;
            nbnxm_sci_t cur = base;
            int acc = 0;
            for (int g = base.cjPackedBegin; g < base.cjPackedEnd; g++)
            {
                const int n = __builtin_popcount(e->innerMask[g]);
                if (acc > 0 && acc + 0.5 * n > target)
                {
                    cur.cjPackedEnd = g;
                    items.push_back(cur);
                    itemSys.push_back(k);
                    itemCost.push_back(acc);
                    cur.cjPackedBegin = g;
                    acc = 0;
                }
                acc += n;
            }
            cur.cjPackedEnd = base.cjPackedEnd;
            items.push_back(cur);
            itemSys.push_back(k);
            itemCost.push_back(acc);
        }
    }
    // Batched order: descending cost, stable.
    std::vector<int> order(items.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return itemCost[a] > itemCost[b]; });
    e->work.resize(items.size());
    e->workSys.resize(items.size());
    std::vector<int> cost(items.size());
    for (size_t i = 0; i < order.size(); i++)
    {
        e->work[i] = items[order[i]];
        e->workSys[i] = itemSys[order[i]];
        cost[i] = itemCost[order[i]];
    }
    const int nwork = static_cast<int>(e->work.size());
    // Plain order: by system, descending cost within a system (stable).
    std::vector<int> plain(nwork);
    std::iota(plain.begin(), plain.end(), 0);
    std::stable_sort(plain.begin(), plain.end(), [&](int a, int b) { return e->workSys[a] < e->workSys[b]; });
    for (int k = 0, p = 0; k < nsys; k++)
    {
        e->sys[k].workBegin = p;
        while (p < nwork && e->workSys[plain[p]] == k)
        {
            p++;
        }
        e->sys[k].workEnd = p;
    }
    // FP64 reduction maps: work items per system, and per (system, shift).
    std::vector<int> entryRanges(nsys + 1, 0);
    std::vector<int> shiftRanges(NBNXM_NUM_SHIFTS * nsys + 1, 0);
    for (int i = 0; i < nwork; i++)
    {
        entryRanges[e->workSys[i] + 1]++;
        shiftRanges[e->work[i].shift + 1]++;
    }
    for (int k = 0; k < nsys; k++)
    {
        entryRanges[k + 1] += entryRanges[k];
    }
    for (int k = 0; k < NBNXM_NUM_SHIFTS * nsys; k++)
    {
        shiftRanges[k + 1] += shiftRanges[k];
    }
    std::vector<int> entryList(std::max(nwork, 1)), shiftList(std::max(nwork, 1));
    {
        std::vector<int> fe(entryRanges.begin(), entryRanges.end() - 1);
        std::vector<int> fs(shiftRanges.begin(), shiftRanges.end() - 1);
        for (int i = 0; i < nwork; i++)
        {
            entryList[fe[e->workSys[i]]++] = i;
            shiftList[fs[e->work[i].shift]++] = i;
        }
    }
    const int drc = dropGraph(e);
    if (drc != NBNXM_OK)
    {
        return drc;
    }
    NBNXM_HIP_CHECK(hipStreamSynchronize(e->stream));
    NBNXM_HIP_CHECK(e->dWork.resize(std::max(nwork, 1)));
    NBNXM_HIP_CHECK(e->dPlainOrder.resize(std::max(nwork, 1)));
    NBNXM_HIP_CHECK(e->partials.resize(8 * static_cast<size_t>(std::max(nwork, 1))));
    NBNXM_HIP_CHECK(e->entryRanges.resize(nsys + 1));
    NBNXM_HIP_CHECK(e->shiftRanges.resize(shiftRanges.size()));
    NBNXM_HIP_CHECK(e->entryList.resize(entryList.size()));
    NBNXM_HIP_CHECK(e->shiftEntryList.resize(shiftList.size()));
    if (nwork > 0)
    {
        NBNXM_HIP_CHECK(hipMemcpy(e->dWork.ptr, e->work.data(), sizeof(nbnxm_sci_t) * nwork, hipMemcpyHostToDevice));
        NBNXM_HIP_CHECK(hipMemcpy(e->dPlainOrder.ptr, plain.data(), sizeof(int) * nwork, hipMemcpyHostToDevice));
    }
    NBNXM_HIP_CHECK(hipMemcpy(e->entryRanges.ptr, entryRanges.data(), sizeof(int) * entryRanges.size(), hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(hipMemcpy(e->shiftRanges.ptr, shiftRanges.data(), sizeof(int) * shiftRanges.size(), hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(hipMemcpy(e->entryList.ptr, entryList.data(), sizeof(int) * entryList.size(), hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(hipMemcpy(e->shiftEntryList.ptr, shiftList.data(), sizeof(int) * shiftList.size(), hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(launchFillI32(reinterpret_cast<int*>(e->partials.ptr), 16 * static_cast<size_t>(std::max(nwork, 1)), 0, e->stream));
    NBNXM_HIP_CHECK(hipStreamSynchronize(e->stream));
    return NBNXM_OK;
}

NbKernelArgs kernelArgs(nbnxm_ensemble* e)
{
    NbKernelArgs a;
    a.xq = e->xq.ptr;
    a.atomTypes = e->types.ptr;
    a.ljComb = e->ljComb.ptr;
    a.nbfp = e->nbfp.ptr;
    a.numTypes = e->numTypes;
    a.shiftVec = e->shiftVec.ptr;
    a.sci = e->dWork.ptr;
    a.sciOrder = nullptr;
    a.nsci = static_cast<int>(e->work.size());
    a.cjPacked = e->dCjPacked.ptr;
    a.excl = e->dExcl.ptr;
    a.sysParamsA = e->paramsA.ptr;
    a.sysParamsB = e->paramsB.ptr;
    a.sysParamsD = e->paramsD.ptr;
    a.fFloat = e->fFloat.ptr;
    a.fFixed = e->fFixed.ptr;
    a.entryPartials = e->partials.ptr;
    a.computeFshift = 1;
    return a;
}

// One evaluation: perturb -> nonbonded -> convert/clear forces, FP64 reductions
// and step++ (one launch in the batched mode).
int enqueueStep(nbnxm_ensemble* e, int mode, float sigma, uint64_t seed, bool computeEnergy, hipEvent_t nbStart = nullptr, hipEvent_t nbStop = nullptr)
{
    const int nsys = static_cast<int>(e->sys.size());
    NbKernelArgs a = kernelArgs(e);
    const bool fixed = e->opt.accumulation != NBNXM_ACCUM_FLOAT;
    if (mode == NBNXM_MODE_BATCHED)
    {
        NBNXM_HIP_CHECK(launchPerturb(e->x0.ptr, e->xq.ptr, 0, e->totalSlots, sigma, seed, e->step.ptr, e->stream));
        if (!fixed)
        {
            NBNXM_HIP_CHECK(launchFillI32(reinterpret_cast<int*>(e->fFloat.ptr), 3 * static_cast<size_t>(e->totalSlots), 0, e->stream));
        }
        if (nbStart)
        {
            NBNXM_HIP_CHECK(hipEventRecord(nbStart, e->stream));
        }
        NBNXM_HIP_CHECK(launchNonbonded(a, e->opt.accumulation, e->vdwType, computeEnergy, e->launchConfig, e->stream));
        if (nbStop)
        {
            NBNXM_HIP_CHECK(hipEventRecord(nbStop, e->stream));
        }
        // Force conversion, energy reduction and the step counter in one launch.
        NBNXM_HIP_CHECK(launchFinishStep(fixed ? e->fFixed.ptr : nullptr, e->totalSlots, nullptr, e->fOutDouble.ptr, e->partials.ptr, e->entryRanges.ptr,
                                         e->entryList.ptr, e->shiftRanges.ptr, e->shiftEntryList.ptr, nsys, e->energies.ptr, e->fshift.ptr, 1,
                                         e->step.ptr, e->stream));
        return NBNXM_OK;
    }
    else
    {
        for (int k = 0; k < nsys; k++)
        {
            const SystemHost& s = e->sys[k];
            NBNXM_HIP_CHECK(launchPerturb(e->x0.ptr, e->xq.ptr, s.slotOffset, s.numSlots, sigma, seed, e->step.ptr, e->stream));
            if (!fixed)
            {
                NBNXM_HIP_CHECK(launchFillI32(reinterpret_cast<int*>(e->fFloat.ptr) + 3 * static_cast<size_t>(s.slotOffset), 3 * static_cast<size_t>(s.numSlots), 0, e->stream));
            }
            NbKernelArgs ak = a;
            ak.sciOrder = e->dPlainOrder.ptr + s.workBegin;
            ak.nsci = s.workEnd - s.workBegin;
            // Per-system launches choose the i-cluster placement by their own entry count.
            NbLaunchConfig pc = e->launchConfig;
            pc.iClusterInLds = -1;
            NBNXM_HIP_CHECK(launchNonbonded(ak, e->opt.accumulation, e->vdwType, computeEnergy, pc, e->stream));
            if (fixed)
            {
                NBNXM_HIP_CHECK(launchConvertForcesRange(e->fFixed.ptr, s.slotOffset, s.numSlots, nullptr, e->fOutDouble.ptr, e->stream));
            }
            NBNXM_HIP_CHECK(launchReduceEnergies(e->partials.ptr, e->entryRanges.ptr + k, e->entryList.ptr, e->shiftRanges.ptr + NBNXM_NUM_SHIFTS * k, e->shiftEntryList.ptr, 1, e->energies.ptr + 2 * k, e->fshift.ptr + 3 * NBNXM_NUM_SHIFTS * k, 1, e->stream));
        }
    }
    NBNXM_HIP_CHECK(launchAdvanceStep(e->step.ptr, e->stream));
    return NBNXM_OK;
}

int buildGraph(nbnxm_ensemble* e, int steps, float sigma, uint64_t seed, bool computeEnergy)
{
    int rc = dropGraph(e);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    hipGraph_t graph = nullptr;
    NBNXM_HIP_CHECK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < steps && rc == NBNXM_OK; s++)
    {
        rc = enqueueStep(e, NBNXM_MODE_BATCHED, sigma, seed, computeEnergy);
    }
    hipError_t endErr = hipStreamEndCapture(e->stream, &graph);
    if (rc != NBNXM_OK)
    {
        if (graph)
        {
            (void)hipGraphDestroy(graph);
        }
        return rc;
    }
    NBNXM_HIP_CHECK(endErr);
    hipError_t instErr = hipGraphInstantiate(&e->graphExec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    NBNXM_HIP_CHECK(instErr);
    e->graphSteps = steps;
    e->graphComputeEnergy = computeEnergy ? 1 : 0;
    e->graphSigma = sigma;
    e->graphSeed = seed;
    return NBNXM_OK;
}

int runSteps(nbnxm_ensemble* e, int mode, int numSteps, float sigma, uint64_t seed, bool computeEnergy, bool useGraph, int stepsPerGraph)
{
    if (mode != NBNXM_MODE_BATCHED && mode != NBNXM_MODE_PLAIN)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_run: unknown mode");
    }
    if (numSteps < 0 || !(sigma >= 0.0f))
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_run: numSteps >= 0 and sigma >= 0 required");
    }
    e->lastComputeEnergy = computeEnergy;
    if (useGraph && mode == NBNXM_MODE_BATCHED && numSteps > 0)
    {
        const int per = std::max(1, std::min(stepsPerGraph > 0 ? stepsPerGraph : numSteps, numSteps));
        if (!e->graphExec || e->graphSteps != per || e->graphComputeEnergy != (computeEnergy ? 1 : 0) || e->graphSigma != sigma || e->graphSeed != seed)
        {
            const int rc = buildGraph(e, per, sigma, seed, computeEnergy);
            if (rc != NBNXM_OK)
            {
                return rc;
            }
        }
        int done = 0;
        while (done + per <= numSteps)
        {
            NBNXM_HIP_CHECK(hipGraphLaunch(e->graphExec, e->stream));
            done += per;
        }
        for (; done < numSteps; done++)
        {
            const int rc = enqueueStep(e, mode, sigma, seed, computeEnergy);
            if (rc != NBNXM_OK)
            {
                return rc;
            }
        }
        return NBNXM_OK;
    }
    for (int s = 0; s < numSteps; s++)
    {
        const int rc = enqueueStep(e, mode, sigma, seed, computeEnergy);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
    }
    return NBNXM_OK;
}

// Mean step time of `reps` batched steps after one warm-up step; the minimum
// over `blocks` such timings (the least disturbed by clock changes).
int timeBatched(nbnxm_ensemble* e, int reps, double* ms, int blocks = 1, bool computeEnergy = false)
{
    HipEvent a, b;
    NBNXM_HIP_CHECK(hipEventCreate(&a.e));
    NBNXM_HIP_CHECK(hipEventCreate(&b.e));
    int rc = runSteps(e, NBNXM_MODE_BATCHED, 1, 0.0f, 1, computeEnergy, false, 0);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    double best = 0.0;
    for (int k = 0; k < blocks; k++)
    {
        NBNXM_HIP_CHECK(hipEventRecord(a.e, e->stream));
        rc = runSteps(e, NBNXM_MODE_BATCHED, reps, 0.0f, 1, computeEnergy, false, 0);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
        NBNXM_HIP_CHECK(hipEventRecord(b.e, e->stream));
        NBNXM_HIP_CHECK(hipEventSynchronize(b.e));
        float t = 0.0f;
        NBNXM_HIP_CHECK(hipEventElapsedTime(&t, a.e, b.e));
        const double m = t / reps;
        best = (k == 0) ? m : std::min(best, m);
    }
    *ms = best;
    return NBNXM_OK;
}

} // namespace

extern "C" {

int nbnxm_ensemble_create(int numSystems, const nbnxm_system_desc_t* systems, const nbnxm_ensemble_options_t* opt, nbnxm_ensemble_t** out)
{
    if (numSystems < 1 || !systems || !opt || !out)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_create: need systems, options and an output pointer");
    }
    *out = nullptr;
    if (opt->numTypes < 1 || !opt->c6c12)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_create: need the C6/C12 type table");
    }
    if (opt->vdwType < NBNXM_VDW_CUT_TYPE_TABLE || opt->vdwType > NBNXM_VDW_CUT_COMB_LB)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "ensemble_create: unknown vdwType");
    }
    if (opt->vdwType == NBNXM_VDW_CUT_TYPE_TABLE && opt->numTypes + 1 > c_maxLdsTableTypes)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "ensemble_create: the type table kernel stages at most 64 types in LDS");
    }
    if (opt->accumulation < NBNXM_ACCUM_FLOAT || opt->accumulation > NBNXM_ACCUM_FP64)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_create: unknown accumulation mode");
    }
    if (!(opt->rcoulomb > 0) || !(opt->rlistInner >= opt->rcoulomb) || !(opt->rlistOuter >= opt->rlistInner))
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_create: need 0 < rc <= rlistInner <= rlistOuter");
    }
    for (int k = 0; k < numSystems; k++)
    {
        const nbnxm_system_desc_t& d = systems[k];
        if (d.numAtoms < 1 || !d.x || !d.q || !d.type || ((d.exclRanges == nullptr) != (d.exclElements == nullptr)))
        {
            return fail(NBNXM_ERR_INVALID, "ensemble_create: system " + std::to_string(k) + " incomplete");
        }
        for (int i = 0; i < d.numAtoms; i++)
        {
            if (d.type[i] < 0 || d.type[i] >= opt->numTypes)
            {
                return fail(NBNXM_ERR_INVALID, "ensemble_create: type out of range in system " + std::to_string(k));
            }
        }
        for (int dim = 0; dim < 3; dim++)
        {
            if (!(d.box[dim] > opt->rlistOuter))
            {
                return fail(NBNXM_ERR_INVALID, "ensemble_create: every box edge must exceed rlistOuter (system " + std::to_string(k) + ")");
            }
        }
    }
    std::unique_ptr<nbnxm_ensemble> e(new (std::nothrow) nbnxm_ensemble());
    if (!e)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_create: out of host memory");
    }
    e->opt = *opt;
    e->opt.c6c12 = nullptr;
    e->vdwType = opt->vdwType;
    e->numTypes = opt->numTypes + 1;
    int rc = nbnxm_init_interaction_const(opt->rcoulomb, opt->ewaldRtol, opt->epsilonR, &e->ic);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    e->launchConfig.wavesPerBlock = (opt->wavesPerBlock >= 1 && opt->wavesPerBlock <= 4) ? opt->wavesPerBlock : 4;
    NBNXM_HIP_CHECK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));

    // Grids and unsplit lists: on the device in one pass, or on host threads.
    e->sys.resize(numSystems);
    if (opt->gpuSearch)
    {
        nbnxm_gpu_search_t* gs = nullptr;
        rc = nbnxm_gpu_search_create(&gs);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
        std::unique_ptr<nbnxm_gpu_search_t, void (*)(nbnxm_gpu_search_t*)> guard(gs, nbnxm_gpu_search_free);
        rc = nbnxm_gpu_search_build(gs, numSystems, systems, opt->rlistOuter, 0);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
        for (int k = 0; k < numSystems; k++)
        {
            SystemHost& s = e->sys[k];
            s.numAtoms = systems[k].numAtoms;
            s.buildStatus = gpuSearchExportSystem(gs, k, s.grid);
        }
    }
    else
    {
        const int nthreads = std::max(1, std::min(numSystems, opt->numHostThreads > 0 ? opt->numHostThreads : static_cast<int>(std::max(1u, std::thread::hardware_concurrency()))));
        std::vector<std::thread> pool;
        for (int t = 0; t < nthreads; t++)
        {
            pool.emplace_back([&, t]() {
                for (int k = t; k < numSystems; k += nthreads)
                {
                    SystemHost& s = e->sys[k];
                    const nbnxm_system_desc_t& d = systems[k];
                    s.numAtoms = d.numAtoms;
                    s.buildStatus = s.grid.putAtomsOnGrid(d.box, d.numAtoms, d.x);
                    if (s.buildStatus == NBNXM_OK)
                    {
                        s.buildStatus = s.grid.constructPairlist(opt->rlistOuter, d.exclRanges, d.exclElements, 0);
                    }
                }
            });
        }
        for (auto& th : pool)
        {
            th.join();
        }
    }
    for (int k = 0; k < numSystems; k++)
    {
        if (e->sys[k].buildStatus != NBNXM_OK)
        {
            return fail(e->sys[k].buildStatus, "ensemble_create: pair search failed for system " + std::to_string(k));
        }
    }

    // Concatenate: global slots, clusters, groups, exclusion entries, shifts.
    int slot = 0;
    size_t ngroups = 0, nexcl = 1, nsci = 0;
    for (auto& s : e->sys)
    {
        s.slotOffset = slot;
        s.numSlots = s.grid.numGridAtoms();
        slot += s.numSlots;
        ngroups += s.grid.cjPacked().size();
        nexcl += s.grid.excl().size() - 1;
        nsci += s.grid.sci().size();
        e->totalAtoms += s.numAtoms;
    }
    e->totalSlots = slot;
    e->baseSci.reserve(nsci);
    e->baseSys.reserve(nsci);
    e->cjPacked.reserve(ngroups);
    e->excl.reserve(nexcl);
    e->excl.push_back(e->sys[0].grid.excl()[0]);
    for (int k = 0; k < numSystems; k++)
    {
        SystemHost& s = e->sys[k];
        const int clusterOffset = s.slotOffset / c_cl;
        const int superOffset = s.slotOffset / (c_cl * c_ncps);
        const int groupOffset = static_cast<int>(e->cjPacked.size());
        const int exclOffset = static_cast<int>(e->excl.size()) - 1; // local entry i>0 -> exclOffset+i
        s.baseBegin = static_cast<int>(e->baseSci.size());
        for (nbnxm_sci_t sc : s.grid.sci())
        {
            sc.sci += superOffset;
            sc.shift += NBNXM_NUM_SHIFTS * k;
            sc.cjPackedBegin += groupOffset;
            sc.cjPackedEnd += groupOffset;
            e->baseSci.push_back(sc);
            e->baseSys.push_back(k);
        }
        s.baseEnd = static_cast<int>(e->baseSci.size());
        for (nbnxm_cj_packed_t p : s.grid.cjPacked())
        {
            for (int j = 0; j < NBNXM_JGROUP_SIZE; j++)
            {
                p.cj[j] += clusterOffset;
            }
            if (p.imei[0].excl_ind > 0)
            {
                p.imei[0].excl_ind += exclOffset;
            }
            e->cjPacked.push_back(p);
        }
        for (size_t x = 1; x < s.grid.excl().size(); x++)
        {
            e->excl.push_back(s.grid.excl()[x]);
        }
    }

    // Atom data in grid order.
    const int ntReal = opt->numTypes;
    const int nt = e->numTypes;
    std::vector<float4> hx(e->totalSlots);
    std::vector<int> htypes(e->totalSlots);
    std::vector<float2> hcomb(e->totalSlots);
    std::vector<float2> hnbfp(static_cast<size_t>(nt) * nt, make_float2(0.0f, 0.0f));
    for (int i = 0; i < ntReal; i++)
    {
        for (int j = 0; j < ntReal; j++)
        {
            hnbfp[i * nt + j] = make_float2(static_cast<float>(6.0 * opt->c6c12[2 * (i * ntReal + j)]), static_cast<float>(12.0 * opt->c6c12[2 * (i * ntReal + j) + 1]));
        }
    }
    std::vector<float2> perTypeComb(nt, make_float2(0.0f, 0.0f));
    for (int i = 0; i < ntReal; i++)
    {
        const double c6 = 6.0 * opt->c6c12[2 * (i * ntReal + i)];
        const double c12 = 12.0 * opt->c6c12[2 * (i * ntReal + i) + 1];
        if (opt->vdwType == NBNXM_VDW_CUT_COMB_GEOM)
        {
            perTypeComb[i] = make_float2(static_cast<float>(std::sqrt(c6)), static_cast<float>(std::sqrt(c12)));
        }
        else if (c6 > 0 && c12 > 0)
        {
            perTypeComb[i] = make_float2(static_cast<float>(0.5 * std::pow(c12 / c6, 1.0 / 6.0)), static_cast<float>(std::sqrt(c6 * c6 / c12)));
        }
    }
    std::vector<float4> hshift(static_cast<size_t>(NBNXM_NUM_SHIFTS) * numSystems);
    std::vector<float4> hpa(numSystems);
    std::vector<float2> hpb(numSystems);
    std::vector<double> hpd(8 * static_cast<size_t>(numSystems));
    const float rcf = static_cast<float>(e->ic.rcoulomb);
    for (int k = 0; k < numSystems; k++)
    {
        const SystemHost& s = e->sys[k];
        const nbnxm_system_desc_t& d = systems[k];
        std::vector<float> xq(4 * static_cast<size_t>(s.numSlots));
        s.grid.fillXq(static_cast<const double*>(nullptr), d.q, xq.data());
        for (int i = 0; i < s.numSlots; i++)
        {
            const int a = s.grid.atomIndices()[i];
            hx[s.slotOffset + i] = make_float4(xq[4 * i], xq[4 * i + 1], xq[4 * i + 2], xq[4 * i + 3]);
            htypes[s.slotOffset + i] = (a < 0) ? nt - 1 : d.type[a];
            hcomb[s.slotOffset + i] = (a < 0) ? make_float2(0.0f, 0.0f) : perTypeComb[d.type[a]];
        }
        for (int m = 0; m < NBNXM_NUM_SHIFTS; m++)
        {
            const float* sv = s.grid.shiftVec() + 3 * m;
            hshift[NBNXM_NUM_SHIFTS * k + m] = make_float4(sv[0], sv[1], sv[2], 0.0f);
        }
        hpa[k] = make_float4(rcf * rcf, static_cast<float>(e->ic.epsfac), static_cast<float>(e->ic.ewaldcoeff_q), static_cast<float>(e->ic.sh_ewald));
        hpb[k] = make_float2(static_cast<float>(e->ic.repulsion_shift_cpot), static_cast<float>(e->ic.dispersion_shift_cpot));
        const double pd[8] = { e->ic.rcoulomb * e->ic.rcoulomb, e->ic.epsfac, e->ic.ewaldcoeff_q, e->ic.sh_ewald,
                               e->ic.repulsion_shift_cpot, e->ic.dispersion_shift_cpot, 0.0, 0.0 };
        std::copy(pd, pd + 8, hpd.begin() + 8 * k);
    }

    const size_t ns = e->totalSlots;
    NBNXM_HIP_CHECK(e->x0.resize(ns));
    NBNXM_HIP_CHECK(e->xq.resize(ns));
    NBNXM_HIP_CHECK(e->types.resize(ns));
    NBNXM_HIP_CHECK(e->ljComb.resize(ns));
    NBNXM_HIP_CHECK(e->nbfp.resize(hnbfp.size()));
    NBNXM_HIP_CHECK(e->shiftVec.resize(hshift.size()));
    NBNXM_HIP_CHECK(e->paramsA.resize(numSystems));
    NBNXM_HIP_CHECK(e->paramsB.resize(numSystems));
    NBNXM_HIP_CHECK(e->paramsD.resize(hpd.size()));
    NBNXM_HIP_CHECK(e->dBaseSci.resize(std::max<size_t>(e->baseSci.size(), 1)));
    NBNXM_HIP_CHECK(e->dCjPacked.resize(std::max<size_t>(e->cjPacked.size(), 1)));
    NBNXM_HIP_CHECK(e->dExcl.resize(e->excl.size()));
    NBNXM_HIP_CHECK(e->dImaskOuter.resize(std::max<size_t>(e->cjPacked.size(), 1)));
    NBNXM_HIP_CHECK(e->fFixed.resize(3 * ns));
    NBNXM_HIP_CHECK(e->fFloat.resize(3 * ns));
    NBNXM_HIP_CHECK(e->fOutDouble.resize(3 * ns));
    NBNXM_HIP_CHECK(e->energies.resize(2 * static_cast<size_t>(numSystems)));
    NBNXM_HIP_CHECK(e->fshift.resize(3 * NBNXM_NUM_SHIFTS * static_cast<size_t>(numSystems)));
    NBNXM_HIP_CHECK(e->step.resize(1));

    NBNXM_HIP_CHECK(hipMemcpy(e->x0.ptr, hx.data(), sizeof(float4) * ns, hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(hipMemcpy(e->xq.ptr, hx.data(), sizeof(float4) * ns, hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(hipMemcpy(e->types.ptr, htypes.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(hipMemcpy(e->ljComb.ptr, hcomb.data(), sizeof(float2) * ns, hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(hipMemcpy(e->nbfp.ptr, hnbfp.data(), sizeof(float2) * hnbfp.size(), hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(hipMemcpy(e->shiftVec.ptr, hshift.data(), sizeof(float4) * hshift.size(), hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(hipMemcpy(e->paramsA.ptr, hpa.data(), sizeof(float4) * numSystems, hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(hipMemcpy(e->paramsB.ptr, hpb.data(), sizeof(float2) * numSystems, hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(hipMemcpy(e->paramsD.ptr, hpd.data(), sizeof(double) * hpd.size(), hipMemcpyHostToDevice));
    if (!e->baseSci.empty())
    {
        NBNXM_HIP_CHECK(hipMemcpy(e->dBaseSci.ptr, e->baseSci.data(), sizeof(nbnxm_sci_t) * e->baseSci.size(), hipMemcpyHostToDevice));
    }
    if (!e->cjPacked.empty())
    {
        NBNXM_HIP_CHECK(hipMemcpy(e->dCjPacked.ptr, e->cjPacked.data(), sizeof(nbnxm_cj_packed_t) * e->cjPacked.size(), hipMemcpyHostToDevice));
    }
    NBNXM_HIP_CHECK(hipMemcpy(e->dExcl.ptr, e->excl.data(), sizeof(nbnxm_excl_t) * e->excl.size(), hipMemcpyHostToDevice));
    NBNXM_HIP_CHECK(launchFillU64(e->fFixed.ptr, 3 * ns, e->stream));
    NBNXM_HIP_CHECK(launchFillU64(e->step.ptr, 1, e->stream));

    // Fresh-list prune of every group: rlistOuter mask and rlistInner mask.
    const int nbase = static_cast<int>(e->baseSci.size());
    const int rollingSize = ((nbase + 3) / 4 + 1) * 4;
    NBNXM_HIP_CHECK(e->dRollingPart.resize(rollingSize));
    NBNXM_HIP_CHECK(launchFillI32(e->dRollingPart.ptr, rollingSize, 0, e->stream));
    PruneArgs pa;
    pa.xq = e->xq.ptr;
    pa.shiftVec = e->shiftVec.ptr;
    pa.sci = e->dBaseSci.ptr;
    pa.nsci = nbase;
    pa.cjPacked = e->dCjPacked.ptr;
    pa.imaskOuter = e->dImaskOuter.ptr;
    pa.rollingPart = e->dRollingPart.ptr;
    const float ro = static_cast<float>(opt->rlistOuter);
    const float ri = static_cast<float>(opt->rlistInner);
    pa.rlistOuter_sq = ro * ro;
    pa.rlistInner_sq = ri * ri;
    pa.numParts = 1;
    pa.numGroups = static_cast<int>(e->cjPacked.size());
    NBNXM_HIP_CHECK(launchPrune(pa, true, e->stream));
    NBNXM_HIP_CHECK(hipStreamSynchronize(e->stream));
    e->innerMask.resize(e->cjPacked.size());
    if (!e->cjPacked.empty())
    {
        std::vector<nbnxm_cj_packed_t> pruned(e->cjPacked.size());
        NBNXM_HIP_CHECK(hipMemcpy(pruned.data(), e->dCjPacked.ptr, sizeof(nbnxm_cj_packed_t) * pruned.size(), hipMemcpyDeviceToHost));
        for (size_t g = 0; g < pruned.size(); g++)
        {
            e->innerMask[g] = pruned[g].imei[0].imask;
        }
    }
    const int initialSplit = opt->maxPairsPerEntry > 0 ? opt->maxPairsPerEntry : 64;
    rc = applySplit(e.get(), {}, { initialSplit });
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    *out = e.release();
    return NBNXM_OK;
}

int nbnxm_ensemble_set_split(nbnxm_ensemble_t* e, int numBuckets, const int* bucketEdges, const int* maxPairs, int wavesPerBlock)
{
    if (!e || numBuckets < 1 || !maxPairs || (numBuckets > 1 && !bucketEdges))
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_set_split: need numBuckets >= 1, edges and maxPairs");
    }
    std::vector<int> edges(bucketEdges, bucketEdges + (numBuckets - 1));
    if (!std::is_sorted(edges.begin(), edges.end()))
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_set_split: bucket edges must ascend");
    }
    if (wavesPerBlock > 0)
    {
        if (wavesPerBlock > 4)
        {
            return fail(NBNXM_ERR_INVALID, "ensemble_set_split: wavesPerBlock 1..4");
        }
        e->launchConfig.wavesPerBlock = wavesPerBlock;
    }
    return applySplit(e, edges, std::vector<int>(maxPairs, maxPairs + numBuckets));
}

int nbnxm_ensemble_set_i_cluster_in_lds(nbnxm_ensemble_t* e, int mode)
{
    if (!e || mode < -1 || mode > 1)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_set_i_cluster_in_lds: mode -1, 0 or 1");
    }
    if (e->launchConfig.iClusterInLds != mode)
    {
        e->launchConfig.iClusterInLds = mode;
        const int rc = dropGraph(e);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
    }
    return NBNXM_OK;
}

int nbnxm_ensemble_get_launch_config(nbnxm_ensemble_t* e, int* wavesPerBlock, int* iClusterInLds)
{
    if (!e)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_get_launch_config: null ensemble");
    }
    if (wavesPerBlock)
    {
        *wavesPerBlock = e->launchConfig.wavesPerBlock;
    }
    if (iClusterInLds)
    {
        *iClusterInLds = iClusterDataInLds(e->launchConfig, kernelArgs(e).nsci) ? 1 : 0;
    }
    return NBNXM_OK;
}

int nbnxm_ensemble_autotune(nbnxm_ensemble_t* e, int numBuckets, const int* bucketEdges, int computeEnergy, int* chosenMaxPairs, int* chosenWavesPerBlock)
{
    if (!e || numBuckets < 1 || (numBuckets > 1 && !bucketEdges))
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_autotune: need numBuckets >= 1 and edges");
    }
    std::vector<int> edges(bucketEdges, bucketEdges + (numBuckets - 1));
    const int candidates[] = { 0, 512, 256, 128, 96, 64, 48, 32 };
    std::vector<int> best(numBuckets, 64);
    int rc = applySplit(e, edges, best);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    // Each candidate: minimum over 3 timings of 5 steps; a candidate replaces
    // the current choice only when it is faster by more than 0.2 %.
    const int reps = 5;
    const int blocks = 3;
    const double margin = 0.998;
    double bestMs = 0.0;
    rc = timeBatched(e, reps, &bestMs, blocks, computeEnergy != 0);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    for (int sweep = 0; sweep < 2; sweep++)
    {
        for (int b = 0; b < numBuckets; b++)
        {
            for (int c : candidates)
            {
                if (c == best[b])
                {
                    continue;
                }
                std::vector<int> trial = best;
                trial[b] = c;
                rc = applySplit(e, edges, trial);
                if (rc != NBNXM_OK)
                {
                    return rc;
                }
                double ms = 0.0;
                rc = timeBatched(e, reps, &ms, blocks, computeEnergy != 0);
                if (rc != NBNXM_OK)
                {
                    return rc;
                }
                if (ms < bestMs * margin)
                {
                    bestMs = ms;
                    best = trial;
                }
            }
        }
    }
    rc = applySplit(e, edges, best);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    int bestW = e->launchConfig.wavesPerBlock;
    for (int w = 1; w <= 4; w *= 2)
    {
        if (w == bestW)
        {
            continue;
        }
        e->launchConfig.wavesPerBlock = w;
        double ms = 0.0;
        rc = timeBatched(e, reps, &ms, blocks, computeEnergy != 0);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
        if (ms < bestMs * margin)
        {
            bestMs = ms;
            bestW = w;
        }
    }
    e->launchConfig.wavesPerBlock = bestW;
    // The other i-cluster placement of the tuned kernel than the automatic one.
    const int autoLds = iClusterDataInLds(e->launchConfig, kernelArgs(e).nsci) ? 1 : 0;
    e->launchConfig.iClusterInLds = 1 - autoLds;
    {
        double ms = 0.0;
        rc = timeBatched(e, reps, &ms, blocks, computeEnergy != 0);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
        if (!(ms < bestMs * margin))
        {
            e->launchConfig.iClusterInLds = autoLds;
        }
        else
        {
            bestMs = ms;
        }
    }
    rc = dropGraph(e);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    if (chosenMaxPairs)
    {
        std::copy(best.begin(), best.end(), chosenMaxPairs);
    }
    if (chosenWavesPerBlock)
    {
        *chosenWavesPerBlock = bestW;
    }
    return NBNXM_OK;
}

int nbnxm_ensemble_run(nbnxm_ensemble_t* e, int mode, int numSteps, float sigma, uint64_t seed, int computeEnergy, int useGraph, int stepsPerGraph)
{
    if (!e)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_run: null handle");
    }
    return runSteps(e, mode, numSteps, sigma, seed, computeEnergy != 0, useGraph != 0, stepsPerGraph);
}

int nbnxm_ensemble_set_step(nbnxm_ensemble_t* e, uint64_t step)
{
    if (!e)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_set_step: null handle");
    }
    const unsigned long long v = step;
    NBNXM_HIP_CHECK(hipStreamSynchronize(e->stream));
    NBNXM_HIP_CHECK(hipMemcpy(e->step.ptr, &v, sizeof(v), hipMemcpyHostToDevice));
    return NBNXM_OK;
}

int nbnxm_ensemble_sync(nbnxm_ensemble_t* e)
{
    if (!e)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_sync: null handle");
    }
    NBNXM_HIP_CHECK(hipStreamSynchronize(e->stream));
    return NBNXM_OK;
}

int nbnxm_ensemble_time(nbnxm_ensemble_t* e, int mode, int numSteps, float sigma, uint64_t seed, int computeEnergy, int useGraph, double* totalMs, double* kernelMs)
{
    if (!e || numSteps < 1 || !totalMs)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_time: need a handle, numSteps >= 1 and totalMs");
    }
    HipEvent a, b;
    NBNXM_HIP_CHECK(hipEventCreate(&a.e));
    NBNXM_HIP_CHECK(hipEventCreate(&b.e));
    NBNXM_HIP_CHECK(hipStreamSynchronize(e->stream));
    NBNXM_HIP_CHECK(hipEventRecord(a.e, e->stream));
    int rc = runSteps(e, mode, numSteps, sigma, seed, computeEnergy != 0, useGraph != 0, numSteps);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    NBNXM_HIP_CHECK(hipEventRecord(b.e, e->stream));
    NBNXM_HIP_CHECK(hipEventSynchronize(b.e));
    float t = 0.0f;
    NBNXM_HIP_CHECK(hipEventElapsedTime(&t, a.e, b.e));
    *totalMs = t;
    if (kernelMs)
    {
        // Events around every nonbonded launch, on the stream it runs on.
        const int n = std::min(numSteps, 64);
        std::vector<HipEvent> ev(2 * n);
        for (auto& x : ev)
        {
            NBNXM_HIP_CHECK(hipEventCreate(&x.e));
        }
        for (int s = 0; s < n; s++)
        {
            rc = enqueueStep(e, NBNXM_MODE_BATCHED, sigma, seed, computeEnergy != 0, ev[2 * s].e, ev[2 * s + 1].e);
            if (rc != NBNXM_OK)
            {
                return rc;
            }
        }
        NBNXM_HIP_CHECK(hipStreamSynchronize(e->stream));
        double sum = 0.0;
        for (int s = 0; s < n; s++)
        {
            float k = 0.0f;
            NBNXM_HIP_CHECK(hipEventElapsedTime(&k, ev[2 * s].e, ev[2 * s + 1].e));
            sum += k;
        }
        *kernelMs = sum / n;
    }
    return NBNXM_OK;
}

int nbnxm_ensemble_get_forces(nbnxm_ensemble_t* e, int system, double* f)
{
    if (!e || !f || system < 0 || system >= static_cast<int>(e->sys.size()))
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_get_forces: bad arguments");
    }
    const SystemHost& s = e->sys[system];
    NBNXM_HIP_CHECK(hipStreamSynchronize(e->stream));
    std::vector<double> grid(3 * static_cast<size_t>(s.numSlots));
    if (e->opt.accumulation == NBNXM_ACCUM_FLOAT)
    {
        std::vector<float> g(grid.size());
        NBNXM_HIP_CHECK(hipMemcpy(g.data(), e->fFloat.ptr + 3 * static_cast<size_t>(s.slotOffset), sizeof(float) * g.size(), hipMemcpyDeviceToHost));
        std::copy(g.begin(), g.end(), grid.begin());
    }
    else
    {
        NBNXM_HIP_CHECK(hipMemcpy(grid.data(), e->fOutDouble.ptr + 3 * static_cast<size_t>(s.slotOffset), sizeof(double) * grid.size(), hipMemcpyDeviceToHost));
    }
    const auto& idx = s.grid.atomIndices();
    for (int i = 0; i < s.numSlots; i++)
    {
        if (idx[i] >= 0)
        {
            f[3 * idx[i] + 0] = grid[3 * i + 0];
            f[3 * idx[i] + 1] = grid[3 * i + 1];
            f[3 * idx[i] + 2] = grid[3 * i + 2];
        }
    }
    return NBNXM_OK;
}

int nbnxm_ensemble_get_energies(nbnxm_ensemble_t* e, double* energies, double* fshift)
{
    if (!e || !energies)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_get_energies: bad arguments");
    }
    NBNXM_HIP_CHECK(hipStreamSynchronize(e->stream));
    NBNXM_HIP_CHECK(hipMemcpy(energies, e->energies.ptr, sizeof(double) * 2 * e->sys.size(), hipMemcpyDeviceToHost));
    if (fshift)
    {
        NBNXM_HIP_CHECK(hipMemcpy(fshift, e->fshift.ptr, sizeof(double) * 3 * NBNXM_NUM_SHIFTS * e->sys.size(), hipMemcpyDeviceToHost));
    }
    return NBNXM_OK;
}

int nbnxm_ensemble_get_positions(nbnxm_ensemble_t* e, int system, double* x)
{
    if (!e || !x || system < 0 || system >= static_cast<int>(e->sys.size()))
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_get_positions: bad arguments");
    }
    const SystemHost& s = e->sys[system];
    NBNXM_HIP_CHECK(hipStreamSynchronize(e->stream));
    std::vector<float4> g(s.numSlots);
    NBNXM_HIP_CHECK(hipMemcpy(g.data(), e->xq.ptr + s.slotOffset, sizeof(float4) * g.size(), hipMemcpyDeviceToHost));
    const auto& idx = s.grid.atomIndices();
    for (int i = 0; i < s.numSlots; i++)
    {
        if (idx[i] >= 0)
        {
            x[3 * idx[i] + 0] = g[i].x;
            x[3 * idx[i] + 1] = g[i].y;
            x[3 * idx[i] + 2] = g[i].z;
        }
    }
    return NBNXM_OK;
}

int nbnxm_ensemble_stats(nbnxm_ensemble_t* e, int64_t* stats)
{
    if (!e || !stats)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_stats: bad arguments");
    }
    int64_t listPairs = 0, innerPairs = 0;
    for (size_t g = 0; g < e->cjPacked.size(); g++)
    {
        listPairs += __builtin_popcount(e->cjPacked[g].imei[0].imask);
        innerPairs += __builtin_popcount(e->innerMask[g]);
    }
    stats[0] = e->totalAtoms;
    stats[1] = e->totalSlots;
    stats[2] = static_cast<int64_t>(e->work.size());
    stats[3] = static_cast<int64_t>(e->cjPacked.size());
    stats[4] = listPairs;
    stats[5] = innerPairs;
    return NBNXM_OK;
}

int nbnxm_ensemble_system_list(nbnxm_ensemble_t* e, int system, nbnxm_pairlist_gpu_t* list, int* atomIndices)
{
    if (!e || system < 0 || system >= static_cast<int>(e->sys.size()))
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_system_list: bad arguments");
    }
    const SystemHost& s = e->sys[system];
    if (list)
    {
        *list = s.grid.listView();
    }
    if (atomIndices)
    {
        std::copy(s.grid.atomIndices().begin(), s.grid.atomIndices().end(), atomIndices);
    }
    return NBNXM_OK;
}

int nbnxm_ensemble_count_pairs(nbnxm_ensemble_t* e, double rc, int64_t* perSystem, int64_t* total)
{
    if (!e || !(rc > 0) || !total)
    {
        return fail(NBNXM_ERR_INVALID, "ensemble_count_pairs: bad arguments");
    }
    const int nsys = static_cast<int>(e->sys.size());
    std::vector<int64_t> counts(nsys, 0);
    const int nthreads = std::max(1, std::min(nsys, static_cast<int>(std::max(1u, std::thread::hardware_concurrency()))));
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++)
    {
        pool.emplace_back([&, t]() {
            for (int k = t; k < nsys; k += nthreads)
            {
                counts[k] = e->sys[k].grid.countPairsWithin(rc);
            }
        });
    }
    for (auto& th : pool)
    {
        th.join();
    }
    int64_t sum = 0;
    for (int k = 0; k < nsys; k++)
    {
        sum += counts[k];
        if (perSystem)
        {
            perSystem[k] = counts[k];
        }
    }
    *total = sum;
    return NBNXM_OK;
}

void nbnxm_ensemble_free(nbnxm_ensemble_t* e)
{
    delete e;
}

} // extern "C"
