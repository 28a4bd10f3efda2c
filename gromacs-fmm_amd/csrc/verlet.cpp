// The nonbonded part of a device-resident MD step: what nonbonded_verlet_t
// does for do_force on one GPU and one domain (src/gromacs/mdlib/sim_util.cpp:
// 1619-1680 for search / x buffer ops / launch, 925-946 for the rolling prune
// and the output clearing), with the list lifetime and the dynamic-pruning
// schedule of PairlistSets (src/gromacs/nbnxm/pairlistsets.h:91-113,
// src/gromacs/nbnxm/pairlist_tuning.cpp:505-600,690-708,
// src/gromacs/nbnxm/prunekerneldispatch.cpp:99-114).  Built on the host search
// and the nbnxm_gpu_* task of this library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "capi_common.h"
#include "gpu_search.h"
#include "nbnxm_kernels.h"
#include "pairsearch.h"

using namespace nbnxm_amd;

namespace
{

// Rolling prune interval of the GPU lists (c_nbnxnGpuRollingListPruningInterval,
// src/gromacs/nbnxm/pairlistparams.h:108): every part is pruned every second step.
constexpr int c_rollingPruningInterval = 2;

} // namespace

struct nbnxm_verlet
{
    int numAtoms = 0;
    int numTypes = 0;
    std::vector<int> types;
    std::vector<float> charges;
    std::vector<int> exclRanges; // empty = no exclusions
    std::vector<int> exclElements;
    nbnxm_interaction_const_t ic{};
    nbnxm_verlet_params_t params{};
    int vdwType = NBNXM_VDW_CUT_TYPE_TABLE;
    std::vector<float> nbfpTable;   // (numTypes+1)^2 pairs with the dummy type
    std::vector<float> perTypeComb; // combination parameters of the detected rule

    bool dynamicPruning = false;
    int numRollingParts = 1;
    int lifetime = 0;

    GridSearch search;
    nbnxm_gpu_search_t* gpuSearch = nullptr; // device search, when asked for
    std::vector<double> hxd;                 // its FP64 host input (first search)
    // Per-atom inputs of device searches after the first: charges, types and
    // per-type combination parameters on the device.
    DeviceBuffer<float> dCharges;
    DeviceBuffer<int> dTypes;
    DeviceBuffer<float2> dTypeComb;
    bool exclUploaded = false;
    // `search` holds the current list (false after a device search, until a
    // caller asks for the list).
    bool hostListCurrent = false;
    nbnxm_gpu_t* gpu = nullptr;
    hipEvent_t waitEvent = nullptr;
    bool haveList = false;
    double box[3] = { 0, 0, 0 };
    long long outerListCreationStep = 0;
    long long numSearches = 0;
    long long numPrunes = 0;
    long long numSteps = 0;
    // Entry split target of the last search (0: no split), reused by the next
    // search so that a search builds its list once.
    int splitTarget = -1;

    std::vector<float> hx;
    std::vector<float> xq;
    std::vector<int> slotTypes;
    std::vector<float> slotComb;

    double eLJ = 0, eCoul = 0;
    double virial[9] = {};
    bool haveEnergies = false;
    bool haveVirial = false;

    ~nbnxm_verlet()
    {
        if (gpu)
        {
            nbnxm_gpu_free(gpu);
        }
        if (gpuSearch)
        {
            nbnxm_gpu_search_free(gpuSearch);
        }
        if (waitEvent)
        {
            (void)hipEventDestroy(waitEvent);
        }
    }

    // isDynamicPruningStepGpu with one domain and no multiple time stepping.
    bool isPruneStep(long long step) const
    {
        const long long age = step - outerListCreationStep;
        return dynamicPruning && age > 0 && age < lifetime && age % c_rollingPruningInterval == 0;
    }

    int searchAndUpload(float* d_x, const double* newBox);
    int searchOnDevice(float* d_x, const double* newBox);
    int ensureHostList();
    int buildList(int maxPairsPerEntry, bool regrid);
};

// Grid and list into `search`: on the host, or by the device search (which
// wraps the coordinates itself) handed over to the host object.  `regrid`
// false rebuilds the host list on the grid of the previous call (same
// coordinates, another entry split); the device search always regrids.
int nbnxm_verlet::buildList(int maxPairsPerEntry, bool regrid)
{
    const int* er = exclRanges.empty() ? nullptr : exclRanges.data();
    const int* ee = exclRanges.empty() ? nullptr : exclElements.data();
    int rc;
    if (params.gpuSearch)
    {
        if (!gpuSearch && (rc = nbnxm_gpu_search_create(&gpuSearch)) != NBNXM_OK)
        {
            gpuSearch = nullptr;
            return rc;
        }
        hxd.assign(hx.begin(), hx.end());
        nbnxm_system_desc_t d{};
        d.numAtoms = numAtoms;
        d.x = hxd.data();
        d.type = types.data();
        d.exclRanges = er;
        d.exclElements = ee;
        std::copy(box, box + 3, d.box);
        if ((rc = nbnxm_gpu_search_build(gpuSearch, 1, &d, params.rlistOuter, maxPairsPerEntry)) != NBNXM_OK)
        {
            return rc;
        }
        return gpuSearchExportSystem(gpuSearch, 0, search);
    }
    if (regrid && (rc = search.putAtomsOnGrid(box, numAtoms, hx.data())) != NBNXM_OK)
    {
        return fail(rc, "verlet_step: grid construction failed");
    }
    rc = search.constructPairlist(params.rlistOuter, er, ee, maxPairsPerEntry);
    return rc == NBNXM_OK ? rc : fail(rc, "verlet_step: pair list construction failed");
}

int nbnxm_verlet::searchAndUpload(float* d_x, const double* newBox)
{
    if (params.gpuSearch && gpu && gpuSearch && numAtoms > 0)
    {
        return searchOnDevice(d_x, newBox);
    }
    const size_t n3 = 3 * static_cast<size_t>(numAtoms);
    hx.resize(n3);
    if (numAtoms > 0)
    {
        NBNXM_HIP_CHECK(hipMemcpy(hx.data(), d_x, sizeof(float) * n3, hipMemcpyDeviceToHost));
    }
    std::copy(newBox, newBox + 3, box);
    // After the first search the split target of the previous list is used
    // directly (the reference estimates it from the grid before building,
    // get_nsubpair_target, src/gromacs/nbnxm/pairlist.cpp:2548-2640; here it
    // comes from the previous list's cluster-pair count, which changes little
    // between searches of one system).
    int rc = buildList(splitTarget > 0 ? splitTarget : 0, true);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    // Atoms are put in the box on search steps (do_md, put_atoms_in_box_omp):
    // the state coordinates become the wrapped ones the grid was built from.
    const std::vector<float>& xw = search.wrappedX();
    if (numAtoms > 0)
    {
        NBNXM_HIP_CHECK(hipMemcpy(d_x, xw.data(), sizeof(float) * n3, hipMemcpyHostToDevice));
    }
    const int ng = search.numGridAtoms();
    xq.resize(4 * static_cast<size_t>(ng));
    search.fillXq(xw.data(), charges.data(), xq.data());
    slotTypes.resize(ng);
    search.fillTypes(types.data(), numTypes, slotTypes.data());
    slotComb.assign(2 * static_cast<size_t>(ng), 0.0f);
    if (vdwType != NBNXM_VDW_CUT_TYPE_TABLE)
    {
        search.fillLjComb(types.data(), perTypeComb.data(), slotComb.data());
    }
    nbnxm_atomdata_t nbat{};
    nbat.numAtoms = ng;
    nbat.xq = xq.data();
    nbat.numTypes = numTypes + 1;
    nbat.type = slotTypes.data();
    nbat.nbfp = nbfpTable.data();
    nbat.vdwType = vdwType;
    nbat.ljComb = slotComb.data();
    nbat.shiftVec = search.shiftVec();
    if (!gpu)
    {
        nbnxm_pairlist_params_t lp{};
        lp.rlistOuter = params.rlistOuter;
        lp.rlistInner = dynamicPruning ? params.rlistInner : params.rlistOuter;
        lp.useDynamicPruning = dynamicPruning ? 1 : 0;
        lp.numRollingPruningParts = numRollingParts;
        rc = nbnxm_gpu_init(&ic, &lp, &nbat, params.accumulation, &gpu);
        if (rc != NBNXM_OK)
        {
            gpu = nullptr;
            return rc;
        }
    }
    else
    {
        if ((rc = nbnxm_gpu_init_atomdata(gpu, &nbat)) != NBNXM_OK || (rc = nbnxm_gpu_upload_shiftvec(gpu, nbat.shiftVec)) != NBNXM_OK)
        {
            return rc;
        }
    }
    // Entry splitting toward min_ci_balanced entries (get_nsubpair_target,
    // src/gromacs/nbnxm/pairlist.cpp:2548-2640, simplified as in the calculator).
    const int minCi = nbnxm_gpu_min_ci_balanced(gpu);
    nbnxm_pairlist_gpu_t view = search.listView();
    const int64_t pairs = search.numClusterPairs();
    const int target = (minCi > 0 && pairs > 0) ? static_cast<int>(std::max<int64_t>(36, (pairs + minCi - 1) / minCi)) : 0;
    if (splitTarget < 0)
    {
        // First search: the unsplit list gave the entry and pair counts.
        splitTarget = 0;
        if (minCi > 0 && view.nsci > 0 && view.nsci < minCi)
        {
            if ((rc = buildList(target, false)) != NBNXM_OK)
            {
                return rc;
            }
            view = search.listView();
            splitTarget = target;
        }
    }
    else if (splitTarget > 0)
    {
        // Splitting keeps the cluster-pair count: refresh the target from it.
        splitTarget = target;
    }
    if ((rc = nbnxm_gpu_init_pairlist(gpu, &view)) != NBNXM_OK)
    {
        return rc;
    }
    const std::vector<int>& atomIndices = search.atomIndices();
    if ((rc = nbnxm_gpu_init_x_to_nbat_x(gpu, numAtoms, ng, atomIndices.data())) != NBNXM_OK)
    {
        return rc;
    }
    haveList = true;
    hostListCurrent = true;
    numSearches++;
    return NBNXM_OK;
}

// Searches after the first with the device search: coordinates, list and
// grid-order atom data stay on the device (the list is built from d_x, the
// wrapped coordinates are written back into it, the task adopts the list and
// gathers its atom data); only the entries pass through the host.  The entry
// split target of the previous search is kept.
int nbnxm_verlet::searchOnDevice(float* d_x, const double* newBox)
{
    if (dCharges.size == 0)
    {
        NBNXM_HIP_CHECK(dCharges.resize(numAtoms));
        NBNXM_HIP_CHECK(dTypes.resize(numAtoms));
        NBNXM_HIP_CHECK(hipMemcpy(dCharges.ptr, charges.data(), sizeof(float) * numAtoms, hipMemcpyHostToDevice));
        NBNXM_HIP_CHECK(hipMemcpy(dTypes.ptr, types.data(), sizeof(int) * numAtoms, hipMemcpyHostToDevice));
        if (vdwType != NBNXM_VDW_CUT_TYPE_TABLE)
        {
            NBNXM_HIP_CHECK(dTypeComb.resize(numTypes));
            NBNXM_HIP_CHECK(hipMemcpy(dTypeComb.ptr, perTypeComb.data(), sizeof(float2) * numTypes, hipMemcpyHostToDevice));
        }
    }
    std::copy(newBox, newBox + 3, box);
    const int* er = exclRanges.empty() ? nullptr : exclRanges.data();
    const int* ee = exclRanges.empty() ? nullptr : exclElements.data();
    DeviceListView view;
    int rc = gpuSearchBuildDevice(gpuSearch, numAtoms, d_x, nullptr, box, er, ee, !exclUploaded, params.rlistOuter,
                                  splitTarget > 0 ? splitTarget : 0, &view);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    exclUploaded = true;
    // Atoms are put in the box on search steps (do_md, put_atoms_in_box_omp).
    NBNXM_HIP_CHECK(launchWriteCoordinates(view.xw, numAtoms, d_x, gpuSearchStream(gpuSearch)));
    if ((rc = gpuTaskAdoptDeviceList(gpu, view, dCharges.ptr, dTypes.ptr, dTypeComb.ptr, numTypes, box)) != NBNXM_OK)
    {
        return rc;
    }
    haveList = true;
    hostListCurrent = false;
    numSearches++;
    return NBNXM_OK;
}

// The host copy of the current list (after a device search: copied from the
// device search and handed to `search`).
int nbnxm_verlet::ensureHostList()
{
    if (hostListCurrent || !haveList)
    {
        return NBNXM_OK;
    }
    int rc = gpuSearchDownloadViews(gpuSearch);
    if (rc == NBNXM_OK)
    {
        rc = gpuSearchExportSystem(gpuSearch, 0, search);
    }
    if (rc == NBNXM_OK)
    {
        hostListCurrent = true;
    }
    return rc;
}

namespace
{

// Per-type combination parameters from the nbfp diagonal (as the force
// calculator derives them).
void combFromDiagonal(int numTypes, const double* nbfp, int rule, std::vector<float>& out)
{
    out.assign(2 * static_cast<size_t>(numTypes), 0.0f);
    for (int i = 0; i < numTypes; i++)
    {
        const double c6 = nbfp[2 * (static_cast<size_t>(i) * numTypes + i)];
        const double c12 = nbfp[2 * (static_cast<size_t>(i) * numTypes + i) + 1];
        if (rule == NBNXM_VDW_CUT_COMB_GEOM)
        {
            out[2 * i] = static_cast<float>(std::sqrt(c6));
            out[2 * i + 1] = static_cast<float>(std::sqrt(c12));
        }
        else if (c6 > 0 && c12 > 0)
        {
            out[2 * i] = static_cast<float>(0.5 * std::pow(c12 / c6, 1.0 / 6.0));
            out[2 * i + 1] = static_cast<float>(std::sqrt(c6 * c6 / c12));
        }
    }
}

} // namespace

extern "C" {

int nbnxm_verlet_create(int numAtoms,
                        const int* types,
                        int numTypes,
                        const double* nbfp,
                        const double* charges,
                        const int* exclRanges,
                        const int* exclElements,
                        const nbnxm_interaction_const_t* ic,
                        const nbnxm_verlet_params_t* params,
                        nbnxm_verlet_t** out)
{
    if (!out || !ic || !params || numAtoms < 0 || numTypes < 1 || !nbfp || (numAtoms > 0 && (!types || !charges)))
    {
        return fail(NBNXM_ERR_INVALID, "verlet_create: null pointer or bad size");
    }
    *out = nullptr;
    const double cutoff = std::max(ic->rcoulomb, ic->rvdw);
    if (params->nstlist < 1 || !(params->rlistOuter >= cutoff) || !(params->rlistInner >= cutoff)
        || params->rlistInner > params->rlistOuter)
    {
        return fail(NBNXM_ERR_INVALID, "verlet_create: need nstlist >= 1 and cut-off <= rlistInner <= rlistOuter");
    }
    if (params->accumulation < NBNXM_ACCUM_FLOAT || params->accumulation > NBNXM_ACCUM_FP64)
    {
        return fail(NBNXM_ERR_INVALID, "verlet_create: unknown accumulation mode");
    }
    for (int i = 0; i < numAtoms; i++)
    {
        if (types[i] < 0 || types[i] >= numTypes)
        {
            return fail(NBNXM_ERR_INVALID, "verlet_create: atom type out of range");
        }
    }
    auto v = std::make_unique<nbnxm_verlet>();
    v->numAtoms = numAtoms;
    v->numTypes = numTypes;
    v->types.assign(types, types + numAtoms);
    v->charges.resize(numAtoms);
    for (int i = 0; i < numAtoms; i++)
    {
        v->charges[i] = static_cast<float>(charges[i]);
    }
    if (exclRanges)
    {
        v->exclRanges.assign(exclRanges, exclRanges + numAtoms + 1);
        const int numElements = exclRanges[numAtoms] - exclRanges[0];
        if (numElements < 0 || (numElements > 0 && !exclElements))
        {
            return fail(NBNXM_ERR_INVALID, "verlet_create: bad exclusion ranges");
        }
        if (numElements > 0)
        {
            v->exclElements.assign(exclElements + exclRanges[0], exclElements + exclRanges[numAtoms]);
            for (int& r : v->exclRanges)
            {
                r -= exclRanges[0];
            }
            for (int e : v->exclElements)
            {
                if (e < 0 || e >= numAtoms)
                {
                    return fail(NBNXM_ERR_INVALID, "verlet_create: exclusion element out of range");
                }
            }
        }
        else
        {
            v->exclRanges.clear();
        }
    }
    v->ic = *ic;
    v->params = *params;
    // Combination rule of the type table (enbnxninitcombruleDETECT); LJ-PME
    // needs the table for its grid part, so it keeps the table kernel.
    int rule = nbnxm_detect_comb_rule(numTypes, nbfp, 1e-5);
    if (rule < 0)
    {
        return rule;
    }
    v->vdwType = ic->ljPme ? NBNXM_VDW_CUT_TYPE_TABLE : rule;
    if (v->vdwType != NBNXM_VDW_CUT_TYPE_TABLE)
    {
        combFromDiagonal(numTypes, nbfp, v->vdwType, v->perTypeComb);
    }
    const int ntp = numTypes + 1;
    v->nbfpTable.assign(2 * static_cast<size_t>(ntp) * ntp, 0.0f);
    for (int i = 0; i < numTypes; i++)
    {
        for (int j = 0; j < numTypes; j++)
        {
            for (int k = 0; k < 2; k++)
            {
                v->nbfpTable[2 * (static_cast<size_t>(i) * ntp + j) + k] = static_cast<float>(nbfp[2 * (static_cast<size_t>(i) * numTypes + j) + k]);
            }
        }
    }
    // List lifetime and pruning setup (pairlist_tuning.cpp:515, 586-591, 690-708).
    v->lifetime = params->nstlist - 1;
    v->dynamicPruning = params->rlistInner < params->rlistOuter && params->nstlistPrune >= c_rollingPruningInterval
                        && params->nstlistPrune < v->lifetime;
    v->numRollingParts = v->dynamicPruning ? std::max(1, params->nstlistPrune / c_rollingPruningInterval) : 1;
    NBNXM_HIP_CHECK(hipEventCreateWithFlags(&v->waitEvent, hipEventDisableTiming));
    *out = v.release();
    return NBNXM_OK;
}

int nbnxm_verlet_step(nbnxm_verlet_t* v,
                      long long step,
                      const double box[3],
                      float* d_x,
                      void* waitStream,
                      int computeEnergy,
                      int computeVirial,
                      float* d_f,
                      int accumulateForces)
{
    if (!v || !box || (v->numAtoms > 0 && (!d_x || !d_f)))
    {
        return fail(NBNXM_ERR_INVALID, "verlet_step: null pointer");
    }
    const double rlist = v->params.rlistOuter;
    for (int d = 0; d < 3; d++)
    {
        if (!(box[d] > rlist))
        {
            return fail(NBNXM_ERR_UNSUPPORTED, "verlet_step: every box edge must exceed rlistOuter");
        }
    }
    if (v->haveList && step < v->outerListCreationStep)
    {
        return fail(NBNXM_ERR_STATE, "verlet_step: step numbers must not decrease");
    }
    const bool boxChanged = box[0] != v->box[0] || box[1] != v->box[1] || box[2] != v->box[2];
    const bool searchStep = !v->haveList || boxChanged || (step - v->outerListCreationStep) >= v->params.nstlist;
    int rc = NBNXM_OK;
    if (searchStep)
    {
        // The host reads the coordinates: their producer must have finished.
        if (waitStream)
        {
            NBNXM_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(waitStream)));
        }
        if ((rc = v->searchAndUpload(d_x, box)) != NBNXM_OK)
        {
            return rc;
        }
        v->outerListCreationStep = step;
    }
    else if (waitStream)
    {
        NBNXM_HIP_CHECK(hipEventRecord(v->waitEvent, static_cast<hipStream_t>(waitStream)));
        NBNXM_HIP_CHECK(hipStreamWaitEvent(gpuTaskStream(v->gpu), v->waitEvent, 0));
    }
    if (v->params.accumulation == NBNXM_ACCUM_FLOAT && (rc = nbnxm_gpu_clear_outputs(v->gpu, computeVirial)) != NBNXM_OK)
    {
        return rc;
    }
    if ((rc = nbnxm_gpu_x_to_nbat_x(v->gpu, d_x)) != NBNXM_OK
        || (rc = nbnxm_gpu_launch_kernel(v->gpu, computeEnergy, computeVirial)) != NBNXM_OK)
    {
        return rc;
    }
    // The rolling prune follows the nonbonded launch (sim_util.cpp:925-939).
    if (v->isPruneStep(step))
    {
        if ((rc = nbnxm_gpu_launch_kernel_pruneonly(v->gpu, v->numRollingParts)) != NBNXM_OK)
        {
            return rc;
        }
        v->numPrunes++;
    }
    if ((rc = nbnxm_gpu_add_nbat_f_to_f(v->gpu, d_f, nullptr, accumulateForces)) != NBNXM_OK
        || (rc = nbnxm_gpu_launch_cpyback(v->gpu, nullptr, nullptr)) != NBNXM_OK)
    {
        return rc;
    }
    double fshift[3 * NBNXM_NUM_SHIFTS] = {};
    double eLJ = 0, eCoul = 0;
    if ((rc = nbnxm_gpu_wait_finish_task(v->gpu, &eLJ, &eCoul, fshift)) != NBNXM_OK)
    {
        return rc;
    }
    v->haveEnergies = computeEnergy != 0;
    if (computeEnergy)
    {
        v->eLJ = eLJ;
        v->eCoul = eCoul;
    }
    v->haveVirial = computeVirial != 0;
    if (computeVirial && (rc = nbnxm_gpu_get_virial(v->gpu, v->virial)) != NBNXM_OK)
    {
        return rc;
    }
    v->numSteps++;
    return NBNXM_OK;
}

int nbnxm_verlet_get_energies(const nbnxm_verlet_t* v, double* eLJ, double* eCoul)
{
    if (!v || !eLJ || !eCoul)
    {
        return fail(NBNXM_ERR_INVALID, "verlet_get_energies: null pointer");
    }
    if (!v->haveEnergies)
    {
        return fail(NBNXM_ERR_STATE, "verlet_get_energies: the last step did not compute energies");
    }
    *eLJ = v->eLJ;
    *eCoul = v->eCoul;
    return NBNXM_OK;
}

int nbnxm_verlet_get_virial(const nbnxm_verlet_t* v, double* virial)
{
    if (!v || !virial)
    {
        return fail(NBNXM_ERR_INVALID, "verlet_get_virial: null pointer");
    }
    if (!v->haveVirial)
    {
        return fail(NBNXM_ERR_STATE, "verlet_get_virial: the last step did not compute the virial");
    }
    std::copy(v->virial, v->virial + 9, virial);
    return NBNXM_OK;
}

int nbnxm_verlet_get_stats(const nbnxm_verlet_t* v, long long* counts, int* flags)
{
    if (!v)
    {
        return fail(NBNXM_ERR_INVALID, "verlet_get_stats: null handle");
    }
    if (counts)
    {
        // Refreshing the host copy of the list changes no observable state.
        const int rcList = const_cast<nbnxm_verlet*>(v)->ensureHostList();
        if (rcList != NBNXM_OK)
        {
            return rcList;
        }
        counts[0] = v->numSearches;
        counts[1] = v->numPrunes;
        counts[2] = v->numSteps;
        counts[3] = v->haveList ? v->search.numClusterPairs() : 0;
        counts[4] = v->outerListCreationStep;
    }
    if (flags)
    {
        flags[0] = v->dynamicPruning ? 1 : 0;
        flags[1] = v->numRollingParts;
    }
    return NBNXM_OK;
}

int nbnxm_verlet_get_list(nbnxm_verlet_t* v, nbnxm_sci_t* sci, nbnxm_cj_packed_t* cjPacked, nbnxm_excl_t* excl, int* atomIndices, int* sizes)
{
    if (!v)
    {
        return fail(NBNXM_ERR_INVALID, "verlet_get_list: null handle");
    }
    if (!v->haveList)
    {
        return fail(NBNXM_ERR_STATE, "verlet_get_list: no list yet");
    }
    int rcList = v->ensureHostList();
    if (rcList != NBNXM_OK)
    {
        return rcList;
    }
    const nbnxm_pairlist_gpu_t view = v->search.listView();
    if (sizes)
    {
        sizes[0] = v->search.numGridAtoms();
        sizes[1] = view.nsci;
        sizes[2] = view.ncjPacked;
        sizes[3] = view.nexcl;
    }
    if (sci)
    {
        std::copy(view.sci, view.sci + view.nsci, sci);
    }
    if (excl)
    {
        std::copy(view.excl, view.excl + view.nexcl, excl);
    }
    if (atomIndices)
    {
        const std::vector<int>& ai = v->search.atomIndices();
        std::copy(ai.begin(), ai.end(), atomIndices);
    }
    if (cjPacked)
    {
        std::copy(view.cjPacked, view.cjPacked + view.ncjPacked, cjPacked);
        // The masks the kernel uses now: the inner (pruned) ones on the device.
        std::vector<unsigned int> inner(view.ncjPacked);
        int rc = nbnxm_gpu_get_masks(v->gpu, inner.data(), nullptr);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
        for (int g = 0; g < view.ncjPacked; g++)
        {
            cjPacked[g].imei[0].imask = inner[g];
        }
    }
    return NBNXM_OK;
}

void nbnxm_verlet_free(nbnxm_verlet_t* v)
{
    delete v;
}

} // extern "C"
