// C ABI: interaction constants, the host pair search and the single-system GPU
// task object that mirrors the reference's Nbnxm::gpu_* API
// (src/gromacs/nbnxm/nbnxm_gpu.h:66-302, src/gromacs/nbnxm/gpu_data_mgmt.h).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "capi_common.h"
#include "nbnxm_kernels.h"
#include "pairsearch.h"

namespace nbnxm_amd
{

static thread_local std::string t_lastError;

void setError(const std::string& msg)
{
    t_lastError = msg;
}

int fail(int code, const std::string& msg)
{
    t_lastError = msg;
    return code;
}

int hipFail(hipError_t e, const char* what)
{
    t_lastError = std::string(what) + ": " + hipGetErrorString(e);
    return NBNXM_ERR_HIP;
}

int validateAtomdata(const nbnxm_atomdata_t* nbat)
{
    if (!nbat || nbat->numAtoms < 0 || nbat->numAtoms % NBNXM_CLUSTER_SIZE != 0)
    {
        return fail(NBNXM_ERR_INVALID, "atomdata: numAtoms must be a non-negative multiple of 8");
    }
    if (nbat->numAtoms > 0 && !nbat->xq)
    {
        return fail(NBNXM_ERR_INVALID, "atomdata: xq is null");
    }
    if (nbat->vdwType == NBNXM_VDW_CUT_TYPE_TABLE)
    {
        if (nbat->numTypes < 1 || !nbat->nbfp || (nbat->numAtoms > 0 && !nbat->type))
        {
            return fail(NBNXM_ERR_INVALID, "atomdata: type table kernel needs numTypes, type and nbfp");
        }
        for (int i = 0; i < nbat->numAtoms; i++)
        {
            if (nbat->type[i] < 0 || nbat->type[i] >= nbat->numTypes)
            {
                return fail(NBNXM_ERR_INVALID, "atomdata: type of slot " + std::to_string(i) + " out of range");
            }
        }
    }
    else if (nbat->vdwType == NBNXM_VDW_CUT_COMB_GEOM || nbat->vdwType == NBNXM_VDW_CUT_COMB_LB)
    {
        if (nbat->numAtoms > 0 && !nbat->ljComb)
        {
            return fail(NBNXM_ERR_INVALID, "atomdata: combination-rule kernel needs ljComb");
        }
    }
    else
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "atomdata: unknown vdwType");
    }
    return NBNXM_OK;
}

// The flavours the kernels implement (nbnxm_cuda.cu:171-254 selects among the
// same set): Ewald or reaction field, LJ PotShift / ForceSwitch / PotSwitch,
// geometric LJ-PME with PotShift, rvdw <= rcoulomb.
int validateFlavour(const nbnxm_interaction_const_t* ic)
{
    if (ic->eeltype != NBNXM_ELEC_EWALD && ic->eeltype != NBNXM_ELEC_RF)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "interaction constants: unknown Coulomb type");
    }
    if (ic->vdwModifier < NBNXM_VDW_MOD_POTSHIFT || ic->vdwModifier > NBNXM_VDW_MOD_POTSWITCH)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "interaction constants: unknown LJ modifier");
    }
    if (ic->ljPme < NBNXM_LJPME_NONE || ic->ljPme > NBNXM_LJPME_LB)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "interaction constants: unknown LJ-PME grid rule");
    }
    if (ic->ljPme && ic->vdwModifier != NBNXM_VDW_MOD_POTSHIFT)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "interaction constants: LJ-PME supports the potential-shift modifier only");
    }
    if (ic->tabulatedEwald && (ic->eeltype != NBNXM_ELEC_EWALD || !(ic->ewaldcoeff_q > 0)))
    {
        return fail(NBNXM_ERR_INVALID, "interaction constants: the tabulated correction needs Ewald with beta > 0");
    }
    if (!(ic->rcoulomb > 0) || !(ic->rvdw > 0) || ic->rvdw > ic->rcoulomb)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "interaction constants: need 0 < rvdw <= rcoulomb");
    }
    return NBNXM_OK;
}

// LJ-PME reads per-slot types and the type table for the grid C6 of every
// combination kernel.
int validateLjPmeAtomdata(const nbnxm_atomdata_t* nbat)
{
    if (nbat->numTypes < 1 || !nbat->nbfp || (nbat->numAtoms > 0 && !nbat->type))
    {
        return fail(NBNXM_ERR_INVALID, "atomdata: LJ-PME needs numTypes, type and nbfp");
    }
    for (int i = 0; i < nbat->numAtoms; i++)
    {
        if (nbat->type[i] < 0 || nbat->type[i] >= nbat->numTypes)
        {
            return fail(NBNXM_ERR_INVALID, "atomdata: type of slot " + std::to_string(i) + " out of range");
        }
    }
    return NBNXM_OK;
}

int validatePairlist(const nbnxm_pairlist_gpu_t* list, int numAtoms)
{
    if (!list || list->na_ci != NBNXM_CLUSTER_SIZE || list->nsci < 0 || list->ncjPacked < 0 || list->nexcl < 1)
    {
        return fail(NBNXM_ERR_INVALID, "pairlist: bad sizes (na_ci must be 8, nexcl >= 1)");
    }
    if ((list->nsci > 0 && !list->sci) || (list->ncjPacked > 0 && !list->cjPacked) || !list->excl)
    {
        return fail(NBNXM_ERR_INVALID, "pairlist: null array");
    }
    const int numClusters = numAtoms / NBNXM_CLUSTER_SIZE;
    for (int i = 0; i < list->nsci; i++)
    {
        const nbnxm_sci_t& s = list->sci[i];
        if (s.shift < 0 || s.shift >= NBNXM_NUM_SHIFTS || s.sci < 0 || (s.sci + 1) * NBNXM_NUM_CLUSTER_PER_SUPERCLUSTER > numClusters
            || s.cjPackedBegin < 0 || s.cjPackedEnd < s.cjPackedBegin || s.cjPackedEnd > list->ncjPacked)
        {
            return fail(NBNXM_ERR_INVALID, "pairlist: sci entry " + std::to_string(i) + " out of range");
        }
    }
    for (int g = 0; g < list->ncjPacked; g++)
    {
        const nbnxm_cj_packed_t& p = list->cjPacked[g];
        if (p.imei[0].excl_ind < 0 || p.imei[0].excl_ind >= list->nexcl)
        {
            return fail(NBNXM_ERR_INVALID, "pairlist: excl_ind out of range in group " + std::to_string(g));
        }
        for (int k = 0; k < NBNXM_JGROUP_SIZE; k++)
        {
            const bool used = (p.imei[0].imask >> (k * NBNXM_NUM_CLUSTER_PER_SUPERCLUSTER)) & 0xFFu;
            if ((used || p.cj[k] != 0) && (p.cj[k] < 0 || p.cj[k] >= numClusters))
            {
                return fail(NBNXM_ERR_INVALID, "pairlist: cj out of range in group " + std::to_string(g));
            }
        }
    }
    return NBNXM_OK;
}

} // namespace nbnxm_amd

using namespace nbnxm_amd;

// ---------------------------------------------------------------------------
// The single-system GPU task (NbnxmGpu, src/gromacs/nbnxm/cuda/nbnxm_cuda_types.h:67-138)
// ---------------------------------------------------------------------------
struct nbnxm_gpu
{
    hipStream_t stream = nullptr;
    nbnxm_interaction_const_t ic{};
    nbnxm_pairlist_params_t listParams{};
    int accumulation = NBNXM_ACCUM_FIXED64;
    int vdwType = NBNXM_VDW_CUT_TYPE_TABLE;
    int numAtoms = 0;
    int numTypes = 0;
    NbLaunchConfig launchConfig;

    DeviceBuffer<float4> xq;
    DeviceBuffer<int> types;
    DeviceBuffer<float2> ljComb;
    DeviceBuffer<float2> nbfp;
    DeviceBuffer<float4> shiftVec;
    DeviceBuffer<float4> paramsA;
    DeviceBuffer<float2> paramsB;
    DeviceBuffer<double> paramsD;
    NbFlavourParams flavour;       // used unless the flavour is the tuned default
    DeviceBuffer<float> ljPmeGrid;  // LJ-PME: grid C6 per type pair
    DeviceBuffer<double> ljPmeGridD;
    DeviceBuffer<float> ewaldTable; // tabulated Ewald: F(r), one padding entry
    DeviceBuffer<double> ewaldTableD;

    int nsci = 0;
    int ncjPacked = 0;
    DeviceBuffer<nbnxm_sci_t> sci;
    DeviceBuffer<int> sciOrder;
    DeviceBuffer<nbnxm_cj_packed_t> cjPacked;
    DeviceBuffer<nbnxm_excl_t> excl;
    DeviceBuffer<unsigned int> imaskOuter;
    DeviceBuffer<int> rollingPart;
    DeviceBuffer<int> sciCount;
    DeviceBuffer<int> sciCountSorted;
    DeviceBuffer<int> iota;
    DeviceBuffer<char> sortTemp;
    bool haveFreshList = false;
    bool haveList = false; // gpu_init_pairlist has run (the list may be empty)
    bool haveSortedOrder = false;

    DeviceBuffer<float> fFloat;
    DeviceBuffer<unsigned long long> fFixed;
    DeviceBuffer<float> fOut;
    DeviceBuffer<double> fOutDouble;
    DeviceBuffer<double> partials;
    DeviceBuffer<double> energies;
    DeviceBuffer<double> fshift;
    DeviceBuffer<int> entryRanges;
    DeviceBuffer<int> entryList;
    DeviceBuffer<int> shiftRanges;
    DeviceBuffer<int> shiftEntryList;
    DeviceBuffer<int> shiftKeys;          // a device list's entry shifts, and sorted
    DeviceBuffer<int> shiftKeysSorted;
    DeviceBuffer<int> slotRange;          // {0, numAtoms}
    bool shiftVecOfBox = false;           // shiftVec holds the shifts of shiftVecBox
    double shiftVecBox[3] = { 0.0, 0.0, 0.0 };
    DeviceBuffer<int> atomIndex;          // grid slot -> atom or -1 (buffer ops)
    DeviceBuffer<int> cellIndex;          // atom -> grid slot (buffer ops)
    int numRealAtoms = -1;                // atoms of the buffer-op map, -1 = none
    bool forcesConsumed = false;          // add_nbat_f_to_f took the forces of the last launch
    bool haveLaunched = false;
    DeviceBuffer<double> virialPartials;  // 9 per block
    DeviceBuffer<double> virial;          // 9
    int virialBlocks = 1;

    double* hEnergies = nullptr; // pinned, 2
    double* hFshift = nullptr;   // pinned, 3*45
    double* hVirial = nullptr;   // pinned, 9
    bool lastComputeEnergy = false;
    bool lastComputeVirial = false;
    hipEvent_t adoptEvent = nullptr; // orders a device list's producer before this task

    // gpu_get_f: every launch also writes its (accumulated) forces as floats
    // into fOut, with the virial partials taken before (fixed-point modes).
    bool externalForces = false;

    // Event timing (gpu_get_timings, GpuTimers of the reference): start/stop
    // events of each timed launch or transfer since the last finish, accounted
    // into `timings` when a finish call finds the task complete.
    struct TimingRecord
    {
        int kind;
        hipEvent_t start;
        hipEvent_t stop;
    };
    bool doTime = false;
    nbnxm_gpu_timings_t timings{};
    std::vector<hipEvent_t> eventPool;
    std::vector<TimingRecord> pendingTimings;

    ~nbnxm_gpu()
    {
        for (const TimingRecord& r : pendingTimings)
        {
            (void)hipEventDestroy(r.start);
            (void)hipEventDestroy(r.stop);
        }
        for (hipEvent_t e : eventPool)
        {
            (void)hipEventDestroy(e);
        }
        if (hEnergies)
        {
            (void)hipHostFree(hEnergies);
        }
        if (hFshift)
        {
            (void)hipHostFree(hFshift);
        }
        if (hVirial)
        {
            (void)hipHostFree(hVirial);
        }
        if (stream)
        {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (adoptEvent)
        {
            (void)hipEventDestroy(adoptEvent);
        }
    }
};

hipStream_t nbnxm_amd::gpuTaskStream(const nbnxm_gpu_t* nb)
{
    return nb ? nb->stream : nullptr;
}

void nbnxm_amd::systemParamBlocks(const nbnxm_interaction_const_t& ic, float4* pa, float2* pb, double pd[8])
{
    const float rc = static_cast<float>(ic.rcoulomb);
    *pa = make_float4(rc * rc, static_cast<float>(ic.epsfac), static_cast<float>(ic.ewaldcoeff_q), static_cast<float>(ic.sh_ewald));
    *pb = make_float2(static_cast<float>(ic.repulsion_shift_cpot), static_cast<float>(ic.dispersion_shift_cpot));
    const double d[8] = { ic.rcoulomb * ic.rcoulomb, ic.epsfac, ic.ewaldcoeff_q, ic.sh_ewald, ic.repulsion_shift_cpot, ic.dispersion_shift_cpot, 0.0, 0.0 };
    for (int k = 0; k < 8; k++)
    {
        pd[k] = d[k];
    }
}

namespace
{

// Kinds of timed work: the nonbonded kernel at c_timeNonbonded + 2*fresh +
// energies (ktime[fresh][energies]), then the two prunes and the transfers.
constexpr int c_timeNonbonded = 0;
constexpr int c_timePrune = 4;
constexpr int c_timeRollingPrune = 5;
constexpr int c_timeXqH2d = 6;
constexpr int c_timeD2h = 7;
constexpr int c_timeListH2d = 8;

hipError_t pooledEvent(nbnxm_gpu* nb, hipEvent_t* e)
{
    if (!nb->eventPool.empty())
    {
        *e = nb->eventPool.back();
        nb->eventPool.pop_back();
        return hipSuccess;
    }
    return hipEventCreate(e);
}

// Opens a timed region on the task's stream (nothing when timing is off).
int timingBegin(nbnxm_gpu* nb, int kind)
{
    if (!nb->doTime)
    {
        return NBNXM_OK;
    }
    nbnxm_gpu::TimingRecord r{ kind, nullptr, nullptr };
    NBNXM_HIP_CHECK(pooledEvent(nb, &r.start));
    NBNXM_HIP_CHECK(pooledEvent(nb, &r.stop));
    nb->pendingTimings.push_back(r);
    NBNXM_HIP_CHECK(hipEventRecord(r.start, nb->stream));
    return NBNXM_OK;
}

// Closes the region timingBegin opened last.
int timingEnd(nbnxm_gpu* nb)
{
    if (!nb->doTime || nb->pendingTimings.empty())
    {
        return NBNXM_OK;
    }
    NBNXM_HIP_CHECK(hipEventRecord(nb->pendingTimings.back().stop, nb->stream));
    return NBNXM_OK;
}

// gpu_accumulate_timings / countPruneKernelTime (gpu_common.h:90-250): the
// regions recorded since the last finish, once the stream has completed them.
int accountTimings(nbnxm_gpu* nb)
{
    bool anyNonbonded = false;
    for (const nbnxm_gpu::TimingRecord& r : nb->pendingTimings)
    {
        float ms = 0.0f;
        NBNXM_HIP_CHECK(hipEventElapsedTime(&ms, r.start, r.stop));
        nbnxm_gpu_timings_t& t = nb->timings;
        if (r.kind < c_timePrune)
        {
            nbnxm_kernel_timing_t& k = t.ktime[(r.kind - c_timeNonbonded) >> 1][(r.kind - c_timeNonbonded) & 1];
            k.t += ms;
            k.c++;
            anyNonbonded = true;
        }
        else if (r.kind == c_timePrune)
        {
            t.pruneTime.t += ms;
            t.pruneTime.c++;
        }
        else if (r.kind == c_timeRollingPrune)
        {
            t.dynamicPruneTime.t += ms;
            t.dynamicPruneTime.c++;
        }
        else if (r.kind == c_timeXqH2d)
        {
            t.nb_h2d_t += ms;
        }
        else if (r.kind == c_timeD2h)
        {
            t.nb_d2h_t += ms;
        }
        else
        {
            t.pl_h2d_t += ms;
            t.pl_h2d_c++;
        }
        nb->eventPool.push_back(r.start);
        nb->eventPool.push_back(r.stop);
    }
    nb->pendingTimings.clear();
    if (anyNonbonded)
    {
        nb->timings.nb_c++;
    }
    return NBNXM_OK;
}

} // namespace

namespace
{

int uploadParams(nbnxm_gpu* nb)
{
    float4 pa;
    float2 pb;
    double pd[8];
    systemParamBlocks(nb->ic, &pa, &pb, pd);
    NBNXM_HIP_CHECK(nb->paramsA.resize(1));
    NBNXM_HIP_CHECK(nb->paramsB.resize(1));
    NBNXM_HIP_CHECK(hipMemcpyAsync(nb->paramsA.ptr, &pa, sizeof(pa), hipMemcpyHostToDevice, nb->stream));
    NBNXM_HIP_CHECK(hipMemcpyAsync(nb->paramsB.ptr, &pb, sizeof(pb), hipMemcpyHostToDevice, nb->stream));
    NBNXM_HIP_CHECK(nb->paramsD.resize(8));
    NBNXM_HIP_CHECK(hipMemcpyAsync(nb->paramsD.ptr, pd, sizeof(pd), hipMemcpyHostToDevice, nb->stream));
    NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    return NBNXM_OK;
}

// The interaction constants of a task: the flavour of its launches, the
// tabulated Ewald correction when selected (init_ewald_coulomb_force_table,
// nbnxm_gpu_data_mgmt.cpp:83-95) and the parameter blocks; the LJ-PME grid
// tables (per type pair, from the atom data) are kept.  The stream is idle
// when this returns.
int setFlavour(nbnxm_gpu* nb, const nbnxm_interaction_const_t& ic)
{
    NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    NbFlavourParams flav = flavourParams(ic);
    flav.ljPmeGrid = nb->ljPmeGrid.ptr;
    flav.ljPmeGridD = nb->ljPmeGridD.ptr;
    if (ic.tabulatedEwald)
    {
        // The extra last entry repeats the end point so that r a rounding step
        // above rc still reads inside the table.
        double scale = 0;
        const int np = nbnxm_ewald_correction_table(ic.ewaldcoeff_q, ic.rcoulomb, nullptr, nullptr, 0, &scale);
        if (np < 2)
        {
            return np < 0 ? np : fail(NBNXM_ERR_INVALID, "gpu_init: Ewald table too short");
        }
        std::vector<float> t(np + 1);
        std::vector<double> td(np + 1);
        nbnxm_ewald_correction_table(ic.ewaldcoeff_q, ic.rcoulomb, t.data(), td.data(), np, &scale);
        t[np] = t[np - 1];
        td[np] = td[np - 1];
        NBNXM_HIP_CHECK(nb->ewaldTable.resize(np + 1));
        NBNXM_HIP_CHECK(nb->ewaldTableD.resize(np + 1));
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->ewaldTable.ptr, t.data(), sizeof(float) * (np + 1), hipMemcpyHostToDevice, nb->stream));
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->ewaldTableD.ptr, td.data(), sizeof(double) * (np + 1), hipMemcpyHostToDevice, nb->stream));
        NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
        flav.ewaldTableScale = scale;
        flav.ewaldTable = nb->ewaldTable.ptr;
        flav.ewaldTableD = nb->ewaldTableD.ptr;
        flav.ewaldTableSize = np + 1;
    }
    nb->ic = ic;
    nb->flavour = flav;
    return uploadParams(nb);
}

} // namespace

namespace
{

// The list of a task from host or device arrays (kind), with the per-shift
// entry lists derived from the host copy of the entries, or on the device
// when there is none (hostSci null: a list that stays on the device).
int setPairlist(nbnxm_gpu* nb,
                const nbnxm_sci_t* hostSci,
                const nbnxm_sci_t* sciSrc,
                int nsci,
                const nbnxm_cj_packed_t* cjSrc,
                int ncj,
                const nbnxm_excl_t* exclSrc,
                int nexcl,
                hipMemcpyKind kind)
{
    // Until the upload below completes there is no usable list.
    nb->haveList = false;
    nb->nsci = nsci;
    nb->ncjPacked = ncj;
    NBNXM_HIP_CHECK(nb->sci.resize(nsci));
    NBNXM_HIP_CHECK(nb->cjPacked.resize(ncj));
    NBNXM_HIP_CHECK(nb->excl.resize(nexcl));
    NBNXM_HIP_CHECK(nb->imaskOuter.resize(ncj));
    NBNXM_HIP_CHECK(nb->sciOrder.resize(nsci));
    NBNXM_HIP_CHECK(nb->sciCount.resize(nsci));
    NBNXM_HIP_CHECK(nb->sciCountSorted.resize(nsci));
    NBNXM_HIP_CHECK(nb->iota.resize(nsci));
    NBNXM_HIP_CHECK(nb->partials.resize(8 * static_cast<size_t>(std::max(nsci, 1))));
    const int wavesPerBlockPrune = 4;
    const int rollingSize = ((nsci + wavesPerBlockPrune - 1) / wavesPerBlockPrune + 1) * wavesPerBlockPrune;
    NBNXM_HIP_CHECK(nb->rollingPart.resize(rollingSize));

    if (nsci > 0)
    {
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->sci.ptr, sciSrc, sizeof(nbnxm_sci_t) * nsci, kind, nb->stream));
    }
    if (ncj > 0)
    {
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->cjPacked.ptr, cjSrc, sizeof(nbnxm_cj_packed_t) * ncj, kind, nb->stream));
    }
    NBNXM_HIP_CHECK(hipMemcpyAsync(nb->excl.ptr, exclSrc, sizeof(nbnxm_excl_t) * nexcl, kind, nb->stream));

    // Identity work order and the per-shift entry lists for the FP64
    // reductions: entry i in shift list s(i), in entry order within a list.
    NBNXM_HIP_CHECK(nb->entryRanges.resize(2));
    NBNXM_HIP_CHECK(nb->entryList.resize(std::max(nsci, 1)));
    NBNXM_HIP_CHECK(nb->shiftRanges.resize(NBNXM_NUM_SHIFTS + 1));
    NBNXM_HIP_CHECK(nb->shiftEntryList.resize(std::max(nsci, 1)));
    size_t shiftSortBytes = 0;
    if (hostSci)
    {
        std::vector<int> iota(nsci);
        std::iota(iota.begin(), iota.end(), 0);
        std::vector<int> shiftRanges(NBNXM_NUM_SHIFTS + 1, 0);
        for (int i = 0; i < nsci; i++)
        {
            shiftRanges[hostSci[i].shift + 1]++;
        }
        for (int k = 0; k < NBNXM_NUM_SHIFTS; k++)
        {
            shiftRanges[k + 1] += shiftRanges[k];
        }
        std::vector<int> shiftList(nsci);
        {
            std::vector<int> fillPos(shiftRanges.begin(), shiftRanges.end() - 1);
            for (int i = 0; i < nsci; i++)
            {
                shiftList[fillPos[hostSci[i].shift]++] = i;
            }
        }
        const int entryRanges[2] = { 0, nsci };
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->entryRanges.ptr, entryRanges, sizeof(entryRanges), hipMemcpyHostToDevice, nb->stream));
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->shiftRanges.ptr, shiftRanges.data(), sizeof(int) * shiftRanges.size(), hipMemcpyHostToDevice, nb->stream));
        if (nsci > 0)
        {
            NBNXM_HIP_CHECK(hipMemcpyAsync(nb->entryList.ptr, iota.data(), sizeof(int) * nsci, hipMemcpyHostToDevice, nb->stream));
            NBNXM_HIP_CHECK(hipMemcpyAsync(nb->sciOrder.ptr, iota.data(), sizeof(int) * nsci, hipMemcpyHostToDevice, nb->stream));
            NBNXM_HIP_CHECK(hipMemcpyAsync(nb->iota.ptr, iota.data(), sizeof(int) * nsci, hipMemcpyHostToDevice, nb->stream));
            NBNXM_HIP_CHECK(hipMemcpyAsync(nb->shiftEntryList.ptr, shiftList.data(), sizeof(int) * nsci, hipMemcpyHostToDevice, nb->stream));
        }
    }
    else
    {
        // The same arrays from the device copy of the entries: identity
        // orders and shift keys, a stable sort of the entries by shift (the
        // entry order within a shift kept), the ranges from the sorted keys.
        NBNXM_HIP_CHECK(nb->shiftKeys.resize(std::max(nsci, 1)));
        NBNXM_HIP_CHECK(nb->shiftKeysSorted.resize(std::max(nsci, 1)));
        NBNXM_HIP_CHECK(launchEntryArrays(nb->sci.ptr, nsci, nb->entryList.ptr, nb->sciOrder.ptr, nb->iota.ptr, nb->shiftKeys.ptr, nb->entryRanges.ptr,
                                          nb->stream));
        int shiftBits = 1;
        while ((1 << shiftBits) < NBNXM_NUM_SHIFTS)
        {
            shiftBits++;
        }
        NBNXM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, shiftSortBytes, nb->shiftKeys.ptr, nb->shiftKeysSorted.ptr, nb->iota.ptr,
                                                           nb->shiftEntryList.ptr, std::max(nsci, 1), 0, shiftBits, nb->stream));
        // Sized once for this sort and the sci sort, before either runs.
        size_t sciSortBytes = 0;
        NBNXM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sciSortBytes, nb->sciCount.ptr, nb->sciCountSorted.ptr, nb->iota.ptr, nb->sciOrder.ptr,
                                                           std::max(nsci, 1), 0, 14, nb->stream));
        NBNXM_HIP_CHECK(nb->sortTemp.resize(std::max<size_t>(std::max(shiftSortBytes, sciSortBytes), 1)));
        if (nsci > 0)
        {
            NBNXM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nb->sortTemp.ptr, shiftSortBytes, nb->shiftKeys.ptr, nb->shiftKeysSorted.ptr, nb->iota.ptr,
                                                               nb->shiftEntryList.ptr, nsci, 0, shiftBits, nb->stream));
        }
        NBNXM_HIP_CHECK(launchShiftRanges(nb->shiftKeysSorted.ptr, nsci, nb->shiftRanges.ptr, nb->stream));
    }
    NBNXM_HIP_CHECK(launchFillI32(nb->rollingPart.ptr, rollingSize, 0, nb->stream));
    // Outer masks start empty: the prune kernels write those of the groups
    // with cluster pairs, so a group uploaded without any keeps 0 rather than
    // what an earlier list left there.
    NBNXM_HIP_CHECK(launchFillI32(reinterpret_cast<int*>(nb->imaskOuter.ptr), ncj, 0, nb->stream));
    NBNXM_HIP_CHECK(launchFillI32(reinterpret_cast<int*>(nb->partials.ptr), 2 * 8 * static_cast<size_t>(std::max(nsci, 1)), 0, nb->stream));

    // Radix-sort scratch for the sci ordering (the shift sort above, if any,
    // ran before it on the same stream).
    size_t tempBytes = 0;
    NBNXM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tempBytes, nb->sciCount.ptr, nb->sciCountSorted.ptr, nb->iota.ptr, nb->sciOrder.ptr, std::max(nsci, 1), 0, 14, nb->stream));
    NBNXM_HIP_CHECK(nb->sortTemp.resize(std::max<size_t>(std::max(tempBytes, shiftSortBytes), 1)));
    if (hostSci)
    {
        // The copies above read host arrays of this call.
        NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    }
    nb->haveFreshList = true;
    nb->haveSortedOrder = false;
    nb->haveList = true;
    return NBNXM_OK;
}

} // namespace

extern "C" {

const char* nbnxm_last_error(void)
{
    return t_lastError.c_str();
}

const char* nbnxm_version(void)
{
    return "nbnxm_amd 0.1 gfx950";
}

double nbnxm_calc_ewaldcoeff_q(double rc, double rtol)
{
    // Same doubling-then-bisection as calc_ewaldcoeff_q (src/gromacs/ewald/ewald_utils.cpp:43-71).
    double beta = 5.0;
    int i = 0;
    do
    {
        i++;
        beta *= 2.0;
    } while (std::erfc(beta * rc) > rtol);
    const int n = i + 60;
    double low = 0.0, high = beta;
    for (int k = 0; k < n; k++)
    {
        beta = 0.5 * (low + high);
        if (std::erfc(beta * rc) > rtol)
        {
            low = beta;
        }
        else
        {
            high = beta;
        }
    }
    return beta;
}

int nbnxm_init_interaction_const(double rc, double ewaldRtol, double epsilonR, nbnxm_interaction_const_t* ic)
{
    if (!ic || !(rc > 0) || !(ewaldRtol > 0) || !(ewaldRtol < 1))
    {
        return fail(NBNXM_ERR_INVALID, "init_interaction_const: need rc > 0 and 0 < rtol < 1");
    }
    nbnxm_ic_options_t o{};
    o.eeltype = NBNXM_ELEC_EWALD;
    o.vdwModifier = NBNXM_VDW_MOD_POTSHIFT;
    o.rcoulomb = rc;
    o.rvdw = rc;
    o.ewaldRtol = ewaldRtol;
    o.epsilonR = epsilonR;
    return nbnxm_init_interaction_const_ex(&o, ic);
}

double nbnxm_calc_ewaldcoeff_lj(double rc, double rtol)
{
    // Root of exp(-x^2)(1 + x^2 + x^4/2) = rtol at x = beta*rc, found like
    // calc_ewaldcoeff_lj (src/gromacs/ewald/ewald_utils.cpp:73-112): double
    // beta from 10 until below rtol, then bisect with i + 60 halvings.
    const auto ljFunction = [rc](double beta) {
        const double x2 = (beta * rc) * (beta * rc);
        return std::exp(-x2) * (1.0 + x2 + 0.5 * x2 * x2);
    };
    double beta = 5.0;
    int i = 0;
    do
    {
        i++;
        beta *= 2.0;
    } while (ljFunction(beta) > rtol);
    const int n = i + 60;
    double low = 0.0, high = beta;
    for (int k = 0; k < n; k++)
    {
        beta = 0.5 * (low + high);
        if (ljFunction(beta) > rtol)
        {
            low = beta;
        }
        else
        {
            high = beta;
        }
    }
    return beta;
}

int nbnxm_ewald_correction_table(double beta, double rc, float* tableF, double* tableFD, int capacity, double* scale)
{
    if (!(beta > 0) || !(rc > 0))
    {
        return fail(NBNXM_ERR_INVALID, "ewald_correction_table: need beta > 0 and rc > 0");
    }
    // Scale: the larger of the FP32 force-accuracy and energy-accuracy
    // spacings for a quadratic spline of erf(x)/x (max third derivative
    // 1.0522), energy tolerance 0.1 erfc(beta rc), in the mixed-precision
    // build's FP32 arithmetic for beta and the result.
    const float betaF = static_cast<float>(beta);
    const double floatEps = static_cast<double>(1.19209290e-07F);
    const double thirdDerivMax = 1.0522;
    const float erfcRc = std::erfc(betaF * static_cast<float>(rc));
    const double funcTol = std::max(0.1 * erfcRc, floatEps);
    const double scDeriv = std::sqrt(thirdDerivMax / (6 * 4 * floatEps * betaF)) * betaF;
    const double scFunc = std::cbrt(thirdDerivMax / (6 * 12 * std::sqrt(3.0) * funcTol)) * betaF;
    const float scaleF = static_cast<float>(std::max(scDeriv, scFunc));
    const int numPoints = static_cast<int>(static_cast<float>(rc) * scaleF) + 2;
    if (scale)
    {
        *scale = scaleF;
    }
    if (!tableF && !tableFD)
    {
        return numPoints;
    }
    if (capacity < numPoints)
    {
        return fail(NBNXM_ERR_INVALID, "ewald_correction_table: capacity below the number of points");
    }
    // Quadratic splines through v at the interval ends whose derivative matches
    // the secant plus half the mid-point second difference; the force at each
    // point is the average spline derivative of its two intervals.  erf(br)/r
    // is finite on the whole range, so the out-of-range continuation of the
    // general generator never applies.
    const double dx = 1.0 / static_cast<double>(scaleF);
    const auto v = [betaF](double r) {
        return r == 0 ? static_cast<double>(betaF) * 2 / std::sqrt(M_PI) : std::erf(static_cast<double>(betaF) * r) / r;
    };
    std::vector<double> f(numPoints, 0.0);
    double dc = 0;
    for (int i = numPoints - 1; i >= 1; i--)
    {
        const double xr0 = i * dx;
        const double vr0 = v(xr0);
        const double vr1 = v((i - 1) * dx);
        double a2dx = (vr0 + vr1 - 2 * v(xr0 - 0.5 * dx)) / (0.25 * dx);
        dc = (vr0 - vr1) / dx + 0.5 * a2dx;
        if (i == numPoints - 1)
        {
            f[i] = -dc;
        }
        else
        {
            f[i] += -0.5 * dc;
        }
        const double a0 = vr0;
        const double a1 = dc;
        a2dx = (a1 * dx + vr1 - a0) * 2 / dx;
        dc = a1 - a2dx;
        f[i - 1] = -0.5 * dc;
    }
    f[0] *= 2;
    for (int i = 0; i < numPoints; i++)
    {
        if (tableF)
        {
            tableF[i] = static_cast<float>(f[i]);
        }
        if (tableFD)
        {
            tableFD[i] = f[i];
        }
    }
    return numPoints;
}

int nbnxm_init_interaction_const_ex(const nbnxm_ic_options_t* o, nbnxm_interaction_const_t* ic)
{
    if (!o || !ic)
    {
        return fail(NBNXM_ERR_INVALID, "init_interaction_const_ex: null pointer");
    }
    const double rc = o->rcoulomb;
    const double rvdw = (o->rvdw > 0) ? o->rvdw : rc;
    if (!(rc > 0) || !(rvdw <= rc))
    {
        return fail(NBNXM_ERR_INVALID, "init_interaction_const_ex: need rcoulomb > 0 and 0 < rvdw <= rcoulomb");
    }
    if (o->eeltype != NBNXM_ELEC_EWALD && o->eeltype != NBNXM_ELEC_RF)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "init_interaction_const_ex: unknown Coulomb type");
    }
    if (o->vdwModifier < NBNXM_VDW_MOD_POTSHIFT || o->vdwModifier > NBNXM_VDW_MOD_POTSWITCH)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "init_interaction_const_ex: unknown LJ modifier");
    }
    if (o->vdwModifier != NBNXM_VDW_MOD_POTSHIFT && !(o->rvdw_switch >= 0 && o->rvdw_switch < rvdw))
    {
        return fail(NBNXM_ERR_INVALID, "init_interaction_const_ex: switched LJ needs 0 <= rvdw_switch < rvdw");
    }
    if (o->ljPme < NBNXM_LJPME_NONE || o->ljPme > NBNXM_LJPME_LB)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "init_interaction_const_ex: unknown LJ-PME grid rule");
    }
    // LJ-PME exists with the potential shift only (src/gromacs/nbnxm/simd_kernel.h:227-228).
    if (o->ljPme && o->vdwModifier != NBNXM_VDW_MOD_POTSHIFT)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "init_interaction_const_ex: LJ-PME supports the potential-shift modifier only");
    }
    if (o->eeltype == NBNXM_ELEC_EWALD && !(o->ewaldRtol > 0 && o->ewaldRtol < 1))
    {
        return fail(NBNXM_ERR_INVALID, "init_interaction_const_ex: Ewald needs 0 < ewaldRtol < 1");
    }
    if (o->ljPme && !(o->ewaldRtolLj > 0 && o->ewaldRtolLj < 1))
    {
        return fail(NBNXM_ERR_INVALID, "init_interaction_const_ex: LJ-PME needs 0 < ewaldRtolLj < 1");
    }
    nbnxm_interaction_const_t c{};
    c.rcoulomb = rc;
    c.rvdw = rvdw;
    c.eeltype = o->eeltype;
    c.vdwModifier = o->vdwModifier;
    c.ljPme = o->ljPme;
    c.rvdw_switch = (o->vdwModifier == NBNXM_VDW_MOD_POTSHIFT) ? 0.0 : o->rvdw_switch;
    c.epsfac = (o->epsilonR != 0) ? one4PiEps0() / o->epsilonR : 0.0;

    // Lennard-Jones modifier constants (interaction_const.cpp:199-246, 284-305).
    const double rsw = c.rvdw_switch;
    const auto forceSwitch = [rsw, rvdw](double p, double* c2, double* c3, double* cpot) {
        const double d = rvdw - rsw;
        *c2 = ((p + 1) * rsw - (p + 4) * rvdw) / (std::pow(rvdw, p + 2) * d * d);
        *c3 = -((p + 1) * rsw - (p + 3) * rvdw) / (std::pow(rvdw, p + 2) * d * d * d);
        *cpot = -std::pow(rvdw, -p) + p * (*c2) / 3 * d * d * d + p * (*c3) / 4 * d * d * d * d;
    };
    switch (o->vdwModifier)
    {
        case NBNXM_VDW_MOD_POTSHIFT:
            c.dispersion_shift_cpot = -1.0 / std::pow(rvdw, 6);
            c.repulsion_shift_cpot = -1.0 / std::pow(rvdw, 12);
            break;
        case NBNXM_VDW_MOD_FORCESWITCH:
            forceSwitch(6.0, &c.dispersion_shift_c2, &c.dispersion_shift_c3, &c.dispersion_shift_cpot);
            forceSwitch(12.0, &c.repulsion_shift_c2, &c.repulsion_shift_c3, &c.repulsion_shift_cpot);
            break;
        default:
        {
            const double d = rvdw - rsw;
            c.vdw_switch_c3 = -10.0 / (d * d * d);
            c.vdw_switch_c4 = 15.0 / (d * d * d * d);
            c.vdw_switch_c5 = -6.0 / (d * d * d * d * d);
            break;
        }
    }
    if (c.ljPme)
    {
        // initVdwEwaldParameters (interaction_const.cpp:122-148).
        c.ewaldcoeff_lj = nbnxm_calc_ewaldcoeff_lj(rvdw, o->ewaldRtolLj);
        const double crc2 = (c.ewaldcoeff_lj * rvdw) * (c.ewaldcoeff_lj * rvdw);
        c.sh_lj_ewald = (std::exp(-crc2) * (1 + crc2 + 0.5 * crc2 * crc2) - 1) / std::pow(rvdw, 6);
    }

    // Coulomb: Ewald with the potential shift (interaction_const.cpp:64-120), or
    // reaction field / plain cut-off (:320-345 with calc_rffac, rf_util.cpp:50-62).
    if (o->tabulatedEwald && o->eeltype != NBNXM_ELEC_EWALD)
    {
        return fail(NBNXM_ERR_INVALID, "init_interaction_const_ex: tabulated correction needs Ewald");
    }
    c.tabulatedEwald = o->tabulatedEwald ? 1 : 0;
    if (o->eeltype == NBNXM_ELEC_EWALD)
    {
        c.ewaldcoeff_q = nbnxm_calc_ewaldcoeff_q(rc, o->ewaldRtol);
        c.sh_ewald = std::erfc(c.ewaldcoeff_q * rc) / rc;
        c.k_rf = 0.0;
        c.c_rf = 1.0 / rc;
    }
    else if (o->plainCutoff)
    {
        c.k_rf = 0.0;
        c.c_rf = 1.0 / rc;
    }
    else
    {
        c.k_rf = (o->epsilonRf == 0) ? 1.0 / (2 * rc * rc * rc)
                                      : (o->epsilonRf - o->epsilonR) / (2 * o->epsilonRf + o->epsilonR) / (rc * rc * rc);
        c.c_rf = 1.0 / rc + c.k_rf * rc * rc;
    }
    *ic = c;
    return NBNXM_OK;
}

int nbnxm_calc_shifts(const double box[3], float* shiftVec)
{
    if (!box || !shiftVec)
    {
        return fail(NBNXM_ERR_INVALID, "calc_shifts: null pointer");
    }
    for (int tz = -1, k = 0; tz <= 1; tz++)
    {
        for (int ty = -1; ty <= 1; ty++)
        {
            for (int tx = -2; tx <= 2; tx++, k++)
            {
                shiftVec[3 * k + 0] = static_cast<float>(tx) * static_cast<float>(box[0]);
                shiftVec[3 * k + 1] = static_cast<float>(ty) * static_cast<float>(box[1]);
                shiftVec[3 * k + 2] = static_cast<float>(tz) * static_cast<float>(box[2]);
            }
        }
    }
    return NBNXM_OK;
}

// ---------------------------------------------------------------------------
// Pair search (producer of the list and of the grid-order atom data)
// ---------------------------------------------------------------------------
struct nbnxm_search
{
    GridSearch grid;
};

nbnxm_search_t* nbnxm_search_create(void)
{
    return new (std::nothrow) nbnxm_search();
}

void nbnxm_search_free(nbnxm_search_t* s)
{
    delete s;
}

int nbnxm_search_put_atoms_on_grid(nbnxm_search_t* s, const double box[3], int numAtoms, const double* x)
{
    if (!s)
    {
        return fail(NBNXM_ERR_INVALID, "search: null handle");
    }
    const int rc = s->grid.putAtomsOnGrid(box, numAtoms, x);
    return rc == NBNXM_OK ? rc : fail(rc, "search: invalid box or coordinates");
}

int nbnxm_search_construct_pairlist(nbnxm_search_t* s, double rlist, const int* exclRanges, const int* exclElements, int maxPairsPerEntry)
{
    if (!s)
    {
        return fail(NBNXM_ERR_INVALID, "search: null handle");
    }
    const int rc = s->grid.constructPairlist(rlist, exclRanges, exclElements, maxPairsPerEntry);
    return rc == NBNXM_OK ? rc : fail(rc, "search: every box edge must exceed rlist > 0, exclusions need ranges and elements");
}

int nbnxm_search_get_pairlist(const nbnxm_search_t* s, nbnxm_pairlist_gpu_t* list)
{
    if (!s || !list)
    {
        return fail(NBNXM_ERR_INVALID, "search: null pointer");
    }
    *list = s->grid.listView();
    return NBNXM_OK;
}

int nbnxm_search_num_grid_atoms(const nbnxm_search_t* s)
{
    return s ? s->grid.numGridAtoms() : NBNXM_ERR_INVALID;
}

int nbnxm_search_atom_indices(const nbnxm_search_t* s, int* atomIndices)
{
    if (!s || !atomIndices)
    {
        return fail(NBNXM_ERR_INVALID, "search: null pointer");
    }
    std::copy(s->grid.atomIndices().begin(), s->grid.atomIndices().end(), atomIndices);
    return NBNXM_OK;
}

int nbnxm_search_shift_vec(const nbnxm_search_t* s, float* shiftVec)
{
    if (!s || !shiftVec)
    {
        return fail(NBNXM_ERR_INVALID, "search: null pointer");
    }
    std::memcpy(shiftVec, s->grid.shiftVec(), sizeof(float) * 3 * NBNXM_NUM_SHIFTS);
    return NBNXM_OK;
}

int nbnxm_search_fill_xq(const nbnxm_search_t* s, const double* x, const double* q, float* xq)
{
    if (!s || !xq)
    {
        return fail(NBNXM_ERR_INVALID, "search: null pointer");
    }
    s->grid.fillXq(x, q, xq);
    return NBNXM_OK;
}

int nbnxm_search_fill_types(const nbnxm_search_t* s, const int* type, int dummyType, int* slotType)
{
    if (!s || !type || !slotType)
    {
        return fail(NBNXM_ERR_INVALID, "search: null pointer");
    }
    s->grid.fillTypes(type, dummyType, slotType);
    return NBNXM_OK;
}

int nbnxm_search_fill_ljcomb(const nbnxm_search_t* s, const int* type, const float* perTypeComb, float* slotComb)
{
    if (!s || !type || !perTypeComb || !slotComb)
    {
        return fail(NBNXM_ERR_INVALID, "search: null pointer");
    }
    s->grid.fillLjComb(type, perTypeComb, slotComb);
    return NBNXM_OK;
}

int nbnxm_search_count_pairs(const nbnxm_search_t* s, double rc, int64_t* count)
{
    if (!s || !count)
    {
        return fail(NBNXM_ERR_INVALID, "search_count_pairs: null pointer");
    }
    if (!(rc > 0) || rc > s->grid.rlist() * (1.0 + 1e-6))
    {
        return fail(NBNXM_ERR_INVALID, "search_count_pairs: need 0 < rc <= the list radius");
    }
    *count = s->grid.countPairsWithin(rc);
    return NBNXM_OK;
}

// ---------------------------------------------------------------------------
// GPU task
// ---------------------------------------------------------------------------
int nbnxm_gpu_init(const nbnxm_interaction_const_t* ic,
                   const nbnxm_pairlist_params_t* listParams,
                   const nbnxm_atomdata_t* nbat,
                   int accumulation,
                   nbnxm_gpu_t** out)
{
    if (!ic || !listParams || !nbat || !out)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_init: null pointer");
    }
    *out = nullptr;
    if (accumulation < NBNXM_ACCUM_FLOAT || accumulation > NBNXM_ACCUM_FP64)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_init: unknown accumulation mode");
    }
    int frc = validateFlavour(ic);
    if (frc != NBNXM_OK)
    {
        return frc;
    }
    if (!(listParams->rlistOuter >= ic->rcoulomb) || !(listParams->rlistInner >= ic->rcoulomb)
        || listParams->rlistInner > listParams->rlistOuter || listParams->numRollingPruningParts < 1)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_init: need rc <= rlistInner <= rlistOuter and numRollingPruningParts >= 1");
    }
    int rc = validateAtomdata(nbat);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    int device = 0;
    NBNXM_HIP_CHECK(hipGetDevice(&device));
    std::unique_ptr<nbnxm_gpu> nb(new (std::nothrow) nbnxm_gpu());
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_init: out of host memory");
    }
    NBNXM_HIP_CHECK(hipStreamCreateWithFlags(&nb->stream, hipStreamNonBlocking));
    NBNXM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&nb->hEnergies), 2 * sizeof(double), hipHostMallocDefault));
    NBNXM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&nb->hFshift), 3 * NBNXM_NUM_SHIFTS * sizeof(double), hipHostMallocDefault));
    NBNXM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&nb->hVirial), 9 * sizeof(double), hipHostMallocDefault));
    nb->listParams = *listParams;
    nb->accumulation = accumulation;
    nb->doTime = std::getenv("GMX_ENABLE_GPU_TIMING") != nullptr;
    rc = setFlavour(nb.get(), *ic);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    NBNXM_HIP_CHECK(nb->energies.resize(2));
    NBNXM_HIP_CHECK(nb->fshift.resize(3 * NBNXM_NUM_SHIFTS));
    rc = nbnxm_gpu_init_atomdata(nb.get(), nbat);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    if (nbat->shiftVec)
    {
        rc = nbnxm_gpu_upload_shiftvec(nb.get(), nbat->shiftVec);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
    }
    NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    *out = nb.release();
    return NBNXM_OK;
}

int nbnxm_gpu_init_atomdata(nbnxm_gpu_t* nb, const nbnxm_atomdata_t* nbat)
{
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_init_atomdata: null handle");
    }
    int rc = validateAtomdata(nbat);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    if (nb->ic.ljPme)
    {
        rc = validateLjPmeAtomdata(nbat);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
    }
    const int n = nbat->numAtoms;
    if (n != nb->numAtoms)
    {
        // The list, the buffer-op maps and the pending forces index the old
        // grid: until gpu_init_pairlist / gpu_init_x_to_nbat_x run again the
        // calls that read them return NBNXM_ERR_STATE.
        nb->haveList = false;
        nb->haveFreshList = false;
        nb->haveSortedOrder = false;
        nb->numRealAtoms = -1;
        nb->haveLaunched = false;
    }
    nb->numAtoms = n;
    nb->vdwType = nbat->vdwType;
    nb->numTypes = nbat->numTypes;
    rc = timingBegin(nb, c_timeListH2d);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    NBNXM_HIP_CHECK(nb->xq.resize(n));
    NBNXM_HIP_CHECK(nb->fFixed.resize(3 * static_cast<size_t>(n)));
    NBNXM_HIP_CHECK(nb->fFloat.resize(3 * static_cast<size_t>(n)));
    NBNXM_HIP_CHECK(nb->fOut.resize(3 * static_cast<size_t>(n)));
    NBNXM_HIP_CHECK(nb->fOutDouble.resize(3 * static_cast<size_t>(n)));
    if (n > 0)
    {
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->xq.ptr, nbat->xq, sizeof(float4) * n, hipMemcpyHostToDevice, nb->stream));
    }
    if (nbat->vdwType == NBNXM_VDW_CUT_TYPE_TABLE || nb->ic.ljPme)
    {
        NBNXM_HIP_CHECK(nb->types.resize(n));
        NBNXM_HIP_CHECK(nb->nbfp.resize(static_cast<size_t>(nbat->numTypes) * nbat->numTypes));
        if (n > 0)
        {
            NBNXM_HIP_CHECK(hipMemcpyAsync(nb->types.ptr, nbat->type, sizeof(int) * n, hipMemcpyHostToDevice, nb->stream));
        }
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->nbfp.ptr, nbat->nbfp, sizeof(float2) * nbat->numTypes * nbat->numTypes, hipMemcpyHostToDevice, nb->stream));
    }
    if (nbat->vdwType != NBNXM_VDW_CUT_TYPE_TABLE)
    {
        NBNXM_HIP_CHECK(nb->ljComb.resize(n));
        if (n > 0)
        {
            NBNXM_HIP_CHECK(hipMemcpyAsync(nb->ljComb.ptr, nbat->ljComb, sizeof(float2) * n, hipMemcpyHostToDevice, nb->stream));
        }
    }
    if (nb->ic.ljPme)
    {
        // Grid C6 of every type pair by the grid rule, from the nbfp diagonal
        // (atomdata.cpp:296-330).
        const size_t nt2 = static_cast<size_t>(nbat->numTypes) * nbat->numTypes;
        std::vector<float> g(nt2);
        std::vector<double> gd(nt2);
        ljPmeGridTables(nb->ic.ljPme, nbat->numTypes, nbat->nbfp, g.data(), gd.data());
        NBNXM_HIP_CHECK(nb->ljPmeGrid.resize(nt2));
        NBNXM_HIP_CHECK(nb->ljPmeGridD.resize(nt2));
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->ljPmeGrid.ptr, g.data(), sizeof(float) * nt2, hipMemcpyHostToDevice, nb->stream));
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->ljPmeGridD.ptr, gd.data(), sizeof(double) * nt2, hipMemcpyHostToDevice, nb->stream));
        // The copies read host vectors that go out of scope: finish them here.
        NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
        nb->flavour.ljPmeGrid = nb->ljPmeGrid.ptr;
        nb->flavour.ljPmeGridD = nb->ljPmeGridD.ptr;
    }
    NBNXM_HIP_CHECK(launchFillU64(nb->fFixed.ptr, 3 * static_cast<size_t>(n), nb->stream));
    NBNXM_HIP_CHECK(launchFillI32(reinterpret_cast<int*>(nb->fFloat.ptr), 3 * static_cast<size_t>(n), 0, nb->stream));
    // Virial reduction geometry: about 1024 slots per block, at most 512 blocks.
    nb->virialBlocks = std::min(512, std::max(1, (n + 1023) / 1024));
    const int range[2] = { 0, n };
    NBNXM_HIP_CHECK(nb->slotRange.resize(2));
    NBNXM_HIP_CHECK(nb->virialPartials.resize(9 * static_cast<size_t>(nb->virialBlocks)));
    NBNXM_HIP_CHECK(nb->virial.resize(9));
    NBNXM_HIP_CHECK(hipMemcpyAsync(nb->slotRange.ptr, range, sizeof(range), hipMemcpyHostToDevice, nb->stream));
    rc = timingEnd(nb);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    // Host buffers of the caller may be reused after return.
    NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    return NBNXM_OK;
}

int nbnxm_gpu_upload_shiftvec(nbnxm_gpu_t* nb, const float* shiftVec)
{
    if (!nb || !shiftVec)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_upload_shiftvec: null pointer");
    }
    float4 sv[NBNXM_NUM_SHIFTS];
    for (int k = 0; k < NBNXM_NUM_SHIFTS; k++)
    {
        sv[k] = make_float4(shiftVec[3 * k], shiftVec[3 * k + 1], shiftVec[3 * k + 2], 0.0f);
    }
    NBNXM_HIP_CHECK(nb->shiftVec.resize(NBNXM_NUM_SHIFTS));
    NBNXM_HIP_CHECK(hipMemcpyAsync(nb->shiftVec.ptr, sv, sizeof(sv), hipMemcpyHostToDevice, nb->stream));
    NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    nb->shiftVecOfBox = false;
    return NBNXM_OK;
}

int nbnxm_gpu_init_pairlist(nbnxm_gpu_t* nb, const nbnxm_pairlist_gpu_t* list)
{
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_init_pairlist: null handle");
    }
    int rc = validatePairlist(list, nb->numAtoms);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    if (!(list->excl[0].pair[0] == 0xFFFFFFFFu))
    {
        return fail(NBNXM_ERR_INVALID, "gpu_init_pairlist: excl[0] must be the all-interaction mask");
    }
    rc = timingBegin(nb, c_timeListH2d);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    rc = setPairlist(nb, list->sci, list->sci, list->nsci, list->cjPacked, list->ncjPacked, list->excl, list->nexcl, hipMemcpyHostToDevice);
    return rc != NBNXM_OK ? rc : timingEnd(nb);
}

int nbnxm_gpu_copy_xq_to_gpu(nbnxm_gpu_t* nb, const float* xq, int numAtoms)
{
    if (!nb || (!xq && numAtoms > 0))
    {
        return fail(NBNXM_ERR_INVALID, "gpu_copy_xq_to_gpu: null pointer");
    }
    if (numAtoms != nb->numAtoms)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_copy_xq_to_gpu: atom count differs from gpu_init_atomdata");
    }
    if (numAtoms > 0)
    {
        int rc = timingBegin(nb, c_timeXqH2d);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->xq.ptr, xq, sizeof(float4) * numAtoms, hipMemcpyHostToDevice, nb->stream));
        rc = timingEnd(nb);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
        NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    }
    return NBNXM_OK;
}

int nbnxm_gpu_clear_outputs(nbnxm_gpu_t* nb, int computeVirial)
{
    (void)computeVirial;
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_clear_outputs: null handle");
    }
    NBNXM_HIP_CHECK(launchFillU64(nb->fFixed.ptr, 3 * static_cast<size_t>(nb->numAtoms), nb->stream));
    NBNXM_HIP_CHECK(launchFillI32(reinterpret_cast<int*>(nb->fFloat.ptr), 3 * static_cast<size_t>(nb->numAtoms), 0, nb->stream));
    return NBNXM_OK;
}

// After the prune of a fresh list (by the prune kernel or the combined one):
// entries ordered by descending pruned pair count (the role of
// nbnxnKernelBucketSciSort, src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel_sci_sort.cuh:66-87);
// a stable radix sort on each entry's bin (NBNXM_SCI_HISTOGRAM_SIZE - 1 -
// count, clamped at 0) keeps the order deterministic.  The reference's
// counting sort needs the bins' histogram; the radix sort does not, so no
// atomic counts are kept (on a list of 15,000 nearly equal entries those
// atomics on a few bins doubled the prune kernel's time).
static int sortEntriesOfFreshList(nbnxm_gpu_t* nb)
{
    if (nb->nsci > 0)
    {
        size_t tempBytes = nb->sortTemp.size;
        NBNXM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nb->sortTemp.ptr, tempBytes, nb->sciCount.ptr, nb->sciCountSorted.ptr, nb->iota.ptr, nb->sciOrder.ptr, nb->nsci, 0, 14, nb->stream));
    }
    nb->haveSortedOrder = true;
    nb->haveFreshList = false;
    // Rolling pruning restarts from part 0 after a fresh list.
    NBNXM_HIP_CHECK(launchFillI32(nb->rollingPart.ptr, nb->rollingPart.size, 0, nb->stream));
    return NBNXM_OK;
}

static int launchPruneImpl(nbnxm_gpu_t* nb, int numParts)
{
    PruneArgs pa;
    pa.xq = nb->xq.ptr;
    pa.shiftVec = nb->shiftVec.ptr;
    pa.sci = nb->sci.ptr;
    pa.sciOrder = (nb->haveFreshList || !nb->haveSortedOrder) ? nullptr : nb->sciOrder.ptr;
    pa.nsci = nb->nsci;
    pa.cjPacked = nb->cjPacked.ptr;
    pa.imaskOuter = nb->imaskOuter.ptr;
    pa.rollingPart = nb->rollingPart.ptr;
    pa.sciCount = nb->sciCount.ptr;
    // Without dynamic pruning the list is pruned to rlistOuter only, as the
    // reference's combined kernel does.
    const float ro = static_cast<float>(nb->listParams.rlistOuter);
    const float ri = static_cast<float>(nb->listParams.useDynamicPruning ? nb->listParams.rlistInner : nb->listParams.rlistOuter);
    pa.rlistOuter_sq = ro * ro;
    pa.rlistInner_sq = ri * ri;
    pa.numParts = numParts;
    pa.numGroups = nb->ncjPacked;
    const bool fresh = nb->haveFreshList;
    if (fresh)
    {
        if (numParts != 1)
        {
            return fail(NBNXM_ERR_INVALID, "gpu_launch_kernel_pruneonly: a fresh list is pruned in one part");
        }
    }
    int trc = timingBegin(nb, fresh ? c_timePrune : c_timeRollingPrune);
    if (trc != NBNXM_OK)
    {
        return trc;
    }
    NBNXM_HIP_CHECK(launchPrune(pa, fresh, nb->stream));
    trc = timingEnd(nb);
    if (trc != NBNXM_OK)
    {
        return trc;
    }
    if (fresh)
    {
        return sortEntriesOfFreshList(nb);
    }
    return NBNXM_OK;
}

int nbnxm_gpu_launch_kernel_pruneonly(nbnxm_gpu_t* nb, int numParts)
{
    if (!nb || numParts < 1)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_launch_kernel_pruneonly: null handle or numParts < 1");
    }
    if (!nb->haveList)
    {
        return fail(NBNXM_ERR_STATE, "gpu_launch_kernel_pruneonly: no pair list uploaded");
    }
    return launchPruneImpl(nb, numParts);
}

int nbnxm_gpu_launch_kernel(nbnxm_gpu_t* nb, int computeEnergy, int computeVirial)
{
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_launch_kernel: null handle");
    }
    if (!nb->shiftVec.ptr || !nb->haveList)
    {
        return fail(NBNXM_ERR_STATE, "gpu_launch_kernel: shift vectors and pair list must be uploaded first");
    }
    // Which kernel runs: the tuned list walk (default flavour), the list walk
    // with a flavour class, or the standalone flavour kernel (FP64 mode of the
    // other flavours, LJ-PME with a combination-rule kernel, type tables
    // beyond the LDS staging).
    // The tuned list walk stages the type table in LDS up to 65 types (64 + dummy).
    const bool tableFitsLds = nb->vdwType != NBNXM_VDW_CUT_TYPE_TABLE || nb->numTypes <= c_maxLdsTableTypes;
    const bool tunedWalk = isDefaultFlavour(nb->ic) && (tableFitsLds || nb->accumulation == NBNXM_ACCUM_FP64);
    const bool flavouredWalk = !tunedWalk && nb->accumulation != NBNXM_ACCUM_FP64 && tableFitsLds
                               && !(nb->ic.ljPme && nb->vdwType != NBNXM_VDW_CUT_TYPE_TABLE);
    // A fresh list: with dynamic pruning it is pruned first by the prune
    // kernel (src/gromacs/nbnxm/cuda/nbnxm_cuda.cu:505-513); without, the
    // FP32 list walks prune it to rlistOuter themselves (the _prune kernels,
    // :565-578), and the entries are sorted after the launch.
    const bool freshList = nb->haveFreshList;
    const bool combinedPrune = freshList && !nb->listParams.useDynamicPruning && (tunedWalk || flavouredWalk)
                               && nb->accumulation != NBNXM_ACCUM_FP64;
    if (freshList && !combinedPrune)
    {
        const int rc = launchPruneImpl(nb, 1);
        if (rc != NBNXM_OK)
        {
            return rc;
        }
    }
    NbKernelArgs a;
    a.xq = nb->xq.ptr;
    a.atomTypes = nb->types.ptr;
    a.ljComb = nb->ljComb.ptr;
    a.nbfp = nb->nbfp.ptr;
    a.numTypes = nb->numTypes;
    a.shiftVec = nb->shiftVec.ptr;
    a.sci = nb->sci.ptr;
    a.sciOrder = nb->haveSortedOrder ? nb->sciOrder.ptr : nullptr;
    a.nsci = nb->nsci;
    a.cjPacked = nb->cjPacked.ptr;
    a.excl = nb->excl.ptr;
    a.sysParamsA = nb->paramsA.ptr;
    a.sysParamsB = nb->paramsB.ptr;
    a.sysParamsD = nb->paramsD.ptr;
    a.fFloat = nb->fFloat.ptr;
    a.fFixed = nb->fFixed.ptr;
    a.entryPartials = nb->partials.ptr;
    a.computeFshift = computeVirial ? 1 : 0;
    if (combinedPrune)
    {
        const float ro = static_cast<float>(nb->listParams.rlistOuter);
        a.sciOrder = nullptr;
        a.pruneCjPacked = nb->cjPacked.ptr;
        a.pruneImaskOuter = nb->imaskOuter.ptr;
        a.pruneSciCount = nb->sciCount.ptr;
        a.pruneRlist_sq = ro * ro;
    }
    const int trc = timingBegin(nb, c_timeNonbonded + 2 * (freshList ? 1 : 0) + (computeEnergy ? 1 : 0));
    if (trc != NBNXM_OK)
    {
        return trc;
    }
    if (tunedWalk)
    {
        NBNXM_HIP_CHECK(launchNonbonded(a, nb->accumulation, nb->vdwType, computeEnergy != 0, nb->launchConfig, nb->stream));
    }
    else if (flavouredWalk)
    {
        // Other FP32 flavours: the tuned list walk with the pair interaction of the flavour's class.
        NBNXM_HIP_CHECK(launchNonbonded(a, nb->accumulation, nb->vdwType, computeEnergy != 0, nb->launchConfig, nb->stream, &nb->flavour));
    }
    else
    {
        NBNXM_HIP_CHECK(launchNonbondedFlavour(a, nb->flavour, nb->accumulation, nb->vdwType, computeEnergy != 0, nb->stream));
    }
    if (combinedPrune)
    {
        // The entry sort of the pruned list, inside the kernel's timed region as in the reference.
        const int src = sortEntriesOfFreshList(nb);
        if (src != NBNXM_OK)
        {
            return src;
        }
    }
    const int trc2 = timingEnd(nb);
    if (trc2 != NBNXM_OK)
    {
        return trc2;
    }
    if (nb->externalForces && nb->accumulation != NBNXM_ACCUM_FLOAT)
    {
        // gpu_get_f consumers read float forces: the virial partials from the
        // accumulator, then its float image (the accumulator is kept for
        // further launches until gpu_clear_outputs, as the reference's f).
        if (computeVirial)
        {
            NBNXM_HIP_CHECK(launchVirialPartials(nb->xq.ptr, nb->fFixed.ptr, nullptr, nb->slotRange.ptr, 1, nb->virialBlocks, nb->virialPartials.ptr, nb->stream));
        }
        NBNXM_HIP_CHECK(launchConvertForces(nb->fFixed.ptr, nb->numAtoms, nb->fOut.ptr, nb->fOutDouble.ptr, nb->stream, false));
    }
    nb->lastComputeEnergy = computeEnergy != 0;
    nb->lastComputeVirial = computeVirial != 0;
    nb->forcesConsumed = false;
    nb->haveLaunched = true;
    return NBNXM_OK;
}

int nbnxm_gpu_launch_cpyback(nbnxm_gpu_t* nb, float* f, double* fDouble)
{
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_launch_cpyback: null handle");
    }
    const size_t n3 = 3 * static_cast<size_t>(nb->numAtoms);
    const int trc = timingBegin(nb, c_timeD2h);
    if (trc != NBNXM_OK)
    {
        return trc;
    }
    const bool external = nb->externalForces && nb->accumulation != NBNXM_ACCUM_FLOAT;
    if (nb->forcesConsumed)
    {
        // add_nbat_f_to_f already reduced the forces and the virial partials;
        // only energies, shift forces and the virial remain.
        if (f || fDouble)
        {
            return fail(NBNXM_ERR_STATE, "gpu_launch_cpyback: the forces of this launch were taken by gpu_add_nbat_f_to_f");
        }
    }
    else if (external)
    {
        // The launch already wrote the float forces (and the virial partials).
        if (f && n3)
        {
            NBNXM_HIP_CHECK(hipMemcpyAsync(f, nb->fOut.ptr, sizeof(float) * n3, hipMemcpyDeviceToHost, nb->stream));
        }
        if (fDouble && n3)
        {
            NBNXM_HIP_CHECK(hipMemcpyAsync(fDouble, nb->fOutDouble.ptr, sizeof(double) * n3, hipMemcpyDeviceToHost, nb->stream));
        }
    }
    else
    {
        if (nb->lastComputeVirial)
        {
            // x (x) f partials before the fixed-point accumulator is cleared.
            const bool fixed = nb->accumulation != NBNXM_ACCUM_FLOAT;
            NBNXM_HIP_CHECK(launchVirialPartials(nb->xq.ptr, fixed ? nb->fFixed.ptr : nullptr, fixed ? nullptr : nb->fFloat.ptr, nb->slotRange.ptr, 1, nb->virialBlocks, nb->virialPartials.ptr, nb->stream));
        }
        if (nb->accumulation == NBNXM_ACCUM_FLOAT)
        {
            if (f && n3)
            {
                NBNXM_HIP_CHECK(hipMemcpyAsync(f, nb->fFloat.ptr, sizeof(float) * n3, hipMemcpyDeviceToHost, nb->stream));
            }
            if (fDouble)
            {
                return fail(NBNXM_ERR_UNSUPPORTED, "gpu_launch_cpyback: double forces need fixed-point accumulation");
            }
        }
        else
        {
            NBNXM_HIP_CHECK(launchConvertForces(nb->fFixed.ptr, nb->numAtoms, nb->fOut.ptr, nb->fOutDouble.ptr, nb->stream));
            if (f && n3)
            {
                NBNXM_HIP_CHECK(hipMemcpyAsync(f, nb->fOut.ptr, sizeof(float) * n3, hipMemcpyDeviceToHost, nb->stream));
            }
            if (fDouble && n3)
            {
                NBNXM_HIP_CHECK(hipMemcpyAsync(fDouble, nb->fOutDouble.ptr, sizeof(double) * n3, hipMemcpyDeviceToHost, nb->stream));
            }
        }
    }
    if (nb->lastComputeEnergy || nb->lastComputeVirial)
    {
        NBNXM_HIP_CHECK(launchReduceEnergies(nb->partials.ptr, nb->entryRanges.ptr, nb->entryList.ptr, nb->shiftRanges.ptr, nb->shiftEntryList.ptr, 1, nb->energies.ptr, nb->fshift.ptr, nb->lastComputeVirial ? 1 : 0, nb->stream));
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->hEnergies, nb->energies.ptr, 2 * sizeof(double), hipMemcpyDeviceToHost, nb->stream));
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->hFshift, nb->fshift.ptr, 3 * NBNXM_NUM_SHIFTS * sizeof(double), hipMemcpyDeviceToHost, nb->stream));
    }
    if (nb->lastComputeVirial)
    {
        NBNXM_HIP_CHECK(launchVirialFinish(nb->virialPartials.ptr, 1, nb->virialBlocks, nb->shiftVec.ptr, nb->fshift.ptr, nb->virial.ptr, nb->stream));
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->hVirial, nb->virial.ptr, 9 * sizeof(double), hipMemcpyDeviceToHost, nb->stream));
    }
    return timingEnd(nb);
}

int nbnxm_gpu_init_x_to_nbat_x(nbnxm_gpu_t* nb, int numAtoms, int numGridAtoms, const int* atomIndices)
{
    if (!nb || numAtoms < 0 || (numGridAtoms > 0 && !atomIndices))
    {
        return fail(NBNXM_ERR_INVALID, "gpu_init_x_to_nbat_x: null pointer or negative size");
    }
    if (numGridAtoms != nb->numAtoms)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_init_x_to_nbat_x: numGridAtoms differs from the uploaded atom data");
    }
    std::vector<int> cells(static_cast<size_t>(numAtoms), -1);
    for (int slot = 0; slot < numGridAtoms; slot++)
    {
        const int a = atomIndices[slot];
        if (a < -1 || a >= numAtoms || (a >= 0 && cells[a] >= 0))
        {
            return fail(NBNXM_ERR_INVALID, "gpu_init_x_to_nbat_x: atomIndices is not a map of slots onto distinct atoms");
        }
        if (a >= 0)
        {
            cells[a] = slot;
        }
    }
    for (int a = 0; a < numAtoms; a++)
    {
        if (cells[a] < 0)
        {
            return fail(NBNXM_ERR_INVALID, "gpu_init_x_to_nbat_x: an atom has no grid slot");
        }
    }
    NBNXM_HIP_CHECK(nb->atomIndex.resize(std::max(numGridAtoms, 1)));
    NBNXM_HIP_CHECK(nb->cellIndex.resize(std::max(numAtoms, 1)));
    if (numGridAtoms > 0)
    {
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->atomIndex.ptr, atomIndices, sizeof(int) * numGridAtoms, hipMemcpyHostToDevice, nb->stream));
    }
    if (numAtoms > 0)
    {
        NBNXM_HIP_CHECK(hipMemcpyAsync(nb->cellIndex.ptr, cells.data(), sizeof(int) * numAtoms, hipMemcpyHostToDevice, nb->stream));
    }
    NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    nb->numRealAtoms = numAtoms;
    return NBNXM_OK;
}

int nbnxm_gpu_x_to_nbat_x(nbnxm_gpu_t* nb, const float* d_x)
{
    if (!nb || (!d_x && nb->numRealAtoms > 0))
    {
        return fail(NBNXM_ERR_INVALID, "gpu_x_to_nbat_x: null pointer");
    }
    if (nb->numRealAtoms < 0)
    {
        return fail(NBNXM_ERR_STATE, "gpu_x_to_nbat_x: call gpu_init_x_to_nbat_x first");
    }
    NBNXM_HIP_CHECK(launchXToXq(d_x, nb->atomIndex.ptr, nb->numAtoms, nb->xq.ptr, nb->stream));
    return NBNXM_OK;
}

int nbnxm_gpu_add_nbat_f_to_f(nbnxm_gpu_t* nb, float* d_f, const float* d_fToAdd, int accumulate)
{
    if (!nb || (!d_f && nb->numRealAtoms > 0))
    {
        return fail(NBNXM_ERR_INVALID, "gpu_add_nbat_f_to_f: null pointer");
    }
    if (nb->numRealAtoms < 0)
    {
        return fail(NBNXM_ERR_STATE, "gpu_add_nbat_f_to_f: call gpu_init_x_to_nbat_x first");
    }
    if (!nb->haveLaunched || nb->forcesConsumed)
    {
        return fail(NBNXM_ERR_STATE, "gpu_add_nbat_f_to_f: no nonbonded launch whose forces are still pending");
    }
    const bool fixed = nb->accumulation != NBNXM_ACCUM_FLOAT;
    if (fixed && nb->externalForces)
    {
        // The launch wrote float forces and the virial partials; the
        // accumulator stays until gpu_clear_outputs.
        NBNXM_HIP_CHECK(launchReduceForces(nullptr, nb->fOut.ptr, nb->cellIndex.ptr, nb->numRealAtoms, d_fToAdd, d_f, accumulate != 0, nb->stream));
        nb->forcesConsumed = true;
        return NBNXM_OK;
    }
    if (nb->lastComputeVirial)
    {
        NBNXM_HIP_CHECK(launchVirialPartials(nb->xq.ptr, fixed ? nb->fFixed.ptr : nullptr, fixed ? nullptr : nb->fFloat.ptr, nb->slotRange.ptr, 1, nb->virialBlocks, nb->virialPartials.ptr, nb->stream));
    }
    NBNXM_HIP_CHECK(launchReduceForces(fixed ? nb->fFixed.ptr : nullptr, nb->fFloat.ptr, nb->cellIndex.ptr, nb->numRealAtoms, d_fToAdd, d_f, accumulate != 0, nb->stream));
    nb->forcesConsumed = true;
    return NBNXM_OK;
}

int nbnxm_gpu_get_virial(nbnxm_gpu_t* nb, double* virial)
{
    if (!nb || !virial)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_get_virial: null pointer");
    }
    if (!nb->lastComputeVirial)
    {
        return fail(NBNXM_ERR_STATE, "gpu_get_virial: the last launch did not compute the virial");
    }
    NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    for (int k = 0; k < 9; k++)
    {
        virial[k] += nb->hVirial[k];
    }
    return NBNXM_OK;
}

} // extern "C"

namespace
{

// The part of gpu_try_finish_task after the stream has completed
// (gpu_common.h:296-325): timings, then the staged outputs.
int finishTask(nbnxm_gpu* nb, double* e_lj, double* e_el, double* fshift)
{
    const int rc = accountTimings(nb);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    // Accumulate into the caller's outputs like gpu_reduce_staged_outputs
    // (src/gromacs/nbnxm/gpu_common.h:138-160).
    if (nb->lastComputeEnergy)
    {
        if (e_lj)
        {
            *e_lj += nb->hEnergies[0];
        }
        if (e_el)
        {
            *e_el += nb->hEnergies[1];
        }
    }
    if (nb->lastComputeVirial && fshift)
    {
        for (int k = 0; k < 3 * NBNXM_NUM_SHIFTS; k++)
        {
            fshift[k] += nb->hFshift[k];
        }
    }
    return NBNXM_OK;
}

} // namespace

extern "C" {

int nbnxm_gpu_wait_finish_task(nbnxm_gpu_t* nb, double* e_lj, double* e_el, double* fshift)
{
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_wait_finish_task: null handle");
    }
    NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    return finishTask(nb, e_lj, e_el, fshift);
}

int nbnxm_gpu_try_finish_task(nbnxm_gpu_t* nb, double* e_lj, double* e_el, double* fshift)
{
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_try_finish_task: null handle");
    }
    const hipError_t q = hipStreamQuery(nb->stream);
    if (q == hipErrorNotReady)
    {
        // Not an error: take it off the thread's last-error state so the next
        // launch check does not report it.
        (void)hipGetLastError();
        return 0;
    }
    NBNXM_HIP_CHECK(q);
    const int rc = finishTask(nb, e_lj, e_el, fshift);
    return rc != NBNXM_OK ? rc : 1;
}

int nbnxm_gpu_pme_loadbal_update_param(nbnxm_gpu_t* nb, const nbnxm_interaction_const_t* ic, const nbnxm_pairlist_params_t* listParams)
{
    if (!nb || !ic)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_pme_loadbal_update_param: null pointer");
    }
    int rc = validateFlavour(ic);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    if (ic->ljPme != nb->ic.ljPme)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "gpu_pme_loadbal_update_param: the LJ-PME grid rule is fixed at gpu_init");
    }
    const nbnxm_pairlist_params_t lp = listParams ? *listParams : nb->listParams;
    if (!(lp.rlistOuter >= ic->rcoulomb) || !(lp.rlistInner >= ic->rcoulomb) || lp.rlistInner > lp.rlistOuter
        || lp.numRollingPruningParts < 1)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_pme_loadbal_update_param: need rc <= rlistInner <= rlistOuter and numRollingPruningParts >= 1");
    }
    // set_cutoff_parameters + nbnxn_gpu_pick_ewald_kernel_type +
    // init_ewald_coulomb_force_table (nbnxm_gpu_data_mgmt.cpp:178-200,511-527).
    rc = setFlavour(nb, *ic);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    nb->listParams = lp;
    return NBNXM_OK;
}

int nbnxm_gpu_set_timing(nbnxm_gpu_t* nb, int enable)
{
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_set_timing: null handle");
    }
    nb->doTime = enable != 0;
    return NBNXM_OK;
}

int nbnxm_gpu_get_timings(const nbnxm_gpu_t* nb, nbnxm_gpu_timings_t* timings)
{
    if (!nb || !timings)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_get_timings: null pointer");
    }
    if (!nb->doTime)
    {
        return 0;
    }
    *timings = nb->timings;
    return 1;
}

int nbnxm_gpu_reset_timings(nbnxm_gpu_t* nb)
{
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_reset_timings: null handle");
    }
    // init_timings (nbnxm_gpu_data_mgmt.cpp:253-272); regions still in flight
    // are dropped with the old counts.
    NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    for (const nbnxm_gpu::TimingRecord& r : nb->pendingTimings)
    {
        nb->eventPool.push_back(r.start);
        nb->eventPool.push_back(r.stop);
    }
    nb->pendingTimings.clear();
    nb->timings = nbnxm_gpu_timings_t{};
    return NBNXM_OK;
}

int nbnxm_gpu_get_f(nbnxm_gpu_t* nb, float** d_f)
{
    if (!nb || !d_f)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_get_f: null pointer");
    }
    if (nb->accumulation == NBNXM_ACCUM_FLOAT)
    {
        *d_f = nb->fFloat.ptr;
        return NBNXM_OK;
    }
    if (!nb->externalForces)
    {
        // Forces of a launch made before the switch are still only in the
        // accumulator: take its virial partials and give it its float image now.
        if (nb->haveLaunched && !nb->forcesConsumed && nb->lastComputeVirial)
        {
            NBNXM_HIP_CHECK(launchVirialPartials(nb->xq.ptr, nb->fFixed.ptr, nullptr, nb->slotRange.ptr, 1, nb->virialBlocks, nb->virialPartials.ptr, nb->stream));
        }
        NBNXM_HIP_CHECK(launchConvertForces(nb->fFixed.ptr, nb->numAtoms, nb->fOut.ptr, nb->fOutDouble.ptr, nb->stream, false));
        nb->externalForces = true;
    }
    *d_f = nb->fOut.ptr;
    return NBNXM_OK;
}

int nbnxm_gpu_get_stream(nbnxm_gpu_t* nb, void** stream)
{
    if (!nb || !stream)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_get_stream: null pointer");
    }
    *stream = nb->stream;
    return NBNXM_OK;
}

int nbnxm_gpu_min_ci_balanced(nbnxm_gpu_t* nb)
{
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_min_ci_balanced: null handle");
    }
    int device = 0, cus = 0;
    NBNXM_HIP_CHECK(hipGetDevice(&device));
    NBNXM_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    // 50 entries per compute unit, the factor of the reference's OpenCL and
    // SYCL back ends (nbnxm_ocl_data_mgmt.cpp:86, nbnxm_sycl_data_mgmt.cpp:68-80;
    // CUDA uses 44-61 per SM): on a 96,000-atom water list 12,000-22,000
    // entries ran the force launch in about half the time of the 2,500
    // unsplit or 4,200 entries, and a list that already has more entries
    // than this (3 M atoms: 56,000) is not split.
    return 50 * cus;
}

int nbnxm_gpu_is_kernel_ewald_analytical(const nbnxm_gpu_t* nb)
{
    // EwaldAna or EwaldAnaTwin (nbnxm_gpu_data_mgmt.cpp:823-827).
    return nb ? ((nb->ic.eeltype == NBNXM_ELEC_EWALD && !nb->ic.tabulatedEwald) ? 1 : 0) : NBNXM_ERR_INVALID;
}

int nbnxm_gpu_get_masks(nbnxm_gpu_t* nb, unsigned int* imaskInner, unsigned int* imaskOuter)
{
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_get_masks: null handle");
    }
    NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    std::vector<nbnxm_cj_packed_t> tmp(nb->ncjPacked);
    if (nb->ncjPacked > 0)
    {
        NBNXM_HIP_CHECK(hipMemcpy(tmp.data(), nb->cjPacked.ptr, sizeof(nbnxm_cj_packed_t) * nb->ncjPacked, hipMemcpyDeviceToHost));
        if (imaskOuter)
        {
            NBNXM_HIP_CHECK(hipMemcpy(imaskOuter, nb->imaskOuter.ptr, sizeof(unsigned int) * nb->ncjPacked, hipMemcpyDeviceToHost));
        }
    }
    if (imaskInner)
    {
        for (int g = 0; g < nb->ncjPacked; g++)
        {
            imaskInner[g] = tmp[g].imei[0].imask;
        }
    }
    return NBNXM_OK;
}

int nbnxm_gpu_get_sci_order(nbnxm_gpu_t* nb, int* order, int* sciCount)
{
    if (!nb)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_get_sci_order: null handle");
    }
    NBNXM_HIP_CHECK(hipStreamSynchronize(nb->stream));
    if (nb->nsci > 0)
    {
        if (order)
        {
            NBNXM_HIP_CHECK(hipMemcpy(order, nb->sciOrder.ptr, sizeof(int) * nb->nsci, hipMemcpyDeviceToHost));
        }
        if (sciCount)
        {
            NBNXM_HIP_CHECK(hipMemcpy(sciCount, nb->sciCount.ptr, sizeof(int) * nb->nsci, hipMemcpyDeviceToHost));
        }
    }
    return NBNXM_OK;
}

int nbnxm_gpu_set_waves_per_block(nbnxm_gpu_t* nb, int wavesPerBlock)
{
    if (!nb || wavesPerBlock < 1 || wavesPerBlock > 4)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_set_waves_per_block: need 1..4");
    }
    nb->launchConfig.wavesPerBlock = wavesPerBlock;
    return NBNXM_OK;
}

int nbnxm_gpu_set_i_cluster_in_lds(nbnxm_gpu_t* nb, int mode)
{
    if (!nb || mode < -1 || mode > 1)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_set_i_cluster_in_lds: mode -1, 0 or 1");
    }
    nb->launchConfig.iClusterInLds = mode;
    return NBNXM_OK;
}

void nbnxm_gpu_free(nbnxm_gpu_t* nb)
{
    delete nb;
}

} // extern "C"

int nbnxm_amd::gpuTaskAdoptDeviceList(nbnxm_gpu_t* nb,
                                      const DeviceListView& v,
                                      const float* charges,
                                      const int* types,
                                      const float2* typeComb,
                                      int dummyType,
                                      const double box[3])
{
    const bool needTypes = nb && (nb->vdwType == NBNXM_VDW_CUT_TYPE_TABLE || nb->ic.ljPme);
    const bool needComb = nb && nb->vdwType != NBNXM_VDW_CUT_TYPE_TABLE;
    if (!nb || !box || v.numAtoms < 1 || v.numSlots < v.numAtoms || !v.atomIndex || !v.xw || !charges || !types
        || (needComb && !typeComb) || v.nsci < 0 || v.ncjPacked < 0 || v.nexcl < 1 || (v.nsci > 0 && !v.sci)
        || (v.ncjPacked > 0 && !v.cjPacked) || !v.excl)
    {
        return fail(NBNXM_ERR_INVALID, "gpu adopt device list: null pointer or bad size");
    }
    // The list and the wrapped coordinates come from another stream.
    if (!nb->adoptEvent)
    {
        NBNXM_HIP_CHECK(hipEventCreateWithFlags(&nb->adoptEvent, hipEventDisableTiming));
    }
    NBNXM_HIP_CHECK(hipEventRecord(nb->adoptEvent, v.stream));
    NBNXM_HIP_CHECK(hipStreamWaitEvent(nb->stream, nb->adoptEvent, 0));

    // Grid-order atom data (gpu_init_atomdata with fillXq/fillTypes/fillLjComb
    // done by a gather kernel) and the buffer-op maps (gpu_init_x_to_nbat_x).
    const int n = v.numSlots;
    nb->numAtoms = n;
    NBNXM_HIP_CHECK(nb->xq.resize(n));
    NBNXM_HIP_CHECK(nb->fFixed.resize(3 * static_cast<size_t>(n)));
    NBNXM_HIP_CHECK(nb->fFloat.resize(3 * static_cast<size_t>(n)));
    NBNXM_HIP_CHECK(nb->fOut.resize(3 * static_cast<size_t>(n)));
    NBNXM_HIP_CHECK(nb->fOutDouble.resize(3 * static_cast<size_t>(n)));
    if (needTypes)
    {
        NBNXM_HIP_CHECK(nb->types.resize(n));
    }
    if (needComb)
    {
        NBNXM_HIP_CHECK(nb->ljComb.resize(n));
    }
    NBNXM_HIP_CHECK(nb->atomIndex.resize(n));
    NBNXM_HIP_CHECK(nb->cellIndex.resize(v.numAtoms));
    NBNXM_HIP_CHECK(hipMemcpyAsync(nb->atomIndex.ptr, v.atomIndex, sizeof(int) * n, hipMemcpyDeviceToDevice, nb->stream));
    NBNXM_HIP_CHECK(launchGatherAtomData(v.atomIndex, n, v.xw, charges, types, typeComb, dummyType, nb->xq.ptr, needTypes ? nb->types.ptr : nullptr,
                                         needComb ? nb->ljComb.ptr : nullptr, nb->cellIndex.ptr, nb->stream));
    nb->numRealAtoms = v.numAtoms;
    NBNXM_HIP_CHECK(launchFillU64(nb->fFixed.ptr, 3 * static_cast<size_t>(n), nb->stream));
    NBNXM_HIP_CHECK(launchFillI32(reinterpret_cast<int*>(nb->fFloat.ptr), 3 * static_cast<size_t>(n), 0, nb->stream));
    nb->virialBlocks = std::min(512, std::max(1, (n + 1023) / 1024));
    NBNXM_HIP_CHECK(nb->slotRange.resize(2));
    NBNXM_HIP_CHECK(nb->virialPartials.resize(9 * static_cast<size_t>(nb->virialBlocks)));
    NBNXM_HIP_CHECK(nb->virial.resize(9));
    NBNXM_HIP_CHECK(launchStoreInt2(nb->slotRange.ptr, 0, n, nb->stream));

    // Shift vectors of the box (uploaded when the box is not the one they
    // were last computed for), then the list.
    if (!nb->shiftVecOfBox || box[0] != nb->shiftVecBox[0] || box[1] != nb->shiftVecBox[1] || box[2] != nb->shiftVecBox[2])
    {
        float shiftVec[3 * NBNXM_NUM_SHIFTS];
        int rc = nbnxm_calc_shifts(box, shiftVec);
        if (rc == NBNXM_OK)
        {
            rc = nbnxm_gpu_upload_shiftvec(nb, shiftVec);
        }
        if (rc != NBNXM_OK)
        {
            return rc;
        }
        nb->shiftVecOfBox = true;
        std::copy(box, box + 3, nb->shiftVecBox);
    }
    return setPairlist(nb, v.hostSci, v.sci, v.nsci, v.cjPacked, v.ncjPacked, v.excl, v.nexcl, hipMemcpyDeviceToDevice);
}
