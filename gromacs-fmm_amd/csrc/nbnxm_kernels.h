// Internal launch interface of the gfx950 nbnxm kernels (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "nbnxm_amd.h"

namespace nbnxm_amd
{

// Everything one nonbonded launch reads and writes.  Device pointers only.
// Systems are concatenated: cluster indices are global, sci.shift is
// 45*system + shift index and shiftVec holds 45 float4 per system.
struct NbKernelArgs
{
    const float4* xq = nullptr;
    const int* atomTypes = nullptr;
    const float2* ljComb = nullptr;
    const float2* nbfp = nullptr;
    int numTypes = 0;
    const float4* shiftVec = nullptr;

    const nbnxm_sci_t* sci = nullptr;
    const int* sciOrder = nullptr; // optional permutation of entries (work order)
    int nsci = 0;
    const nbnxm_cj_packed_t* cjPacked = nullptr;
    const nbnxm_excl_t* excl = nullptr;

    // Per-system parameters (indexed by sci.shift / 45).
    const float4* sysParamsA = nullptr; // rcoulomb_sq, epsfac, beta, sh_ewald
    const float2* sysParamsB = nullptr; // repulsion cpot, dispersion cpot
    const double* sysParamsD = nullptr; // FP64 mode: 8 per system, rc^2, epsfac, beta, sh_ewald, cpot rep, cpot disp

    // Outputs.
    float* fFloat = nullptr;                  // NBNXM_ACCUM_FLOAT: 3 per atom
    unsigned long long* fFixed = nullptr;     // NBNXM_ACCUM_FIXED64 / FP64: 3 per atom
    double* entryPartials = nullptr;          // 8 doubles per entry: Elj, Eel, fsx, fsy, fsz
    int computeFshift = 0;

    // Combined prune of a fresh list (set only for that launch; entries in
    // list order): masks written back in place and to the outer masks, the
    // bin of each entry (its sort key) for the entry sort.
    nbnxm_cj_packed_t* pruneCjPacked = nullptr;
    unsigned int* pruneImaskOuter = nullptr;
    int* pruneSciCount = nullptr;
    float pruneRlist_sq = 0.0f;
};

enum class KernelFlavor : int
{
    F = 0,
    VF = 1
};

// Launch tuning chosen per size bucket by the autotuner.
struct NbLaunchConfig
{
    int wavesPerBlock = 4;
    // Compile-time flavour classes: i-cluster data in LDS at 4 waves per
    // SIMD (1; instantiations that would spill keep registers), in registers
    // (0), or chosen by the entry count (-1).
    int iClusterInLds = -1;
};

// Entries from which the automatic choice puts the i-cluster data in LDS
// (two rounds of 4,096 resident waves on 256 CUs): with several entries per
// wave slot the extra wave per SIMD pays for the extra LDS read per cluster
// pair; with one or two entries per slot the shorter register path wins.
// Measured on water boxes: registers 3 % faster at 7,200 entries, LDS 3-6 %
// faster from 8,400.
constexpr int c_iClusterLdsMinEntries = 8192;

inline bool iClusterDataInLds(const NbLaunchConfig& cfg, int nsci)
{
    return cfg.iClusterInLds >= 0 ? cfg.iClusterInLds != 0 : nsci >= c_iClusterLdsMinEntries;
}

// Flavour classes of the tuned list walk: the tuned flavour, the run-time
// flavour (any combination, read from NbFlavourParams), and combinations
// compiled in their own instantiation because they are common: reaction field
// with the LJ potential shift (plain cut-off included), analytical Ewald with
// the LJ force switch, tabulated Ewald with the LJ potential shift, analytical
// Ewald with the LJ potential switch (these four with rvdw = rcoulomb and no
// LJ-PME), analytical Ewald + LJ potential shift with a shorter LJ cut-off
// (twin cut-off), and analytical Ewald with LJ-PME (rvdw = rcoulomb).
enum FlavourClass : int
{
    c_flavTuned = 0,
    c_flavRuntime = 1,
    c_flavRfPotShift = 2,
    c_flavEwaldForceSwitch = 3,
    c_flavEwaldTabPotShift = 4,
    c_flavEwaldPotSwitch = 5,
    c_flavEwaldTwin = 6,
    c_flavEwaldLjPme = 7
};

// Flavour of a non-tuned launch: the interaction constants of one launch
// (nbnxm_interaction_const_t, wave-uniform), plus the LJ-PME grid C6 of every
// type pair (numTypes^2, row-major by the i type) on the device in FP32 and
// FP64, formed on the host by the grid rule exactly as the reference kernels
// form it per pair (ljPmeGridTables).
struct NbFlavourParams
{
    int eeltype = NBNXM_ELEC_EWALD;
    int vdwModifier = NBNXM_VDW_MOD_POTSHIFT;
    int ljPme = 0; // NBNXM_LJPME_*
    int vdwCutoffCheck = 0; // rvdw < rcoulomb
    double rcoulomb = 0, rvdw = 0, epsfac = 0, beta = 0, sh_ewald = 0, k_rf = 0, c_rf = 0;
    double cpot_rep = 0, cpot_disp = 0, rvdw_switch = 0;
    double disp_c2 = 0, disp_c3 = 0, rep_c2 = 0, rep_c3 = 0, sw_c3 = 0, sw_c4 = 0, sw_c5 = 0;
    double ewaldcoeff_lj = 0, sh_lj_ewald = 0;
    const float* ljPmeGrid = nullptr;
    const double* ljPmeGridD = nullptr;
    // Tabulated Ewald correction: F(r) table with one padding entry.
    int tabulatedEwald = 0;
    double ewaldTableScale = 0;
    const float* ewaldTable = nullptr;
    const double* ewaldTableD = nullptr;
    int ewaldTableSize = 0; // entries, padding included
    int flavourClass = c_flavRuntime;
};

// True for the flavour of the tuned kernels: analytical Ewald + LJ potential
// shift, no LJ-PME, rvdw == rcoulomb.
inline bool isDefaultFlavour(const nbnxm_interaction_const_t& ic)
{
    return ic.eeltype == NBNXM_ELEC_EWALD && !ic.tabulatedEwald && ic.vdwModifier == NBNXM_VDW_MOD_POTSHIFT && !ic.ljPme
           && ic.rvdw == ic.rcoulomb;
}

NbFlavourParams flavourParams(const nbnxm_interaction_const_t& ic);

// The tuned list walk for a flavour other than the tuned one (FP32
// accumulation modes; the class comes from flav.flavourClass).
hipError_t launchNonbondedFlavoured(const NbKernelArgs& a,
                                    const NbFlavourParams& flav,
                                    int accumulation,
                                    int vdwType,
                                    bool computeEnergy,
                                    const NbLaunchConfig& cfg,
                                    hipStream_t stream);

// Classes 5-7 of the tuned list walk (nbnxm_kernels_flavoured_b.hip).
hipError_t launchNonbondedFlavouredB(const NbKernelArgs& a,
                                     const NbFlavourParams& flav,
                                     int accumulation,
                                     int vdwType,
                                     bool computeEnergy,
                                     const NbLaunchConfig& cfg,
                                     hipStream_t stream);

// The run-time flavour class of the tuned list walk (nbnxm_kernels_runtime.hip).
hipError_t launchNonbondedRuntimeFlavour(const NbKernelArgs& a,
                                         const NbFlavourParams& flav,
                                         int accumulation,
                                         int vdwType,
                                         bool computeEnergy,
                                         const NbLaunchConfig& cfg,
                                         hipStream_t stream);

// Every flavour (one wave per entry, FP32 or, with NBNXM_ACCUM_FP64, FP64).
hipError_t launchNonbondedFlavour(const NbKernelArgs& a, const NbFlavourParams& flav, int accumulation, int vdwType, bool computeEnergy, hipStream_t stream);

// Types (with the dummy type) of the largest LJ table the tuned kernel stages in LDS.
constexpr int c_maxLdsTableTypes = 65;

// LJ-PME grid C6 per type pair (numTypes^2, row-major) from the nbfp diagonal
// (6 C6, 12 C12): the per-type combination parameters of nbfp_comb
// (src/gromacs/nbnxm/atomdata.cpp:296-330) combined as the reference kernels
// combine them per pair, geometric sqrt(6C6_ii) sqrt(6C6_jj)
// (calculate_lj_ewald_c6grid) or Lorentz-Berthelot eps (s_i + s_j)^6
// (calculate_lj_ewald_comb_LB_F_E, cuda/nbnxm_cuda_kernel_utils.cuh:290-337),
// in FP32 in the reference's operation order; the FP64 table does the same in
// double from the FP32 nbfp values.
void ljPmeGridTables(int rule, int numTypes, const float* nbfp, float* grid, double* gridD);


// flav == nullptr: the tuned flavour (per-system parameters from sysParams*);
// otherwise the FP32 instantiation of flav's flavour class (single parameter
// set, FLOAT or FIXED64 accumulation, LJ-PME with the type table).
hipError_t launchNonbonded(const NbKernelArgs& a,
                           int accumulation,
                           int vdwType,
                           bool computeEnergy,
                           const NbLaunchConfig& cfg,
                           hipStream_t stream,
                           const NbFlavourParams* flav = nullptr);

struct PruneArgs
{
    const float4* xq = nullptr;
    const float4* shiftVec = nullptr;
    const nbnxm_sci_t* sci = nullptr;
    const int* sciOrder = nullptr;
    int nsci = 0;
    nbnxm_cj_packed_t* cjPacked = nullptr; // imei[0].imask updated (inner mask)
    unsigned int* imaskOuter = nullptr;    // per packed group
    int* rollingPart = nullptr;            // per block part index (device), as the reference
    int* sciCount = nullptr;               // bin per entry, the sort key (fresh list)
    float rlistOuter_sq = 0;
    float rlistInner_sq = 0;
    int numParts = 1;
    int numGroups = 0; // packed groups of the list (sets the waves per entry)
};

hipError_t launchPrune(const PruneArgs& a, bool haveFreshList, hipStream_t stream);

// Converts the fixed-point forces to float (optionally double) and, with
// clear, clears the accumulator for the next evaluation.
hipError_t launchConvertForces(unsigned long long* fFixed, int numAtoms, float* fOut, double* fOutDouble, hipStream_t stream, bool clear = true);
// Converts the fixed-point forces of slots [first, first+count).
hipError_t launchConvertForcesRange(unsigned long long* fFixed, int first, int count, float* fOut, double* fOutDouble, hipStream_t stream);

// Per-system FP64 reduction of entry partials.  entryRanges: numSystems+1
// offsets into entryList (entries of each system), shiftRanges: 45*numSystems+1
// offsets into shiftEntryList.  out: per system Elj, Eel; fshift 45*3 per system.
hipError_t launchReduceEnergies(const double* entryPartials,
                                const int* entryRanges,
                                const int* entryList,
                                const int* shiftRanges,
                                const int* shiftEntryList,
                                int numSystems,
                                double* energiesOut,
                                double* fshiftOut,
                                int computeFshift,
                                hipStream_t stream);

// Single-sum virial in FP64.  Partials: x (x) f summed over the slots
// [slotRanges[s], slotRanges[s+1]) of each system s by blocksPerSystem blocks
// (fixed-point forces if fFixed, else fFloat; run before the fixed-point
// accumulator is cleared).  Finish: virial[9*s] = -1/2 (sum x (x) f +
// sum_k shiftVec[45*s+k] (x) fshift[45*s+k]).
hipError_t launchVirialPartials(const float4* xq,
                                const unsigned long long* fFixed,
                                const float* fFloat,
                                const int* slotRanges,
                                int numSystems,
                                int blocksPerSystem,
                                double* partials,
                                hipStream_t stream);
hipError_t launchVirialFinish(const double* partials, int numSystems, int blocksPerSystem, const float4* shiftVec, const double* fshift, double* virial, hipStream_t stream);

// Buffer ops around the kernel.  xq.xyz of real slots from coordinates in atom
// order (3 floats per atom; atomIndices: slot -> atom or -1), w kept.
hipError_t launchXToXq(const float* x, const int* atomIndices, int numSlots, float4* xq, hipStream_t stream);
// Grid-order xq (fillers far away, q = 0), types (dummy for fillers) and
// combination parameters of a device-built list, and cells[atom] = slot.
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
                                hipStream_t stream);
// x[3i..3i+2] = xw[i].xyz.
hipError_t launchWriteCoordinates(const float4* xw, int numAtoms, float* x, hipStream_t stream);
// f[3a..] (+)= nonbonded force of slot cells[a] (+ fToAdd[3a..] if non-null);
// fixed-point slots (fFixed non-null) are converted and cleared, else fFloat is read.
hipError_t launchReduceForces(unsigned long long* fFixed, const float* fFloat, const int* cells, int numAtoms, const float* fToAdd, float* f, bool accumulate, hipStream_t stream);

// xq = x0 + N(0, sigma) for real atoms (q and fillers kept) over slots
// [first, first+count); the step number is read from *step on the device so a
// captured graph can be replayed; Philox4x32-10 counter-based generator keyed
// by (seed), counter (slot, step).
hipError_t launchPerturb(const float4* x0, float4* xq, int first, int count, float sigma, uint64_t seed, const unsigned long long* step, hipStream_t stream);
// *step += 1 (one thread), ends an evaluation in a captured graph.
hipError_t launchAdvanceStep(unsigned long long* step, hipStream_t stream);

// Fill kernels (graph-safe clears).
hipError_t launchFillU64(unsigned long long* p, size_t n, hipStream_t stream);
hipError_t launchFillI32(int* p, size_t n, int value, hipStream_t stream);
// p[0] = a, p[1] = b (one thread).
hipError_t launchStoreInt2(int* p, int a, int b, hipStream_t stream);
// Work-order arrays of a list whose n entries are in device memory:
// entryList, sciOrder and iota = 0..n-1, shiftKeys[i] = sci[i].shift and
// entryRanges = {0, n}.
hipError_t launchEntryArrays(const nbnxm_sci_t* sci, int n, int* entryList, int* sciOrder, int* iota, int* shiftKeys, int* entryRanges, hipStream_t stream);
// shiftRanges[k] (k = 0..NBNXM_NUM_SHIFTS) = number of the n sorted shift
// keys below k: the ranges of the per-shift entry lists.
hipError_t launchShiftRanges(const int* sortedKeys, int n, int* shiftRanges, hipStream_t stream);
// Convert and clear the fixed-point forces (fFixed may be null: no forces to
// convert), reduce energies and shift forces per system and advance the step
// counter (step may be null), in one launch.
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
                            hipStream_t stream);

} // namespace nbnxm_amd
