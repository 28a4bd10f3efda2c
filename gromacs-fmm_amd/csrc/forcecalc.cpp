// nblib-style force calculator over the GPU kernel: the interface of
// nblib::GmxNBForceCalculatorCpu (api/nblib/include/nblib/gmxcalculatorcpu.h:67-105,
// api/nblib/gmxcalculatorcpu.cpp:108-251) with the work done by the search
// and the GPU task of this library.  Also the combination-rule detection of
// nbnxn_atomdata_t::Params (src/gromacs/nbnxm/atomdata.cpp:421-540).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "capi_common.h"
#include "pairsearch.h"

using namespace nbnxm_amd;

namespace
{

// |a - b| <= tol * (|a| + |b|) / 2, the relative comparison of gmx_within_tol.
bool withinTol(double a, double b, double tol)
{
    return std::fabs(a - b) <= tol * 0.5 * (std::fabs(a) + std::fabs(b));
}

// Per-type combination parameters from the nbfp diagonal, as the kernels
// take them: geometric (sqrt(6C6), sqrt(12C12)); Lorentz-Berthelot
// (sigma/2, 2 sqrt(eps)) = (0.5 (12C12/6C6)^(1/6), sqrt((6C6)^2 / 12C12)).
void combParams(int numTypes, const double* nbfp, int rule, std::vector<float>& out)
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

extern "C" int nbnxm_detect_comb_rule(int numTypes, const double* nbfp, double tol)
{
    if (numTypes < 1 || !nbfp || !(tol >= 0))
    {
        return fail(NBNXM_ERR_INVALID, "detect_comb_rule: need numTypes >= 1, nbfp and tol >= 0");
    }
    const int nt = numTypes;
    auto c6At = [&](int i, int j) { return nbfp[2 * (static_cast<size_t>(i) * nt + j)]; };
    auto c12At = [&](int i, int j) { return nbfp[2 * (static_cast<size_t>(i) * nt + j) + 1]; };
    bool geom = true;
    bool lb = true;
    // Per-type sigma and epsilon for the LB test (C6 = 4 eps sigma^6, C12 = 4 eps sigma^12).
    std::vector<double> sigma(nt, 0.0), eps(nt, 0.0);
    for (int i = 0; i < nt; i++)
    {
        const double c6 = c6At(i, i) / 6.0;
        const double c12 = c12At(i, i) / 12.0;
        if (c6 > 0 && c12 > 0)
        {
            sigma[i] = std::pow(c12 / c6, 1.0 / 6.0);
            eps[i] = 0.25 * c6 * c6 / c12;
        }
        else if (!(c6 == 0 && c12 == 0))
        {
            lb = false; // only dispersion or only repulsion: no LB parameters
        }
    }
    for (int i = 0; i < nt; i++)
    {
        for (int j = 0; j < nt; j++)
        {
            const double c6 = c6At(i, j);
            const double c12 = c12At(i, j);
            geom = geom && withinTol(c6 * c6, c6At(i, i) * c6At(j, j), tol) && withinTol(c12 * c12, c12At(i, i) * c12At(j, j), tol);
            const double C6 = c6 / 6.0;
            const double C12 = c12 / 12.0;
            const bool zeroPair = (C6 == 0 && C12 == 0 && (eps[i] == 0 || eps[j] == 0));
            const bool lbPair = (C6 > 0 && C12 > 0 && withinTol(std::pow(C12 / C6, 1.0 / 6.0), 0.5 * (sigma[i] + sigma[j]), tol)
                                 && withinTol(0.25 * C6 * C6 / C12, std::sqrt(eps[i] * eps[j]), tol));
            lb = lb && (zeroPair || lbPair);
        }
    }
    if (geom)
    {
        return NBNXM_VDW_CUT_COMB_GEOM;
    }
    return lb ? NBNXM_VDW_CUT_COMB_LB : NBNXM_VDW_CUT_TYPE_TABLE;
}

struct nbnxm_force_calculator
{
    int numParticles = 0;
    int numTypes = 0;
    std::vector<int> types;
    std::vector<double> nbfp;    // numTypes^2 pairs, (6C6, 12C12)
    std::vector<double> charges;
    std::vector<int> exclRanges; // empty = no exclusions
    std::vector<int> exclElements;
    nbnxm_kernel_options_t options{};
    nbnxm_interaction_const_t ic{};
    double rlist = 0;
    int vdwType = NBNXM_VDW_CUT_TYPE_TABLE;

    GridSearch search;
    nbnxm_gpu_t* gpu = nullptr;
    bool havePairlist = false;
    // Entry split target of the previous list (0: none; -1: no list yet),
    // reused so that an update builds its list once.
    int splitTarget = -1;
    double box[3] = { 0, 0, 0 };

    // Grid-order host staging.
    std::vector<float> xq;
    std::vector<int> slotTypes;
    std::vector<float> slotComb;
    std::vector<float> nbfpTable; // (numTypes+1)^2 pairs with the dummy type
    std::vector<double> fGrid;

    ~nbnxm_force_calculator()
    {
        if (gpu)
        {
            nbnxm_gpu_free(gpu);
        }
    }
};

extern "C" int nbnxm_force_calculator_create(int numParticles,
                                             const int* particleTypeIdOfAllParticles,
                                             int numTypes,
                                             const double* nonBondedParams,
                                             const double* charges,
                                             const int* exclusionRanges,
                                             const int* exclusionElements,
                                             const nbnxm_kernel_options_t* options,
                                             nbnxm_force_calculator_t** out)
{
    if (!out || !options || numParticles < 0 || numTypes < 1 || !nonBondedParams || (numParticles > 0 && (!particleTypeIdOfAllParticles || !charges)))
    {
        return fail(NBNXM_ERR_INVALID, "force_calculator_create: null pointer or bad size");
    }
    *out = nullptr;
    // The reference throws for these (gmxcalculatorcpu.cpp:278-281 for useGpu,
    // nbnxmsetuphelpers.cpp:226-233 for the Coulomb type).
    if (!options->useGpu)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "force_calculator_create: this calculator runs on the GPU only (useGpu must be set)");
    }
    if (options->coulombType < NBNXM_COULOMB_PME || options->coulombType > NBNXM_COULOMB_REACTION_FIELD)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, "force_calculator_create: unknown Coulomb type");
    }
    if (!(options->pairlistCutoff > 0) || (options->rlist != 0 && options->rlist < options->pairlistCutoff))
    {
        return fail(NBNXM_ERR_INVALID, "force_calculator_create: pairlistCutoff must be > 0 and rlist 0 or >= pairlistCutoff");
    }
    if (options->accumulation < NBNXM_ACCUM_FLOAT || options->accumulation > NBNXM_ACCUM_FP64)
    {
        return fail(NBNXM_ERR_INVALID, "force_calculator_create: unknown accumulation mode");
    }
    for (int i = 0; i < numParticles; i++)
    {
        if (particleTypeIdOfAllParticles[i] < 0 || particleTypeIdOfAllParticles[i] >= numTypes)
        {
            return fail(NBNXM_ERR_INVALID, "force_calculator_create: particle type id out of range");
        }
    }
    if (exclusionRanges && exclusionRanges[numParticles] > exclusionRanges[0] && !exclusionElements)
    {
        return fail(NBNXM_ERR_INVALID, "force_calculator_create: exclusion ranges without elements");
    }
    auto c = std::make_unique<nbnxm_force_calculator>();
    c->numParticles = numParticles;
    c->numTypes = numTypes;
    c->types.assign(particleTypeIdOfAllParticles, particleTypeIdOfAllParticles + numParticles);
    c->nbfp.assign(nonBondedParams, nonBondedParams + 2 * static_cast<size_t>(numTypes) * numTypes);
    c->charges.assign(charges, charges + numParticles);
    if (exclusionRanges)
    {
        c->exclRanges.assign(exclusionRanges, exclusionRanges + numParticles + 1);
        for (int i = 0; i < numParticles; i++)
        {
            if (c->exclRanges[i + 1] < c->exclRanges[i])
            {
                return fail(NBNXM_ERR_INVALID, "force_calculator_create: exclusion ranges must be non-decreasing");
            }
        }
        const int numElements = c->exclRanges[numParticles] - c->exclRanges[0];
        if (numElements > 0)
        {
            c->exclElements.assign(exclusionElements + c->exclRanges[0], exclusionElements + c->exclRanges[numParticles]);
            for (int& v : c->exclRanges)
            {
                v -= exclusionRanges[0];
            }
            for (int e : c->exclElements)
            {
                if (e < 0 || e >= numParticles)
                {
                    return fail(NBNXM_ERR_INVALID, "force_calculator_create: exclusion element out of range");
                }
            }
        }
        else
        {
            c->exclRanges.clear();
        }
    }
    c->options = *options;
    // createInteractionConst (api/nblib/nbnxmsetuphelpers.cpp:219-276): LJ cut-off
    // with the potential shift at rvdw = rcoulomb = pairlistCutoff; PME without
    // the Coulomb potential shift ("we ignore the potential shift", sh_ewald = 0);
    // Cutoff and ReactionField through calc_rffac with epsilon_rf = 1, which
    // gives k_rf = 0 and c_rf = 1/rc for both.
    nbnxm_ic_options_t io{};
    io.eeltype = (options->coulombType == NBNXM_COULOMB_PME) ? NBNXM_ELEC_EWALD : NBNXM_ELEC_RF;
    io.vdwModifier = NBNXM_VDW_MOD_POTSHIFT;
    io.rcoulomb = options->pairlistCutoff;
    io.rvdw = options->pairlistCutoff;
    io.ewaldRtol = options->ewaldRtol > 0 ? options->ewaldRtol : 1e-5;
    io.epsilonR = options->epsilonR > 0 ? options->epsilonR : 1.0;
    io.epsilonRf = 1.0;
    io.tabulatedEwald = (options->coulombType == NBNXM_COULOMB_PME && options->useTabulatedEwaldCorr) ? 1 : 0;
    int rc = nbnxm_init_interaction_const_ex(&io, &c->ic);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    c->ic.sh_ewald = 0.0;
    c->rlist = options->rlist > 0 ? options->rlist : options->pairlistCutoff;
    rc = nbnxm_detect_comb_rule(numTypes, c->nbfp.data(), 1e-5);
    if (rc < 0)
    {
        return rc;
    }
    c->vdwType = rc;
    // nbfp with the trailing dummy type (atomdata.cpp:417, 497-503).
    const int ntp = numTypes + 1;
    c->nbfpTable.assign(2 * static_cast<size_t>(ntp) * ntp, 0.0f);
    for (int i = 0; i < numTypes; i++)
    {
        for (int j = 0; j < numTypes; j++)
        {
            c->nbfpTable[2 * (static_cast<size_t>(i) * ntp + j)] = static_cast<float>(c->nbfp[2 * (static_cast<size_t>(i) * numTypes + j)]);
            c->nbfpTable[2 * (static_cast<size_t>(i) * ntp + j) + 1] = static_cast<float>(c->nbfp[2 * (static_cast<size_t>(i) * numTypes + j) + 1]);
        }
    }
    *out = c.release();
    return NBNXM_OK;
}

extern "C" int nbnxm_force_calculator_update_pairlist(nbnxm_force_calculator_t* c, int numCoordinates, double* coordinates, const double box[3])
{
    if (!c || !box || (numCoordinates > 0 && !coordinates))
    {
        return fail(NBNXM_ERR_INVALID, "force_calculator_update_pairlist: null pointer");
    }
    if (numCoordinates != c->numParticles)
    {
        return fail(NBNXM_ERR_INVALID, "force_calculator_update_pairlist: coordinate array containing different number of entries than particles in the system");
    }
    for (int d = 0; d < 3; d++)
    {
        if (!(box[d] > c->rlist))
        {
            return fail(NBNXM_ERR_UNSUPPORTED, "force_calculator_update_pairlist: every box edge must exceed the list radius");
        }
    }
    // Put the particles in the box, in place (put_atoms_in_box_omp,
    // src/gromacs/pbcutil/pbc.cpp:1421-1530, rectangular case), such that the
    // FP32 value the kernel sees is inside [0, L) too.
    for (int i = 0; i < numCoordinates; i++)
    {
        for (int d = 0; d < 3; d++)
        {
            const double L = box[d];
            double v = coordinates[3 * i + d];
            v -= std::floor(v / L) * L;
            if (!(v >= 0) || !(static_cast<float>(v) < static_cast<float>(L)))
            {
                v = 0.0;
            }
            coordinates[3 * i + d] = v;
        }
    }
    std::copy(box, box + 3, c->box);
    int rc = c->search.putAtomsOnGrid(box, numCoordinates, coordinates);
    if (rc != NBNXM_OK)
    {
        return fail(rc, "force_calculator_update_pairlist: grid construction failed");
    }
    const int* er = c->exclRanges.empty() ? nullptr : c->exclRanges.data();
    const int* ee = c->exclRanges.empty() ? nullptr : c->exclElements.data();
    rc = c->search.constructPairlist(c->rlist, er, ee, c->splitTarget > 0 ? c->splitTarget : 0);
    if (rc != NBNXM_OK)
    {
        return fail(rc, "force_calculator_update_pairlist: pair list construction failed");
    }

    // Grid-order atom data (nonbonded_verlet_t::setAtomProperties).
    const int ng = c->search.numGridAtoms();
    c->xq.resize(4 * static_cast<size_t>(ng));
    c->search.fillXq(coordinates, c->charges.data(), c->xq.data());
    c->slotTypes.resize(ng);
    c->search.fillTypes(c->types.data(), c->numTypes, c->slotTypes.data());
    c->slotComb.assign(2 * static_cast<size_t>(ng), 0.0f);
    if (c->vdwType != NBNXM_VDW_CUT_TYPE_TABLE)
    {
        std::vector<float> comb;
        combParams(c->numTypes, c->nbfp.data(), c->vdwType, comb);
        c->search.fillLjComb(c->types.data(), comb.data(), c->slotComb.data());
    }
    nbnxm_atomdata_t nbat{};
    nbat.numAtoms = ng;
    nbat.xq = c->xq.data();
    nbat.numTypes = c->numTypes + 1;
    nbat.type = c->slotTypes.data();
    nbat.nbfp = c->nbfpTable.data();
    nbat.vdwType = c->vdwType;
    nbat.ljComb = c->slotComb.data();
    nbat.shiftVec = c->search.shiftVec();

    if (!c->gpu)
    {
        nbnxm_pairlist_params_t lp{};
        lp.rlistOuter = c->rlist;
        lp.rlistInner = c->rlist;
        lp.useDynamicPruning = 0;
        lp.numRollingPruningParts = 1;
        rc = nbnxm_gpu_init(&c->ic, &lp, &nbat, c->options.accumulation, &c->gpu);
        if (rc != NBNXM_OK)
        {
            c->gpu = nullptr;
            return rc;
        }
    }
    else
    {
        rc = nbnxm_gpu_init_atomdata(c->gpu, &nbat);
        if (rc == NBNXM_OK)
        {
            rc = nbnxm_gpu_upload_shiftvec(c->gpu, nbat.shiftVec);
        }
        if (rc != NBNXM_OK)
        {
            return rc;
        }
    }

    // Split entries when there are too few for the GPU (get_nsubpair_target,
    // src/gromacs/nbnxm/pairlist.cpp:2548-2640): aim at min_ci_balanced
    // entries, with at least 36 cluster pairs each.  The reference estimates
    // the target from the grid before building; here the first update builds
    // the list unsplit to count its pairs and then split, and later updates
    // build once with the target of the previous list.
    const int minCi = nbnxm_gpu_min_ci_balanced(c->gpu);
    nbnxm_pairlist_gpu_t view = c->search.listView();
    const int64_t pairs = c->search.numClusterPairs();
    const int target = (minCi > 0 && pairs > 0) ? static_cast<int>(std::max<int64_t>(36, (pairs + minCi - 1) / minCi)) : 0;
    if (c->splitTarget < 0)
    {
        c->splitTarget = 0;
        if (minCi > 0 && view.nsci > 0 && view.nsci < minCi)
        {
            rc = c->search.constructPairlist(c->rlist, er, ee, target);
            if (rc != NBNXM_OK)
            {
                return fail(rc, "force_calculator_update_pairlist: pair list construction failed");
            }
            view = c->search.listView();
            c->splitTarget = target;
        }
    }
    else if (c->splitTarget > 0)
    {
        c->splitTarget = target;
    }
    rc = nbnxm_gpu_init_pairlist(c->gpu, &view);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    c->havePairlist = true;
    return NBNXM_OK;
}

extern "C" int nbnxm_force_calculator_compute(nbnxm_force_calculator_t* c,
                                              int numCoordinates,
                                              const double* coordinates,
                                              const double box[3],
                                              double* forces,
                                              double* virial,
                                              double* energies)
{
    if (!c || !box || !forces || (numCoordinates > 0 && !coordinates))
    {
        return fail(NBNXM_ERR_INVALID, "force_calculator_compute: null pointer");
    }
    if (numCoordinates != c->numParticles)
    {
        return fail(NBNXM_ERR_INVALID, "force_calculator_compute: coordinate array and force buffer size mismatch");
    }
    if (!c->havePairlist)
    {
        return fail(NBNXM_ERR_STATE, "force_calculator_compute: compute called without updating pairlist at least once");
    }
    int rc = NBNXM_OK;
    // A changed box updates the shift vectors (updateForcerec, gmxcalculatorcpu.cpp:189-194).
    if (box[0] != c->box[0] || box[1] != c->box[1] || box[2] != c->box[2])
    {
        for (int d = 0; d < 3; d++)
        {
            if (!(box[d] > c->rlist))
            {
                return fail(NBNXM_ERR_UNSUPPORTED, "force_calculator_compute: every box edge must exceed the list radius");
            }
        }
        float sv[3 * NBNXM_NUM_SHIFTS];
        rc = nbnxm_calc_shifts(box, sv);
        if (rc == NBNXM_OK)
        {
            rc = nbnxm_gpu_upload_shiftvec(c->gpu, sv);
        }
        if (rc != NBNXM_OK)
        {
            return rc;
        }
        std::copy(box, box + 3, c->box);
    }
    const bool computeVirial = virial != nullptr;
    const bool computeEnergy = energies != nullptr;
    // convertCoordinates: coordinates as given, in grid order.
    c->search.fillXq(coordinates, c->charges.data(), c->xq.data());
    const int ng = c->search.numGridAtoms();
    if ((rc = nbnxm_gpu_copy_xq_to_gpu(c->gpu, c->xq.data(), ng)) != NBNXM_OK)
    {
        return rc;
    }
    if ((rc = nbnxm_gpu_clear_outputs(c->gpu, computeVirial)) != NBNXM_OK)
    {
        return rc;
    }
    if ((rc = nbnxm_gpu_launch_kernel(c->gpu, computeEnergy, computeVirial)) != NBNXM_OK)
    {
        return rc;
    }
    c->fGrid.assign(3 * static_cast<size_t>(ng), 0.0);
    std::vector<float> fFloat;
    const bool floatForces = c->options.accumulation == NBNXM_ACCUM_FLOAT;
    if (floatForces)
    {
        fFloat.assign(3 * static_cast<size_t>(ng), 0.0f);
    }
    if ((rc = nbnxm_gpu_launch_cpyback(c->gpu, floatForces ? fFloat.data() : nullptr, floatForces ? nullptr : c->fGrid.data())) != NBNXM_OK)
    {
        return rc;
    }
    double eLj = 0.0, eEl = 0.0;
    double fshift[3 * NBNXM_NUM_SHIFTS] = {};
    if ((rc = nbnxm_gpu_wait_finish_task(c->gpu, &eLj, &eEl, fshift)) != NBNXM_OK)
    {
        return rc;
    }
    if (floatForces)
    {
        for (size_t k = 0; k < fFloat.size(); k++)
        {
            c->fGrid[k] = fFloat[k];
        }
    }
    // atomdata_add_nbat_f_to_f (src/gromacs/nbnxm/atomdata.cpp:1224-1262): add, grid -> original order.
    const std::vector<int>& atomIndex = c->search.atomIndices();
    for (int s = 0; s < ng; s++)
    {
        const int a = atomIndex[s];
        if (a >= 0)
        {
            forces[3 * a + 0] += c->fGrid[3 * static_cast<size_t>(s) + 0];
            forces[3 * a + 1] += c->fGrid[3 * static_cast<size_t>(s) + 1];
            forces[3 * a + 2] += c->fGrid[3 * static_cast<size_t>(s) + 2];
        }
    }
    if (computeVirial)
    {
        std::fill(virial, virial + 9, 0.0);
        if ((rc = nbnxm_gpu_get_virial(c->gpu, virial)) != NBNXM_OK)
        {
            return rc;
        }
    }
    if (computeEnergy)
    {
        std::fill(energies, energies + NBNXM_NUM_ENERGY_TERMS, 0.0);
        energies[0] = eEl; // CoulombSR
        energies[1] = eLj; // LJSR
    }
    return NBNXM_OK;
}

extern "C" void nbnxm_force_calculator_free(nbnxm_force_calculator_t* c)
{
    delete c;
}
