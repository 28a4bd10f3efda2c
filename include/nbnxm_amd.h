/*
 * nbnxm_amd.h — C ABI of the MI355X (gfx950) nbnxm short-range cluster-pair library.
 *
 * Everything here is plain C: opaque handles, plain pointers and sizes.  The
 * data layouts are the ones GROMACS uses on its GPU path, so a GROMACS build
 * (or any other host) can hand its own buffers over without conversion:
 *
 *   - pairlist: sci / cj_packed / excl exactly as NbnxnPairlistGpu
 *     (src/gromacs/nbnxm/pairlist.h:201-357) with the cluster-pair split
 *     disabled, i.e. the layout GROMACS itself selects for 64-wide AMD
 *     devices (GMX_GPU_NB_DISABLE_CLUSTER_PAIR_SPLIT,
 *     src/gromacs/nbnxm/pairlistparams.h:88-98): one imask/excl_ind per
 *     packed j-group and 64 interaction words per exclusion entry.
 *   - atom data: xq float4 (x, y, z, q) per atom in grid order, padded to
 *     whole clusters (src/gromacs/nbnxm/atomdata.cpp:629-638,736-790),
 *     per-atom type index with numTypes-1 the dummy type, nbfp holding
 *     (6*C6, 12*C12) per type pair (src/gromacs/nbnxm/atomdata.h:190).
 *   - shift vectors: 45 x 3 floats, index 5*(3*(z+1)+y+1)+x+2
 *     (src/gromacs/pbcutil/include/gromacs/pbcutil/ishift.h:43-62).
 *
 * All functions return 0 on success and a negative NBNXM_ERR_* code on
 * failure; nbnxm_last_error() gives the message.  The reference aborts via
 * gmx_fatal / CU_RET_ERR instead (src/gromacs/nbnxm/cuda/nbnxm_cuda.cu:140-148).
 */
#ifndef NBNXM_AMD_H
#define NBNXM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry constants (src/gromacs/nbnxm/pairlistparams.h:59-99, pairlist.h:141-183) ---- */
#define NBNXM_CLUSTER_SIZE 8              /* atoms per i- and j-cluster */
#define NBNXM_NUM_CLUSTER_PER_SUPERCLUSTER 8 /* 2x2x2 clusters per super-cluster */
#define NBNXM_JGROUP_SIZE 4               /* j-clusters per packed group (32 mask bits / 8) */
#define NBNXM_EXCL_SIZE 64                /* interaction words per exclusion entry, split 1 */
#define NBNXM_NUM_SHIFTS 45
#define NBNXM_CENTRAL_SHIFT 22
#define NBNXM_MIN_DISTANCE_SQUARED 3.82e-07f /* src/gromacs/nbnxm/pairlist.h:169 */
#define NBNXM_SCI_HISTOGRAM_SIZE 8192     /* src/gromacs/nbnxm/gpu_types_common.h:72 */

/* ---- error codes ---- */
#define NBNXM_OK 0
#define NBNXM_ERR_INVALID -1   /* bad argument (null pointer, size, layout) */
#define NBNXM_ERR_HIP -2       /* HIP runtime failure */
#define NBNXM_ERR_STATE -3     /* call out of order (e.g. launch before list upload) */
#define NBNXM_ERR_UNSUPPORTED -4

/* ---- list types: NbnxnPairlistGpu, split = 1 (src/gromacs/nbnxm/pairlist.h:201-283) ---- */
typedef struct nbnxm_sci
{
    int sci;           /* i-super-cluster index */
    int shift;         /* shift vector index (0..44); the ensemble adds 45*system */
    int cjPackedBegin; /* first packed j-group */
    int cjPackedEnd;   /* one past the last packed j-group */
} nbnxm_sci_t;

typedef struct nbnxm_im_ei
{
    unsigned int imask; /* bit jm*8+im: j-cluster jm of the group x i-cluster im of the sci */
    int excl_ind;       /* index into the exclusion array, 0 = no exclusions */
} nbnxm_im_ei_t;

typedef struct nbnxm_cj_packed
{
    int cj[NBNXM_JGROUP_SIZE];
    nbnxm_im_ei_t imei[1];
} nbnxm_cj_packed_t;

typedef struct nbnxm_excl
{
    unsigned int pair[NBNXM_EXCL_SIZE]; /* word jInner*8+iInner, bit jm*8+im */
} nbnxm_excl_t;

/* Host view of a GPU pair list (NbnxnPairlistGpu, src/gromacs/nbnxm/pairlist.h:324-357). */
typedef struct nbnxm_pairlist_gpu
{
    int na_ci; /* must be NBNXM_CLUSTER_SIZE */
    float rlist;
    int nsci;
    const nbnxm_sci_t* sci;
    int ncjPacked;
    const nbnxm_cj_packed_t* cjPacked;
    int nexcl;
    const nbnxm_excl_t* excl; /* excl[0] must be all ones */
} nbnxm_pairlist_gpu_t;

/* ---- interaction parameters ---- */
enum nbnxm_vdw_type
{
    NBNXM_VDW_CUT_TYPE_TABLE = 0, /* LJ cut-off, per-type-pair nbfp table (VdwType::Cut) */
    NBNXM_VDW_CUT_COMB_GEOM = 1,  /* LJ cut-off, geometric combination (VdwType::CutCombGeom) */
    NBNXM_VDW_CUT_COMB_LB = 2     /* LJ cut-off, Lorentz-Berthelot (VdwType::CutCombLB) */
};

/* Fixed-point force accumulation: forces are summed as int64 multiples of 2^-NBNXM_FIXED_FRAC_BITS. */
#define NBNXM_FIXED_FRAC_BITS 32

enum nbnxm_accumulation
{
    NBNXM_ACCUM_FLOAT = 0,   /* FP32 atomics, order-dependent (the reference behaviour) */
    NBNXM_ACCUM_FIXED64 = 1, /* FP32 pair math, int64 fixed-point forces, FP64 energy/fshift reduction */
    NBNXM_ACCUM_FP64 = 2     /* FP64 pair math with exact erfc/exp (GPU oracle mode) */
};

/* Coulomb flavour (CoulombInteractionType of the kernel). */
enum nbnxm_elec_type
{
    NBNXM_ELEC_EWALD = 0, /* Ewald real space, analytical correction (PME) */
    NBNXM_ELEC_RF = 1     /* reaction field; with k_rf = 0 the plain cut-off */
};

/* LJ-PME grid combination rule (interaction_const_t::ljpme_comb_rule,
 * LongRangeVdW; the GPU kernels VdwType::EwaldGeom / EwaldLB,
 * src/gromacs/nbnxm/nbnxm_gpu_data_mgmt.cpp:327-346). */
#define NBNXM_LJPME_NONE 0
#define NBNXM_LJPME_GEOM 1 /* grid C6 = sqrt(6C6_ii 6C6_jj) */
#define NBNXM_LJPME_LB 2   /* grid C6 from per-type sigma and epsilon, Lorentz-Berthelot
                            * (calculate_lj_ewald_comb_LB_F_E, cuda/nbnxm_cuda_kernel_utils.cuh:290-337) */

/* LJ interaction modifier (InteractionModifiers of the kernel). */
enum nbnxm_vdw_modifier
{
    NBNXM_VDW_MOD_POTSHIFT = 0,
    NBNXM_VDW_MOD_FORCESWITCH = 1,
    NBNXM_VDW_MOD_POTSWITCH = 2
};

/* Subset of interaction_const_t used by the kernel
 * (src/gromacs/mdtypes/interaction_const.cpp:106-115,277-278,309). */
typedef struct nbnxm_interaction_const
{
    double rcoulomb;
    double rvdw;              /* <= rcoulomb; smaller gives the twin cut-off (LJ cut-off check) */
    double ewaldcoeff_q;      /* beta, from nbnxm_calc_ewaldcoeff_q */
    double sh_ewald;          /* erfc(beta*rc)/rc with PotShift, else 0 */
    double epsfac;            /* c_one4PiEps0 / epsilon_r */
    double dispersion_shift_cpot; /* -rvdw^-6 with PotShift; the force-switch constant with ForceSwitch */
    double repulsion_shift_cpot;  /* -rvdw^-12 with PotShift; the force-switch constant with ForceSwitch */
    /* Kernel flavour; nbnxm_init_interaction_const sets Ewald, PotShift, no LJ-PME. */
    int eeltype;              /* enum nbnxm_elec_type */
    int vdwModifier;          /* enum nbnxm_vdw_modifier */
    int ljPme;                /* NBNXM_LJPME_*: LJ-PME grid correction and its grid combination rule */
    double rvdw_switch;
    double ewaldcoeff_lj;     /* calc_ewaldcoeff_lj, src/gromacs/ewald/ewald_utils.cpp:84-112 */
    double sh_lj_ewald;       /* LJ-PME potential shift (interaction_const.cpp:140-143) */
    double dispersion_shift_c2, dispersion_shift_c3; /* force_switch_constants, interaction_const.cpp:214-230 */
    double repulsion_shift_c2, repulsion_shift_c3;
    double vdw_switch_c3, vdw_switch_c4, vdw_switch_c5; /* potential_switch_constants, :232-246 */
    double k_rf;              /* reaction-field coefficient (calc_rffac, src/gromacs/mdlib/rf_util.cpp:50-62) */
    double c_rf;              /* reaction-field potential shift */
    /* Nonzero: the Ewald correction force comes from the quadratic-spline table
     * of init_interaction_const_tables (src/gromacs/mdlib/forcerec.cpp:586-640),
     * ElecType::EwaldTab of nbnxm_gpu_data_mgmt.cpp:128-175; zero: analytical. */
    int tabulatedEwald;
} nbnxm_interaction_const_t;

/* Flavour choices for nbnxm_init_interaction_const_ex (t_inputrec fields). */
typedef struct nbnxm_ic_options
{
    int eeltype;        /* enum nbnxm_elec_type */
    int vdwModifier;    /* enum nbnxm_vdw_modifier */
    int ljPme;          /* NBNXM_LJPME_*: LJ-PME with the geometric or Lorentz-Berthelot grid rule */
    double rcoulomb;
    double rvdw;        /* 0 -> rcoulomb */
    double rvdw_switch; /* ForceSwitch / PotSwitch */
    double ewaldRtol;   /* Coulomb Ewald tolerance */
    double ewaldRtolLj; /* LJ-PME tolerance */
    double epsilonR;
    double epsilonRf;   /* reaction field: 0 = infinity */
    int plainCutoff;    /* with NBNXM_ELEC_RF: plain cut-off (k_rf = 0, c_rf = 1/rc) */
    int tabulatedEwald; /* with NBNXM_ELEC_EWALD: tabulated instead of analytical correction */
} nbnxm_ic_options_t;

/* Pair-list parameters (src/gromacs/nbnxm/pairlistparams.h, PairlistParams). */
typedef struct nbnxm_pairlist_params
{
    double rlistOuter;
    double rlistInner;
    int useDynamicPruning;
    int numRollingPruningParts;
} nbnxm_pairlist_params_t;

/* Host view of the kernel inputs of nbnxn_atomdata_t (src/gromacs/nbnxm/atomdata.h:176-310). */
typedef struct nbnxm_atomdata
{
    int numAtoms;         /* grid atoms including fillers, multiple of 8 */
    const float* xq;      /* 4*numAtoms */
    int numTypes;         /* including the trailing dummy type */
    const int* type;      /* numAtoms; may be NULL with a combination rule */
    const float* nbfp;    /* 2*numTypes*numTypes: 6*C6, 12*C12 */
    int vdwType;          /* enum nbnxm_vdw_type */
    const float* ljComb;  /* 2*numAtoms, required for COMB_GEOM / COMB_LB */
    const float* shiftVec;/* 3*NBNXM_NUM_SHIFTS */
} nbnxm_atomdata_t;

/* ---------------------------------------------------------------------------
 * Library-wide
 * ------------------------------------------------------------------------- */
/* Message for the last failing call on this thread. */
const char* nbnxm_last_error(void);
/* Library version string. */
const char* nbnxm_version(void);
/* Beta from the Ewald tolerance: restates calc_ewaldcoeff_q
 * (src/gromacs/ewald/ewald_utils.cpp:43-71). */
double nbnxm_calc_ewaldcoeff_q(double rc, double rtol);
/* Fills ic for Ewald (analytical) + LJ cut-off with PotShift on both
 * (src/gromacs/mdtypes/interaction_const.cpp:106-115,266-278,306-314). */
int nbnxm_init_interaction_const(double rc, double ewaldRtol, double epsilonR, nbnxm_interaction_const_t* ic);
/* Any supported flavour: init_interaction_const
 * (src/gromacs/mdtypes/interaction_const.cpp:249-345) for Ewald or reaction
 * field with PotShift Coulomb, LJ with PotShift / ForceSwitch / PotSwitch, an
 * optional LJ-PME grid correction and rvdw <= rcoulomb. */
int nbnxm_init_interaction_const_ex(const nbnxm_ic_options_t* options, nbnxm_interaction_const_t* ic);
/* calc_ewaldcoeff_lj (src/gromacs/ewald/ewald_utils.cpp:73-112). */
double nbnxm_calc_ewaldcoeff_lj(double rc, double rtol);
/* Coulomb Ewald correction force table of the tabulated kernel: the scale of
 * ewald_spline3_table_scale (src/gromacs/tables/forcetable.cpp:296-365, FP32
 * accuracy targets) and tableF of generateEwaldCorrectionTables
 * (:139-276) for v(r) = erf(beta r)/r over [0, rc], int(rc*scale) + 2
 * points.  Returns the number of points (tableF may be NULL to query it;
 * capacity is the room in tableF or tableFD), or a negative error code.  The
 * table values are computed in FP64 and stored in FP32 (tableF) and / or FP64
 * (tableFD); *scale gets the FP32-rounded scale the kernel uses. */
int nbnxm_ewald_correction_table(double beta, double rc, float* tableF, double* tableFD, int capacity, double* scale);
/* Shift vectors of a rectangular box (calc_shifts, src/gromacs/pbcutil/pbc.cpp:1218-1233). */
int nbnxm_calc_shifts(const double box[3], float* shiftVec /* 3*45 */);

/* ---------------------------------------------------------------------------
 * Pair search: grid of 8-atom clusters and the GPU pair list.
 * Replaces nonbonded_verlet_t::putAtomsOnGrid (src/gromacs/nbnxm/nbnxm.cpp) and
 * constructPairlist (src/gromacs/nbnxm/pairlist.cpp:4353) for the GPU list
 * type; the list semantics are those of make_cluster_list_supersub
 * (pairlist.cpp:1122-1270), setExclusionsForIEntry (pairlist.cpp:1948-2046)
 * and split_sci_entry (pairlist.cpp:2150-2262).
 * ------------------------------------------------------------------------- */
typedef struct nbnxm_search nbnxm_search_t;

nbnxm_search_t* nbnxm_search_create(void);
void nbnxm_search_free(nbnxm_search_t* s);
/* Wraps x (3*numAtoms) into the rectangular box and sorts atoms into clusters. */
int nbnxm_search_put_atoms_on_grid(nbnxm_search_t* s, const double box[3], int numAtoms, const double* x);
/* Builds the list for rlist; exclusions as ListOfLists<int> (ranges numAtoms+1,
 * elements), both may be NULL.  maxPairsPerEntry > 0 splits long entries.
 * Every box edge must exceed rlist. */
int nbnxm_search_construct_pairlist(nbnxm_search_t* s, double rlist, const int* exclRanges, const int* exclElements, int maxPairsPerEntry);
/* View of the list; valid until the next construct or free. */
int nbnxm_search_get_pairlist(const nbnxm_search_t* s, nbnxm_pairlist_gpu_t* list);
/* Grid atoms including fillers (multiple of 64). */
int nbnxm_search_num_grid_atoms(const nbnxm_search_t* s);
/* Grid slot -> original atom index, -1 for fillers (gridSet.atomIndices()). */
int nbnxm_search_atom_indices(const nbnxm_search_t* s, int* atomIndices);
int nbnxm_search_shift_vec(const nbnxm_search_t* s, float* shiftVec /* 3*45 */);
/* Grid-order xq (4 per slot) from original-order x (3 per atom, NULL = the
 * wrapped grid coordinates) and q (NULL = 0); fillers at -1e6 with q = 0
 * (copy_rvec_to_nbat_real, src/gromacs/nbnxm/atomdata.cpp:136-259). */
int nbnxm_search_fill_xq(const nbnxm_search_t* s, const double* x, const double* q, float* xq);
int nbnxm_search_fill_types(const nbnxm_search_t* s, const int* type, int dummyType, int* slotType);
int nbnxm_search_fill_ljcomb(const nbnxm_search_t* s, const int* type, const float* perTypeComb, float* slotComb);
/* Useful pairs of the current list: non-excluded atom pairs (all images)
 * within rc, FP64 distances on the wrapped search coordinates; rc must not
 * exceed the list radius.  The exact count the benchmark harness estimates in
 * closed form (src/gromacs/nbnxm/benchmark/bench_setup.cpp:291-295). */
int nbnxm_search_count_pairs(const nbnxm_search_t* s, double rc, int64_t* count);

/* ---------------------------------------------------------------------------
 * Single-system GPU task: the Nbnxm::gpu_* API
 * (src/gromacs/nbnxm/nbnxm_gpu.h:66-302, src/gromacs/nbnxm/gpu_data_mgmt.h).
 * One HIP stream per task; calls are asynchronous except wait_finish_task and
 * the getters.
 * ------------------------------------------------------------------------- */
typedef struct nbnxm_gpu nbnxm_gpu_t;

/* gpu_init (src/gromacs/nbnxm/nbnxm_gpu_data_mgmt.cpp:442): parameters, atom
 * data (as gpu_init_atomdata) and shift vectors when nbat->shiftVec is set.
 * The flavour comes from ic: Ewald + LJ potential shift with rvdw = rcoulomb
 * runs the tuned kernels; every other flavour (reaction field, plain cut-off,
 * twin cut-off, force / potential switch, LJ-PME) runs the generic flavour
 * kernel (nbnxm_cuda.cu:171-254 selects among the same set).  LJ-PME needs
 * nbat->type and nbat->nbfp for every vdwType: the grid C6 of each type pair
 * comes from the nbfp diagonal by ic->ljPme's rule (atomdata.cpp:296-330).
 * Timing (gpu_get_timings) starts enabled when the environment sets
 * GMX_ENABLE_GPU_TIMING, as decideGpuTimingsUsage does for CUDA
 * (src/gromacs/gpu_utils/gpu_utils.cpp:66-76). */
int nbnxm_gpu_init(const nbnxm_interaction_const_t* ic,
                   const nbnxm_pairlist_params_t* listParams,
                   const nbnxm_atomdata_t* nbat,
                   int accumulation,
                   nbnxm_gpu_t** nb);
/* gpu_init_atomdata (nbnxm_gpu_data_mgmt.cpp:689). */
int nbnxm_gpu_init_atomdata(nbnxm_gpu_t* nb, const nbnxm_atomdata_t* nbat);
/* gpu_upload_shiftvec (nbnxm_gpu_data_mgmt.cpp:528). */
int nbnxm_gpu_upload_shiftvec(nbnxm_gpu_t* nb, const float* shiftVec);
/* gpu_init_pairlist (nbnxm_gpu_data_mgmt.cpp:548); marks the list fresh. */
int nbnxm_gpu_init_pairlist(nbnxm_gpu_t* nb, const nbnxm_pairlist_gpu_t* list);
/* gpu_copy_xq_to_gpu (nbnxm_gpu_data_mgmt.cpp:1006). */
int nbnxm_gpu_copy_xq_to_gpu(nbnxm_gpu_t* nb, const float* xq, int numAtoms);
/* gpu_clear_outputs (nbnxm_gpu_data_mgmt.cpp:792). */
int nbnxm_gpu_clear_outputs(nbnxm_gpu_t* nb, int computeVirial);
/* gpu_launch_kernel (src/gromacs/nbnxm/cuda/nbnxm_cuda.cu:478): prunes a fresh
 * list first, then the nonbonded kernel (F or VF). */
int nbnxm_gpu_launch_kernel(nbnxm_gpu_t* nb, int computeEnergy, int computeVirial);
/* gpu_launch_kernel_pruneonly (nbnxm_cuda.cu:606): fresh list -> outer and inner
 * masks plus the sci ordering; otherwise one rolling part of numParts. */
int nbnxm_gpu_launch_kernel_pruneonly(nbnxm_gpu_t* nb, int numParts);
/* gpu_launch_cpyback (nbnxm_gpu_data_mgmt.cpp:853): forces in grid order to f
 * (3*numAtoms floats) and/or fDouble; either may be NULL. */
int nbnxm_gpu_launch_cpyback(nbnxm_gpu_t* nb, float* f, double* fDouble);
/* gpu_wait_finish_task (src/gromacs/nbnxm/gpu_common.h:345): waits and adds the
 * energies and shift forces (3*45) of the last launch to the outputs. */
int nbnxm_gpu_wait_finish_task(nbnxm_gpu_t* nb, double* e_lj, double* e_el, double* fshift);
/* gpu_try_finish_task with GpuTaskCompletion::Check
 * (src/gromacs/nbnxm/nbnxm_gpu.h:199, gpu_common.h:253-333): returns 0 at
 * once while work queued on the task's stream is still running, otherwise
 * does what nbnxm_gpu_wait_finish_task does (energies and shift forces added
 * to the outputs, timings accounted) and returns 1; negative on error. */
int nbnxm_gpu_try_finish_task(nbnxm_gpu_t* nb, double* e_lj, double* e_el, double* fshift);
/* gpu_pme_loadbal_update_param (src/gromacs/nbnxm/gpu_data_mgmt.h:92,
 * nbnxm_gpu_data_mgmt.cpp:511-527), called by PME load balancing
 * (src/gromacs/ewald/pme_load_balancing.cpp:856): new cut-off, Ewald
 * coefficient, potential shifts and the flavour they select (twin cut-off
 * when rcoulomb grows past rvdw; analytical or tabulated Ewald with the
 * correction table regenerated), keeping the list and the atom data.
 * listParams (may be NULL: unchanged) are the new list radii, as
 * set_cutoff_parameters reads them from the pair-list sets; the caller builds
 * the next list at the new radius.  The LJ-PME rule cannot change here. */
int nbnxm_gpu_pme_loadbal_update_param(nbnxm_gpu_t* nb, const nbnxm_interaction_const_t* ic, const nbnxm_pairlist_params_t* listParams);

/* gmx_kernel_timing_data_t (src/gromacs/timing/include/gromacs/timing/gpu_timing.h:48-52). */
typedef struct nbnxm_kernel_timing
{
    double t; /* accumulated time, ms */
    int c;    /* number of calls */
} nbnxm_kernel_timing_t;

/* gmx_wallclock_gpu_nbnxn_t (gpu_timing.h:80-94), filled from HIP events on
 * the task's stream and accounted when a finish call completes the task. */
typedef struct nbnxm_gpu_timings
{
    nbnxm_kernel_timing_t ktime[2][2];    /* [fresh list][energies]: nonbonded launches */
    nbnxm_kernel_timing_t pruneTime;      /* first-pass prune of a fresh list */
    nbnxm_kernel_timing_t dynamicPruneTime; /* rolling prune */
    double nb_h2d_t;                      /* xq uploads */
    double nb_d2h_t;                      /* force, energy and shift-force downloads */
    int nb_c;                             /* finished tasks with a nonbonded launch */
    double pl_h2d_t;                      /* pair-list and atom-data uploads */
    int pl_h2d_c;
} nbnxm_gpu_timings_t;

/* Turns event timing of the task's launches and transfers on (1) or off (0);
 * no counterpart (the reference decides once from the environment). */
int nbnxm_gpu_set_timing(nbnxm_gpu_t* nb, int enable);
/* gpu_get_timings (gpu_data_mgmt.h:109, nbnxm_gpu_data_mgmt.cpp:809-812):
 * returns 1 and copies the accumulated timings when timing is on, 0 (and
 * leaves *timings alone) when it is off, as the reference returns null. */
int nbnxm_gpu_get_timings(const nbnxm_gpu_t* nb, nbnxm_gpu_timings_t* timings);
/* gpu_reset_timings (gpu_data_mgmt.h:114, nbnxm_gpu_data_mgmt.cpp:815-821). */
int nbnxm_gpu_reset_timings(nbnxm_gpu_t* nb);
/* gpu_get_f (gpu_data_mgmt.h:133): the device nbat force buffer, 3 floats per
 * grid slot, that GpuForceReduction registers (registerNbnxmForce,
 * src/gromacs/mdlib/sim_util.cpp:1167,1248) and reduces into the atom-order
 * force buffer through the grid's cell indices.  With fixed-point
 * accumulation the integer accumulator cannot be registered, so once this is
 * called every nbnxm_gpu_launch_kernel ends by writing the accumulated forces
 * there as floats (converted exactly from fixed point; the virial partials
 * are taken before) on the task's stream; with FLOAT accumulation it is the
 * accumulator itself.  Either way the forces then accumulate over launches
 * until nbnxm_gpu_clear_outputs, as the reference's nbat force buffer does
 * (gpu_clear_outputs, nbnxm_gpu_data_mgmt.cpp:792); nbnxm_gpu_launch_cpyback
 * and nbnxm_gpu_add_nbat_f_to_f read the same buffer and no longer clear it.
 * A consumer orders its reads after the launch on the task's stream
 * (nbnxm_gpu_get_stream).  The address is valid until the next
 * nbnxm_gpu_init_atomdata, after which the reference registers it again
 * (sim_util.cpp:1155-1167 at every search step). */
int nbnxm_gpu_get_f(nbnxm_gpu_t* nb, float** d_f);
/* The task's HIP stream (the reference's local nonbonded DeviceStream,
 * nbnxm_cuda_types.h:106), as a hipStream_t. */
int nbnxm_gpu_get_stream(nbnxm_gpu_t* nb, void** stream);
/* Device inputs of nbnxm_gpu_launch_kernel_noalloc: what gpu_init_atomdata,
 * gpu_upload_shiftvec and gpu_init_pairlist would upload, already in DEVICE
 * memory owned by the caller, in the layouts of nbnxm_atomdata_t and
 * nbnxm_pairlist_gpu_t. */
typedef struct nbnxm_device_inputs
{
    int numAtoms;                      /* grid slots, multiple of 8 */
    const float* xq;                   /* 4*numAtoms */
    int numTypes;                      /* type table: including the dummy type */
    const int* type;                   /* numAtoms, type table kernel, else NULL */
    const float* nbfp;                 /* 2*numTypes^2, type table kernel, else NULL */
    int vdwType;                       /* enum nbnxm_vdw_type */
    const float* ljComb;               /* 2*numAtoms, combination rules, else NULL */
    const float* shiftVec;             /* 4*NBNXM_NUM_SHIFTS: x, y, z, 0 per shift */
    int nsci;                          /* entries, at most NBNXM_NOALLOC_MAX_ENTRIES */
    const nbnxm_sci_t* sci;
    const nbnxm_cj_packed_t* cjPacked; /* masks as evaluated (a pruned list is not pruned again) */
    const nbnxm_excl_t* excl;          /* excl[0] all ones */
} nbnxm_device_inputs_t;

/* Largest number of pair-list entries of one nbnxm_gpu_launch_kernel_noalloc
 * call: its scratch holds 64 bytes per entry (4 MiB in all, per GPU). */
#define NBNXM_NOALLOC_MAX_ENTRIES 65536

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the nonbonded evaluation of nbnxm_gpu_launch_kernel on a
 * list that needs no further pruning, followed by the energy and shift-force
 * reduction of nbnxm_gpu_launch_cpyback, on caller-owned device buffers,
 * asynchronously on `stream` (a hipStream_t, NULL = the null stream).  It
 * allocates nothing: its scratch (parameter blocks and the FP64 partials of
 * every entry) is a fixed library array on each GPU, so it accepts at most
 * NBNXM_NOALLOC_MAX_ENTRIES entries (NBNXM_ERR_INVALID beyond).  Calls that
 * overlap on one GPU share that array and race: order them on one stream.
 * The tuned flavour only (analytical Ewald + LJ potential shift, rvdw =
 * rcoulomb; NBNXM_ERR_UNSUPPORTED otherwise), any accumulation mode.  Forces
 * are ADDED to d_fFixed (FIXED64 / FP64: int64, 2^-32 units, 3 per slot) or
 * d_fFloat (FLOAT); d_energies (2: E_lj, E_el) and d_fshift (3*45) are
 * written, and are only needed with computeEnergy / computeFshift.  Forces,
 * energies and shift forces equal, bit for bit, those of a task created by
 * nbnxm_gpu_init with the same constants, list and atom data. */
int nbnxm_gpu_launch_kernel_noalloc(const nbnxm_interaction_const_t* ic,
                                    const nbnxm_device_inputs_t* in,
                                    int accumulation,
                                    int computeEnergy,
                                    int computeFshift,
                                    unsigned long long* d_fFixed,
                                    float* d_fFloat,
                                    double* d_energies,
                                    double* d_fshift,
                                    void* stream);
/* Adds the single-sum virial (9, row-major) of the last launch made with
 * computeVirial to virial, after waiting.  Replaces the host calc_vir over the
 * nbat forces and shift forces (src/gromacs/mdlib/calcvir.cpp:89-133, as
 * api/nblib/virials.cpp:55-84 uses it): here the sum over atoms and shifts is
 * reduced on the GPU in FP64 in a fixed order. */
int nbnxm_gpu_get_virial(nbnxm_gpu_t* nb, double* virial);
/* GPU buffer ops, for callers whose coordinates and forces live on the device.
 * nbnxn_gpu_init_x_to_nbat_x (src/gromacs/nbnxm/nbnxm_gpu_data_mgmt.cpp:1076):
 * uploads the grid's slot -> atom map (numGridAtoms entries, -1 for fillers;
 * nbnxm_search_atom_indices gives it) and derives the atom -> slot map. */
int nbnxm_gpu_init_x_to_nbat_x(nbnxm_gpu_t* nb, int numAtoms, int numGridAtoms, const int* atomIndices);
/* nbnxn_gpu_x_to_nbat_x (src/gromacs/nbnxm/nbnxm_gpu_buffer_ops.cpp:63,
 * kernel cuda/nbnxm_gpu_buffer_ops_internal.cu:66-105): d_x is a DEVICE array
 * of 3*numAtoms floats in atom order; writes x, y, z of every real slot of xq
 * and keeps the charges.  Asynchronous on the task's stream. */
int nbnxm_gpu_x_to_nbat_x(nbnxm_gpu_t* nb, const float* d_x);
/* GPU force reduction (reduceKernel, src/gromacs/mdlib/gpuforcereduction_impl_internal.cu:62-120):
 * d_f, a DEVICE array of 3*numAtoms floats in atom order, is set to (accumulate
 * = 0) or incremented by the nonbonded forces of the last launch, then d_fToAdd
 * (DEVICE, may be NULL) is added.  Takes the forces of that launch: a following
 * gpu_launch_cpyback then only reduces energies, shift forces and virial. */
int nbnxm_gpu_add_nbat_f_to_f(nbnxm_gpu_t* nb, float* d_f, const float* d_fToAdd, int accumulate);
/* gpu_min_ci_balanced (gpu_data_mgmt.h:119). */
int nbnxm_gpu_min_ci_balanced(nbnxm_gpu_t* nb);
/* gpu_is_kernel_ewald_analytical (gpu_data_mgmt.h:123,
 * nbnxm_gpu_data_mgmt.cpp:823-827): 1 for analytical Ewald (twin or not), 0 for
 * tabulated Ewald, reaction field and plain cut-off. */
int nbnxm_gpu_is_kernel_ewald_analytical(const nbnxm_gpu_t* nb);
/* Downloads the per-group masks (cjPacked imask and the rlistOuter mask,
 * gpu_plist::imask); no counterpart in the reference, used by tests. */
int nbnxm_gpu_get_masks(nbnxm_gpu_t* nb, unsigned int* imaskInner, unsigned int* imaskOuter);
/* Downloads the entry work order and histogram bins after a fresh prune
 * (gpuPlistSorting::sciSorted / sciCount). */
int nbnxm_gpu_get_sci_order(nbnxm_gpu_t* nb, int* order, int* sciCount);
/* Launch shape of the nonbonded kernel: waves (entries) per workgroup, 1..4. */
int nbnxm_gpu_set_waves_per_block(nbnxm_gpu_t* nb, int wavesPerBlock);
/* Where the list walk keeps the i-cluster coordinates and charges: 1 in LDS
 * at 4 waves per SIMD (every F kernel of the compile-time flavour classes;
 * F+E of the tuned, reaction-field and tabulated-Ewald classes with a
 * combination rule; the other kernels keep registers), 0 in registers (3
 * waves per SIMD), -1 chosen by the entry count (LDS from 8192 entries).
 * Launch tuning with no counterpart in the reference; fixed-point results
 * are bit-identical either way. */
int nbnxm_gpu_set_i_cluster_in_lds(nbnxm_gpu_t* nb, int mode);
/* gpu_free. */
void nbnxm_gpu_free(nbnxm_gpu_t* nb);

/* ---------------------------------------------------------------------------
 * Ensemble: many independent systems of uneven size in one launch.
 * No single reference counterpart: it replaces one nonbonded_verlet_t + one
 * gpu_launch_kernel per system (src/gromacs/nbnxm/cuda/nbnxm_cuda.cu:478) with
 * one list over all systems, one nonbonded launch per evaluation and a hipGraph
 * over many evaluations.  Systems share the LJ type table.
 * ------------------------------------------------------------------------- */
typedef struct nbnxm_ensemble nbnxm_ensemble_t;

typedef struct nbnxm_system_desc
{
    int numAtoms;
    const double* x;          /* 3*numAtoms */
    const double* q;          /* numAtoms, e */
    const int* type;          /* numAtoms, index into the shared type table */
    const int* exclRanges;    /* numAtoms+1, or NULL */
    const int* exclElements;  /* exclusion partners (self included or not) */
    double box[3];            /* rectangular, every edge > rlistOuter */
} nbnxm_system_desc_t;

typedef struct nbnxm_ensemble_options
{
    int numTypes;             /* real types; a dummy type is appended internally */
    const double* c6c12;      /* 2*numTypes*numTypes: C6, C12 per type pair */
    int vdwType;              /* enum nbnxm_vdw_type; combination rules use the diagonal */
    double rcoulomb;          /* = rvdw */
    double ewaldRtol;
    double epsilonR;
    double rlistOuter;        /* list cut-off */
    double rlistInner;        /* dynamic-prune cut-off, rc <= rlistInner <= rlistOuter */
    int accumulation;         /* enum nbnxm_accumulation */
    int maxPairsPerEntry;     /* initial entry split (cluster pairs), 0 = 64 */
    int wavesPerBlock;        /* 1..4, 0 = 4 */
    int numHostThreads;       /* list construction threads, 0 = hardware concurrency */
    int gpuSearch;            /* nonzero: grids and lists built on the device (nbnxm_gpu_search_*) */
} nbnxm_ensemble_options_t;

#define NBNXM_MODE_BATCHED 0  /* one launch per kernel for all systems */
#define NBNXM_MODE_PLAIN 1    /* one launch per kernel per system, in turn */

/* Builds every system's grid and list on the host, uploads all data and
 * prunes the fresh lists (rlistOuter -> rlistInner) with the entry ordering. */
int nbnxm_ensemble_create(int numSystems, const nbnxm_system_desc_t* systems, const nbnxm_ensemble_options_t* opt, nbnxm_ensemble_t** out);
/* Re-splits the entries of the systems in each size bucket: bucketEdges
 * (numBuckets-1 ascending atom counts) and maxPairs (numBuckets, 0 = no split). */
int nbnxm_ensemble_set_split(nbnxm_ensemble_t* e, int numBuckets, const int* bucketEdges, const int* maxPairs, int wavesPerBlock);
/* Times candidate splits per size bucket and keeps the fastest, timing the
 * batched launch the caller will run (computeEnergy: F+E, else F); writes the
 * chosen maxPairs per bucket and waves per block when the pointers are set. */
int nbnxm_ensemble_autotune(nbnxm_ensemble_t* e, int numBuckets, const int* bucketEdges, int computeEnergy, int* chosenMaxPairs, int* chosenWavesPerBlock);
/* The i-cluster placement of the batched launch (as nbnxm_gpu_set_i_cluster_in_lds;
 * the autotuner sets it).  Per-system launches always choose by entry count. */
int nbnxm_ensemble_set_i_cluster_in_lds(nbnxm_ensemble_t* e, int mode);
/* The batched launch's waves per block and i-cluster placement setting,
 * resolved for its entry count (0 or 1; kernels without an LDS variant keep
 * the data in registers whatever the setting). */
int nbnxm_ensemble_get_launch_config(nbnxm_ensemble_t* e, int* wavesPerBlock, int* iClusterInLds);
/* Runs numSteps evaluations: x = x_search + N(0, sigma) per atom and step
 * (sigma 0: the search positions), nonbonded forces, FP64 energies and shift
 * forces.  useGraph: capture stepsPerGraph evaluations in a hipGraph and replay
 * it (batched mode only).  Asynchronous; nbnxm_ensemble_sync waits. */
int nbnxm_ensemble_run(nbnxm_ensemble_t* e, int mode, int numSteps, float sigma, uint64_t seed, int computeEnergy, int useGraph, int stepsPerGraph);
int nbnxm_ensemble_sync(nbnxm_ensemble_t* e);
/* Sets the step counter that keys the perturbation of the next evaluation. */
int nbnxm_ensemble_set_step(nbnxm_ensemble_t* e, uint64_t step);
/* Times numSteps evaluations between HIP events on the ensemble's stream;
 * writes total milliseconds, and the mean nonbonded-kernel duration in ms when
 * kernelMs is set (batched mode, measured in a separate pass of events around
 * each nonbonded launch). */
int nbnxm_ensemble_time(nbnxm_ensemble_t* e, int mode, int numSteps, float sigma, uint64_t seed, int computeEnergy, int useGraph, double* totalMs, double* kernelMs);
/* Forces of one system after the last evaluation, original atom order (3*numAtoms). */
int nbnxm_ensemble_get_forces(nbnxm_ensemble_t* e, int system, double* f);
/* Energies per system: 2*numSystems (E_lj, E_el); fshift: 3*45*numSystems (may be NULL). */
int nbnxm_ensemble_get_energies(nbnxm_ensemble_t* e, double* energies, double* fshift);
/* Coordinates used by the last evaluation, original order (3*numAtoms). */
int nbnxm_ensemble_get_positions(nbnxm_ensemble_t* e, int system, double* x);
/* Counts: [0] atoms, [1] grid atoms, [2] entries, [3] packed groups,
 * [4] cluster pairs in the list, [5] cluster pairs after the inner prune. */
int nbnxm_ensemble_stats(nbnxm_ensemble_t* e, int64_t* stats);
/* Pair list of one system as uploaded (before pruning), for checks. */
int nbnxm_ensemble_system_list(nbnxm_ensemble_t* e, int system, nbnxm_pairlist_gpu_t* list, int* atomIndices);
/* Useful pairs at the search positions: non-excluded atom pairs (every
 * periodic image) with r < rc, counted in FP64 over each system's list
 * (the pair counts of src/gromacs/nbnxm/pairlist.cpp:4143-4172 restricted to
 * the cut-off).  perSystem may be NULL. */
int nbnxm_ensemble_count_pairs(nbnxm_ensemble_t* e, double rc, int64_t* perSystem, int64_t* total);
void nbnxm_ensemble_free(nbnxm_ensemble_t* e);

/* ---------------------------------------------------------------------------
 * nblib-style force calculator: nblib::GmxNBForceCalculatorCpu
 * (api/nblib/include/nblib/gmxcalculatorcpu.h:67-105,
 * api/nblib/gmxcalculatorcpu.cpp:108-251) backed by the GPU kernel, the
 * counterpart the reference only sketches in createNbnxmGPU
 * (api/nblib/nbnxmsetuphelpers.cpp:313-357).
 * ------------------------------------------------------------------------- */

/* Combination-rule detection of nbnxn_atomdata_t::Params
 * (src/gromacs/nbnxm/atomdata.cpp:421-540, enbnxninitcombruleDETECT):
 * nbfp holds (6*C6, 12*C12) per type pair, numTypes^2 pairs; tol is the
 * relative tolerance (the reference's default is 1e-5).  Returns the
 * nbnxm_vdw_type (geometric preferred over Lorentz-Berthelot), or a negative
 * error code. */
int nbnxm_detect_comb_rule(int numTypes, const double* nbfp, double tol);

/* nblib::CoulombType (api/nblib/include/nblib/kerneloptions.h:73-79). */
enum nbnxm_coulomb_type
{
    NBNXM_COULOMB_PME = 0,           /* Ewald real space, analytical correction, no potential shift */
    NBNXM_COULOMB_CUTOFF = 1,        /* plain cut-off: k_rf = 0, c_rf = 1/rc */
    NBNXM_COULOMB_REACTION_FIELD = 2 /* reaction field with epsilon_rf = 1: the same constants */
};

/* Energy terms of the output, in NonBondedEnergyTerms order
 * (src/gromacs/mdtypes/enerdata.h:52-60): CoulombSR, LJSR, BuckinghamSR,
 * Coulomb14, LJ14 (one energy-group pair on the GPU). */
#define NBNXM_NUM_ENERGY_TERMS 5

/* nblib::NBKernelOptions (api/nblib/include/nblib/kerneloptions.h:84-102),
 * the fields that apply to the GPU kernel.  The nblib defaults are
 * pairlistCutoff 1.0, coulombType PME, analytical Ewald, Ewald rtol 1e-5,
 * epsilon_r 1; the interaction constants follow createInteractionConst
 * (api/nblib/nbnxmsetuphelpers.cpp:219-276): LJ potential shift, PME without
 * the Coulomb potential shift, cut-off and reaction field with
 * epsilon_rf = 1. */
typedef struct nbnxm_kernel_options
{
    int useGpu;                /* must be nonzero: the kernel runs on the GPU only */
    int numOpenMPThreads;      /* host threads for the search (0 = hardware) */
    double pairlistCutoff;     /* rc = rvdw = rcoulomb; also the list radius unless rlist > 0 */
    int coulombType;           /* enum nbnxm_coulomb_type */
    int useTabulatedEwaldCorr; /* PME: tabulated instead of analytical Ewald correction */
    double ewaldRtol;          /* 0 -> 1e-5 */
    double epsilonR;           /* 0 -> 1 */
    double rlist;              /* 0 -> pairlistCutoff (nblib); larger gives a buffered list */
    int accumulation;          /* enum nbnxm_accumulation */
} nbnxm_kernel_options_t;

typedef struct nbnxm_force_calculator nbnxm_force_calculator_t;

/* GmxNBForceCalculatorCpu(particleTypeIdOfAllParticles, nonBondedParams,
 * charges, particleInteractionFlags, exclusionRanges, exclusionElements,
 * options).  nonBondedParams: (6*C6, 12*C12) per type pair, numTypes^2
 * pairs (createNonBondedParameters, api/nblib/nbnxmsetuphelpers.cpp:153-177).
 * Exclusions as a list of lists: exclusionRanges[numParticles+1] offsets into
 * exclusionElements (either may be NULL for none).  One energy group: the
 * reference's particleInteractionFlags carry energy groups, which the GPU path
 * does not support (nbnxmsetuphelpers.cpp:328), so there is no such
 * argument.  The combination rule is detected as in nbnxm_detect_comb_rule. */
int nbnxm_force_calculator_create(int numParticles,
                                  const int* particleTypeIdOfAllParticles,
                                  int numTypes,
                                  const double* nonBondedParams,
                                  const double* charges,
                                  const int* exclusionRanges,
                                  const int* exclusionElements,
                                  const nbnxm_kernel_options_t* options,
                                  nbnxm_force_calculator_t** out);
/* updatePairlist(coordinates, box) (gmxcalculatorcpu.cpp:125-171): puts the
 * coordinates in the rectangular box in place (put_atoms_in_box_omp), builds
 * the grid and the GPU pair list and uploads them.  box: the three edges;
 * each edge must exceed the list radius. */
int nbnxm_force_calculator_update_pairlist(nbnxm_force_calculator_t* c, int numCoordinates, double* coordinates, const double box[3]);
/* compute(coordinates, box, forces[, virial[, energies]])
 * (gmxcalculatorcpu.cpp:173-251): forces (3*numParticles) are ADDED to, as
 * atomdata_add_nbat_f_to_f does; virial (9, or NULL) is overwritten with the
 * single-sum virial; energies (NBNXM_NUM_ENERGY_TERMS, or NULL) are
 * overwritten.  Fails with NBNXM_ERR_STATE before the first update_pairlist. */
int nbnxm_force_calculator_compute(nbnxm_force_calculator_t* c,
                                   int numCoordinates,
                                   const double* coordinates,
                                   const double box[3],
                                   double* forces,
                                   double* virial,
                                   double* energies);
void nbnxm_force_calculator_free(nbnxm_force_calculator_t* c);

/* ---------------------------------------------------------------------------
 * GPU pair search: the grid and the GPU cluster-pair list built on the device,
 * for one or many systems in one pass.  Replaces the host-side
 * putAtomsOnGrid / constructPairlist of the GPU list (src/gromacs/nbnxm/grid.cpp:1101-1240,
 * src/gromacs/nbnxm/pairlist.cpp:914-939,1122-1270,1948-2046,2150-2262) and
 * produces, for every system, exactly the list nbnxm_search_* builds.
 * ------------------------------------------------------------------------- */
typedef struct nbnxm_gpu_search nbnxm_gpu_search_t;
int nbnxm_gpu_search_create(nbnxm_gpu_search_t** out);
/* Grids and lists of numSystems systems (coordinates, boxes and exclusions as
 * in nbnxm_ensemble_create; q and type are not used).  Every box edge must
 * exceed rlist; maxPairsPerEntry > 0 splits entries as nbnxm_search does.
 * Fails with NBNXM_ERR_UNSUPPORTED when an entry exceeds 1024 j-clusters or
 * needs more than 32 exclusion masks. */
int nbnxm_gpu_search_build(nbnxm_gpu_search_t* s, int numSystems, const nbnxm_system_desc_t* systems, double rlist, int maxPairsPerEntry);
/* Totals over all systems: [0] grid slots, [1] entries, [2] packed groups,
 * [3] exclusion masks (including the shared all-ones mask 0). */
int nbnxm_gpu_search_num_entries(const nbnxm_gpu_search_t* s, int64_t* counts);
/* One system's list in single-system numbering (cells, groups and masks from
 * 0, its own all-ones mask 0) and its slot -> atom map.  sizes[4] receives the
 * grid slots, entries, groups and masks; any output pointer may be NULL to
 * query the sizes first. */
int nbnxm_gpu_search_system_list(const nbnxm_gpu_search_t* s, int system, nbnxm_sci_t* sci, nbnxm_cj_packed_t* cjPacked, nbnxm_excl_t* excl, int* atomIndices, int* sizes);
void nbnxm_gpu_search_free(nbnxm_gpu_search_t* s);

/* ---------------------------------------------------------------------------
 * Nonbonded part of a device-resident MD step: the pair-search / list /
 * kernel / force-reduction sequence of nonbonded_verlet_t as do_force drives
 * it (src/gromacs/mdlib/sim_util.cpp:1619-1680, 925-946), with the
 * reference's list lifetime and dynamic-pruning schedule
 * (src/gromacs/nbnxm/pairlistsets.h:91-113, pairlist_tuning.cpp:505-600,
 * 690-708, prunekerneldispatch.cpp:99-114) on one GPU, one domain.
 * ------------------------------------------------------------------------- */
typedef struct nbnxm_verlet nbnxm_verlet_t;

/* PairlistParams of one domain (src/gromacs/nbnxm/pairlistparams.h:117-160). */
typedef struct nbnxm_verlet_params
{
    int nstlist;         /* steps per outer list (inputrec nstlist); lifetime = nstlist - 1 */
    double rlistOuter;   /* outer list radius, >= max(rcoulomb, rvdw) */
    double rlistInner;   /* pruned radius in [cut-off, rlistOuter]; rlistOuter = no dynamic pruning */
    int nstlistPrune;    /* steps between prunes of the whole list (even, >= 2) */
    int accumulation;    /* enum nbnxm_accumulation */
    int gpuSearch;       /* nonzero: grid and list built by the device search (nbnxm_gpu_search_*);
                          * searches after the first keep coordinates and list on the device */
} nbnxm_verlet_params_t;

/* Sets up the system: per-atom types (numTypes types, nbfp (6C6, 12C12)
 * per type pair as nblib's createNonBondedParameters), charges, exclusions as
 * list of lists (exclRanges[numAtoms+1] offsets into exclElements, both may
 * be NULL), the interaction constants (any flavour) and the list parameters.
 * Dynamic pruning is used when rlistInner < rlistOuter and nstlistPrune <
 * nstlist - 1, with nstlistPrune / 2 rolling parts (pairlist_tuning.cpp:586-591,
 * 690-708). */
int nbnxm_verlet_create(int numAtoms,
                        const int* types,
                        int numTypes,
                        const double* nbfp,
                        const double* charges,
                        const int* exclRanges,
                        const int* exclElements,
                        const nbnxm_interaction_const_t* ic,
                        const nbnxm_verlet_params_t* params,
                        nbnxm_verlet_t** out);
/* The nonbonded work of step `step` on device arrays in atom order:
 *  - on the first call, every nstlist steps and when the box changes: copies
 *    d_x to the host, puts the atoms in the box IN PLACE (as do_md does on
 *    search steps, d_x is rewritten with the wrapped coordinates), builds the
 *    grid and the outer list (rlistOuter) and uploads it; the first kernel
 *    of a fresh list prunes it to rlistInner;
 *  - every step: d_x -> xq (nbnxn_gpu_x_to_nbat_x), the nonbonded kernel
 *    (energies / shift forces and virial when asked), the rolling prune of
 *    one part when isDynamicPruningStepGpu holds (age > 0, age < lifetime,
 *    age even), and d_f (+)= nonbonded forces (accumulateForces selects add or
 *    set).
 * waitStream (a hipStream_t, may be NULL): the step's work starts after the
 * work already queued on it (the coordinates' producer).  Returns after the
 * step's device work has finished. */
int nbnxm_verlet_step(nbnxm_verlet_t* v,
                      long long step,
                      const double box[3],
                      float* d_x,
                      void* waitStream,
                      int computeEnergy,
                      int computeVirial,
                      float* d_f,
                      int accumulateForces);
/* Energies (LJ, Coulomb) and the virial (9, row-major, -1/2 sum x (x) f) of
 * the last step that computed them. */
int nbnxm_verlet_get_energies(const nbnxm_verlet_t* v, double* eLJ, double* eCoul);
int nbnxm_verlet_get_virial(const nbnxm_verlet_t* v, double* virial);
/* counts[0] searches, [1] rolling prune launches, [2] steps, [3] outer-list
 * cluster pairs, [4] step of the last search; flags[0] dynamic pruning used,
 * [1] rolling parts. */
int nbnxm_verlet_get_stats(const nbnxm_verlet_t* v, long long* counts, int* flags);
/* The current list with its inner (pruned) masks, its exclusion masks and
 * the slot -> atom map, for inspection: sizes[4] receives the grid slots,
 * entries, packed groups and exclusion masks; any output may be NULL. */
int nbnxm_verlet_get_list(nbnxm_verlet_t* v, nbnxm_sci_t* sci, nbnxm_cj_packed_t* cjPacked, nbnxm_excl_t* excl, int* atomIndices, int* sizes);
void nbnxm_verlet_free(nbnxm_verlet_t* v);

#ifdef __cplusplus
}
#endif

#endif /* NBNXM_AMD_H */
