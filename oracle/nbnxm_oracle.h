/*
 * nbnxm_oracle.h — CPU oracle for the nbnxm cluster-pair kernel.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; tests, the smoke check and the CPU-baseline leg of the benchmark
 * use it as the checker.  Every function restates a reference function and
 * cites the file:line it follows (paths relative to the GROMACS-FMM tree).
 *
 * Pinning status (what fixes each function independently of its own formula):
 *   oracle_calc_ewaldcoeff_q   closed form erfc(beta*rc) == rtol; beta(1.0,1e-5)
 *   oracle_ewald_table         interpolated table against the exact derivative
 *                              -d/dr erf(beta r)/r (FP32 spline error bound); the
 *                              12 Ewald golden cases with the tabulated kernel
 *   oracle_pmecorrF            closed form 2exp(-z^2)/(sqrt(pi)z^2) - erf(z)/z^3, |err|<1e-6
 *   oracle_kernel_f64          refdata XML golden forces/energies (243-atom box) of
 *                              all 18 Coulomb x LJ flavours, brute force all-image
 *                              sum per flavour, Newton's third law
 *   oracle_kernel_f32          same golden vectors at the reference FP32 tolerance;
 *                              agreement with oracle_kernel_f64 per flavour
 *   oracle_bruteforce_f64      exact pair sums on 2-3 atom systems (hand values):
 *                              Ewald, reaction field (+ excluded pair), LJ-PME as
 *                              the real-space kernel -C6 g(beta r)/r^6, force and
 *                              potential switch continuity and dV/dr = -F;
 *                              translation invariance; all 18 golden XML cases to
 *                              1e-5 (forces) and 1e-8 (LJ energy)
 *   oracle_prune               double-precision any-pair-within-r predicate away
 *                              from the threshold; masks subset of input masks
 *   oracle_sci_sort            counts non-increasing, permutation of input
 *   oracle_bruteforce_virial_f64  hand value for two atoms; equality with the
 *                              single-sum virial from oracle_kernel_f64 forces
 *                              and shift forces (an independent formulation)
 */
#ifndef NBNXM_ORACLE_H
#define NBNXM_ORACLE_H

#include "../include/nbnxm_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

double oracle_calc_ewaldcoeff_q(double rc, double rtol);

/* Ewald correction force table of the tabulated kernel (forcetable.cpp:139-365,
 * forcerec.cpp:586-624): returns the point count, fills tableF / tableFD when
 * given (capacity entries), *scale = the FP32 table scale. */
int oracle_ewald_table(double beta, double rc, float* tableF, double* tableFD, int capacity, double* scale);
float oracle_pmecorrF(float z2);

/* Every kernel flavour follows ic: Ewald or reaction field (plain cut-off is
 * reaction field with k_rf = 0); LJ with potential shift, force switch or
 * potential switch; optional LJ-PME grid correction with the geometric grid
 * rule (needs nbat->type and the nbfp table); rvdw < rcoulomb adds the LJ
 * cut-off check of the twin-cut-off kernels.
 *
 * Restatement of the CUDA kernels nbnxn_kernel_Elec*_Vdw*_{F,VF}_cuda in FP32
 * (src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel.cuh:145-741), loop structure of
 * nbnxn_kernel_gpu_ref (src/gromacs/nbnxm/kernels_reference/kernel_gpu_ref.cpp:57-354).
 * Forces are accumulated (f += ...) in grid order; fshift[45*3] accumulated. */
int oracle_kernel_f32(const nbnxm_pairlist_gpu_t* list,
                      const nbnxm_atomdata_t* nbat,
                      const nbnxm_interaction_const_t* ic,
                      int computeEnergy,
                      float* f,
                      float* fshift,
                      double* Vvdw,
                      double* Vc);

/* Same semantics in FP64 with exact erfc/exp (the FP64 oracle). */
int oracle_kernel_f64(const nbnxm_pairlist_gpu_t* list,
                      const nbnxm_atomdata_t* nbat,
                      const nbnxm_interaction_const_t* ic,
                      int computeEnergy,
                      double* f,
                      double* fshift,
                      double* Vvdw,
                      double* Vc);

/* All-image O(N^2) FP64 sum in original atom order, independent of any list.
 * Box must be rectangular with every edge > rc.  Excluded pairs get the Ewald
 * or reaction-field exclusion correction (and the LJ-PME grid correction, with
 * grid C6 = sqrt(nbfp_ii nbfp_jj)) in every image within rc (the list semantics of
 * src/gromacs/nbnxm/pairlist.cpp:1948-2046), self energy included
 * (src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel.cuh:381-418). */
int oracle_bruteforce_f64(int n,
                          const double* x,
                          const double* q,
                          const int* type,
                          int numTypes,
                          const double* nbfp, /* 6C6,12C12 per type pair */
                          const int* exclRanges,
                          const int* exclElements,
                          const double box[3],
                          const nbnxm_interaction_const_t* ic,
                          double* f,
                          double* Vvdw,
                          double* Vc,
                          long long* numPairsInRange);

/* Same all-image sum, accumulating forces and the pair virial
 * -1/2 sum_pairs r_ij (x) F_ij (9, row-major; added to virial), the quantity
 * the single-sum virial of src/gromacs/mdlib/calcvir.cpp:89-133 equals. */
int oracle_bruteforce_virial_f64(int n,
                                 const double* x,
                                 const double* q,
                                 const int* type,
                                 int numTypes,
                                 const double* nbfp,
                                 const int* exclRanges,
                                 const int* exclElements,
                                 const double box[3],
                                 const nbnxm_interaction_const_t* ic,
                                 double* f,
                                 double* virial);

/* Prune restatement (src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel_pruneonly.cuh:104-343)
 * with split 1.  imaskOuter / imaskInner are per packed group.  On a fresh
 * list imaskInner must hold the list's masks on entry; both are written.
 * sciCount (nsci entries, may be NULL) gets the histogram bin of each sci
 * processed on a fresh list. */
int oracle_prune(const nbnxm_pairlist_gpu_t* list,
                 const float* xq,
                 const float* shiftVec,
                 float rlistOuter_sq,
                 float rlistInner_sq,
                 int haveFreshList,
                 int part,
                 int numParts,
                 unsigned int* imaskOuter,
                 unsigned int* imaskInner,
                 int* sciCount);

/* Counting sort of sci entries by bin index (ascending bin = descending pair
 * count), stable within a bin (nbnxnKernelBucketSciSort,
 * src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel_sci_sort.cuh:66-87). */
int oracle_sci_sort(int nsci, const nbnxm_sci_t* sci, const int* sciCount, nbnxm_sci_t* sorted);

/* Useful pairs: non-excluded pairs i<j (all images) with r < rc, FP64. */
long long oracle_count_pairs_f64(const nbnxm_pairlist_gpu_t* list,
                                 const nbnxm_atomdata_t* nbat,
                                 double rc);

#ifdef __cplusplus
}
#endif

#endif
