/*
 * nbnxm_oracle.c — CPU oracle for the nbnxm cluster-pair kernel.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests, the smoke check and the
 * benchmark's CPU-baseline leg.  The product library never links this file.
 * Built with -ffp-contract=off so that every fused multiply-add below is the
 * explicit fmaf()/fma() written in the source and nothing else.
 *
 * Citations are relative to the GROMACS-FMM tree.
 */
#include "nbnxm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CL NBNXM_CLUSTER_SIZE
#define NCPS NBNXM_NUM_CLUSTER_PER_SUPERCLUSTER
#define JG NBNXM_JGROUP_SIZE

/* c_one4PiEps0 is not needed here: epsfac arrives in the interaction constants. */
static const double c_oneSqrtPi = 0.56418958354775628695; /* 1/sqrt(pi) */
static const double c_pi = 3.14159265358979323846;

/* Restates calc_ewaldcoeff_q (src/gromacs/ewald/ewald_utils.cpp:43-71):
 * double beta until erfc(beta*rc) <= rtol, then bisect for i+60 steps. */
double oracle_calc_ewaldcoeff_q(double rc, double rtol)
{
    double beta = 5.0;
    int i = 0;
    do
    {
        i++;
        beta *= 2.0;
    } while (erfc(beta * rc) > rtol);

    const int n = i + 60;
    double low = 0.0;
    double high = beta;
    for (int k = 0; k < n; k++)
    {
        beta = 0.5 * (low + high);
        if (erfc(beta * rc) > rtol)
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

/* Restates pmecorrF (src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel_utils.cuh:407-443),
 * the single-precision rational of pmeForceCorrection
 * (src/gromacs/simd/include/gromacs/simd/simd_math.h:1611-1648), in the same
 * evaluation order with every a*b+c written as an explicit fused multiply-add.
 * The division is IEEE 1/x here; the HIP kernel uses the hardware reciprocal. */
float oracle_pmecorrF(float z2)
{
    const float FN6 = -1.7357322914161492954e-8F;
    const float FN5 = 1.4703624142580877519e-6F;
    const float FN4 = -0.000053401640219807709149F;
    const float FN3 = 0.0010054721316683106153F;
    const float FN2 = -0.019278317264888380590F;
    const float FN1 = 0.069670166153766424023F;
    const float FN0 = -0.75225204789749321333F;

    const float FD4 = 0.0011193462567257629232F;
    const float FD3 = 0.014866955030185295499F;
    const float FD2 = 0.11583842382862377919F;
    const float FD1 = 0.50736591960530292870F;
    const float FD0 = 1.0F;

    const float z4 = z2 * z2;

    float polyFD0 = fmaf(FD4, z4, FD2);
    float polyFD1 = fmaf(FD3, z4, FD1);
    polyFD0 = fmaf(polyFD0, z4, FD0);
    polyFD0 = fmaf(polyFD1, z2, polyFD0);
    polyFD0 = 1.0F / polyFD0;

    float polyFN0 = fmaf(FN6, z4, FN4);
    float polyFN1 = fmaf(FN5, z4, FN3);
    polyFN0 = fmaf(polyFN0, z4, FN2);
    polyFN1 = fmaf(polyFN1, z4, FN1);
    polyFN0 = fmaf(polyFN0, z4, FN0);
    polyFN0 = fmaf(polyFN1, z2, polyFN0);

    return polyFN0 * polyFD0;
}

/* Restates ewald_spline3_table_scale for Coulomb (src/gromacs/tables/forcetable.cpp:296-340
 * with spline3_table_scale :280-294, mixed-precision FP32 for beta, the
 * erfc argument and the result) and generateEwaldCorrectionTables (:139-276,
 * the F column) for v(r) = erf(beta r)/r, the point count of
 * init_ewald_f_table (src/gromacs/mdlib/forcerec.cpp:586-624).  Returns the
 * number of points; tables may be NULL to query it. */
int oracle_ewald_table(double beta, double rc, float* tableF, double* tableFD, int capacity, double* scaleOut)
{
    const float betaF = (float)beta;
    const double derivTol = (double)1.19209290e-07F;
    const double xd3 = 1.0522;
    const float erfcRc = erfcf(betaF * (float)rc);
    double funcTol = 0.1 * erfcRc;
    if (funcTol < derivTol)
    {
        funcTol = derivTol;
    }
    const double scDeriv = sqrt(xd3 / (6 * 4 * derivTol * betaF)) * betaF;
    const double scFunc = cbrt(xd3 / (6 * 12 * sqrt(3.0) * funcTol)) * betaF;
    const float scale = (float)(scDeriv > scFunc ? scDeriv : scFunc);
    const int n = (int)((float)rc * scale) + 2;
    if (scaleOut)
    {
        *scaleOut = scale;
    }
    if (!tableF && !tableFD)
    {
        return n;
    }
    if (capacity < n)
    {
        return NBNXM_ERR_INVALID;
    }
    const double dx = 1.0 / (double)scale;
    double* f = (double*)calloc((size_t)n, sizeof(double));
    if (!f)
    {
        return NBNXM_ERR_INVALID;
    }
    double dc = 0.0;
    for (int i = n - 1; i >= 1; i--)
    {
        const double x0 = i * dx;
        const double x1 = (i - 1) * dx;
        const double xm = x0 - 0.5 * dx;
        const double v0 = erf(betaF * x0) / x0;
        const double v1 = (i - 1 == 0) ? betaF * 2 / sqrt(c_pi) : erf(betaF * x1) / x1;
        const double vm = erf(betaF * xm) / xm;
        /* mean second derivative times dx over [x1, x0]; spline slope at x0 */
        double a2dx = (v0 + v1 - 2 * vm) / (0.25 * dx);
        dc = (v0 - v1) / dx + 0.5 * a2dx;
        if (i == n - 1)
        {
            f[i] = -dc;
        }
        else
        {
            f[i] += -0.5 * dc;
        }
        /* spline a0 + a1 (x - x0) + a2/2 (x - x0)^2 through v0, v1 with slope dc
         * at x0; its slope at x1 starts the next interval */
        a2dx = (dc * dx + v1 - v0) * 2 / dx;
        dc = dc - a2dx;
        f[i - 1] = -0.5 * dc;
    }
    f[0] *= 2;
    for (int i = 0; i < n; i++)
    {
        if (tableF)
        {
            tableF[i] = (float)f[i];
        }
        if (tableFD)
        {
            tableFD[i] = f[i];
        }
    }
    free(f);
    return n;
}

/* The tables of one evaluation, with one padding entry repeating the end point. */
typedef struct
{
    int n;
    double scale;
    float* f32;
    double* f64;
} ewald_table_t;

static int ewald_table_make(const nbnxm_interaction_const_t* ic, ewald_table_t* t)
{
    memset(t, 0, sizeof(*t));
    if (!ic->tabulatedEwald)
    {
        return NBNXM_OK;
    }
    const int n = oracle_ewald_table(ic->ewaldcoeff_q, ic->rcoulomb, NULL, NULL, 0, &t->scale);
    t->f32 = (float*)malloc(sizeof(float) * (size_t)(n + 1));
    t->f64 = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    if (!t->f32 || !t->f64 || oracle_ewald_table(ic->ewaldcoeff_q, ic->rcoulomb, t->f32, t->f64, n, &t->scale) != n)
    {
        free(t->f32);
        free(t->f64);
        return NBNXM_ERR_INVALID;
    }
    t->f32[n] = t->f32[n - 1];
    t->f64[n] = t->f64[n - 1];
    t->n = n;
    return NBNXM_OK;
}

static void ewald_table_free(ewald_table_t* t)
{
    free(t->f32);
    free(t->f64);
    memset(t, 0, sizeof(*t));
}

static int check_inputs(const nbnxm_pairlist_gpu_t* list, const nbnxm_atomdata_t* nbat)
{
    if (!list || !nbat || !nbat->xq || !nbat->shiftVec)
    {
        return NBNXM_ERR_INVALID;
    }
    if (list->na_ci != CL || (list->nsci > 0 && (!list->sci || !list->cjPacked || !list->excl)))
    {
        return NBNXM_ERR_INVALID;
    }
    if (nbat->vdwType == NBNXM_VDW_CUT_TYPE_TABLE && (!nbat->type || !nbat->nbfp))
    {
        return NBNXM_ERR_INVALID;
    }
    if (nbat->vdwType != NBNXM_VDW_CUT_TYPE_TABLE && !nbat->ljComb)
    {
        return NBNXM_ERR_INVALID;
    }
    return NBNXM_OK;
}

/* ------------------------------------------------------------------------
 * Pair interaction of every kernel flavour, in the order of the CUDA kernel
 * body (src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel.cuh:514-665): clamp r^2,
 * LJ with the PotShift constants, force switch (with the unmasked c6/c12 of
 * :564-566, as the GPU kernels do), LJ-PME grid correction (:570-596),
 * potential switch (:599-605), the separate LJ cut-off check of twin cut-offs
 * (:607-617), then Coulomb: Ewald (:638-641, :655-660) or reaction field
 * (:633-635, :651-654; plain cut-off is reaction field with k_rf = 0).
 * ---------------------------------------------------------------------- */

/* The FP32 parameters as the CUDA kernel holds them (NBParamGpu floats,
 * src/gromacs/nbnxm/nbnxm_gpu_data_mgmt.cpp:93-159) and derives them
 * (nbnxm_cuda_kernel.cuh:311-376). */
typedef struct
{
    int eeltype, vdwModifier, ljPme, vdwCutoffCheck;
    float rcoulomb_sq, rvdw_sq, epsfac, beta, beta2, beta3, sh_ewald, two_k_rf, c_rf;
    float cpot_rep, cpot_disp, rvdw_switch;
    float disp_c2, disp_c3, rep_c2, rep_c3, sw_c3, sw_c4, sw_c5;
    float lje_coeff2, lje_coeff6_6, sh_lj_ewald;
    int tabulated;
    float tabScale;
    const float* tab;
} flav32_t;

static flav32_t flav32_from_ic(const nbnxm_interaction_const_t* ic, const ewald_table_t* tab)
{
    flav32_t p;
    memset(&p, 0, sizeof(p));
    p.tabulated = ic->tabulatedEwald;
    p.tabScale = (float)tab->scale;
    p.tab = tab->f32;
    p.eeltype = ic->eeltype;
    p.vdwModifier = ic->vdwModifier;
    p.ljPme = ic->ljPme;
    p.vdwCutoffCheck = ic->rvdw < ic->rcoulomb;
    const float rc = (float)ic->rcoulomb;
    const float rvdw = (float)ic->rvdw;
    p.rcoulomb_sq = rc * rc;
    p.rvdw_sq = rvdw * rvdw;
    p.epsfac = (float)ic->epsfac;
    p.beta = (float)ic->ewaldcoeff_q;
    p.beta2 = p.beta * p.beta;
    p.beta3 = p.beta * p.beta * p.beta;
    p.sh_ewald = (float)ic->sh_ewald;
    p.two_k_rf = (float)(2.0 * ic->k_rf);
    p.c_rf = (float)ic->c_rf;
    p.cpot_rep = (float)ic->repulsion_shift_cpot;
    p.cpot_disp = (float)ic->dispersion_shift_cpot;
    p.rvdw_switch = (float)ic->rvdw_switch;
    p.disp_c2 = (float)ic->dispersion_shift_c2;
    p.disp_c3 = (float)ic->dispersion_shift_c3;
    p.rep_c2 = (float)ic->repulsion_shift_c2;
    p.rep_c3 = (float)ic->repulsion_shift_c3;
    p.sw_c3 = (float)ic->vdw_switch_c3;
    p.sw_c4 = (float)ic->vdw_switch_c4;
    p.sw_c5 = (float)ic->vdw_switch_c5;
    const float ljc = (float)ic->ewaldcoeff_lj;
    p.lje_coeff2 = ljc * ljc;
    p.lje_coeff6_6 = p.lje_coeff2 * p.lje_coeff2 * p.lje_coeff2 * (1.0F / 6.0F);
    p.sh_lj_ewald = (float)ic->sh_lj_ewald;
    return p;
}

/* One pair in FP32.  qq = qi*epsfac*qj.  lbForceOnly selects the
 * sigma/epsilon form of :555-562 (LB combination, no energies, no switch).
 * Returns F/r; adds the pair's energies to *eLJ / *eEl when computeEnergy. */
static float pair_f32(const flav32_t* p,
                      int computeEnergy,
                      float r2,
                      float int_bit,
                      float qq,
                      float c6,
                      float c12,
                      int lbForceOnly,
                      float sigma,
                      float epsilon,
                      float c6grid,
                      float* eLJ,
                      float* eEl)
{
    const float oneSixth = 1.0F / 6.0F;
    const float oneTwelveth = 1.0F / 12.0F;
    r2 = fmaxf(r2, NBNXM_MIN_DISTANCE_SQUARED);
    const float inv_r = 1.0F / sqrtf(r2);
    const float inv_r2 = inv_r * inv_r;
    float F_invr;
    float E_lj_p = 0.0F;
    float inv_r6 = inv_r2 * inv_r2 * inv_r2;
    inv_r6 *= int_bit;
    if (lbForceOnly)
    {
        const float sig_r = sigma * inv_r;
        const float sig_r2 = sig_r * sig_r;
        float sig_r6 = sig_r2 * sig_r2 * sig_r2;
        sig_r6 *= int_bit;
        F_invr = epsilon * sig_r6 * (sig_r6 - 1.0F) * inv_r2;
    }
    else
    {
        F_invr = inv_r6 * (c12 * inv_r6 - c6) * inv_r2;
    }
    if (computeEnergy || p->vdwModifier == NBNXM_VDW_MOD_POTSWITCH)
    {
        E_lj_p = int_bit * (c12 * (inv_r6 * inv_r6 + p->cpot_rep) * oneTwelveth - c6 * (inv_r6 + p->cpot_disp) * oneSixth);
    }
    if (p->vdwModifier == NBNXM_VDW_MOD_FORCESWITCH)
    {
        /* calculate_force_switch_F[_E] (nbnxm_cuda_kernel_utils.cuh:89-145) */
        const float r = r2 * inv_r;
        float r_switch = r - p->rvdw_switch;
        r_switch = r_switch >= 0.0F ? r_switch : 0.0F;
        F_invr += -c6 * (p->disp_c2 + p->disp_c3 * r_switch) * r_switch * r_switch * inv_r
                  + c12 * (p->rep_c2 + p->rep_c3 * r_switch) * r_switch * r_switch * inv_r;
        if (computeEnergy)
        {
            const float disp_F2 = p->disp_c2 / 3;
            const float disp_F3 = p->disp_c3 / 4;
            const float repu_F2 = p->rep_c2 / 3;
            const float repu_F3 = p->rep_c3 / 4;
            E_lj_p += c6 * (disp_F2 + disp_F3 * r_switch) * r_switch * r_switch * r_switch
                      - c12 * (repu_F2 + repu_F3 * r_switch) * r_switch * r_switch * r_switch;
        }
    }
    if (p->ljPme)
    {
        /* calculate_lj_ewald_comb_geom_F[_E] (nbnxm_cuda_kernel_utils.cuh:233-288) and
         * calculate_lj_ewald_comb_LB_F_E (:290-337): the same terms, c6grid by the rule */
        const float inv_r6_nm = inv_r2 * inv_r2 * inv_r2;
        const float cr2 = p->lje_coeff2 * r2;
        const float expmcr2 = expf(-cr2);
        const float poly = 1.0F + cr2 + 0.5F * cr2 * cr2;
        F_invr += c6grid * (inv_r6_nm - expmcr2 * (inv_r6_nm * poly + p->lje_coeff6_6)) * inv_r2;
        if (computeEnergy)
        {
            const float sh_mask = p->sh_lj_ewald * int_bit;
            E_lj_p += oneSixth * c6grid * (inv_r6_nm * (1.0F - expmcr2 * poly) + sh_mask);
        }
    }
    if (p->vdwModifier == NBNXM_VDW_MOD_POTSWITCH)
    {
        /* calculate_potential_switch_F[_E] (nbnxm_cuda_kernel_utils.cuh:147-209) */
        const float switch_F2 = 3 * p->sw_c3;
        const float switch_F3 = 4 * p->sw_c4;
        const float switch_F4 = 5 * p->sw_c5;
        const float r = r2 * inv_r;
        float r_switch = r - p->rvdw_switch;
        if (computeEnergy)
        {
            r_switch = r_switch >= 0.0F ? r_switch : 0.0F;
        }
        if (computeEnergy || r_switch > 0.0F)
        {
            const float sw = 1.0F + (p->sw_c3 + (p->sw_c4 + p->sw_c5 * r_switch) * r_switch) * r_switch * r_switch * r_switch;
            const float dsw = (switch_F2 + (switch_F3 + switch_F4 * r_switch) * r_switch) * r_switch * r_switch;
            F_invr = F_invr * sw - inv_r * E_lj_p * dsw;
            if (computeEnergy)
            {
                E_lj_p *= sw;
            }
        }
    }
    if (p->vdwCutoffCheck)
    {
        const float vdw_in_range = (r2 < p->rvdw_sq) ? 1.0F : 0.0F;
        F_invr *= vdw_in_range;
        E_lj_p *= vdw_in_range;
    }
    if (computeEnergy)
    {
        *eLJ += E_lj_p;
    }
    if (p->eeltype == NBNXM_ELEC_RF)
    {
        F_invr += qq * (int_bit * inv_r2 * inv_r - p->two_k_rf);
        if (computeEnergy)
        {
            *eEl += qq * (int_bit * inv_r + 0.5F * p->two_k_rf * r2 - p->c_rf);
        }
    }
    else
    {
        if (p->tabulated)
        {
            /* interpolate_coulomb_force_r (nbnxm_cuda_kernel_utils.cuh:345-385, :640-645) */
            const float normalized = p->tabScale * (r2 * inv_r);
            const int index = (int)normalized;
            const float fraction = normalized - (float)index;
            const float corr = fmaf(fraction, p->tab[index + 1], fmaf(-fraction, p->tab[index], p->tab[index]));
            F_invr += qq * (int_bit * inv_r2 - corr) * inv_r;
        }
        else
        {
            F_invr += qq * (int_bit * inv_r2 * inv_r + oracle_pmecorrF(p->beta2 * r2) * p->beta3);
        }
        if (computeEnergy)
        {
            *eEl += qq * (inv_r * (int_bit - erff(r2 * inv_r * p->beta)) - int_bit * p->sh_ewald);
        }
    }
    return F_invr;
}

/* The same pair in FP64 with exact erf/exp: the Ewald force correction is
 * 2 beta/sqrt(pi) exp(-beta^2 r^2)/r^2 - erf(beta r)/r^3, the exact function
 * pmecorrF approximates. */
static double pair_f64(const nbnxm_interaction_const_t* ic,
                       const ewald_table_t* tab,
                       double r2,
                       double int_bit,
                       double qq,
                       double c6,
                       double c12,
                       double c6grid,
                       double* eLJ,
                       double* eEl)
{
    if (r2 < NBNXM_MIN_DISTANCE_SQUARED)
    {
        r2 = NBNXM_MIN_DISTANCE_SQUARED;
    }
    const double r = sqrt(r2);
    const double inv_r = 1.0 / r;
    const double inv_r2 = inv_r * inv_r;
    const double inv_r6 = int_bit * inv_r2 * inv_r2 * inv_r2;
    double F_invr = inv_r6 * (c12 * inv_r6 - c6) * inv_r2;
    double E_lj_p = int_bit
                    * (c12 * (inv_r6 * inv_r6 + ic->repulsion_shift_cpot) / 12.0
                       - c6 * (inv_r6 + ic->dispersion_shift_cpot) / 6.0);
    if (ic->vdwModifier == NBNXM_VDW_MOD_FORCESWITCH)
    {
        double rs = r - ic->rvdw_switch;
        rs = rs > 0.0 ? rs : 0.0;
        F_invr += (-c6 * (ic->dispersion_shift_c2 + ic->dispersion_shift_c3 * rs)
                   + c12 * (ic->repulsion_shift_c2 + ic->repulsion_shift_c3 * rs))
                  * rs * rs * inv_r;
        E_lj_p += (c6 * (ic->dispersion_shift_c2 / 3.0 + ic->dispersion_shift_c3 / 4.0 * rs)
                   - c12 * (ic->repulsion_shift_c2 / 3.0 + ic->repulsion_shift_c3 / 4.0 * rs))
                  * rs * rs * rs;
    }
    if (ic->ljPme)
    {
        const double b2 = ic->ewaldcoeff_lj * ic->ewaldcoeff_lj;
        const double b6_6 = b2 * b2 * b2 / 6.0;
        const double inv_r6_nm = inv_r2 * inv_r2 * inv_r2;
        const double cr2 = b2 * r2;
        const double e = exp(-cr2);
        const double poly = 1.0 + cr2 + 0.5 * cr2 * cr2;
        F_invr += c6grid * (inv_r6_nm - e * (inv_r6_nm * poly + b6_6)) * inv_r2;
        E_lj_p += c6grid / 6.0 * (inv_r6_nm * (1.0 - e * poly) + ic->sh_lj_ewald * int_bit);
    }
    if (ic->vdwModifier == NBNXM_VDW_MOD_POTSWITCH)
    {
        double rs = r - ic->rvdw_switch;
        rs = rs > 0.0 ? rs : 0.0;
        const double sw = 1.0 + (ic->vdw_switch_c3 + (ic->vdw_switch_c4 + ic->vdw_switch_c5 * rs) * rs) * rs * rs * rs;
        const double dsw = (3.0 * ic->vdw_switch_c3 + (4.0 * ic->vdw_switch_c4 + 5.0 * ic->vdw_switch_c5 * rs) * rs) * rs * rs;
        F_invr = F_invr * sw - inv_r * E_lj_p * dsw;
        E_lj_p *= sw;
    }
    if (ic->rvdw < ic->rcoulomb && !(r2 < ic->rvdw * ic->rvdw))
    {
        F_invr = 0.0;
        E_lj_p = 0.0;
    }
    *eLJ += E_lj_p;
    if (ic->eeltype == NBNXM_ELEC_RF)
    {
        F_invr += qq * (int_bit * inv_r2 * inv_r - 2.0 * ic->k_rf);
        *eEl += qq * (int_bit * inv_r + ic->k_rf * r2 - ic->c_rf);
    }
    else
    {
        const double br = ic->ewaldcoeff_q * r;
        const double erfbr = erf(br);
        if (ic->tabulatedEwald)
        {
            const double normalized = tab->scale * r;
            const int index = (int)normalized;
            const double fraction = normalized - index;
            const double corr = fma(fraction, tab->f64[index + 1], fma(-fraction, tab->f64[index], tab->f64[index]));
            F_invr += qq * (int_bit * inv_r2 - corr) * inv_r;
        }
        else
        {
            F_invr += qq * (int_bit * inv_r2 * inv_r + 2.0 * ic->ewaldcoeff_q * c_oneSqrtPi * exp(-br * br) * inv_r2 - erfbr * inv_r2 * inv_r);
        }
        *eEl += qq * (inv_r * (int_bit - erfbr) - int_bit * ic->sh_ewald);
    }
    return F_invr;
}

/* Self energies on the diagonal entry (nbnxm_cuda_kernel.cuh:381-418):
 * Coulomb -epsfac sum q^2 times beta/sqrt(pi) (Ewald) or c_rf/2 (reaction
 * field); LJ-PME + sum 6C6_ii/12 * beta_lj^6/6. */
static double self_coulomb_factor(const nbnxm_interaction_const_t* ic)
{
    return ic->eeltype == NBNXM_ELEC_RF ? 0.5 * ic->c_rf : ic->ewaldcoeff_q * c_oneSqrtPi;
}

static int check_flavour(const nbnxm_atomdata_t* nbat, const nbnxm_interaction_const_t* ic)
{
    if (ic->eeltype != NBNXM_ELEC_EWALD && ic->eeltype != NBNXM_ELEC_RF)
    {
        return NBNXM_ERR_INVALID;
    }
    if (ic->vdwModifier < NBNXM_VDW_MOD_POTSHIFT || ic->vdwModifier > NBNXM_VDW_MOD_POTSWITCH)
    {
        return NBNXM_ERR_INVALID;
    }
    if (ic->tabulatedEwald && (ic->eeltype != NBNXM_ELEC_EWALD || !(ic->ewaldcoeff_q > 0)))
    {
        return NBNXM_ERR_INVALID;
    }
    if (ic->ljPme < NBNXM_LJPME_NONE || ic->ljPme > NBNXM_LJPME_LB)
    {
        return NBNXM_ERR_INVALID;
    }
    /* LJ-PME reads the type table for the grid C6 (NULL nbat: brute force, table given). */
    if (ic->ljPme && nbat && (!nbat->type || !nbat->nbfp))
    {
        return NBNXM_ERR_INVALID;
    }
    return NBNXM_OK;
}

/* Per-type LJ-PME grid parameters, nbfp_comb of atomdata.cpp:296-330 in the
 * reference's FP32 ("real") arithmetic, two floats per type: geometric
 * (sqrt(6C6_ii), 0); Lorentz-Berthelot (0.5 sixthroot(12C12/6C6),
 * sqrt(6C6^2/12C12)), with sixthroot(x) = sqrt(cbrt(x))
 * (api/legacy/include/gromacs/math/functions.h:241-244), zero without LJ. */
static float* grid_params_f32(int rule, const nbnxm_atomdata_t* nbat)
{
    const int nt = nbat->numTypes;
    float* g = (float*)malloc(sizeof(float) * 2 * (size_t)(nt > 0 ? nt : 1));
    if (g)
    {
        for (int t = 0; t < nt; t++)
        {
            const float c6 = nbat->nbfp[2 * (t * nt + t)];
            const float c12 = nbat->nbfp[2 * (t * nt + t) + 1];
            if (rule == NBNXM_LJPME_LB)
            {
                const int lj = c6 > 0.0F && c12 > 0.0F;
                g[2 * t] = lj ? (float)(0.5 * sqrtf(cbrtf(c12 / c6))) : 0.0F;
                g[2 * t + 1] = lj ? sqrtf(c6 * c6 / c12) : 0.0F;
            }
            else
            {
                g[2 * t] = sqrtf(c6);
                g[2 * t + 1] = 0.0F;
            }
        }
    }
    return g;
}

/* Grid C6 of a pair in FP32 as the CUDA kernels combine it: geometric
 * calculate_lj_ewald_c6grid, Lorentz-Berthelot sigma = s_i + s_j,
 * epsilon = e_i e_j, epsilon sigma2 sigma2 sigma2
 * (nbnxm_cuda_kernel_utils.cuh:307-324). */
static float grid_c6_pair_f32(int rule, const float* g, int ti, int tj)
{
    if (rule == NBNXM_LJPME_LB)
    {
        const float sigma = g[2 * ti] + g[2 * tj];
        const float epsilon = g[2 * ti + 1] * g[2 * tj + 1];
        const float sigma2 = sigma * sigma;
        return epsilon * sigma2 * sigma2 * sigma2;
    }
    return g[2 * ti] * g[2 * tj];
}

/* The same grid C6 in FP64 from the (6C6, 12C12) diagonals of the two types. */
static double grid_c6_pair_f64(int rule, double c6i, double c12i, double c6j, double c12j)
{
    if (rule == NBNXM_LJPME_LB)
    {
        const double si = (c6i > 0 && c12i > 0) ? 0.5 * sqrt(cbrt(c12i / c6i)) : 0.0;
        const double ei = (c6i > 0 && c12i > 0) ? sqrt(c6i * c6i / c12i) : 0.0;
        const double sj = (c6j > 0 && c12j > 0) ? 0.5 * sqrt(cbrt(c12j / c6j)) : 0.0;
        const double ej = (c6j > 0 && c12j > 0) ? sqrt(c6j * c6j / c12j) : 0.0;
        const double sigma = si + sj;
        const double sigma2 = sigma * sigma;
        return ei * ej * sigma2 * sigma2 * sigma2;
    }
    return sqrt(c6i) * sqrt(c6j);
}

/* ------------------------------------------------------------------------
 * FP32 restatement of the CUDA kernel body
 * (src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel.cuh:325-718; LJ parameters as
 * :514-533), every flavour through pair_f32.
 * ---------------------------------------------------------------------- */
int oracle_kernel_f32(const nbnxm_pairlist_gpu_t* list,
                      const nbnxm_atomdata_t* nbat,
                      const nbnxm_interaction_const_t* ic,
                      int computeEnergy,
                      float* f,
                      float* fshift,
                      double* Vvdw,
                      double* Vc)
{
    int rc = check_inputs(list, nbat);
    if (rc != NBNXM_OK || !ic || !f || !fshift || check_flavour(nbat, ic) != NBNXM_OK)
    {
        return NBNXM_ERR_INVALID;
    }
    ewald_table_t tab;
    if (ewald_table_make(ic, &tab) != NBNXM_OK)
    {
        return NBNXM_ERR_INVALID;
    }
    const flav32_t p = flav32_from_ic(ic, &tab);
    float* gridC6 = p.ljPme ? grid_params_f32(p.ljPme, nbat) : NULL;
    if (p.ljPme && !gridC6)
    {
        ewald_table_free(&tab);
        return NBNXM_ERR_INVALID;
    }
    const float* xq = nbat->xq;
    const float epsfac = p.epsfac;
    const float selfFactor = (p.eeltype == NBNXM_ELEC_RF) ? 0.5F * p.c_rf : p.beta * (float)c_oneSqrtPi;
    const int ntypes = nbat->numTypes;
    /* The sigma/epsilon force of :555-562 exists in the LB cut-off kernel only. */
    const int lbForceOnlyKernel = !computeEnergy && nbat->vdwType == NBNXM_VDW_CUT_COMB_LB
                                  && p.vdwModifier == NBNXM_VDW_MOD_POTSHIFT && !p.ljPme;
    double vvdwTot = 0.0;
    double vcTot = 0.0;

    for (int e = 0; e < list->nsci; e++)
    {
        const nbnxm_sci_t s = list->sci[e];
        const float* sh = nbat->shiftVec + 3 * s.shift;
        float E_lj = 0.0F;
        float E_el = 0.0F;

        if (computeEnergy && s.shift == NBNXM_CENTRAL_SHIFT && s.cjPackedEnd > s.cjPackedBegin
            && list->cjPacked[s.cjPackedBegin].cj[0] == s.sci * NCPS)
        {
            float sumq2 = 0.0F;
            float sumc6 = 0.0F;
            for (int a = 0; a < NCPS * CL; a++)
            {
                const int ai = s.sci * NCPS * CL + a;
                const float qi = xq[4 * ai + 3] * epsfac;
                sumq2 += qi * qi;
                if (p.ljPme)
                {
                    sumc6 += nbat->nbfp[2 * (nbat->type[ai] * (ntypes + 1))];
                }
            }
            E_el += -(sumq2 / epsfac) * selfFactor;
            if (p.ljPme)
            {
                E_lj += sumc6 * 0.5F * (1.0F / 6.0F) * p.lje_coeff6_6;
            }
        }

        for (int jp = s.cjPackedBegin; jp < s.cjPackedEnd; jp++)
        {
            const nbnxm_cj_packed_t* cjp = &list->cjPacked[jp];
            const unsigned int imask = cjp->imei[0].imask;
            const nbnxm_excl_t* ex = &list->excl[cjp->imei[0].excl_ind];
            for (int jm = 0; jm < JG; jm++)
            {
                if (!((imask >> (jm * NCPS)) & 0xFFU))
                {
                    continue;
                }
                const int cj = cjp->cj[jm];
                for (int im = 0; im < NCPS; im++)
                {
                    const int bit = jm * NCPS + im;
                    if (!((imask >> bit) & 1U))
                    {
                        continue;
                    }
                    const int ci = s.sci * NCPS + im;
                    for (int ti = 0; ti < CL; ti++)
                    {
                        const int ai = ci * CL + ti;
                        const float xi = xq[4 * ai + 0] + sh[0];
                        const float yi = xq[4 * ai + 1] + sh[1];
                        const float zi = xq[4 * ai + 2] + sh[2];
                        const float qi = xq[4 * ai + 3] * epsfac;
                        float fix = 0.0F, fiy = 0.0F, fiz = 0.0F;
                        for (int tj = 0; tj < CL; tj++)
                        {
                            const int aj = cj * CL + tj;
                            const float dx = xi - xq[4 * aj + 0];
                            const float dy = yi - xq[4 * aj + 1];
                            const float dz = zi - xq[4 * aj + 2];
                            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                            const int nonSelf = !(s.shift == NBNXM_CENTRAL_SHIFT && tj <= ti);
                            if (!(r2 < p.rcoulomb_sq && (nonSelf || ci != cj)))
                            {
                                continue;
                            }
                            const float int_bit = ((ex->pair[tj * CL + ti] >> bit) & 1U) ? 1.0F : 0.0F;
                            const float qj = xq[4 * aj + 3];
                            float c6, c12;
                            float sigma = 0.0F, epsilon = 0.0F;
                            if (nbat->vdwType == NBNXM_VDW_CUT_TYPE_TABLE)
                            {
                                const int idx = ntypes * nbat->type[ai] + nbat->type[aj];
                                c6 = nbat->nbfp[2 * idx];
                                c12 = nbat->nbfp[2 * idx + 1];
                            }
                            else if (nbat->vdwType == NBNXM_VDW_CUT_COMB_GEOM)
                            {
                                c6 = nbat->ljComb[2 * ai] * nbat->ljComb[2 * aj];
                                c12 = nbat->ljComb[2 * ai + 1] * nbat->ljComb[2 * aj + 1];
                            }
                            else
                            {
                                sigma = nbat->ljComb[2 * ai] + nbat->ljComb[2 * aj];
                                epsilon = nbat->ljComb[2 * ai + 1] * nbat->ljComb[2 * aj + 1];
                                /* convert_sigma_epsilon_to_c6_c12 (nbnxm_cuda_kernel_utils.cuh:76-86) */
                                const float sigma2 = sigma * sigma;
                                const float sigma6 = sigma2 * sigma2 * sigma2;
                                c6 = epsilon * sigma6;
                                c12 = c6 * sigma6;
                            }
                            const float c6grid = p.ljPme ? grid_c6_pair_f32(p.ljPme, gridC6, nbat->type[ai], nbat->type[aj]) : 0.0F;
                            const float F_invr = pair_f32(&p, computeEnergy, r2, int_bit, qi * qj, c6, c12,
                                                          lbForceOnlyKernel, sigma, epsilon, c6grid, &E_lj, &E_el);
                            const float tx = dx * F_invr;
                            const float ty = dy * F_invr;
                            const float tz = dz * F_invr;
                            fix += tx;
                            fiy += ty;
                            fiz += tz;
                            f[3 * aj + 0] -= tx;
                            f[3 * aj + 1] -= ty;
                            f[3 * aj + 2] -= tz;
                        }
                        f[3 * ai + 0] += fix;
                        f[3 * ai + 1] += fiy;
                        f[3 * ai + 2] += fiz;
                        if (s.shift != NBNXM_CENTRAL_SHIFT)
                        {
                            fshift[3 * s.shift + 0] += fix;
                            fshift[3 * s.shift + 1] += fiy;
                            fshift[3 * s.shift + 2] += fiz;
                        }
                    }
                }
            }
        }
        vvdwTot += E_lj;
        vcTot += E_el;
    }
    free(gridC6);
    ewald_table_free(&tab);
    if (computeEnergy)
    {
        if (Vvdw)
        {
            *Vvdw += vvdwTot;
        }
        if (Vc)
        {
            *Vc += vcTot;
        }
    }
    return NBNXM_OK;
}

int oracle_kernel_f64(const nbnxm_pairlist_gpu_t* list,
                      const nbnxm_atomdata_t* nbat,
                      const nbnxm_interaction_const_t* ic,
                      int computeEnergy,
                      double* f,
                      double* fshift,
                      double* Vvdw,
                      double* Vc)
{
    int rc = check_inputs(list, nbat);
    if (rc != NBNXM_OK || !ic || !f || !fshift || check_flavour(nbat, ic) != NBNXM_OK)
    {
        return NBNXM_ERR_INVALID;
    }
    ewald_table_t tab;
    if (ewald_table_make(ic, &tab) != NBNXM_OK)
    {
        return NBNXM_ERR_INVALID;
    }
    const float* xq = nbat->xq;
    const double epsfac = ic->epsfac;
    const double rcoulomb_sq = ic->rcoulomb * ic->rcoulomb;
    const int ntypes = nbat->numTypes;
    double vvdwTot = 0.0;
    double vcTot = 0.0;

    for (int e = 0; e < list->nsci; e++)
    {
        const nbnxm_sci_t s = list->sci[e];
        const float* sh = nbat->shiftVec + 3 * s.shift;

        if (computeEnergy && s.shift == NBNXM_CENTRAL_SHIFT && s.cjPackedEnd > s.cjPackedBegin
            && list->cjPacked[s.cjPackedBegin].cj[0] == s.sci * NCPS)
        {
            double sumq2 = 0.0, sumc6 = 0.0;
            for (int a = 0; a < NCPS * CL; a++)
            {
                const int ai = s.sci * NCPS * CL + a;
                const double q = xq[4 * ai + 3];
                sumq2 += q * q;
                if (ic->ljPme)
                {
                    sumc6 += nbat->nbfp[2 * (nbat->type[ai] * (ntypes + 1))];
                }
            }
            vcTot += -epsfac * sumq2 * self_coulomb_factor(ic);
            if (ic->ljPme)
            {
                const double b2 = ic->ewaldcoeff_lj * ic->ewaldcoeff_lj;
                vvdwTot += sumc6 / 12.0 * (b2 * b2 * b2 / 6.0);
            }
        }

        for (int jp = s.cjPackedBegin; jp < s.cjPackedEnd; jp++)
        {
            const nbnxm_cj_packed_t* cjp = &list->cjPacked[jp];
            const unsigned int imask = cjp->imei[0].imask;
            const nbnxm_excl_t* ex = &list->excl[cjp->imei[0].excl_ind];
            for (int jm = 0; jm < JG; jm++)
            {
                const int cj = cjp->cj[jm];
                for (int im = 0; im < NCPS; im++)
                {
                    const int bit = jm * NCPS + im;
                    if (!((imask >> bit) & 1U))
                    {
                        continue;
                    }
                    const int ci = s.sci * NCPS + im;
                    for (int ti = 0; ti < CL; ti++)
                    {
                        const int ai = ci * CL + ti;
                        const double xi = (double)xq[4 * ai + 0] + (double)sh[0];
                        const double yi = (double)xq[4 * ai + 1] + (double)sh[1];
                        const double zi = (double)xq[4 * ai + 2] + (double)sh[2];
                        const double qi = (double)xq[4 * ai + 3] * epsfac;
                        double fix = 0.0, fiy = 0.0, fiz = 0.0;
                        for (int tj = 0; tj < CL; tj++)
                        {
                            const int aj = cj * CL + tj;
                            const double dx = xi - (double)xq[4 * aj + 0];
                            const double dy = yi - (double)xq[4 * aj + 1];
                            const double dz = zi - (double)xq[4 * aj + 2];
                            const double r2 = dx * dx + dy * dy + dz * dz;
                            const int nonSelf = !(s.shift == NBNXM_CENTRAL_SHIFT && tj <= ti);
                            if (!(r2 < rcoulomb_sq && (nonSelf || ci != cj)))
                            {
                                continue;
                            }
                            const double int_bit = ((ex->pair[tj * CL + ti] >> bit) & 1U) ? 1.0 : 0.0;
                            const double qj = xq[4 * aj + 3];
                            double c6, c12;
                            if (nbat->vdwType == NBNXM_VDW_CUT_TYPE_TABLE)
                            {
                                const int idx = ntypes * nbat->type[ai] + nbat->type[aj];
                                c6 = nbat->nbfp[2 * idx];
                                c12 = nbat->nbfp[2 * idx + 1];
                            }
                            else if (nbat->vdwType == NBNXM_VDW_CUT_COMB_GEOM)
                            {
                                c6 = (double)nbat->ljComb[2 * ai] * (double)nbat->ljComb[2 * aj];
                                c12 = (double)nbat->ljComb[2 * ai + 1] * (double)nbat->ljComb[2 * aj + 1];
                            }
                            else
                            {
                                const double sigma = (double)nbat->ljComb[2 * ai] + (double)nbat->ljComb[2 * aj];
                                const double epsilon = (double)nbat->ljComb[2 * ai + 1] * (double)nbat->ljComb[2 * aj + 1];
                                const double sigma6 = sigma * sigma * sigma * sigma * sigma * sigma;
                                c6 = epsilon * sigma6;
                                c12 = c6 * sigma6;
                            }
                            double c6grid = 0.0;
                            if (ic->ljPme)
                            {
                                const float* di = nbat->nbfp + 2 * (nbat->type[ai] * (ntypes + 1));
                                const float* dj = nbat->nbfp + 2 * (nbat->type[aj] * (ntypes + 1));
                                c6grid = grid_c6_pair_f64(ic->ljPme, di[0], di[1], dj[0], dj[1]);
                            }
                            double elj = 0.0, eel = 0.0;
                            const double F_invr = pair_f64(ic, &tab, r2, int_bit, qi * qj, c6, c12, c6grid, &elj, &eel);
                            if (computeEnergy)
                            {
                                vvdwTot += elj;
                                vcTot += eel;
                            }
                            const double tx = dx * F_invr;
                            const double ty = dy * F_invr;
                            const double tz = dz * F_invr;
                            fix += tx;
                            fiy += ty;
                            fiz += tz;
                            f[3 * aj + 0] -= tx;
                            f[3 * aj + 1] -= ty;
                            f[3 * aj + 2] -= tz;
                        }
                        f[3 * ai + 0] += fix;
                        f[3 * ai + 1] += fiy;
                        f[3 * ai + 2] += fiz;
                        if (s.shift != NBNXM_CENTRAL_SHIFT)
                        {
                            fshift[3 * s.shift + 0] += fix;
                            fshift[3 * s.shift + 1] += fiy;
                            fshift[3 * s.shift + 2] += fiz;
                        }
                    }
                }
            }
        }
    }
    ewald_table_free(&tab);
    if (computeEnergy)
    {
        if (Vvdw)
        {
            *Vvdw += vvdwTot;
        }
        if (Vc)
        {
            *Vc += vcTot;
        }
    }
    return NBNXM_OK;
}

static int is_excluded(int i, int j, const int* ranges, const int* elems)
{
    if (!ranges)
    {
        return 0;
    }
    for (int k = ranges[i]; k < ranges[i + 1]; k++)
    {
        if (elems[k] == j)
        {
            return 1;
        }
    }
    return 0;
}

static int bruteforce_f64(int n,
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
                          double* Vvdw,
                          double* Vc,
                          long long* numPairsInRange,
                          double* virial)
{
    if (n < 0 || !x || !q || !type || !nbfp || !box || !ic || !f || check_flavour(NULL, ic) != NBNXM_OK)
    {
        return NBNXM_ERR_INVALID;
    }
    const double rc = ic->rcoulomb;
    for (int d = 0; d < 3; d++)
    {
        if (!(box[d] > rc))
        {
            return NBNXM_ERR_INVALID;
        }
    }
    const double rc2 = rc * rc;
    const double epsfac = ic->epsfac;
    double vvdw = 0.0, vc = 0.0;
    long long npairs = 0;
    ewald_table_t tab;
    if (ewald_table_make(ic, &tab) != NBNXM_OK)
    {
        return NBNXM_ERR_INVALID;
    }

    /* Wrap into the unit cell so every separation is within (-L, L). */
    double* xw = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    if (!xw)
    {
        ewald_table_free(&tab);
        return NBNXM_ERR_INVALID;
    }
    for (int i = 0; i < n; i++)
    {
        for (int d = 0; d < 3; d++)
        {
            double v = fmod(x[3 * i + d], box[d]);
            if (v < 0)
            {
                v += box[d];
            }
            xw[3 * i + d] = v;
        }
    }

    double sumq2 = 0.0, sumc6 = 0.0;
    for (int i = 0; i < n; i++)
    {
        sumq2 += q[i] * q[i];
        sumc6 += nbfp[2 * (numTypes * type[i] + type[i])];
        for (int j = i + 1; j < n; j++)
        {
            const int excluded = is_excluded(i, j, exclRanges, exclElements);
            const double qq = epsfac * q[i] * q[j];
            const double c6 = nbfp[2 * (numTypes * type[i] + type[j])];
            const double c12 = nbfp[2 * (numTypes * type[i] + type[j]) + 1];
            const double* di = nbfp + 2 * (numTypes * type[i] + type[i]);
            const double* dj = nbfp + 2 * (numTypes * type[j] + type[j]);
            const double c6grid = ic->ljPme ? grid_c6_pair_f64(ic->ljPme, di[0], di[1], dj[0], dj[1]) : 0.0;
            for (int sz = -1; sz <= 1; sz++)
            {
                for (int sy = -1; sy <= 1; sy++)
                {
                    for (int sx = -1; sx <= 1; sx++)
                    {
                        const double dx = xw[3 * i + 0] - xw[3 * j + 0] + sx * box[0];
                        const double dy = xw[3 * i + 1] - xw[3 * j + 1] + sy * box[1];
                        const double dz = xw[3 * i + 2] - xw[3 * j + 2] + sz * box[2];
                        const double r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 >= rc2)
                        {
                            continue;
                        }
                        const double int_bit = excluded ? 0.0 : 1.0;
                        if (!excluded)
                        {
                            npairs++;
                        }
                        const double fr = pair_f64(ic, &tab, r2, int_bit, qq, c6, c12, c6grid, &vvdw, &vc);
                        f[3 * i + 0] += fr * dx;
                        f[3 * i + 1] += fr * dy;
                        f[3 * i + 2] += fr * dz;
                        f[3 * j + 0] -= fr * dx;
                        f[3 * j + 1] -= fr * dy;
                        f[3 * j + 2] -= fr * dz;
                        if (virial)
                        {
                            /* pair virial -1/2 r_ij (x) F_ij, F_ij = fr r_ij */
                            const double r[3] = { dx, dy, dz };
                            for (int a = 0; a < 3; a++)
                            {
                                for (int b = 0; b < 3; b++)
                                {
                                    virial[3 * a + b] -= 0.5 * r[a] * fr * r[b];
                                }
                            }
                        }
                    }
                }
            }
        }
    }
    vc += -epsfac * self_coulomb_factor(ic) * sumq2;
    if (ic->ljPme)
    {
        const double b2 = ic->ewaldcoeff_lj * ic->ewaldcoeff_lj;
        vvdw += sumc6 / 12.0 * (b2 * b2 * b2 / 6.0);
    }
    free(xw);
    ewald_table_free(&tab);
    if (Vvdw)
    {
        *Vvdw += vvdw;
    }
    if (Vc)
    {
        *Vc += vc;
    }
    if (numPairsInRange)
    {
        *numPairsInRange = npairs;
    }
    return NBNXM_OK;
}

int oracle_bruteforce_f64(int n,
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
                          double* Vvdw,
                          double* Vc,
                          long long* numPairsInRange)
{
    return bruteforce_f64(n, x, q, type, numTypes, nbfp, exclRanges, exclElements, box, ic, f, Vvdw, Vc, numPairsInRange, NULL);
}

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
                                 double* virial)
{
    if (!virial)
    {
        return NBNXM_ERR_INVALID;
    }
    return bruteforce_f64(n, x, q, type, numTypes, nbfp, exclRanges, exclElements, box, ic, f, NULL, NULL, NULL, virial);
}

/* Restates nbnxn_kernel_prune_cuda<haveFreshList>
 * (src/gromacs/nbnxm/cuda/nbnxm_cuda_kernel_pruneonly.cuh:104-343) for split 1:
 * the any-pair-within-distance test of :283-299 with r2 evaluated as
 * fmaf(dz,dz,fmaf(dy,dy,dx*dx)) after adding the shift to x_i in FP32. */
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
                 int* sciCount)
{
    if (!list || !xq || !shiftVec || !imaskOuter || !imaskInner || numParts < 1 || part < 0 || part >= numParts)
    {
        return NBNXM_ERR_INVALID;
    }
    const int numSciInPart = (list->nsci - part + numParts - 1) / numParts;
    for (int b = 0; b < numSciInPart; b++)
    {
        const int e = b * numParts + part;
        const nbnxm_sci_t s = list->sci[e];
        const float* sh = shiftVec + 3 * s.shift;
        int prunedPairCount = 0;
        for (int jp = s.cjPackedBegin; jp < s.cjPackedEnd; jp++)
        {
            unsigned int imaskFull, imaskCheck, imaskNew;
            if (haveFreshList)
            {
                imaskFull = imaskInner[jp];
                imaskCheck = imaskFull;
                imaskNew = 0;
            }
            else
            {
                imaskFull = imaskOuter[jp];
                imaskNew = imaskInner[jp];
                imaskCheck = imaskNew ^ imaskFull;
            }
            if (!imaskCheck)
            {
                continue;
            }
            for (int jm = 0; jm < JG; jm++)
            {
                const int cj = list->cjPacked[jp].cj[jm];
                for (int im = 0; im < NCPS; im++)
                {
                    const unsigned int mask_ji = 1U << (jm * NCPS + im);
                    if (!(imaskCheck & mask_ji))
                    {
                        continue;
                    }
                    const int ci = s.sci * NCPS + im;
                    int anyOuter = 0, anyInner = 0;
                    for (int tj = 0; tj < CL; tj++)
                    {
                        const int aj = cj * CL + tj;
                        for (int ti = 0; ti < CL; ti++)
                        {
                            const int ai = ci * CL + ti;
                            const float xi = xq[4 * ai + 0] + sh[0];
                            const float yi = xq[4 * ai + 1] + sh[1];
                            const float zi = xq[4 * ai + 2] + sh[2];
                            const float dx = xi - xq[4 * aj + 0];
                            const float dy = yi - xq[4 * aj + 1];
                            const float dz = zi - xq[4 * aj + 2];
                            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                            anyOuter |= (r2 < rlistOuter_sq);
                            anyInner |= (r2 < rlistInner_sq);
                        }
                    }
                    if (haveFreshList && !anyOuter)
                    {
                        imaskFull &= ~mask_ji;
                    }
                    if (anyInner)
                    {
                        imaskNew |= mask_ji;
                    }
                }
            }
            if (haveFreshList)
            {
                imaskOuter[jp] = imaskFull;
                prunedPairCount += __builtin_popcount(imaskNew);
            }
            imaskInner[jp] = imaskNew;
        }
        if (haveFreshList && sciCount)
        {
            int index = NBNXM_SCI_HISTOGRAM_SIZE - prunedPairCount - 1;
            sciCount[e] = index > 0 ? index : 0;
        }
    }
    return NBNXM_OK;
}

int oracle_sci_sort(int nsci, const nbnxm_sci_t* sci, const int* sciCount, nbnxm_sci_t* sorted)
{
    if (nsci < 0 || (nsci > 0 && (!sci || !sciCount || !sorted)))
    {
        return NBNXM_ERR_INVALID;
    }
    int* offset = (int*)calloc(NBNXM_SCI_HISTOGRAM_SIZE + 1, sizeof(int));
    if (!offset)
    {
        return NBNXM_ERR_INVALID;
    }
    for (int i = 0; i < nsci; i++)
    {
        offset[sciCount[i] + 1]++;
    }
    for (int b = 0; b < NBNXM_SCI_HISTOGRAM_SIZE; b++)
    {
        offset[b + 1] += offset[b];
    }
    for (int i = 0; i < nsci; i++)
    {
        sorted[offset[sciCount[i]]++] = sci[i];
    }
    free(offset);
    return NBNXM_OK;
}

long long oracle_count_pairs_f64(const nbnxm_pairlist_gpu_t* list, const nbnxm_atomdata_t* nbat, double rc)
{
    if (!list || !nbat || !nbat->xq || !nbat->shiftVec)
    {
        return -1;
    }
    const float* xq = nbat->xq;
    const double rc2 = rc * rc;
    long long count = 0;
    for (int e = 0; e < list->nsci; e++)
    {
        const nbnxm_sci_t s = list->sci[e];
        const float* sh = nbat->shiftVec + 3 * s.shift;
        for (int jp = s.cjPackedBegin; jp < s.cjPackedEnd; jp++)
        {
            const nbnxm_cj_packed_t* cjp = &list->cjPacked[jp];
            const nbnxm_excl_t* ex = &list->excl[cjp->imei[0].excl_ind];
            for (int jm = 0; jm < JG; jm++)
            {
                for (int im = 0; im < NCPS; im++)
                {
                    const int bit = jm * NCPS + im;
                    if (!((cjp->imei[0].imask >> bit) & 1U))
                    {
                        continue;
                    }
                    const int ci = s.sci * NCPS + im;
                    const int cj = cjp->cj[jm];
                    for (int ti = 0; ti < CL; ti++)
                    {
                        const int ai = ci * CL + ti;
                        if (xq[4 * ai] <= -1.0e5F)
                        {
                            continue; /* filler */
                        }
                        for (int tj = 0; tj < CL; tj++)
                        {
                            const int aj = cj * CL + tj;
                            if (xq[4 * aj] <= -1.0e5F)
                            {
                                continue;
                            }
                            if (s.shift == NBNXM_CENTRAL_SHIFT && ci == cj && tj <= ti)
                            {
                                continue;
                            }
                            if (!((ex->pair[tj * CL + ti] >> bit) & 1U))
                            {
                                continue;
                            }
                            const double dx = (double)xq[4 * ai] + sh[0] - (double)xq[4 * aj];
                            const double dy = (double)xq[4 * ai + 1] + sh[1] - (double)xq[4 * aj + 1];
                            const double dz = (double)xq[4 * ai + 2] + sh[2] - (double)xq[4 * aj + 2];
                            if (dx * dx + dy * dy + dz * dz < rc2)
                            {
                                count++;
                            }
                        }
                    }
                }
            }
        }
    }
    return count;
}
