// CPU baseline: a from-scratch 4xM SIMD restatement of the reference CPU
// kernel nbnxmKernelSimd<KernelLayout::r4xM, EwaldAnalytical, ...>
// (src/gromacs/nbnxm/simd_kernel.h:116-566, simd_kernel_inner.h) for AVX2+FMA:
// 4 i-atoms against one 8-atom j-cluster per inner iteration (M = 8 lanes),
// analytical Ewald force via the rational pmeForceCorrection and energy via
// pmePotentialCorrection (src/gromacs/simd/include/gromacs/simd/simd_math.h:
// 1611-1648, 1687-1720), LJ cut-off with potential shift, exclusion forces,
// OpenMP over i-entries with per-thread force buffers reduced afterwards
// (src/gromacs/nbnxm/kerneldispatch.cpp:241-395, atomdata.cpp:1224-1262).
//
// TEST AND BENCHMARK INFRASTRUCTURE ONLY (the CPU baseline leg).  The 4x8
// list is derived from the GPU list: each 8-atom i-cluster is split into two
// 4-atom i-clusters and every (i-half, j-cluster) pair with an atom pair
// inside the prune radius is kept, with its 32-bit interaction and validity
// masks (bit i*8+j, the CPU layout checked by src/gromacs/nbnxm/tests/exclusions.cpp:170-197).
#include <immintrin.h>
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/nbnxm_amd.h"

namespace
{

struct CiEntry
{
    int ci4;
    int shift;
    int cjBegin;
    int cjEnd;
};

struct CjEntry
{
    int cj;
    uint32_t interact; // topology mask: 1 = full interaction, 0 = excluded (Ewald correction only)
    uint32_t valid;    // 0 for self pairs and j <= i on the central diagonal
};

inline __m256 rsqrtNR(__m256 x)
{
    const __m256 y = _mm256_rsqrt_ps(x);
    const __m256 half = _mm256_set1_ps(0.5f), threeHalf = _mm256_set1_ps(1.5f);
    return _mm256_mul_ps(y, _mm256_fnmadd_ps(_mm256_mul_ps(half, x), _mm256_mul_ps(y, y), threeHalf));
}

inline __m256 rcpNR(__m256 x)
{
    const __m256 y = _mm256_rcp_ps(x);
    return _mm256_mul_ps(y, _mm256_fnmadd_ps(x, y, _mm256_set1_ps(2.0f)));
}

inline __m256 pmeForceCorrection(__m256 z2)
{
    const __m256 z4 = _mm256_mul_ps(z2, z2);
    __m256 d0 = _mm256_fmadd_ps(_mm256_set1_ps(0.0011193462567257629232f), z4, _mm256_set1_ps(0.11583842382862377919f));
    __m256 d1 = _mm256_fmadd_ps(_mm256_set1_ps(0.014866955030185295499f), z4, _mm256_set1_ps(0.50736591960530292870f));
    d0 = _mm256_fmadd_ps(d0, z4, _mm256_set1_ps(1.0f));
    d0 = _mm256_fmadd_ps(d1, z2, d0);
    d0 = rcpNR(d0);
    __m256 n0 = _mm256_fmadd_ps(_mm256_set1_ps(-1.7357322914161492954e-8f), z4, _mm256_set1_ps(-0.000053401640219807709149f));
    __m256 n1 = _mm256_fmadd_ps(_mm256_set1_ps(1.4703624142580877519e-6f), z4, _mm256_set1_ps(0.0010054721316683106153f));
    n0 = _mm256_fmadd_ps(n0, z4, _mm256_set1_ps(-0.019278317264888380590f));
    n1 = _mm256_fmadd_ps(n1, z4, _mm256_set1_ps(0.069670166153766424023f));
    n0 = _mm256_fmadd_ps(n0, z4, _mm256_set1_ps(-0.75225204789749321333f));
    n0 = _mm256_fmadd_ps(n1, z2, n0);
    return _mm256_mul_ps(n0, d0);
}

inline __m256 pmePotentialCorrection(__m256 z2)
{
    const __m256 z4 = _mm256_mul_ps(z2, z2);
    __m256 d1 = _mm256_fmadd_ps(_mm256_set1_ps(0.0066752224023576045451f), z4, _mm256_set1_ps(0.43336185284710920150f));
    __m256 d0 = _mm256_fmadd_ps(_mm256_set1_ps(0.078647795836373922256f), z4, _mm256_set1_ps(1.0f));
    d0 = _mm256_fmadd_ps(d1, z2, d0);
    d0 = rcpNR(d0);
    __m256 n0 = _mm256_fmadd_ps(_mm256_set1_ps(1.9296833005951166339e-8f), z4, _mm256_set1_ps(0.000041603292906656984871f));
    __m256 n1 = _mm256_fmadd_ps(_mm256_set1_ps(-1.4213390571557850962e-6f), z4, _mm256_set1_ps(-0.00013134036773265025626f));
    n0 = _mm256_fmadd_ps(n0, z4, _mm256_set1_ps(0.038657983986041781264f));
    n1 = _mm256_fmadd_ps(n1, z4, _mm256_set1_ps(0.11285044772717598220f));
    n0 = _mm256_fmadd_ps(n0, z4, _mm256_set1_ps(1.1283802385263030286f));
    n0 = _mm256_fmadd_ps(n1, z2, n0);
    return _mm256_mul_ps(n0, d0);
}

// Lane mask from 8 bits of a 32-bit cluster-pair mask.
inline __m256 bitsToMask(uint32_t bits8)
{
    const __m256i sel = _mm256_setr_epi32(1, 2, 4, 8, 16, 32, 64, 128);
    const __m256i v = _mm256_and_si256(_mm256_set1_epi32(static_cast<int>(bits8)), sel);
    return _mm256_castsi256_ps(_mm256_cmpeq_epi32(v, sel));
}

inline float hsum(__m256 v)
{
    __m128 lo = _mm256_castps256_ps128(v);
    __m128 hi = _mm256_extractf128_ps(v, 1);
    lo = _mm_add_ps(lo, hi);
    lo = _mm_hadd_ps(lo, lo);
    lo = _mm_hadd_ps(lo, lo);
    return _mm_cvtss_f32(lo);
}

} // namespace

struct cpu4xm
{
    int numAtoms = 0;
    std::vector<float> x, y, z, q, c6, c12; // SoA, grid order; c6/c12 = sqrt(6C6), sqrt(12C12)
    float shift[3 * NBNXM_NUM_SHIFTS];
    std::vector<CiEntry> ci;
    std::vector<CjEntry> cj;
    nbnxm_interaction_const_t ic;
    int nthreads = 1;
    std::vector<std::vector<float>> fbuf; // per thread 3*numAtoms SoA (x block, y block, z block)
};

extern "C" {

// Builds the 4x8 list from a GPU list (geometric-combination LJ parameters in
// nbat->ljComb) keeping (i-half, j-cluster) pairs with any atom pair within rPrune.
cpu4xm* cpu4xm_create(const nbnxm_pairlist_gpu_t* list, const nbnxm_atomdata_t* nbat, const nbnxm_interaction_const_t* ic, double rPrune, int numThreads)
{
    // Ewald + LJ cut-off with the potential shift and rvdw = rcoulomb only.
    if (!list || !nbat || !ic || nbat->vdwType != NBNXM_VDW_CUT_COMB_GEOM || !nbat->ljComb
        || ic->eeltype != NBNXM_ELEC_EWALD || ic->vdwModifier != NBNXM_VDW_MOD_POTSHIFT || ic->ljPme
        || ic->rvdw != ic->rcoulomb)
    {
        return nullptr;
    }
    cpu4xm* k = new cpu4xm();
    const int n = nbat->numAtoms;
    k->numAtoms = n;
    k->ic = *ic;
    k->x.resize(n);
    k->y.resize(n);
    k->z.resize(n);
    k->q.resize(n);
    k->c6.resize(n);
    k->c12.resize(n);
    for (int a = 0; a < n; a++)
    {
        k->x[a] = nbat->xq[4 * a];
        k->y[a] = nbat->xq[4 * a + 1];
        k->z[a] = nbat->xq[4 * a + 2];
        k->q[a] = nbat->xq[4 * a + 3];
        k->c6[a] = nbat->ljComb[2 * a];
        k->c12[a] = nbat->ljComb[2 * a + 1];
    }
    std::memcpy(k->shift, nbat->shiftVec, sizeof(k->shift));
    const double rp2 = rPrune * rPrune;
    for (int e = 0; e < list->nsci; e++)
    {
        const nbnxm_sci_t s = list->sci[e];
        const float* sh = nbat->shiftVec + 3 * s.shift;
        const bool central = (s.shift == NBNXM_CENTRAL_SHIFT);
        for (int im = 0; im < NBNXM_NUM_CLUSTER_PER_SUPERCLUSTER; im++)
        {
            const int ci8 = s.sci * NBNXM_NUM_CLUSTER_PER_SUPERCLUSTER + im;
            for (int half = 0; half < 2; half++)
            {
                CiEntry entry{ 2 * ci8 + half, s.shift, static_cast<int>(k->cj.size()), 0 };
                for (int g = s.cjPackedBegin; g < s.cjPackedEnd; g++)
                {
                    const nbnxm_cj_packed_t& p = list->cjPacked[g];
                    const nbnxm_excl_t& ex = list->excl[p.imei[0].excl_ind];
                    for (int jm = 0; jm < NBNXM_JGROUP_SIZE; jm++)
                    {
                        const int bit = jm * NBNXM_NUM_CLUSTER_PER_SUPERCLUSTER + im;
                        if (!((p.imei[0].imask >> bit) & 1u))
                        {
                            continue;
                        }
                        const int cj8 = p.cj[jm];
                        uint32_t interact = 0, valid = 0;
                        bool inRange = false;
                        for (int i = 0; i < 4; i++)
                        {
                            const int ti = 4 * half + i;
                            const int ai = ci8 * 8 + ti;
                            for (int tj = 0; tj < 8; tj++)
                            {
                                const int aj = cj8 * 8 + tj;
                                if (central && ci8 == cj8 && tj <= ti)
                                {
                                    continue;
                                }
                                valid |= 1u << (i * 8 + tj);
                                if ((ex.pair[tj * 8 + ti] >> bit) & 1u)
                                {
                                    interact |= 1u << (i * 8 + tj);
                                }
                                const double dx = static_cast<double>(k->x[ai]) + sh[0] - k->x[aj];
                                const double dy = static_cast<double>(k->y[ai]) + sh[1] - k->y[aj];
                                const double dz = static_cast<double>(k->z[ai]) + sh[2] - k->z[aj];
                                inRange |= (dx * dx + dy * dy + dz * dz < rp2);
                            }
                        }
                        if (inRange && valid)
                        {
                            k->cj.push_back(CjEntry{ cj8, interact, valid });
                        }
                    }
                }
                entry.cjEnd = static_cast<int>(k->cj.size());
                if (entry.cjEnd > entry.cjBegin)
                {
                    k->ci.push_back(entry);
                }
            }
        }
    }
    k->nthreads = numThreads > 0 ? numThreads : omp_get_max_threads();
    k->fbuf.assign(k->nthreads, std::vector<float>(3 * static_cast<size_t>(n), 0.0f));
    return k;
}

int cpu4xm_num_threads(const cpu4xm* k)
{
    return k ? k->nthreads : 0;
}

long long cpu4xm_num_cluster_pairs(const cpu4xm* k)
{
    return k ? static_cast<long long>(k->cj.size()) : -1;
}

// One force (+energy) evaluation; f (3*numAtoms, grid order, xyz interleaved)
// is overwritten.  Self energy -epsfac*beta/sqrt(pi)*sum q^2 is included in Vc.
int cpu4xm_compute(cpu4xm* k, int computeEnergy, float* f, double* Vvdw, double* Vc)
{
    if (!k || !f)
    {
        return NBNXM_ERR_INVALID;
    }
    const int n = k->numAtoms;
    const float rc2 = static_cast<float>(k->ic.rcoulomb * k->ic.rcoulomb);
    const float beta = static_cast<float>(k->ic.ewaldcoeff_q);
    const float epsfac = static_cast<float>(k->ic.epsfac);
    const float shEwald = static_cast<float>(k->ic.sh_ewald);
    const float cpRep = static_cast<float>(k->ic.repulsion_shift_cpot);
    const float cpDisp = static_cast<float>(k->ic.dispersion_shift_cpot);
    double vvdw = 0.0, vc = 0.0;
    const int nci = static_cast<int>(k->ci.size());

#pragma omp parallel num_threads(k->nthreads) reduction(+ : vvdw, vc)
    {
        const int t = omp_get_thread_num();
        float* fx = k->fbuf[t].data();
        float* fy = fx + n;
        float* fz = fy + n;
        std::memset(fx, 0, sizeof(float) * 3 * n);
        const __m256 vrc2 = _mm256_set1_ps(rc2);
        const __m256 vbeta = _mm256_set1_ps(beta);
        const __m256 vbeta2 = _mm256_set1_ps(beta * beta);
        const __m256 vbeta3 = _mm256_set1_ps(beta * beta * beta);
        const __m256 vmin = _mm256_set1_ps(NBNXM_MIN_DISTANCE_SQUARED);
        const __m256 one = _mm256_set1_ps(1.0f);
        const __m256 zero = _mm256_setzero_ps();

#pragma omp for schedule(dynamic, 16)
        for (int e = 0; e < nci; e++)
        {
            const CiEntry ce = k->ci[e];
            const float* sh = k->shift + 3 * ce.shift;
            __m256 ix[4], iy[4], iz[4], iq[4], ic6[4], ic12[4];
            __m256 fix[4], fiy[4], fiz[4];
            for (int i = 0; i < 4; i++)
            {
                const int a = ce.ci4 * 4 + i;
                ix[i] = _mm256_set1_ps(k->x[a] + sh[0]);
                iy[i] = _mm256_set1_ps(k->y[a] + sh[1]);
                iz[i] = _mm256_set1_ps(k->z[a] + sh[2]);
                iq[i] = _mm256_set1_ps(k->q[a] * epsfac);
                ic6[i] = _mm256_set1_ps(k->c6[a]);
                ic12[i] = _mm256_set1_ps(k->c12[a]);
                fix[i] = fiy[i] = fiz[i] = zero;
            }
            __m256 vctot = zero, vvdwtot = zero;
            for (int c = ce.cjBegin; c < ce.cjEnd; c++)
            {
                const CjEntry je = k->cj[c];
                const int j0 = je.cj * 8;
                const __m256 jx = _mm256_loadu_ps(&k->x[j0]);
                const __m256 jy = _mm256_loadu_ps(&k->y[j0]);
                const __m256 jz = _mm256_loadu_ps(&k->z[j0]);
                const __m256 jq = _mm256_loadu_ps(&k->q[j0]);
                const __m256 jc6 = _mm256_loadu_ps(&k->c6[j0]);
                const __m256 jc12 = _mm256_loadu_ps(&k->c12[j0]);
                __m256 fjx = zero, fjy = zero, fjz = zero;
                for (int i = 0; i < 4; i++)
                {
                    const __m256 dx = _mm256_sub_ps(ix[i], jx);
                    const __m256 dy = _mm256_sub_ps(iy[i], jy);
                    const __m256 dz = _mm256_sub_ps(iz[i], jz);
                    __m256 r2 = _mm256_fmadd_ps(dz, dz, _mm256_fmadd_ps(dy, dy, _mm256_mul_ps(dx, dx)));
                    const __m256 wco = _mm256_and_ps(_mm256_cmp_ps(r2, vrc2, _CMP_LT_OQ), bitsToMask((je.valid >> (8 * i)) & 0xFFu));
                    if (_mm256_movemask_ps(wco) == 0)
                    {
                        continue;
                    }
                    const __m256 intBit = _mm256_and_ps(bitsToMask((je.interact >> (8 * i)) & 0xFFu), one);
                    r2 = _mm256_max_ps(r2, vmin);
                    const __m256 rinv = _mm256_and_ps(rsqrtNR(r2), wco);
                    const __m256 rinv2 = _mm256_mul_ps(rinv, rinv);
                    const __m256 rinv6 = _mm256_mul_ps(_mm256_mul_ps(_mm256_mul_ps(rinv2, rinv2), rinv2), intBit);
                    const __m256 c6 = _mm256_mul_ps(ic6[i], jc6);
                    const __m256 c12 = _mm256_mul_ps(ic12[i], jc12);
                    const __m256 frLJ = _mm256_mul_ps(rinv6, _mm256_fmsub_ps(c12, rinv6, c6));
                    const __m256 qq = _mm256_mul_ps(iq[i], jq);
                    const __m256 brsq = _mm256_mul_ps(vbeta2, r2);
                    const __m256 fcoul = _mm256_mul_ps(qq, _mm256_fmadd_ps(_mm256_mul_ps(intBit, rinv2), rinv, _mm256_mul_ps(vbeta3, pmeForceCorrection(brsq))));
                    const __m256 fscal = _mm256_and_ps(_mm256_fmadd_ps(frLJ, rinv2, fcoul), wco);
                    const __m256 tx = _mm256_mul_ps(fscal, dx);
                    const __m256 ty = _mm256_mul_ps(fscal, dy);
                    const __m256 tz = _mm256_mul_ps(fscal, dz);
                    fix[i] = _mm256_add_ps(fix[i], tx);
                    fiy[i] = _mm256_add_ps(fiy[i], ty);
                    fiz[i] = _mm256_add_ps(fiz[i], tz);
                    fjx = _mm256_sub_ps(fjx, tx);
                    fjy = _mm256_sub_ps(fjy, ty);
                    fjz = _mm256_sub_ps(fjz, tz);
                    if (computeEnergy)
                    {
                        // V = qq*(int_bit/r - erf(beta r)/r - int_bit*sh_ewald)
                        const __m256 erfOverR = _mm256_mul_ps(vbeta, pmePotentialCorrection(brsq));
                        __m256 v = _mm256_mul_ps(qq, _mm256_sub_ps(_mm256_mul_ps(intBit, _mm256_sub_ps(rinv, _mm256_set1_ps(shEwald))), erfOverR));
                        vctot = _mm256_add_ps(vctot, _mm256_and_ps(v, wco));
                        const __m256 vlj = _mm256_sub_ps(
                                _mm256_mul_ps(_mm256_mul_ps(c12, _mm256_fmadd_ps(rinv6, rinv6, _mm256_mul_ps(intBit, _mm256_set1_ps(cpRep)))), _mm256_set1_ps(1.0f / 12.0f)),
                                _mm256_mul_ps(_mm256_mul_ps(c6, _mm256_fmadd_ps(intBit, _mm256_set1_ps(cpDisp), rinv6)), _mm256_set1_ps(1.0f / 6.0f)));
                        vvdwtot = _mm256_add_ps(vvdwtot, _mm256_and_ps(vlj, wco));
                    }
                }
                _mm256_storeu_ps(&fx[j0], _mm256_add_ps(_mm256_loadu_ps(&fx[j0]), fjx));
                _mm256_storeu_ps(&fy[j0], _mm256_add_ps(_mm256_loadu_ps(&fy[j0]), fjy));
                _mm256_storeu_ps(&fz[j0], _mm256_add_ps(_mm256_loadu_ps(&fz[j0]), fjz));
            }
            for (int i = 0; i < 4; i++)
            {
                const int a = ce.ci4 * 4 + i;
                fx[a] += hsum(fix[i]);
                fy[a] += hsum(fiy[i]);
                fz[a] += hsum(fiz[i]);
            }
            if (computeEnergy)
            {
                vc += hsum(vctot);
                vvdw += hsum(vvdwtot);
            }
        }
        // Reduce the per-thread buffers (reduceForces, atomdata.cpp:1224-1262).
#pragma omp for schedule(static)
        for (int a = 0; a < n; a++)
        {
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            for (int u = 0; u < k->nthreads; u++)
            {
                const float* b = k->fbuf[u].data();
                sx += b[a];
                sy += b[n + a];
                sz += b[2 * n + a];
            }
            f[3 * a] = sx;
            f[3 * a + 1] = sy;
            f[3 * a + 2] = sz;
        }
    }
    if (computeEnergy)
    {
        // Self term -epsfac*beta/sqrt(pi)*q_i^2 of every atom (the do_self
        // part of the reference CPU kernels; fillers carry q = 0).
        double q2 = 0.0;
        for (int a = 0; a < n; a++)
        {
            q2 += static_cast<double>(k->q[a]) * k->q[a];
        }
        vc -= k->ic.epsfac * k->ic.ewaldcoeff_q * 0.56418958354775628695 * q2;
    }
    if (Vvdw)
    {
        *Vvdw = vvdw;
    }
    if (Vc)
    {
        *Vc = vc;
    }
    return NBNXM_OK;
}

void cpu4xm_free(cpu4xm* k)
{
    delete k;
}

} // extern "C"
