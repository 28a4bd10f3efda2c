// GPU pair search: the grid of 8-atom clusters and the GPU cluster-pair list
// built on the device for one or many systems at once.  It produces the same
// list, bit for bit, as the host builder in pairsearch.cpp (the grid and list
// semantics of src/gromacs/nbnxm/grid.cpp:1101-1240 and
// src/gromacs/nbnxm/pairlist.cpp:914-939,1122-1270,1948-2046,2150-2262): every
// floating-point step is evaluated with explicitly rounded operations in the
// same order, and every ordering decision (sort ties, entry order, exclusion
// entry creation order) follows the host code.
//
// Steps, all on the device:
//   1. wrap coordinates into the box (FP64, then FP32) and key each atom by
//      (column, z) -- one stable radix sort orders atoms into z-sorted columns;
//   2. per-column counts and scans give the cells (64 atoms) of each column;
//   3. one wave per cell splits its z-sorted atoms into 8 clusters (halves in
//      z, then y, then x, stable ranks in LDS) and computes bounding boxes;
//   4. one wave per (i-cell, shift) task searches the j-cells, builds the
//      entry's j-cluster masks, the Newton and topology exclusion masks and the
//      entry split, first counting (pass 1) then writing at scanned offsets (pass 2).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "capi_common.h"
#include "gpu_search.h"
#include "nbnxm_kernels.h"
#include "nbnxm_amd.h"
#include "pairsearch.h"

// No contraction of a multiply and an add into an FMA anywhere in this file:
// the rounded operations must be the host builder's, one by one.
#pragma clang fp contract(off)

namespace nbnxm_amd
{
namespace
{

constexpr int c_cl = NBNXM_CLUSTER_SIZE;
constexpr int c_ncps = NBNXM_NUM_CLUSTER_PER_SUPERCLUSTER;
constexpr int c_cellAtoms = c_cl * c_ncps;
constexpr int c_jg = NBNXM_JGROUP_SIZE;
constexpr int c_numHalfShifts = 14;   // shifts with index <= central, tz/ty/tx in [-1, 1]
constexpr int c_maxEntryClusters = 1024; // j-clusters of one (i-cell, shift) entry
constexpr int c_maxEntryGroups = c_maxEntryClusters / c_jg;
constexpr int c_maxEntryExcl = 32;       // exclusion entries created by one entry
constexpr int c_noEvent = 0x7fffffff;

// Per-system grid geometry, computed on the host exactly as GridSearch::placeAtoms.
struct SysGrid
{
    int ncx, ncy;
    float colSize[2];
    float boxF[3];
    double box[3];
    int colOffset;   // global index of the first column
    int atomOffset;  // global index of the first atom
    int numAtoms;
};

struct Bb
{
    float lower[3];
    float upper[3];
};

// Half-shift order of the host search loop: tz, then ty, then tx, index <= 22.
struct ShiftTable
{
    int t[c_numHalfShifts][3];
};

__device__ __forceinline__ int shiftIndexOf(int tx, int ty, int tz)
{
    return 5 * (3 * (tz + 1) + ty + 1) + tx + 2;
}

// ((0 + dx^2) + dy^2) + dz^2 with separately rounded operations, as the host.
__device__ __forceinline__ float bbDist2(const Bb& a, const float* sh, const Bb& b)
{
    float d2 = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; d++)
    {
        const float al = __fadd_rn(a.lower[d], sh[d]);
        const float au = __fadd_rn(a.upper[d], sh[d]);
        const float dl = __fsub_rn(al, b.upper[d]);
        const float du = __fsub_rn(b.lower[d], au);
        const float dd = fmaxf(fmaxf(dl, du), 0.0f);
        d2 = __fadd_rn(d2, __fmul_rn(dd, dd));
    }
    return d2;
}

__device__ __forceinline__ Bb loadBb(const float* p)
{
    Bb b;
    b.lower[0] = p[0];
    b.lower[1] = p[1];
    b.lower[2] = p[2];
    b.upper[0] = p[3];
    b.upper[1] = p[4];
    b.upper[2] = p[5];
    return b;
}

__device__ __forceinline__ int systemOf(const int* offsets, int numSystems, int index)
{
    int lo = 0, hi = numSystems - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) / 2;
        if (offsets[mid] <= index)
        {
            lo = mid;
        }
        else
        {
            hi = mid - 1;
        }
    }
    return lo;
}

// Step 1: wrapped FP32 coordinates and the (column, z) sort key of every atom.
__global__ void wrapAndKeyKernel(const double* x, int numAtomsTotal, const int* sysAtomStart, const SysGrid* grids, int numSystems, float4* xw, unsigned long long* keys, int* vals)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= numAtomsTotal)
    {
        return;
    }
    const int sys = systemOf(sysAtomStart, numSystems, a);
    const SysGrid g = grids[sys];
    float w[3];
#pragma unroll
    for (int d = 0; d < 3; d++)
    {
        const double L = g.box[d];
        double v = x[3 * static_cast<size_t>(a) + d];
        v = __dsub_rn(v, __dmul_rn(floor(__ddiv_rn(v, L)), L));
        float vf = static_cast<float>(v);
        if (!(vf < static_cast<float>(L)) || vf < 0.0f)
        {
            vf = 0.0f;
        }
        w[d] = vf;
    }
    xw[a] = make_float4(w[0], w[1], w[2], 0.0f);
    int cx = static_cast<int>(__fdiv_rn(__fmul_rn(w[0], static_cast<float>(g.ncx)), g.boxF[0]));
    int cy = static_cast<int>(__fdiv_rn(__fmul_rn(w[1], static_cast<float>(g.ncy)), g.boxF[1]));
    cx = min(max(cx, 0), g.ncx - 1);
    cy = min(max(cy, 0), g.ncy - 1);
    const unsigned int col = static_cast<unsigned int>(g.colOffset + cx * g.ncy + cy);
    keys[a] = (static_cast<unsigned long long>(col) << 32) | static_cast<unsigned long long>(__float_as_uint(w[2]));
    vals[a] = a;
}

// Step 2: atoms per column.  The keys are sorted by column, so a wave holds
// a few runs of equal columns: the first lane of each run adds the run's
// length (one atomic per run and wave instead of one per atom, which
// serialised on the counter of each column).  Active lanes are lanes
// 0 .. n-1 of the wave (only the last wave of the grid is partial).
__global__ void columnCountKernel(const unsigned long long* sortedKeys, int numAtomsTotal, int* colCount)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool active = i < numAtomsTotal;
    const int col = active ? static_cast<int>(sortedKeys[i] >> 32) : -1;
    const int prevCol = __shfl_up(col, 1, 64);
    const bool head = active && (lane == 0 || col != prevCol);
    const unsigned long long heads = __builtin_amdgcn_ballot_w64(head);
    const int numActive = __popcll(__builtin_amdgcn_ballot_w64(active));
    if (head)
    {
        const unsigned long long later = heads & ~((2ull << lane) - 1ull);
        const int next = later ? (__ffsll(static_cast<long long>(later)) - 1) : numActive;
        atomicAdd(colCount + col, next - lane);
    }
}

__global__ void cellsPerColumnKernel(const int* colCount, int numCols, int* colCells)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= numCols)
    {
        return;
    }
    colCells[c] = (colCount[c] + c_cellAtoms - 1) / c_cellAtoms;
}

// Step 3: one wave per cell.  Lane p holds the atom at position p of the
// cell's z-sorted run; three levels of stable splits (z by count, then y, then
// x) leave cluster k at positions [8k, 8k+8), so slot = cell*64 + position.
__global__ __launch_bounds__(64) void buildCellsKernel(const int* sortedAtoms,
                                                       const float4* xw,
                                                       const int* colAtomStart,
                                                       const int* colCellStart,
                                                       int numCols,
                                                       const int* colSysStart,
                                                       int numSystems,
                                                       int* atomIndex,
                                                       int* cellOfAtom,
                                                       int* numClustersInCell,
                                                       int* cellSys,
                                                       float* clusterBb,
                                                       float* cellBb)
{
    // Launched for an upper bound of the cell count; the count itself is
    // colCellStart[numCols].
    const int cell = blockIdx.x;
    if (cell >= colCellStart[numCols])
    {
        return;
    }
    const int lane = threadIdx.x;
    // Column of this cell: last column whose first cell is <= cell.
    int lo = 0, hi = numCols - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) / 2;
        if (colCellStart[mid] <= cell)
        {
            lo = mid;
        }
        else
        {
            hi = mid - 1;
        }
    }
    // Empty columns share their start with the following column, so the owner
    // is the last column whose first cell is <= cell.
    const int col = lo;
    const int cnt = colAtomStart[col + 1] - colAtomStart[col];
    const int first = (cell - colCellStart[col]) * c_cellAtoms;
    const int m = min(c_cellAtoms, cnt - first);
    const int nclusters = (m + c_cl - 1) / c_cl;

    __shared__ int sAtom[c_cellAtoms];
    __shared__ float sKey[c_cellAtoms];
    const bool real = lane < m;
    int atom = real ? sortedAtoms[colAtomStart[col] + first + lane] : -1;
    float4 p = real ? xw[atom] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);

    // Segment of this lane's position: [segStart, segStart + segCount) with
    // segClusters clusters; level 0 is the whole run, already z-sorted.
    int segStart = 0, segCount = m, segClusters = nclusters;
    for (int level = 0; level < 3; level++)
    {
        if (level > 0)
        {
            // Stable sort of each segment with more than one cluster by y
            // (level 1) or x (level 2); every lane takes part in the barriers.
            const bool sorting = real && segClusters > 1;
            const float key = (level == 1) ? p.y : p.x;
            sKey[lane] = key;
            sAtom[lane] = atom;
            __syncthreads();
            int rank = 0;
            if (sorting)
            {
                for (int q = segStart; q < segStart + segCount; q++)
                {
                    const float kq = sKey[q];
                    rank += (kq < key || (kq == key && q < lane)) ? 1 : 0;
                }
            }
            __syncthreads();
            if (sorting)
            {
                sAtom[segStart + rank] = atom;
            }
            __syncthreads();
            if (sorting)
            {
                atom = sAtom[lane];
                p = xw[atom];
            }
            __syncthreads();
        }
        if (segClusters > 1)
        {
            const int clustersA = (segClusters + 1) / 2;
            const int countA = min(segCount, clustersA * c_cl);
            if (lane - segStart < countA)
            {
                segCount = countA;
                segClusters = clustersA;
            }
            else
            {
                segStart += countA;
                segCount -= countA;
                segClusters -= clustersA;
            }
        }
    }
    const int slot = cell * c_cellAtoms + lane;
    atomIndex[slot] = atom;
    if (real)
    {
        cellOfAtom[atom] = slot;
    }

    // Bounding boxes: lanes of cluster k = lane / 8, reduced over 8 lanes.
    const float big = 3.402823466e+38f;
    float lox = real ? p.x : big, loy = real ? p.y : big, loz = real ? p.z : big;
    float hix = real ? p.x : -big, hiy = real ? p.y : -big, hiz = real ? p.z : -big;
#pragma unroll
    for (int off = 1; off < c_cl; off <<= 1)
    {
        lox = fminf(lox, __shfl_xor(lox, off));
        loy = fminf(loy, __shfl_xor(loy, off));
        loz = fminf(loz, __shfl_xor(loz, off));
        hix = fmaxf(hix, __shfl_xor(hix, off));
        hiy = fmaxf(hiy, __shfl_xor(hiy, off));
        hiz = fmaxf(hiz, __shfl_xor(hiz, off));
    }
    if ((lane & 7) == 0)
    {
        float* b = clusterBb + 6 * (static_cast<size_t>(cell) * c_ncps + lane / c_cl);
        b[0] = lox;
        b[1] = loy;
        b[2] = loz;
        b[3] = hix;
        b[4] = hiy;
        b[5] = hiz;
    }
#pragma unroll
    for (int off = c_cl; off < 64; off <<= 1)
    {
        lox = fminf(lox, __shfl_xor(lox, off));
        loy = fminf(loy, __shfl_xor(loy, off));
        loz = fminf(loz, __shfl_xor(loz, off));
        hix = fmaxf(hix, __shfl_xor(hix, off));
        hiy = fmaxf(hiy, __shfl_xor(hiy, off));
        hiz = fmaxf(hiz, __shfl_xor(hiz, off));
    }
    if (lane == 0)
    {
        float* b = cellBb + 6 * static_cast<size_t>(cell);
        b[0] = lox;
        b[1] = loy;
        b[2] = loz;
        b[3] = hix;
        b[4] = hiy;
        b[5] = hiz;
        numClustersInCell[cell] = nclusters;
        cellSys[cell] = systemOf(colSysStart, numSystems, col);
    }
}

struct SearchParams
{
    const SysGrid* grids;
    ShiftTable shifts;
    const int* numCells;     // the cell count (device; tasks are launched for an upper bound)
    float rlist;
    float rlist2;
    const int* colCellStart;
    const float4* xw;
    const int* atomIndex;
    const int* cellOfAtom;
    const int* numClustersInCell;
    const int* cellSys;
    const float* clusterBb;
    const float* cellBb;
    const int* exclRanges;   // global CSR over all atoms, may be null
    const int* exclElements; // global atom indices
    int maxPairsPerEntry;
    // pass 1
    int* counts;             // groups, exclusion entries, sci records of task t at t, stride + t, 2 stride + t
    int countStride;         // numTasks + 1 (element numTasks of each array stays 0: the scans' totals)
    // pass 2
    const int* groupOffset;
    const int* exclOffset;
    const int* sciOffset;
    nbnxm_sci_t* sci;
    nbnxm_cj_packed_t* cjPacked;
    nbnxm_excl_t* excl;
    int* error;
};

// Step 4: one wave per (i-cell, half shift).
template<bool write>
__global__ __launch_bounds__(64) void searchKernel(const SearchParams P)
{
    const int task = blockIdx.x;
    const int icell = task / c_numHalfShifts;
    const int sIdx = task % c_numHalfShifts;
    if (icell >= *P.numCells)
    {
        return;
    }
    const int lane = threadIdx.x;
    __shared__ int sCj[c_maxEntryClusters];
    __shared__ unsigned char sMask[c_maxEntryClusters];
    __shared__ int sFirstEvent[c_maxEntryGroups];
    __shared__ int sExclSlot[c_maxEntryGroups];
    __shared__ int sEventGroups[c_maxEntryExcl];
    __shared__ unsigned int sExcl[c_maxEntryExcl][NBNXM_EXCL_SIZE];
    __shared__ int sNumEvents;
    __shared__ int sOverflow;

    const int tx = P.shifts.t[sIdx][0];
    const int ty = P.shifts.t[sIdx][1];
    const int tz = P.shifts.t[sIdx][2];
    const int shift = shiftIndexOf(tx, ty, tz);
    const bool central = (shift == NBNXM_CENTRAL_SHIFT);
    const int sys = P.cellSys[icell];
    const SysGrid g = P.grids[sys];
    const float sh[3] = { __fmul_rn(static_cast<float>(tx), g.boxF[0]), __fmul_rn(static_cast<float>(ty), g.boxF[1]),
                          __fmul_rn(static_cast<float>(tz), g.boxF[2]) };
    const float rl = P.rlist;
    const float rlist2 = P.rlist2;
    const Bb ib = loadBb(P.cellBb + 6 * static_cast<size_t>(icell));
    const float il[3] = { __fadd_rn(ib.lower[0], sh[0]), __fadd_rn(ib.lower[1], sh[1]), __fadd_rn(ib.lower[2], sh[2]) };
    const float iu[3] = { __fadd_rn(ib.upper[0], sh[0]), __fadd_rn(ib.upper[1], sh[1]), __fadd_rn(ib.upper[2], sh[2]) };

    int len = 0;
    if (lane == 0)
    {
        sOverflow = 0;
    }
    int cxf = static_cast<int>(floorf(__fdiv_rn(__fsub_rn(il[0], rl), g.colSize[0])));
    int cxl = static_cast<int>(floorf(__fdiv_rn(__fadd_rn(iu[0], rl), g.colSize[0])));
    int cyf = static_cast<int>(floorf(__fdiv_rn(__fsub_rn(il[1], rl), g.colSize[1])));
    int cyl = static_cast<int>(floorf(__fdiv_rn(__fadd_rn(iu[1], rl), g.colSize[1])));
    const bool anyColumn = !(cxl < 0 || cyl < 0 || cxf >= g.ncx || cyf >= g.ncy);
    cxf = max(cxf, 0);
    cyf = max(cyf, 0);
    cxl = min(cxl, g.ncx - 1);
    cyl = min(cyl, g.ncy - 1);
    const int niClusters = P.numClustersInCell[icell];
    const int ci = lane & 7;
    const int subcLane = lane >> 3;
    const Bb iClusterBb = loadBb(P.clusterBb + 6 * (static_cast<size_t>(icell) * c_ncps + ci));
    __syncthreads();

    if (anyColumn)
    {
        const float zlo = __fsub_rn(il[2], rl);
        const float zhi = __fadd_rn(iu[2], rl);
        for (int cx = cxf; cx <= cxl; cx++)
        {
            for (int cy = cyf; cy <= cyl; cy++)
            {
                const int gcol = g.colOffset + cx * g.ncy + cy;
                const int c0 = P.colCellStart[gcol];
                const int c1 = P.colCellStart[gcol + 1];
                if (c0 == c1 || (central && c1 <= icell))
                {
                    continue;
                }
                int lo = c0, hi = c1;
                while (lo < hi)
                {
                    const int mid = (lo + hi) / 2;
                    if (P.cellBb[6 * static_cast<size_t>(mid) + 5] < zlo)
                    {
                        lo = mid + 1;
                    }
                    else
                    {
                        hi = mid;
                    }
                }
                int firstCell = lo;
                if (central)
                {
                    firstCell = max(firstCell, icell);
                }
                for (int jcell = firstCell; jcell < c1; jcell++)
                {
                    const Bb jb = loadBb(P.cellBb + 6 * static_cast<size_t>(jcell));
                    if (jb.lower[2] > zhi)
                    {
                        break;
                    }
                    if (bbDist2(ib, sh, jb) >= rlist2)
                    {
                        continue;
                    }
                    const bool sameCell = central && jcell == icell;
                    const int ncj = P.numClustersInCell[jcell];
                    const int ciEnd = sameCell ? subcLane + 1 : niClusters;
                    bool inr = false;
                    if (subcLane < ncj && ci < ciEnd)
                    {
                        const Bb jcb = loadBb(P.clusterBb + 6 * (static_cast<size_t>(jcell) * c_ncps + subcLane));
                        inr = bbDist2(iClusterBb, sh, jcb) < rlist2;
                    }
                    const unsigned long long m64 = __ballot(inr);
                    for (int subc = 0; subc < ncj; subc++)
                    {
                        unsigned int mask = static_cast<unsigned int>((m64 >> (subc * c_cl)) & 0xFFull);
                        if (mask == 0u)
                        {
                            continue;
                        }
                        const int cj = jcell * c_ncps + subc;
                        if (__popc(mask) == 1)
                        {
                            // A lone cluster pair is kept only with an atom pair in range.
                            const int last = __ffs(mask) - 1;
                            const int ti = lane & 7;
                            const int tj = lane >> 3;
                            const int a = P.atomIndex[(icell * c_ncps + last) * c_cl + ti];
                            const int b = P.atomIndex[cj * c_cl + tj];
                            bool close = false;
                            if (a >= 0 && b >= 0)
                            {
                                const float4 xa = P.xw[a];
                                const float4 xb = P.xw[b];
                                const float dx = __fsub_rn(__fadd_rn(xa.x, sh[0]), xb.x);
                                const float dy = __fsub_rn(__fadd_rn(xa.y, sh[1]), xb.y);
                                const float dz = __fsub_rn(__fadd_rn(xa.z, sh[2]), xb.z);
                                close = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz)) < rlist2;
                            }
                            if (__ballot(close) == 0ull)
                            {
                                mask = 0u;
                            }
                        }
                        if (mask != 0u)
                        {
                            if (len < c_maxEntryClusters)
                            {
                                if (lane == 0)
                                {
                                    sCj[len] = cj;
                                    sMask[len] = static_cast<unsigned char>(mask);
                                }
                            }
                            else if (lane == 0)
                            {
                                sOverflow = 1;
                            }
                            len++;
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    if (len == 0)
    {
        if (!write && lane == 0)
        {
            P.counts[task] = 0;
            P.counts[P.countStride + task] = 0;
            P.counts[2 * P.countStride + task] = 0;
        }
        return;
    }
    if (sOverflow || len > c_maxEntryClusters)
    {
        if (lane == 0)
        {
            atomicOr(P.error, 1);
            if (!write)
            {
                P.counts[task] = 0;
                P.counts[P.countStride + task] = 0;
                P.counts[2 * P.countStride + task] = 0;
            }
        }
        return;
    }
    const int ngroups = (len + c_jg - 1) / c_jg;

    // Exclusion entries: the first event that needs one, per group.  Events are
    // ordered as the host creates them: Newton masks of diagonal clusters in
    // entry order (time k), then topology exclusions by i-slot and element.
    for (int gi = lane; gi < ngroups; gi += 64)
    {
        sFirstEvent[gi] = c_noEvent;
    }
    __syncthreads();
    auto groupMask = [&](int group) {
        unsigned int m = 0u;
        for (int k = 0; k < c_jg; k++)
        {
            const int idx = group * c_jg + k;
            if (idx < len)
            {
                m |= static_cast<unsigned int>(sMask[idx]) << (k * c_ncps);
            }
        }
        return m;
    };
    if (central)
    {
        for (int k = lane; k < len; k += 64)
        {
            if (sCj[k] / c_ncps == icell)
            {
                atomicMin(&sFirstEvent[k / c_jg], k);
            }
        }
    }
    const int slotI = icell * c_cellAtoms + lane;
    const int iAtom = P.atomIndex[slotI];
    const int im = lane / c_cl;
    const bool haveExcl = P.exclRanges != nullptr && iAtom >= 0;
    const int eBegin = haveExcl ? P.exclRanges[iAtom] : 0;
    const int eEnd = haveExcl ? P.exclRanges[iAtom + 1] : 0;
    const int sysLo = g.atomOffset;
    const int sysHi = g.atomOffset + g.numAtoms;
    // Position of a j-cluster in the entry (the list is in ascending cluster order).
    auto findPos = [&](int cluster) {
        int lo = 0, hi = len;
        while (lo < hi)
        {
            const int mid = (lo + hi) / 2;
            if (sCj[mid] < cluster)
            {
                lo = mid + 1;
            }
            else
            {
                hi = mid;
            }
        }
        return (lo < len && sCj[lo] == cluster) ? lo : -1;
    };
    for (int e = eBegin; e < eEnd; e++)
    {
        const int jAtom = P.exclElements[e];
        if (jAtom == iAtom || jAtom < sysLo || jAtom >= sysHi)
        {
            continue;
        }
        const int slotJ = P.cellOfAtom[jAtom];
        if (central && slotJ <= slotI)
        {
            continue;
        }
        const int pos = findPos(slotJ / c_cl);
        if (pos < 0)
        {
            continue;
        }
        const int group = pos / c_jg;
        const unsigned int bit = 1u << ((pos % c_jg) * c_ncps + im);
        if (!(groupMask(group) & bit))
        {
            continue;
        }
        atomicMin(&sFirstEvent[group], len + lane * 65536 + (e - eBegin));
    }
    __syncthreads();
    // Groups with an exclusion entry, ranked by their first event.
    if (lane == 0)
    {
        sNumEvents = 0;
    }
    __syncthreads();
    for (int gi = lane; gi < ngroups; gi += 64)
    {
        sExclSlot[gi] = -1;
        if (sFirstEvent[gi] != c_noEvent)
        {
            const int k = atomicAdd(&sNumEvents, 1);
            if (k < c_maxEntryExcl)
            {
                sEventGroups[k] = gi;
            }
        }
    }
    __syncthreads();
    const int numEvents = sNumEvents;
    if (numEvents > c_maxEntryExcl)
    {
        if (lane == 0)
        {
            atomicOr(P.error, 2);
            if (!write)
            {
                P.counts[task] = 0;
                P.counts[P.countStride + task] = 0;
                P.counts[2 * P.countStride + task] = 0;
            }
        }
        return;
    }
    if (lane < numEvents)
    {
        const int gl = sEventGroups[lane];
        const int t = sFirstEvent[gl];
        int rank = 0;
        for (int q = 0; q < numEvents; q++)
        {
            rank += (sFirstEvent[sEventGroups[q]] < t) ? 1 : 0;
        }
        sExclSlot[gl] = rank;
    }
    __syncthreads();

    // Entry split (split_sci_entry): sequential over groups, lane 0.
    __shared__ int sCuts[c_maxEntryGroups + 1];
    __shared__ int sNumPieces;
    if (lane == 0)
    {
        int pieces = 1;
        sCuts[0] = 0;
        if (P.maxPairsPerEntry > 0 && ngroups > 1)
        {
            int total = 0;
            for (int gi = 0; gi < ngroups; gi++)
            {
                total += __popc(groupMask(gi));
            }
            if (total > P.maxPairsPerEntry)
            {
                const int numP = (total + P.maxPairsPerEntry - 1) / P.maxPairsPerEntry;
                const double target = static_cast<double>(total) / numP;
                int acc = 0;
                for (int gi = 0; gi < ngroups; gi++)
                {
                    const int n = __popc(groupMask(gi));
                    if (acc > 0 && static_cast<double>(acc) + 0.5 * n > target)
                    {
                        sCuts[pieces++] = gi;
                        acc = 0;
                    }
                    acc += n;
                }
            }
        }
        sCuts[pieces] = ngroups;
        sNumPieces = pieces;
    }
    __syncthreads();

    if (!write)
    {
        if (lane == 0)
        {
            P.counts[task] = ngroups;
            P.counts[P.countStride + task] = numEvents;
            P.counts[2 * P.countStride + task] = sNumPieces;
        }
        return;
    }

    // Pass 2: exclusion words in LDS, then everything to its offsets.
    for (int k = 0; k < numEvents; k++)
    {
        sExcl[k][lane] = 0xFFFFFFFFu;
    }
    __syncthreads();
    if (central)
    {
        for (int k = 0; k < len; k++)
        {
            const int cj = sCj[k];
            if (cj / c_ncps != icell)
            {
                continue;
            }
            const int slotK = sExclSlot[k / c_jg];
            const int subc = cj % c_ncps;
            const unsigned int bit = 1u << ((k % c_jg) * c_ncps + subc);
            // word tj*8 + ti, cleared for ti >= tj
            const int tjw = lane >> 3;
            const int tiw = lane & 7;
            if (tiw >= tjw)
            {
                sExcl[slotK][lane] &= ~bit;
            }
        }
    }
    for (int e = eBegin; e < eEnd; e++)
    {
        const int jAtom = P.exclElements[e];
        if (jAtom == iAtom || jAtom < sysLo || jAtom >= sysHi)
        {
            continue;
        }
        const int slotJ = P.cellOfAtom[jAtom];
        if (central && slotJ <= slotI)
        {
            continue;
        }
        const int pos = findPos(slotJ / c_cl);
        if (pos < 0)
        {
            continue;
        }
        const int group = pos / c_jg;
        const unsigned int bit = 1u << ((pos % c_jg) * c_ncps + im);
        if (!(groupMask(group) & bit))
        {
            continue;
        }
        atomicAnd(&sExcl[sExclSlot[group]][(slotJ % c_cl) * c_cl + (slotI % c_cl)], ~bit);
    }
    __syncthreads();

    const int gBase = P.groupOffset[task];
    const int eBase = 1 + P.exclOffset[task];
    const int sBase = P.sciOffset[task];
    for (int gi = lane; gi < ngroups; gi += 64)
    {
        nbnxm_cj_packed_t pk;
        unsigned int m = 0u;
        for (int k = 0; k < c_jg; k++)
        {
            const int idx = gi * c_jg + k;
            pk.cj[k] = sCj[min(idx, len - 1)];
            if (idx < len)
            {
                m |= static_cast<unsigned int>(sMask[idx]) << (k * c_ncps);
            }
        }
        pk.imei[0].imask = m;
        pk.imei[0].excl_ind = (sExclSlot[gi] >= 0) ? eBase + sExclSlot[gi] : 0;
        P.cjPacked[gBase + gi] = pk;
    }
    for (int k = 0; k < numEvents; k++)
    {
        P.excl[eBase + k].pair[lane] = sExcl[k][lane];
    }
    if (lane < sNumPieces)
    {
        nbnxm_sci_t s;
        s.sci = icell;
        s.shift = shift;
        s.cjPackedBegin = gBase + sCuts[lane];
        s.cjPackedEnd = gBase + sCuts[lane + 1];
        P.sci[sBase + lane] = s;
    }
}

// FP32 device coordinates -> the FP64 input of the search (exact).
__global__ void widenCoordinatesKernel(const float* x, int n3, double* xd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3)
    {
        xd[i] = static_cast<double>(x[i]);
    }
}

__global__ void fillAllOnesKernel(nbnxm_excl_t* excl)
{
    excl[0].pair[threadIdx.x] = 0xFFFFFFFFu;
}

} // namespace
} // namespace nbnxm_amd

using namespace nbnxm_amd;

// After the scans of the per-task counts: the list totals (groups, exclusion
// masks less the all-ones one, entries), the error flag, and the first group,
// mask, entry and cell of every system (its first task's offsets; the entry
// after the last system holds the totals, the cell count among them), in one
// small array for one copy to the host.
__global__ void searchSummaryKernel(const int* groupOffset,
                                    const int* exclOffset,
                                    const int* sciOffset,
                                    int numTasks,
                                    const int* colCellStart,
                                    const int* colSysStart,
                                    int numSystems,
                                    const int* error,
                                    int* summary)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0)
    {
        summary[0] = groupOffset[numTasks];
        summary[1] = exclOffset[numTasks];
        summary[2] = sciOffset[numTasks];
        summary[3] = *error;
    }
    if (k <= numSystems)
    {
        const int cellStart = colCellStart[colSysStart[k]];
        const int t = min(cellStart * c_numHalfShifts, numTasks);
        summary[4 + 4 * k + 0] = groupOffset[t];
        summary[4 + 4 * k + 1] = exclOffset[t];
        summary[4 + 4 * k + 2] = sciOffset[t];
        summary[4 + 4 * k + 3] = cellStart;
    }
}

struct nbnxm_gpu_search
{
    hipStream_t stream = nullptr;
    int numSystems = 0;
    std::vector<SysGrid> hostGrids;
    std::vector<int> sysAtomStart;   // numSystems + 1
    std::vector<int> sysColStart;    // numSystems + 1
    std::vector<int> sysCellStart;   // numSystems + 1 (after the build)
    std::vector<int> sysSciStart, sysGroupStart, sysExclStart; // numSystems + 1
    int numAtoms = 0, numCols = 0, numCells = 0, nsci = 0, ngroups = 0, nexcl = 0;
    float rlistF = 0.0f;

    DeviceBuffer<double> x;
    DeviceBuffer<SysGrid> grids;
    DeviceBuffer<int> atomStart, colSysStart;
    DeviceBuffer<float4> xw;
    DeviceBuffer<unsigned long long> keys, keysSorted;
    DeviceBuffer<int> vals, valsSorted;
    DeviceBuffer<char> temp;
    DeviceBuffer<int> colCount, colCells, colCellStart, colAtomStart;
    DeviceBuffer<int> atomIndex, cellOfAtom, numClustersInCell, cellSys;
    DeviceBuffer<float> clusterBb, cellBb;
    DeviceBuffer<int> exclRanges, exclElements;
    DeviceBuffer<int> counts, groupOffset, exclOffset, sciOffset, countsCol, summary;
    DeviceBuffer<nbnxm_sci_t> sci;
    DeviceBuffer<nbnxm_cj_packed_t> cjPacked;
    DeviceBuffer<nbnxm_excl_t> excl;
    DeviceBuffer<int> error;

    // Exclusions on the device belong to the single system of this many atoms
    // (device-input builds upload them once; -1: none uploaded that way).
    int exclUploadedFor = -1;
    // The host copies below hold the current build (device-input builds copy
    // only the entries until gpuSearchDownloadViews).
    bool viewsCurrent = false;

    // Host copies for the per-system views.
    std::vector<nbnxm_sci_t> hSci;
    std::vector<nbnxm_cj_packed_t> hCj;
    std::vector<nbnxm_excl_t> hExcl;
    std::vector<int> hAtomIndex;
    std::vector<float4> hXw;

    ~nbnxm_gpu_search()
    {
        if (stream)
        {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
};

namespace
{

// Grid geometry of one system, as GridSearch::placeAtoms computes it.
SysGrid hostGrid(const double box[3], int n)
{
    SysGrid g{};
    const double volume = box[0] * box[1] * box[2];
    const double density = n > 0 ? n / volume : 1.0;
    const double edge = std::cbrt(c_cellAtoms / density);
    g.ncx = std::max(1, static_cast<int>(std::floor(box[0] / edge + 0.5)));
    g.ncy = std::max(1, static_cast<int>(std::floor(box[1] / edge + 0.5)));
    g.colSize[0] = static_cast<float>(box[0] / g.ncx);
    g.colSize[1] = static_cast<float>(box[1] / g.ncy);
    for (int d = 0; d < 3; d++)
    {
        g.box[d] = box[d];
        g.boxF[d] = static_cast<float>(box[d]);
    }
    return g;
}

template<typename Tin, typename Tout>
hipError_t exclusiveScan(nbnxm_gpu_search* s, const Tin* in, Tout* out, int n)
{
    size_t bytes = 0;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s->stream);
    if (e != hipSuccess)
    {
        return e;
    }
    if ((e = s->temp.resize(bytes + 16)) != hipSuccess)
    {
        return e;
    }
    return hipcub::DeviceScan::ExclusiveSum(s->temp.ptr, bytes, in, out, n, s->stream);
}

// Host copies of the lists, slot map and wrapped coordinates of the last
// build (what nbnxm_gpu_search_system_list and the export read).
int downloadViewsOf(nbnxm_gpu_search* s)
{
    hipStream_t st = s->stream;
    s->hSci.resize(s->nsci);
    s->hCj.resize(s->ngroups);
    s->hExcl.resize(s->nexcl);
    s->hAtomIndex.resize(static_cast<size_t>(s->numCells) * c_cellAtoms);
    if (s->nsci)
    {
        NBNXM_HIP_CHECK(hipMemcpyAsync(s->hSci.data(), s->sci.ptr, sizeof(nbnxm_sci_t) * s->nsci, hipMemcpyDeviceToHost, st));
    }
    if (s->ngroups)
    {
        NBNXM_HIP_CHECK(hipMemcpyAsync(s->hCj.data(), s->cjPacked.ptr, sizeof(nbnxm_cj_packed_t) * s->ngroups, hipMemcpyDeviceToHost, st));
    }
    NBNXM_HIP_CHECK(hipMemcpyAsync(s->hExcl.data(), s->excl.ptr, sizeof(nbnxm_excl_t) * s->nexcl, hipMemcpyDeviceToHost, st));
    NBNXM_HIP_CHECK(hipMemcpyAsync(s->hAtomIndex.data(), s->atomIndex.ptr, sizeof(int) * s->hAtomIndex.size(), hipMemcpyDeviceToHost, st));
    s->hXw.resize(s->numAtoms);
    NBNXM_HIP_CHECK(hipMemcpyAsync(s->hXw.data(), s->xw.ptr, sizeof(float4) * s->numAtoms, hipMemcpyDeviceToHost, st));
    NBNXM_HIP_CHECK(hipStreamSynchronize(st));
    s->viewsCurrent = true;
    return NBNXM_OK;
}

// The device part of a build, once the systems' host grids (hostGrids,
// sysAtomStart, sysColStart), the FP64 coordinates (x) and, with anyExcl,
// the exclusions are in place: wrap and sort, cells and clusters, the list
// (count, scan on the device, write); downloadViews also copies everything the
// per-system views read to the host.
int runSearch(nbnxm_gpu_search* s, int numSystems, bool anyExcl, double rlist, int maxPairsPerEntry, bool downloadViews)
{
    hipStream_t st = s->stream;
    const int N = s->numAtoms;
    const int ncol = s->numCols;
    // 1. wrap + keys, stable sort by (column, z).
    NBNXM_HIP_CHECK(s->xw.resize(N));
    NBNXM_HIP_CHECK(s->keys.resize(N));
    NBNXM_HIP_CHECK(s->keysSorted.resize(N));
    NBNXM_HIP_CHECK(s->vals.resize(N));
    NBNXM_HIP_CHECK(s->valsSorted.resize(N));
    hipLaunchKernelGGL(wrapAndKeyKernel, dim3((N + 255) / 256), dim3(256), 0, st, s->x.ptr, N, s->atomStart.ptr, s->grids.ptr, numSystems, s->xw.ptr, s->keys.ptr, s->vals.ptr);
    NBNXM_HIP_CHECK(hipGetLastError());
    int colBits = 1;
    while ((1 << colBits) < ncol + 1)
    {
        colBits++;
    }
    {
        size_t bytes = 0;
        NBNXM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, s->keys.ptr, s->keysSorted.ptr, s->vals.ptr, s->valsSorted.ptr, N, 0, 32 + colBits, st));
        NBNXM_HIP_CHECK(s->temp.resize(bytes + 16));
        NBNXM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(s->temp.ptr, bytes, s->keys.ptr, s->keysSorted.ptr, s->vals.ptr, s->valsSorted.ptr, N, 0, 32 + colBits, st));
    }

    // 2. columns -> cells.
    NBNXM_HIP_CHECK(s->colCount.resize(ncol + 1));
    NBNXM_HIP_CHECK(s->colCells.resize(ncol + 1));
    NBNXM_HIP_CHECK(s->colCellStart.resize(ncol + 1));
    NBNXM_HIP_CHECK(s->colAtomStart.resize(ncol + 1));
    NBNXM_HIP_CHECK(launchFillI32(s->colCount.ptr, ncol + 1, 0, st));
    hipLaunchKernelGGL(columnCountKernel, dim3((N + 255) / 256), dim3(256), 0, st, s->keysSorted.ptr, N, s->colCount.ptr);
    NBNXM_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(cellsPerColumnKernel, dim3((ncol + 1 + 255) / 256), dim3(256), 0, st, s->colCount.ptr, ncol + 1, s->colCells.ptr);
    NBNXM_HIP_CHECK(hipGetLastError());
    NBNXM_HIP_CHECK(exclusiveScan(s, s->colCount.ptr, s->colAtomStart.ptr, ncol + 1));
    NBNXM_HIP_CHECK(exclusiveScan(s, s->colCells.ptr, s->colCellStart.ptr, ncol + 1));
    // The cell count stays on the device until the list sizes come back:
    // cells, tasks and their buffers are sized for an upper bound (a column
    // of n atoms has ceil(n / 64) cells), and the kernels skip the cells and
    // tasks beyond colCellStart[ncol].
    const int maxCells = ncol + (N + c_cellAtoms - 1) / c_cellAtoms;

    // 3. cells and clusters.
    NBNXM_HIP_CHECK(s->atomIndex.resize(static_cast<size_t>(maxCells) * c_cellAtoms));
    NBNXM_HIP_CHECK(s->cellOfAtom.resize(N));
    NBNXM_HIP_CHECK(s->numClustersInCell.resize(maxCells));
    NBNXM_HIP_CHECK(s->cellSys.resize(maxCells));
    NBNXM_HIP_CHECK(s->clusterBb.resize(6 * static_cast<size_t>(maxCells) * c_ncps));
    NBNXM_HIP_CHECK(s->cellBb.resize(6 * static_cast<size_t>(maxCells)));
    hipLaunchKernelGGL(buildCellsKernel, dim3(maxCells), dim3(64), 0, st, s->valsSorted.ptr, s->xw.ptr, s->colAtomStart.ptr, s->colCellStart.ptr, ncol,
                       s->colSysStart.ptr, numSystems, s->atomIndex.ptr, s->cellOfAtom.ptr, s->numClustersInCell.ptr, s->cellSys.ptr,
                       s->clusterBb.ptr, s->cellBb.ptr);
    NBNXM_HIP_CHECK(hipGetLastError());

    // 4. search: count, scan, write.
    const int numTasks = maxCells * c_numHalfShifts;
    SearchParams P{};
    P.grids = s->grids.ptr;
    {
        int k = 0;
        for (int tz = -1; tz <= 1; tz++)
        {
            for (int ty = -1; ty <= 1; ty++)
            {
                for (int tx = -1; tx <= 1; tx++)
                {
                    if (5 * (3 * (tz + 1) + ty + 1) + tx + 2 > NBNXM_CENTRAL_SHIFT)
                    {
                        continue;
                    }
                    P.shifts.t[k][0] = tx;
                    P.shifts.t[k][1] = ty;
                    P.shifts.t[k][2] = tz;
                    k++;
                }
            }
        }
    }
    P.numCells = s->colCellStart.ptr + ncol;
    P.rlist = static_cast<float>(rlist);
    P.rlist2 = P.rlist * P.rlist;
    P.colCellStart = s->colCellStart.ptr;
    P.xw = s->xw.ptr;
    P.atomIndex = s->atomIndex.ptr;
    P.cellOfAtom = s->cellOfAtom.ptr;
    P.numClustersInCell = s->numClustersInCell.ptr;
    P.cellSys = s->cellSys.ptr;
    P.clusterBb = s->clusterBb.ptr;
    P.cellBb = s->cellBb.ptr;
    P.exclRanges = anyExcl ? s->exclRanges.ptr : nullptr;
    P.exclElements = anyExcl ? s->exclElements.ptr : nullptr;
    P.maxPairsPerEntry = maxPairsPerEntry;
    const int stride = numTasks + 1;
    NBNXM_HIP_CHECK(s->counts.resize(3 * static_cast<size_t>(stride)));
    NBNXM_HIP_CHECK(s->error.resize(1));
    NBNXM_HIP_CHECK(launchFillI32(s->error.ptr, 1, 0, st));
    NBNXM_HIP_CHECK(launchFillI32(s->counts.ptr, 3 * static_cast<size_t>(stride), 0, st));
    P.counts = s->counts.ptr;
    P.countStride = stride;
    P.error = s->error.ptr;
    hipLaunchKernelGGL(searchKernel<false>, dim3(numTasks), dim3(64), 0, st, P);
    NBNXM_HIP_CHECK(hipGetLastError());

    // Offsets of every task in the three outputs: scans on the device; only
    // the totals, the error flag and the systems' first offsets come back.
    NBNXM_HIP_CHECK(s->groupOffset.resize(stride));
    NBNXM_HIP_CHECK(s->exclOffset.resize(stride));
    NBNXM_HIP_CHECK(s->sciOffset.resize(stride));
    NBNXM_HIP_CHECK(exclusiveScan(s, s->counts.ptr, s->groupOffset.ptr, stride));
    NBNXM_HIP_CHECK(exclusiveScan(s, s->counts.ptr + stride, s->exclOffset.ptr, stride));
    NBNXM_HIP_CHECK(exclusiveScan(s, s->counts.ptr + 2 * stride, s->sciOffset.ptr, stride));
    const int summarySize = 4 + 4 * (numSystems + 1);
    NBNXM_HIP_CHECK(s->summary.resize(summarySize));
    hipLaunchKernelGGL(searchSummaryKernel, dim3((numSystems + 1 + 63) / 64), dim3(64), 0, st, s->groupOffset.ptr, s->exclOffset.ptr, s->sciOffset.ptr,
                       numTasks, s->colCellStart.ptr, s->colSysStart.ptr, numSystems, s->error.ptr, s->summary.ptr);
    NBNXM_HIP_CHECK(hipGetLastError());
    std::vector<int> hSummary(summarySize);
    NBNXM_HIP_CHECK(hipMemcpyAsync(hSummary.data(), s->summary.ptr, sizeof(int) * summarySize, hipMemcpyDeviceToHost, st));
    NBNXM_HIP_CHECK(hipStreamSynchronize(st));
    const int hError = hSummary[3];
    if (hError)
    {
        return fail(NBNXM_ERR_UNSUPPORTED, (hError & 1) ? "gpu_search_build: an entry exceeds 1024 j-clusters" : "gpu_search_build: an entry needs more than 32 exclusion masks");
    }
    s->ngroups = hSummary[0];
    s->nexcl = hSummary[1] + 1;
    s->nsci = hSummary[2];
    // Per-system ranges of entries, groups, exclusion masks and cells.
    s->sysSciStart.assign(numSystems + 1, 0);
    s->sysGroupStart.assign(numSystems + 1, 0);
    s->sysExclStart.assign(numSystems + 1, 0);
    s->sysCellStart.assign(numSystems + 1, 0);
    for (int k = 0; k <= numSystems; k++)
    {
        s->sysGroupStart[k] = hSummary[4 + 4 * k + 0];
        s->sysExclStart[k] = hSummary[4 + 4 * k + 1];
        s->sysSciStart[k] = hSummary[4 + 4 * k + 2];
        s->sysCellStart[k] = hSummary[4 + 4 * k + 3];
    }
    s->numCells = s->sysCellStart[numSystems];
    NBNXM_HIP_CHECK(s->sci.resize(std::max(s->nsci, 1)));
    NBNXM_HIP_CHECK(s->cjPacked.resize(std::max(s->ngroups, 1)));
    NBNXM_HIP_CHECK(s->excl.resize(s->nexcl));
    hipLaunchKernelGGL(fillAllOnesKernel, dim3(1), dim3(NBNXM_EXCL_SIZE), 0, st, s->excl.ptr);
    NBNXM_HIP_CHECK(hipGetLastError());
    P.groupOffset = s->groupOffset.ptr;
    P.exclOffset = s->exclOffset.ptr;
    P.sciOffset = s->sciOffset.ptr;
    P.sci = s->sci.ptr;
    P.cjPacked = s->cjPacked.ptr;
    P.excl = s->excl.ptr;
    hipLaunchKernelGGL(searchKernel<true>, dim3(numTasks), dim3(64), 0, st, P);
    NBNXM_HIP_CHECK(hipGetLastError());

    // Host copies for the per-system views, unless deferred: a list that
    // stays on the device is not waited for here (its users order their
    // work after this stream's).
    s->rlistF = static_cast<float>(rlist);
    s->viewsCurrent = false;
    if (downloadViews)
    {
        return downloadViewsOf(s);
    }
    return NBNXM_OK;
}

} // namespace

extern "C" int nbnxm_gpu_search_create(nbnxm_gpu_search_t** out)
{
    if (!out)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_search_create: null pointer");
    }
    auto s = std::make_unique<nbnxm_gpu_search>();
    NBNXM_HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    *out = s.release();
    return NBNXM_OK;
}

extern "C" void nbnxm_gpu_search_free(nbnxm_gpu_search_t* s)
{
    delete s;
}

extern "C" int nbnxm_gpu_search_build(nbnxm_gpu_search_t* s, int numSystems, const nbnxm_system_desc_t* systems, double rlist, int maxPairsPerEntry)
{
    if (!s || numSystems < 1 || !systems || !(rlist > 0))
    {
        return fail(NBNXM_ERR_INVALID, "gpu_search_build: null pointer, no systems or rlist <= 0");
    }
    s->numSystems = numSystems;
    s->hostGrids.resize(numSystems);
    s->sysAtomStart.assign(numSystems + 1, 0);
    s->sysColStart.assign(numSystems + 1, 0);
    bool anyExcl = false;
    for (int k = 0; k < numSystems; k++)
    {
        const nbnxm_system_desc_t& d = systems[k];
        if (d.numAtoms < 0 || (d.numAtoms > 0 && !d.x))
        {
            return fail(NBNXM_ERR_INVALID, "gpu_search_build: bad system description");
        }
        for (int dd = 0; dd < 3; dd++)
        {
            if (!(d.box[dd] > rlist))
            {
                return fail(NBNXM_ERR_INVALID, "gpu_search_build: every box edge must exceed rlist");
            }
        }
        SysGrid g = hostGrid(d.box, d.numAtoms);
        g.atomOffset = s->sysAtomStart[k];
        g.numAtoms = d.numAtoms;
        g.colOffset = s->sysColStart[k];
        s->hostGrids[k] = g;
        s->sysAtomStart[k + 1] = s->sysAtomStart[k] + d.numAtoms;
        s->sysColStart[k + 1] = s->sysColStart[k] + g.ncx * g.ncy;
        anyExcl = anyExcl || (d.exclRanges != nullptr);
    }
    const int N = s->sysAtomStart[numSystems];
    const int ncol = s->sysColStart[numSystems];
    s->numAtoms = N;
    s->numCols = ncol;
    if (N == 0)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_search_build: no atoms");
    }

    // Concatenated coordinates and exclusions (global atom indices).
    std::vector<double> hx(3 * static_cast<size_t>(N));
    std::vector<int> hr, he;
    if (anyExcl)
    {
        hr.assign(N + 1, 0);
    }
    for (int k = 0; k < numSystems; k++)
    {
        const nbnxm_system_desc_t& d = systems[k];
        std::copy(d.x, d.x + 3 * static_cast<size_t>(d.numAtoms), hx.begin() + 3 * static_cast<size_t>(s->sysAtomStart[k]));
        if (anyExcl)
        {
            for (int i = 0; i < d.numAtoms; i++)
            {
                const int a = s->sysAtomStart[k] + i;
                hr[a] = static_cast<int>(he.size());
                if (d.exclRanges)
                {
                    if (!d.exclElements && d.exclRanges[i + 1] > d.exclRanges[i])
                    {
                        return fail(NBNXM_ERR_INVALID, "gpu_search_build: exclusion ranges without elements");
                    }
                    for (int e = d.exclRanges[i]; e < d.exclRanges[i + 1]; e++)
                    {
                        const int j = d.exclElements[e];
                        // Out-of-range elements are skipped by the search, as on the host.
                        he.push_back((j >= 0 && j < d.numAtoms) ? s->sysAtomStart[k] + j : -1);
                    }
                }
            }
        }
    }
    if (anyExcl)
    {
        hr[N] = static_cast<int>(he.size());
        if (he.empty())
        {
            he.push_back(-1);
        }
    }

    hipStream_t st = s->stream;
    NBNXM_HIP_CHECK(s->x.resize(hx.size()));
    NBNXM_HIP_CHECK(hipMemcpyAsync(s->x.ptr, hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice, st));
    NBNXM_HIP_CHECK(s->grids.resize(numSystems));
    NBNXM_HIP_CHECK(hipMemcpyAsync(s->grids.ptr, s->hostGrids.data(), sizeof(SysGrid) * numSystems, hipMemcpyHostToDevice, st));
    NBNXM_HIP_CHECK(s->atomStart.resize(numSystems + 1));
    NBNXM_HIP_CHECK(hipMemcpyAsync(s->atomStart.ptr, s->sysAtomStart.data(), sizeof(int) * (numSystems + 1), hipMemcpyHostToDevice, st));
    NBNXM_HIP_CHECK(s->colSysStart.resize(numSystems + 1));
    NBNXM_HIP_CHECK(hipMemcpyAsync(s->colSysStart.ptr, s->sysColStart.data(), sizeof(int) * (numSystems + 1), hipMemcpyHostToDevice, st));
    if (anyExcl)
    {
        NBNXM_HIP_CHECK(s->exclRanges.resize(hr.size()));
        NBNXM_HIP_CHECK(s->exclElements.resize(he.size()));
        NBNXM_HIP_CHECK(hipMemcpyAsync(s->exclRanges.ptr, hr.data(), sizeof(int) * hr.size(), hipMemcpyHostToDevice, st));
        NBNXM_HIP_CHECK(hipMemcpyAsync(s->exclElements.ptr, he.data(), sizeof(int) * he.size(), hipMemcpyHostToDevice, st));
    }

    const int rcRun = runSearch(s, numSystems, anyExcl, rlist, maxPairsPerEntry, true);
    s->exclUploadedFor = -1;
    return rcRun;
}

extern "C" int nbnxm_gpu_search_num_entries(const nbnxm_gpu_search_t* s, int64_t* counts)
{
    if (!s || !counts)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_search_num_entries: null pointer");
    }
    counts[0] = s->numCells * static_cast<int64_t>(c_cellAtoms);
    counts[1] = s->nsci;
    counts[2] = s->ngroups;
    counts[3] = s->nexcl;
    return NBNXM_OK;
}

extern "C" int nbnxm_gpu_search_system_list(const nbnxm_gpu_search_t* s, int system, nbnxm_sci_t* sci, nbnxm_cj_packed_t* cjPacked, nbnxm_excl_t* excl, int* atomIndices, int* sizes)
{
    if (!s || !sizes || system < 0 || system >= s->numSystems)
    {
        return fail(NBNXM_ERR_INVALID, "gpu_search_system_list: null pointer or bad system");
    }
    if (!s->viewsCurrent)
    {
        return fail(NBNXM_ERR_STATE, "gpu_search_system_list: the last build kept its lists on the device");
    }
    const int cell0 = s->sysCellStart[system], cell1 = s->sysCellStart[system + 1];
    const int sci0 = s->sysSciStart[system], sci1 = s->sysSciStart[system + 1];
    const int g0 = s->sysGroupStart[system], g1 = s->sysGroupStart[system + 1];
    const int e0 = s->sysExclStart[system], e1 = s->sysExclStart[system + 1];
    const int atom0 = s->sysAtomStart[system];
    sizes[0] = (cell1 - cell0) * c_cellAtoms;
    sizes[1] = sci1 - sci0;
    sizes[2] = g1 - g0;
    sizes[3] = e1 - e0 + 1;
    // Local numbering: cells, groups and masks of this system start at 0; the
    // system's masks follow its own all-ones mask 0, as a single-system list.
    if (sci)
    {
        for (int i = sci0; i < sci1; i++)
        {
            nbnxm_sci_t v = s->hSci[i];
            v.sci -= cell0;
            v.cjPackedBegin -= g0;
            v.cjPackedEnd -= g0;
            sci[i - sci0] = v;
        }
    }
    if (cjPacked)
    {
        for (int i = g0; i < g1; i++)
        {
            nbnxm_cj_packed_t v = s->hCj[i];
            for (int k = 0; k < c_jg; k++)
            {
                v.cj[k] -= cell0 * c_ncps;
            }
            if (v.imei[0].excl_ind != 0)
            {
                v.imei[0].excl_ind -= e0;
            }
            cjPacked[i - g0] = v;
        }
    }
    if (excl)
    {
        excl[0] = s->hExcl[0];
        for (int i = e0; i < e1; i++)
        {
            excl[i - e0 + 1] = s->hExcl[1 + i];
        }
    }
    if (atomIndices)
    {
        for (int i = cell0 * c_cellAtoms; i < cell1 * c_cellAtoms; i++)
        {
            const int a = s->hAtomIndex[i];
            atomIndices[i - cell0 * c_cellAtoms] = (a >= 0) ? a - atom0 : -1;
        }
    }
    return NBNXM_OK;
}

namespace nbnxm_amd
{

hipStream_t gpuSearchStream(const nbnxm_gpu_search_t* s)
{
    return s ? s->stream : nullptr;
}

int gpuSearchDownloadViews(nbnxm_gpu_search_t* s)
{
    if (!s || s->numSystems < 1)
    {
        return fail(NBNXM_ERR_STATE, "gpu_search: no build to copy");
    }
    return s->viewsCurrent ? NBNXM_OK : downloadViewsOf(s);
}

int gpuSearchBuildDevice(nbnxm_gpu_search_t* s,
                         int numAtoms,
                         const float* x,
                         hipEvent_t after,
                         const double box[3],
                         const int* exclRanges,
                         const int* exclElements,
                         bool uploadExclusions,
                         double rlist,
                         int maxPairsPerEntry,
                         DeviceListView* view)
{
    if (!s || numAtoms < 1 || !x || !box || !view || !(rlist > 0))
    {
        return fail(NBNXM_ERR_INVALID, "gpu_search_build_device: null pointer or bad size");
    }
    for (int d = 0; d < 3; d++)
    {
        if (!(box[d] > rlist))
        {
            return fail(NBNXM_ERR_INVALID, "gpu_search_build_device: every box edge must exceed rlist");
        }
    }
    const bool anyExcl = exclRanges != nullptr;
    if (anyExcl && !uploadExclusions && s->exclUploadedFor != numAtoms)
    {
        return fail(NBNXM_ERR_STATE, "gpu_search_build_device: no exclusions uploaded for this system");
    }
    hipStream_t st = s->stream;
    if (after)
    {
        NBNXM_HIP_CHECK(hipStreamWaitEvent(st, after, 0));
    }
    s->numSystems = 1;
    s->hostGrids.resize(1);
    SysGrid g = hostGrid(box, numAtoms);
    g.atomOffset = 0;
    g.numAtoms = numAtoms;
    g.colOffset = 0;
    s->hostGrids[0] = g;
    s->sysAtomStart.assign({ 0, numAtoms });
    s->sysColStart.assign({ 0, g.ncx * g.ncy });
    s->numAtoms = numAtoms;
    s->numCols = g.ncx * g.ncy;
    NBNXM_HIP_CHECK(s->grids.resize(1));
    NBNXM_HIP_CHECK(hipMemcpyAsync(s->grids.ptr, s->hostGrids.data(), sizeof(SysGrid), hipMemcpyHostToDevice, st));
    NBNXM_HIP_CHECK(s->atomStart.resize(2));
    NBNXM_HIP_CHECK(hipMemcpyAsync(s->atomStart.ptr, s->sysAtomStart.data(), 2 * sizeof(int), hipMemcpyHostToDevice, st));
    NBNXM_HIP_CHECK(s->colSysStart.resize(2));
    NBNXM_HIP_CHECK(hipMemcpyAsync(s->colSysStart.ptr, s->sysColStart.data(), 2 * sizeof(int), hipMemcpyHostToDevice, st));
    const int n3 = 3 * numAtoms;
    NBNXM_HIP_CHECK(s->x.resize(n3));
    hipLaunchKernelGGL(widenCoordinatesKernel, dim3((n3 + 255) / 256), dim3(256), 0, st, x, n3, s->x.ptr);
    NBNXM_HIP_CHECK(hipGetLastError());
    if (anyExcl && uploadExclusions)
    {
        const int numElements = exclRanges[numAtoms];
        NBNXM_HIP_CHECK(s->exclRanges.resize(numAtoms + 1));
        NBNXM_HIP_CHECK(s->exclElements.resize(std::max(numElements, 1)));
        NBNXM_HIP_CHECK(hipMemcpyAsync(s->exclRanges.ptr, exclRanges, sizeof(int) * (numAtoms + 1), hipMemcpyHostToDevice, st));
        if (numElements > 0)
        {
            NBNXM_HIP_CHECK(hipMemcpyAsync(s->exclElements.ptr, exclElements, sizeof(int) * numElements, hipMemcpyHostToDevice, st));
        }
        s->exclUploadedFor = numAtoms;
    }
    const int rc = runSearch(s, 1, anyExcl, rlist, maxPairsPerEntry, false);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    view->numAtoms = numAtoms;
    view->numSlots = s->numCells * c_cellAtoms;
    view->atomIndex = s->atomIndex.ptr;
    view->xw = s->xw.ptr;
    view->sci = s->sci.ptr;
    view->hostSci = nullptr; // the task derives its entry lists on the device
    view->nsci = s->nsci;
    view->cjPacked = s->cjPacked.ptr;
    view->ncjPacked = s->ngroups;
    view->excl = s->excl.ptr;
    view->nexcl = s->nexcl;
    view->stream = st;
    return NBNXM_OK;
}

int gpuSearchExportSystem(const nbnxm_gpu_search_t* s, int system, GridSearch& grid)
{
    int sizes[4];
    int rc = nbnxm_gpu_search_system_list(s, system, nullptr, nullptr, nullptr, nullptr, sizes);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    std::vector<nbnxm_sci_t> sci(sizes[1]);
    std::vector<nbnxm_cj_packed_t> cj(sizes[2]);
    std::vector<nbnxm_excl_t> excl(sizes[3]);
    std::vector<int> ai(sizes[0]);
    rc = nbnxm_gpu_search_system_list(s, system, sci.data(), cj.data(), excl.data(), ai.data(), sizes);
    if (rc != NBNXM_OK)
    {
        return rc;
    }
    const SysGrid& g = s->hostGrids[system];
    std::vector<float> xw(3 * static_cast<size_t>(g.numAtoms));
    for (int i = 0; i < g.numAtoms; i++)
    {
        const float4 v = s->hXw[g.atomOffset + i];
        xw[3 * i] = v.x;
        xw[3 * i + 1] = v.y;
        xw[3 * i + 2] = v.z;
    }
    grid.adopt(g.box, g.numAtoms, std::move(xw), sizes[0] / c_cellAtoms, std::move(ai), s->rlistF, std::move(sci), std::move(cj), std::move(excl));
    return NBNXM_OK;
}

} // namespace nbnxm_amd
