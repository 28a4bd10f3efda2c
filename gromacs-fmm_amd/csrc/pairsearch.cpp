// Host-side GPU cluster-pair list builder; see pairsearch.h for the semantics
// and the reference functions each step corresponds to.
#include "pairsearch.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

namespace nbnxm_amd
{

namespace
{

constexpr int c_cl = NBNXM_CLUSTER_SIZE;
constexpr int c_ncps = NBNXM_NUM_CLUSTER_PER_SUPERCLUSTER;
constexpr int c_cellAtoms = c_cl * c_ncps;
constexpr int c_jg = NBNXM_JGROUP_SIZE;
constexpr float c_farAway = -1000000.0f; // filler coordinate, as atomdata.cpp:148

inline int shiftIndex(int tx, int ty, int tz)
{
    // src/gromacs/pbcutil/include/gromacs/pbcutil/ishift.h:58-62 (dBoxX=2, dBoxY=dBoxZ=1)
    return 5 * (3 * (tz + 1) + ty + 1) + tx + 2;
}

inline float bbDistance2(const BoundingBox& a, const float* sh, const BoundingBox& b)
{
    float d2 = 0.0f;
    for (int d = 0; d < 3; d++)
    {
        const float al = a.lower[d] + sh[d];
        const float au = a.upper[d] + sh[d];
        float dl = al - b.upper[d];
        float du = b.lower[d] - au;
        float dd = std::max(std::max(dl, du), 0.0f);
        d2 += dd * dd;
    }
    return d2;
}

} // namespace

template<typename T>
static void wrapCoordinates(const double box[3], int n, const T* x, std::vector<float>& xw)
{
    xw.resize(3 * static_cast<size_t>(n));
    for (int i = 0; i < n; i++)
    {
        for (int d = 0; d < 3; d++)
        {
            const double L = box[d];
            double v = static_cast<double>(x[3 * i + d]);
            v -= std::floor(v / L) * L;
            float vf = static_cast<float>(v);
            if (!(vf < static_cast<float>(L)) || vf < 0.0f)
            {
                vf = 0.0f;
            }
            xw[3 * i + d] = vf;
        }
    }
}

int GridSearch::putAtomsOnGrid(const double box[3], int n, const double* x)
{
    if (!box || n < 0 || (n > 0 && !x))
    {
        return NBNXM_ERR_INVALID;
    }
    for (int d = 0; d < 3; d++)
    {
        if (!(box[d] > 0))
        {
            return NBNXM_ERR_INVALID;
        }
        box_[d] = box[d];
    }
    numAtoms_ = n;
    wrapCoordinates(box, n, x, xw_);
    return placeAtoms();
}

int GridSearch::putAtomsOnGrid(const double box[3], int n, const float* x)
{
    if (!box || n < 0 || (n > 0 && !x))
    {
        return NBNXM_ERR_INVALID;
    }
    for (int d = 0; d < 3; d++)
    {
        if (!(box[d] > 0))
        {
            return NBNXM_ERR_INVALID;
        }
        box_[d] = box[d];
    }
    numAtoms_ = n;
    wrapCoordinates(box, n, x, xw_);
    return placeAtoms();
}

// Recursively splits a z-sorted run of atoms into clusters of 8: halves along
// z, then y, then x (the 2x2x2 super-cluster of the GPU grid, grid.cpp:1101-1240).
static void splitIntoClusters(int* atoms, int count, int numClusters, int level, const std::vector<float>& xw, std::vector<std::vector<int>>& out)
{
    if (count <= 0)
    {
        return;
    }
    if (numClusters <= 1 || level >= 3)
    {
        out.emplace_back(atoms, atoms + count);
        return;
    }
    const int dim = 2 - level; // z, y, x
    if (level > 0)
    {
        std::stable_sort(atoms, atoms + count, [&](int a, int b) { return xw[3 * a + dim] < xw[3 * b + dim]; });
    }
    const int clustersA = (numClusters + 1) / 2;
    const int countA = std::min(count, clustersA * c_cl);
    splitIntoClusters(atoms, countA, clustersA, level + 1, xw, out);
    splitIntoClusters(atoms + countA, count - countA, numClusters - clustersA, level + 1, xw, out);
}

int GridSearch::placeAtoms()
{
    const int n = numAtoms_;
    const double volume = box_[0] * box_[1] * box_[2];
    const double density = n > 0 ? n / volume : 1.0;
    // Column edge such that a super-cluster of 64 atoms is about cubic.
    const double edge = std::cbrt(c_cellAtoms / density);
    ncx_ = std::max(1, static_cast<int>(std::floor(box_[0] / edge + 0.5)));
    ncy_ = std::max(1, static_cast<int>(std::floor(box_[1] / edge + 0.5)));
    colSize_[0] = static_cast<float>(box_[0] / ncx_);
    colSize_[1] = static_cast<float>(box_[1] / ncy_);

    // Shift vectors (calc_shifts, src/gromacs/pbcutil/pbc.cpp:1218-1233), in FP32.
    for (int tz = -1, k = 0; tz <= 1; tz++)
    {
        for (int ty = -1; ty <= 1; ty++)
        {
            for (int tx = -2; tx <= 2; tx++, k++)
            {
                shiftVec_[3 * k + 0] = static_cast<float>(tx) * static_cast<float>(box_[0]);
                shiftVec_[3 * k + 1] = static_cast<float>(ty) * static_cast<float>(box_[1]);
                shiftVec_[3 * k + 2] = static_cast<float>(tz) * static_cast<float>(box_[2]);
            }
        }
    }

    const int ncol = ncx_ * ncy_;
    std::vector<int> colOf(n);
    std::vector<int> colCount(ncol + 1, 0);
    for (int i = 0; i < n; i++)
    {
        int cx = static_cast<int>(xw_[3 * i] * ncx_ / static_cast<float>(box_[0]));
        int cy = static_cast<int>(xw_[3 * i + 1] * ncy_ / static_cast<float>(box_[1]));
        cx = std::min(std::max(cx, 0), ncx_ - 1);
        cy = std::min(std::max(cy, 0), ncy_ - 1);
        colOf[i] = cx * ncy_ + cy;
        colCount[colOf[i] + 1]++;
    }
    std::vector<int> colStart(ncol + 1, 0);
    for (int c = 0; c < ncol; c++)
    {
        colStart[c + 1] = colStart[c] + colCount[c + 1];
    }
    std::vector<int> order(n);
    {
        std::vector<int> fill(colStart.begin(), colStart.end() - 1);
        for (int i = 0; i < n; i++)
        {
            order[fill[colOf[i]]++] = i;
        }
    }

    columnCellStart_.assign(ncol + 1, 0);
    for (int c = 0; c < ncol; c++)
    {
        const int cnt = colStart[c + 1] - colStart[c];
        columnCellStart_[c + 1] = columnCellStart_[c] + (cnt + c_cellAtoms - 1) / c_cellAtoms;
    }
    numCells_ = columnCellStart_[ncol];
    atomIndices_.assign(static_cast<size_t>(numCells_) * c_cellAtoms, -1);
    cells_.assign(n, -1);
    numClustersInCell_.assign(numCells_, 0);

    std::vector<std::vector<int>> clusters;
    for (int c = 0; c < ncol; c++)
    {
        int* colAtoms = order.data() + colStart[c];
        const int cnt = colStart[c + 1] - colStart[c];
        std::stable_sort(colAtoms, colAtoms + cnt, [&](int a, int b) { return xw_[3 * a + 2] < xw_[3 * b + 2]; });
        for (int cell = columnCellStart_[c]; cell < columnCellStart_[c + 1]; cell++)
        {
            const int first = (cell - columnCellStart_[c]) * c_cellAtoms;
            const int m = std::min(c_cellAtoms, cnt - first);
            const int nclusters = (m + c_cl - 1) / c_cl;
            clusters.clear();
            splitIntoClusters(colAtoms + first, m, nclusters, 0, xw_, clusters);
            numClustersInCell_[cell] = static_cast<int>(clusters.size());
            for (size_t k = 0; k < clusters.size(); k++)
            {
                for (size_t t = 0; t < clusters[k].size(); t++)
                {
                    const int slot = cell * c_cellAtoms + static_cast<int>(k) * c_cl + static_cast<int>(t);
                    atomIndices_[slot] = clusters[k][t];
                    cells_[clusters[k][t]] = slot;
                }
            }
        }
    }
    computeBoundingBoxes();
    return NBNXM_OK;
}

void GridSearch::computeBoundingBoxes()
{
    const float big = std::numeric_limits<float>::max();
    clusterBB_.resize(static_cast<size_t>(numCells_) * c_ncps);
    cellBB_.resize(numCells_);
    for (int cell = 0; cell < numCells_; cell++)
    {
        BoundingBox cb;
        for (int d = 0; d < 3; d++)
        {
            cb.lower[d] = big;
            cb.upper[d] = -big;
        }
        for (int k = 0; k < c_ncps; k++)
        {
            BoundingBox bb;
            for (int d = 0; d < 3; d++)
            {
                bb.lower[d] = big;
                bb.upper[d] = -big;
            }
            for (int t = 0; t < c_cl; t++)
            {
                const int a = atomIndices_[cell * c_cellAtoms + k * c_cl + t];
                if (a < 0)
                {
                    continue;
                }
                for (int d = 0; d < 3; d++)
                {
                    bb.lower[d] = std::min(bb.lower[d], xw_[3 * a + d]);
                    bb.upper[d] = std::max(bb.upper[d], xw_[3 * a + d]);
                }
            }
            clusterBB_[cell * c_ncps + k] = bb;
            for (int d = 0; d < 3; d++)
            {
                cb.lower[d] = std::min(cb.lower[d], bb.lower[d]);
                cb.upper[d] = std::max(cb.upper[d], bb.upper[d]);
            }
        }
        cellBB_[cell] = cb;
    }
}

bool GridSearch::clusterPairInRange(int ci, const float* sh, int cj, float rlist2) const
{
    for (int ti = 0; ti < c_cl; ti++)
    {
        const int a = atomIndices_[ci * c_cl + ti];
        if (a < 0)
        {
            continue;
        }
        const float xi = xw_[3 * a] + sh[0];
        const float yi = xw_[3 * a + 1] + sh[1];
        const float zi = xw_[3 * a + 2] + sh[2];
        for (int tj = 0; tj < c_cl; tj++)
        {
            const int b = atomIndices_[cj * c_cl + tj];
            if (b < 0)
            {
                continue;
            }
            const float dx = xi - xw_[3 * b];
            const float dy = yi - xw_[3 * b + 1];
            const float dz = zi - xw_[3 * b + 2];
            if (dx * dx + dy * dy + dz * dz < rlist2)
            {
                return true;
            }
        }
    }
    return false;
}

int GridSearch::constructPairlist(double rlist, const int* exclRanges, const int* exclElements, int maxPairsPerEntry)
{
    if (!(rlist > 0))
    {
        return NBNXM_ERR_INVALID;
    }
    for (int d = 0; d < 3; d++)
    {
        // With every edge longer than rlist, images beyond +-1 box are never
        // within rlist, so the 45 shift vectors cover the list.
        if (!(box_[d] > rlist))
        {
            return NBNXM_ERR_INVALID;
        }
    }
    if ((exclRanges == nullptr) != (exclElements == nullptr))
    {
        return NBNXM_ERR_INVALID;
    }
    rlist_ = static_cast<float>(rlist);
    sci_.clear();
    cjPacked_.clear();
    excl_.assign(1, nbnxm_excl_t{});
    for (int w = 0; w < NBNXM_EXCL_SIZE; w++)
    {
        excl_[0].pair[w] = 0xFFFFFFFFu;
    }
    numClusterPairs_ = 0;
    clusterPos_.assign(static_cast<size_t>(numCells_) * c_ncps, -1);

    for (int icell = 0; icell < numCells_; icell++)
    {
        // Half of the shifts: index <= central (c_pbcShiftBackward, pairlist.cpp:3376-3380).
        for (int tz = -1; tz <= 1; tz++)
        {
            for (int ty = -1; ty <= 1; ty++)
            {
                for (int tx = -1; tx <= 1; tx++)
                {
                    if (shiftIndex(tx, ty, tz) > NBNXM_CENTRAL_SHIFT)
                    {
                        continue;
                    }
                    searchCell(icell, tx, ty, tz, exclRanges, exclElements, maxPairsPerEntry);
                }
            }
        }
    }
    return NBNXM_OK;
}

void GridSearch::searchCell(int icell, int tx, int ty, int tz, const int* exclRanges, const int* exclElements, int maxPairsPerEntry)
{
    const int shift = shiftIndex(tx, ty, tz);
    const bool central = (shift == NBNXM_CENTRAL_SHIFT);
    const float* sh = shiftVec_ + 3 * shift;
    const float rl = rlist_;
    const float rlist2 = rl * rl;
    const BoundingBox& ib = cellBB_[icell];
    const float il[3] = { ib.lower[0] + sh[0], ib.lower[1] + sh[1], ib.lower[2] + sh[2] };
    const float iu[3] = { ib.upper[0] + sh[0], ib.upper[1] + sh[1], ib.upper[2] + sh[2] };

    int cxf = static_cast<int>(std::floor((il[0] - rl) / colSize_[0]));
    int cxl = static_cast<int>(std::floor((iu[0] + rl) / colSize_[0]));
    int cyf = static_cast<int>(std::floor((il[1] - rl) / colSize_[1]));
    int cyl = static_cast<int>(std::floor((iu[1] + rl) / colSize_[1]));
    if (cxl < 0 || cyl < 0 || cxf >= ncx_ || cyf >= ncy_)
    {
        return;
    }
    cxf = std::max(cxf, 0);
    cyf = std::max(cyf, 0);
    cxl = std::min(cxl, ncx_ - 1);
    cyl = std::min(cyl, ncy_ - 1);

    entryClusters_.clear();
    entryMask_.clear();
    const int niClusters = numClustersInCell_[icell];

    for (int cx = cxf; cx <= cxl; cx++)
    {
        for (int cy = cyf; cy <= cyl; cy++)
        {
            const int col = cx * ncy_ + cy;
            const int c0 = columnCellStart_[col];
            const int c1 = columnCellStart_[col + 1];
            if (c0 == c1 || (central && c1 <= icell))
            {
                continue;
            }
            // Cells in a column are z-ordered: skip those entirely below range.
            const float zlo = il[2] - rl;
            const float zhi = iu[2] + rl;
            int first = c0;
            {
                int lo = c0, hi = c1;
                while (lo < hi)
                {
                    const int mid = (lo + hi) / 2;
                    if (cellBB_[mid].upper[2] < zlo)
                    {
                        lo = mid + 1;
                    }
                    else
                    {
                        hi = mid;
                    }
                }
                first = lo;
            }
            if (central)
            {
                first = std::max(first, icell);
            }
            for (int jcell = first; jcell < c1; jcell++)
            {
                if (cellBB_[jcell].lower[2] > zhi)
                {
                    break;
                }
                if (bbDistance2(ib, sh, cellBB_[jcell]) >= rlist2)
                {
                    continue;
                }
                const bool sameCell = central && jcell == icell;
                for (int subc = 0; subc < numClustersInCell_[jcell]; subc++)
                {
                    const int cj = jcell * c_ncps + subc;
                    const int ci1 = sameCell ? subc + 1 : niClusters;
                    unsigned int mask = 0;
                    int npair = 0;
                    int last = -1;
                    for (int ci = 0; ci < ci1; ci++)
                    {
                        if (bbDistance2(clusterBB_[icell * c_ncps + ci], sh, clusterBB_[cj]) < rlist2)
                        {
                            mask |= 1u << ci;
                            npair++;
                            last = ci;
                        }
                    }
                    // A lone cluster pair is checked on atom distances so that
                    // j-groups are not filled with empty pairs (pairlist.cpp:1234-1244).
                    if (npair == 1 && !clusterPairInRange(icell * c_ncps + last, sh, cj, rlist2))
                    {
                        mask = 0;
                    }
                    if (mask)
                    {
                        entryClusters_.push_back(cj);
                        entryMask_.push_back(static_cast<unsigned char>(mask));
                    }
                }
            }
        }
    }
    if (entryClusters_.empty())
    {
        return;
    }

    nbnxm_sci_t e;
    e.sci = icell;
    e.shift = shift;
    e.cjPackedBegin = static_cast<int>(cjPacked_.size());
    const int len = static_cast<int>(entryClusters_.size());
    const int ngroups = (len + c_jg - 1) / c_jg;
    e.cjPackedEnd = e.cjPackedBegin + ngroups;
    for (int g = 0; g < ngroups; g++)
    {
        nbnxm_cj_packed_t p;
        p.imei[0].imask = 0;
        p.imei[0].excl_ind = 0;
        for (int k = 0; k < c_jg; k++)
        {
            const int idx = std::min(g * c_jg + k, len - 1);
            p.cj[k] = entryClusters_[idx];
            if (g * c_jg + k < len)
            {
                p.imei[0].imask |= static_cast<unsigned int>(entryMask_[idx]) << (k * c_ncps);
            }
        }
        cjPacked_.push_back(p);
    }
    for (int k = 0; k < len; k++)
    {
        clusterPos_[entryClusters_[k]] = k;
        numClusterPairs_ += __builtin_popcount(entryMask_[k]);
    }

    // Self and Newton exclusions on diagonal cluster pairs (pairlist.cpp:914-939).
    if (central)
    {
        for (int k = 0; k < len; k++)
        {
            const int cj = entryClusters_[k];
            if (cj / c_ncps != icell)
            {
                continue;
            }
            const int subc = cj % c_ncps;
            const int group = e.cjPackedBegin + k / c_jg;
            const int jm = k % c_jg;
            nbnxm_im_ei_t& imei = cjPacked_[group].imei[0];
            if (imei.excl_ind == 0)
            {
                imei.excl_ind = static_cast<int>(excl_.size());
                excl_.push_back(excl_[0]);
            }
            nbnxm_excl_t& ex = excl_[imei.excl_ind];
            for (int tj = 0; tj < c_cl; tj++)
            {
                for (int ti = tj; ti < c_cl; ti++)
                {
                    ex.pair[tj * c_cl + ti] &= ~(1u << (jm * c_ncps + subc));
                }
            }
        }
    }

    sci_.push_back(e);
    if (exclRanges)
    {
        setExclusions(sci_.size() - 1, exclRanges, exclElements);
    }
    for (int k = 0; k < len; k++)
    {
        clusterPos_[entryClusters_[k]] = -1;
    }
    if (maxPairsPerEntry > 0)
    {
        splitEntry(sci_.size() - 1, maxPairsPerEntry);
    }
}

// Topology exclusions for the last entry (setExclusionsForIEntry, pairlist.cpp:1948-2046).
void GridSearch::setExclusions(size_t entry, const int* exclRanges, const int* exclElements)
{
    const nbnxm_sci_t e = sci_[entry];
    const bool central = (e.shift == NBNXM_CENTRAL_SHIFT);
    for (int i = 0; i < c_cellAtoms; i++)
    {
        const int slotI = e.sci * c_cellAtoms + i;
        const int iAtom = atomIndices_[slotI];
        if (iAtom < 0)
        {
            continue;
        }
        const int im = i / c_cl;
        for (int k = exclRanges[iAtom]; k < exclRanges[iAtom + 1]; k++)
        {
            const int jAtom = exclElements[k];
            if (jAtom == iAtom || jAtom < 0 || jAtom >= numAtoms_)
            {
                continue;
            }
            const int slotJ = cells_[jAtom];
            if (central && slotJ <= slotI)
            {
                continue;
            }
            const int pos = clusterPos_[slotJ / c_cl];
            if (pos < 0)
            {
                continue;
            }
            const int group = e.cjPackedBegin + pos / c_jg;
            const int jm = pos % c_jg;
            const unsigned int bit = 1u << (jm * c_ncps + im);
            nbnxm_im_ei_t& imei = cjPacked_[group].imei[0];
            if (!(imei.imask & bit))
            {
                continue;
            }
            if (imei.excl_ind == 0)
            {
                imei.excl_ind = static_cast<int>(excl_.size());
                excl_.push_back(excl_[0]);
            }
            excl_[imei.excl_ind].pair[(slotJ % c_cl) * c_cl + (slotI % c_cl)] &= ~bit;
        }
    }
}

// Splits the last entry into consecutive pieces of at most about
// maxPairsPerEntry cluster pairs (the role of split_sci_entry, pairlist.cpp:2150-2262).
void GridSearch::splitEntry(size_t entry, int maxPairsPerEntry)
{
    const nbnxm_sci_t e = sci_[entry];
    if (e.cjPackedEnd - e.cjPackedBegin <= 1)
    {
        return;
    }
    int total = 0;
    for (int g = e.cjPackedBegin; g < e.cjPackedEnd; g++)
    {
        total += __builtin_popcount(cjPacked_[g].imei[0].imask);
    }
    if (total <= maxPairsPerEntry)
    {
        return;
    }
    const int pieces = (total + maxPairsPerEntry - 1) / maxPairsPerEntry;
    const double target = static_cast<double>(total) / pieces;
    sci_.pop_back();
    nbnxm_sci_t cur = e;
    int acc = 0;
    for (int g = e.cjPackedBegin; g < e.cjPackedEnd; g++)
    {
        const int n = __builtin_popcount(cjPacked_[g].imei[0].imask);
        if (acc > 0 && acc + 0.5 * n > target)
        {
            cur.cjPackedEnd = g;
            sci_.push_back(cur);
            cur.cjPackedBegin = g;
            acc = 0;
        }
        acc += n;
    }
    cur.cjPackedEnd = e.cjPackedEnd;
    sci_.push_back(cur);
}

void GridSearch::fillXq(const double* x, const double* q, float* xq) const
{
    for (int s = 0; s < numGridAtoms(); s++)
    {
        const int a = atomIndices_[s];
        if (a < 0)
        {
            xq[4 * s + 0] = c_farAway;
            xq[4 * s + 1] = c_farAway;
            xq[4 * s + 2] = c_farAway;
            xq[4 * s + 3] = 0.0f;
        }
        else
        {
            xq[4 * s + 0] = x ? static_cast<float>(x[3 * a + 0]) : xw_[3 * a + 0];
            xq[4 * s + 1] = x ? static_cast<float>(x[3 * a + 1]) : xw_[3 * a + 1];
            xq[4 * s + 2] = x ? static_cast<float>(x[3 * a + 2]) : xw_[3 * a + 2];
            xq[4 * s + 3] = q ? static_cast<float>(q[a]) : 0.0f;
        }
    }
}

void GridSearch::fillXq(const float* x, const float* q, float* xq) const
{
    for (int s = 0; s < numGridAtoms(); s++)
    {
        const int a = atomIndices_[s];
        if (a < 0)
        {
            xq[4 * s + 0] = c_farAway;
            xq[4 * s + 1] = c_farAway;
            xq[4 * s + 2] = c_farAway;
            xq[4 * s + 3] = 0.0f;
        }
        else
        {
            xq[4 * s + 0] = x ? x[3 * a + 0] : xw_[3 * a + 0];
            xq[4 * s + 1] = x ? x[3 * a + 1] : xw_[3 * a + 1];
            xq[4 * s + 2] = x ? x[3 * a + 2] : xw_[3 * a + 2];
            xq[4 * s + 3] = q ? q[a] : 0.0f;
        }
    }
}

void GridSearch::fillTypes(const int* type, int dummyType, int* slotType) const
{
    for (int s = 0; s < numGridAtoms(); s++)
    {
        const int a = atomIndices_[s];
        slotType[s] = (a < 0) ? dummyType : type[a];
    }
}

void GridSearch::fillLjComb(const int* type, const float* perTypeComb, float* slotComb) const
{
    for (int s = 0; s < numGridAtoms(); s++)
    {
        const int a = atomIndices_[s];
        slotComb[2 * s + 0] = (a < 0) ? 0.0f : perTypeComb[2 * type[a] + 0];
        slotComb[2 * s + 1] = (a < 0) ? 0.0f : perTypeComb[2 * type[a] + 1];
    }
}

int64_t GridSearch::countPairsWithin(double rc) const
{
    const double rc2 = rc * rc;
    int64_t count = 0;
    for (const nbnxm_sci_t& s : sci_)
    {
        const float* sh = shiftVec_ + 3 * s.shift;
        const bool central = (s.shift == NBNXM_CENTRAL_SHIFT);
        for (int g = s.cjPackedBegin; g < s.cjPackedEnd; g++)
        {
            const nbnxm_cj_packed_t& p = cjPacked_[g];
            const nbnxm_excl_t& ex = excl_[p.imei[0].excl_ind];
            for (int jm = 0; jm < c_jg; jm++)
            {
                for (int im = 0; im < c_ncps; im++)
                {
                    const int bit = jm * c_ncps + im;
                    if (!((p.imei[0].imask >> bit) & 1u))
                    {
                        continue;
                    }
                    const int ci = s.sci * c_ncps + im;
                    const int cj = p.cj[jm];
                    for (int ti = 0; ti < c_cl; ti++)
                    {
                        const int a = atomIndices_[ci * c_cl + ti];
                        if (a < 0)
                        {
                            continue;
                        }
                        for (int tj = 0; tj < c_cl; tj++)
                        {
                            const int b = atomIndices_[cj * c_cl + tj];
                            if (b < 0 || (central && ci == cj && tj <= ti) || !((ex.pair[tj * c_cl + ti] >> bit) & 1u))
                            {
                                continue;
                            }
                            const double dx = static_cast<double>(xw_[3 * a]) + sh[0] - xw_[3 * b];
                            const double dy = static_cast<double>(xw_[3 * a + 1]) + sh[1] - xw_[3 * b + 1];
                            const double dz = static_cast<double>(xw_[3 * a + 2]) + sh[2] - xw_[3 * b + 2];
                            count += (dx * dx + dy * dy + dz * dz < rc2);
                        }
                    }
                }
            }
        }
    }
    return count;
}

void GridSearch::adopt(const double box[3],
                       int numAtoms,
                       std::vector<float>&& wrappedX,
                       int numCells,
                       std::vector<int>&& atomIndices,
                       float rlist,
                       std::vector<nbnxm_sci_t>&& sci,
                       std::vector<nbnxm_cj_packed_t>&& cjPacked,
                       std::vector<nbnxm_excl_t>&& excl)
{
    for (int d = 0; d < 3; d++)
    {
        box_[d] = box[d];
    }
    for (int tz = -1, k = 0; tz <= 1; tz++)
    {
        for (int ty = -1; ty <= 1; ty++)
        {
            for (int tx = -2; tx <= 2; tx++, k++)
            {
                shiftVec_[3 * k + 0] = static_cast<float>(tx) * static_cast<float>(box_[0]);
                shiftVec_[3 * k + 1] = static_cast<float>(ty) * static_cast<float>(box_[1]);
                shiftVec_[3 * k + 2] = static_cast<float>(tz) * static_cast<float>(box_[2]);
            }
        }
    }
    numAtoms_ = numAtoms;
    xw_ = std::move(wrappedX);
    numCells_ = numCells;
    atomIndices_ = std::move(atomIndices);
    cells_.assign(numAtoms, -1);
    for (size_t slot = 0; slot < atomIndices_.size(); slot++)
    {
        if (atomIndices_[slot] >= 0)
        {
            cells_[atomIndices_[slot]] = static_cast<int>(slot);
        }
    }
    rlist_ = rlist;
    sci_ = std::move(sci);
    cjPacked_ = std::move(cjPacked);
    excl_ = std::move(excl);
    numClusterPairs_ = 0;
    for (const nbnxm_cj_packed_t& p : cjPacked_)
    {
        numClusterPairs_ += __builtin_popcount(p.imei[0].imask);
    }
}

nbnxm_pairlist_gpu_t GridSearch::listView() const
{
    nbnxm_pairlist_gpu_t v;
    v.na_ci = c_cl;
    v.rlist = rlist_;
    v.nsci = static_cast<int>(sci_.size());
    v.sci = sci_.data();
    v.ncjPacked = static_cast<int>(cjPacked_.size());
    v.cjPacked = cjPacked_.data();
    v.nexcl = static_cast<int>(excl_.size());
    v.excl = excl_.data();
    return v;
}

} // namespace nbnxm_amd
