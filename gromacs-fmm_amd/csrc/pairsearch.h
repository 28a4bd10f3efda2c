// Host-side producer of the GPU cluster-pair list: grid of 8-atom clusters in
// 2x2x2 super-clusters, bounding-box search over half of the periodic images,
// topology and Newton exclusion masks, and entry splitting for load balance.
//
// Produces exactly the NbnxnPairlistGpu layout (split 1) that the HIP kernels
// consume.  Semantics follow the reference list builder:
//   cluster pairs by bounding-box distance < rlist
//     (src/gromacs/nbnxm/pairlist.cpp:1122-1270, make_cluster_list_supersub),
//   half-shift enumeration with sub-diagonal removal for the central shift
//     (pairlist.cpp:3290-3420),
//   self / Newton masks on diagonal cluster pairs (pairlist.cpp:914-939),
//   topology exclusions for every listed image (pairlist.cpp:1948-2046),
//   entry splitting to a maximum number of cluster pairs (pairlist.cpp:2150-2262).
// The grid geometry and the order of j-clusters are this library's own; the
// kernel results depend on them only through FP32 summation order.
#pragma once

#include <cstdint>
#include <vector>

#include "nbnxm_amd.h"

namespace nbnxm_amd
{

struct BoundingBox
{
    float lower[3];
    float upper[3];
};

class GridSearch
{
public:
    // Wraps x (3*n) into the rectangular box, sorts atoms into the cluster
    // grid and computes bounding boxes.  Every box edge must exceed rlist,
    // which is checked when the list is constructed.
    int putAtomsOnGrid(const double box[3], int n, const double* x);
    int putAtomsOnGrid(const double box[3], int n, const float* x);

    // Builds the list for pairs within rlist.  exclRanges (n+1) and
    // exclElements are the per-atom exclusion lists (ListOfLists<int>), may be
    // null.  maxPairsPerEntry > 0 splits entries with more cluster pairs.
    int constructPairlist(double rlist, const int* exclRanges, const int* exclElements, int maxPairsPerEntry);

    // Grid-order kernel inputs.  xq: 4 per grid slot; fillers at -1e6, q=0.
    void fillXq(const double* x, const double* q, float* xq) const;
    void fillXq(const float* x, const float* q, float* xq) const;
    // Fills per-slot types (fillers get dummyType) and, when perTypeComb is
    // non-null (2 per type), per-slot combination parameters.
    void fillTypes(const int* type, int dummyType, int* slotType) const;
    void fillLjComb(const int* type, const float* perTypeComb, float* slotComb) const;

    int numGridAtoms() const { return numCells_ * 64; }
    int numCells() const { return numCells_; }
    int numAtoms() const { return numAtoms_; }
    const std::vector<int>& atomIndices() const { return atomIndices_; } // slot -> atom or -1
    const std::vector<int>& cells() const { return cells_; }             // atom -> slot
    const float* shiftVec() const { return shiftVec_; }
    const double* box() const { return box_; }

    // Wrapped coordinates of the last putAtomsOnGrid, in original atom order.
    const std::vector<float>& wrappedX() const { return xw_; }

    const std::vector<nbnxm_sci_t>& sci() const { return sci_; }
    const std::vector<nbnxm_cj_packed_t>& cjPacked() const { return cjPacked_; }
    const std::vector<nbnxm_excl_t>& excl() const { return excl_; }
    float rlist() const { return rlist_; }
    int64_t numClusterPairs() const { return numClusterPairs_; }

    nbnxm_pairlist_gpu_t listView() const;

    // Non-excluded atom pairs within rc in the current list (FP64 distances on
    // the wrapped search coordinates); requires the list to cover rc.
    int64_t countPairsWithin(double rc) const;
    // Takes over a grid and list built elsewhere (the device search): the box,
    // wrapped FP32 coordinates in atom order, the slot -> atom map of numCells
    // cells and the list; afterwards every accessor behaves as after
    // putAtomsOnGrid + constructPairlist with the same inputs.
    void adopt(const double box[3],
               int numAtoms,
               std::vector<float>&& wrappedX,
               int numCells,
               std::vector<int>&& atomIndices,
               float rlist,
               std::vector<nbnxm_sci_t>&& sci,
               std::vector<nbnxm_cj_packed_t>&& cjPacked,
               std::vector<nbnxm_excl_t>&& excl);

private:
    int placeAtoms();
    void computeBoundingBoxes();
    void searchCell(int icell, int tx, int ty, int tz, const int* exclRanges, const int* exclElements, int maxPairsPerEntry);
    bool clusterPairInRange(int ci, const float* shift, int cj, float rlist2) const;
    void setExclusions(size_t entry, const int* exclRanges, const int* exclElements);
    void splitEntry(size_t entry, int maxPairsPerEntry);

    double box_[3] = { 0, 0, 0 };
    float shiftVec_[3 * NBNXM_NUM_SHIFTS] = {};
    int numAtoms_ = 0;
    std::vector<float> xw_; // wrapped coordinates, 3 per atom

    int ncx_ = 0, ncy_ = 0;
    float colSize_[2] = { 0, 0 };
    std::vector<int> columnCellStart_; // first cell of each column, size ncol+1
    int numCells_ = 0;
    std::vector<int> atomIndices_;
    std::vector<int> cells_;
    std::vector<int> numClustersInCell_;
    std::vector<BoundingBox> clusterBB_;
    std::vector<BoundingBox> cellBB_;

    float rlist_ = 0;
    std::vector<nbnxm_sci_t> sci_;
    std::vector<nbnxm_cj_packed_t> cjPacked_;
    std::vector<nbnxm_excl_t> excl_;
    int64_t numClusterPairs_ = 0;

    // scratch for building one entry
    std::vector<int> entryClusters_;
    std::vector<unsigned char> entryMask_; // per j-cluster of the entry: 8 i-bits
    std::vector<int> clusterPos_;          // j-cluster -> position in the entry, -1 if absent
};

} // namespace nbnxm_amd
