// Internal interface of the device pair search for other parts of the library.
#pragma once

#include "capi_common.h"
#include "nbnxm_amd.h"
#include "pairsearch.h"

namespace nbnxm_amd
{

// Hands system `system` of a finished device build to a host GridSearch, which
// then holds the same grid and list as after its own putAtomsOnGrid +
// constructPairlist on that system.
int gpuSearchExportSystem(const nbnxm_gpu_search_t* s, int system, GridSearch& grid);

// One system from device coordinates (three floats per atom), read on the
// search stream once `after` has completed (null: no wait).  Its exclusions
// (local ranges and elements, validated by the caller) are uploaded when
// uploadExclusions is set and otherwise reused from the previous call with
// the same atom count.  The grid and list stay on the device (`view`); only
// the entries are copied to the host.  The export above needs
// gpuSearchDownloadViews first.
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
                         DeviceListView* view);

// Host copies of the last build's lists, slot map and wrapped coordinates.
int gpuSearchDownloadViews(nbnxm_gpu_search_t* s);

// The search's stream (device-input builds write their view on it).
hipStream_t gpuSearchStream(const nbnxm_gpu_search_t* s);

} // namespace nbnxm_amd
