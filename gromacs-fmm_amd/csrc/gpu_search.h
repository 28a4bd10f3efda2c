// Internal interface of the device pair search for other parts of the library.
#pragma once

#include "nbnxm_amd.h"
#include "pairsearch.h"

namespace nbnxm_amd
{

// Hands system `system` of a finished device build to a host GridSearch, which
// then holds the same grid and list as after its own putAtomsOnGrid +
// constructPairlist on that system.
int gpuSearchExportSystem(const nbnxm_gpu_search_t* s, int system, GridSearch& grid);

} // namespace nbnxm_amd
