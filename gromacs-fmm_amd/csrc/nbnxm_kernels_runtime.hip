// Instantiations of the tuned list walk for the run-time flavour class: every
// flavour combination without a class of its own, read from NbFlavourParams
// with wave-uniform branches.  A translation unit of its own so that these
// instantiations build in parallel with the others.
#include "nbnxm_kernel_body.h"

namespace nbnxm_amd
{

hipError_t launchNonbondedRuntimeFlavour(const NbKernelArgs& a,
                                         const NbFlavourParams& flav,
                                         int accumulation,
                                         int vdwType,
                                         bool computeEnergy,
                                         const NbLaunchConfig& cfg,
                                         hipStream_t stream)
{
    return launchTunedFlavour<c_flavRuntime>(a, flav, accumulation, vdwType, computeEnergy, cfg, stream);
}

} // namespace nbnxm_amd
