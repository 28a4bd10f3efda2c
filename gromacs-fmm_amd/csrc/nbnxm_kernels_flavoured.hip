// Instantiations of the tuned list walk for the flavours other than the tuned
// one: the classes compiled with their flavour fixed (reaction field + LJ
// potential shift, analytical Ewald + LJ force switch, tabulated Ewald + LJ
// potential shift).  A translation unit of its own so that these
// instantiations build in parallel with the tuned and run-time ones.
#include "nbnxm_kernel_body.h"

namespace nbnxm_amd
{

hipError_t launchNonbondedFlavoured(const NbKernelArgs& a,
                                    const NbFlavourParams& flav,
                                    int accumulation,
                                    int vdwType,
                                    bool computeEnergy,
                                    const NbLaunchConfig& cfg,
                                    hipStream_t stream)
{
    switch (flav.flavourClass)
    {
        case c_flavRfPotShift:
            return launchTunedFlavour<c_flavRfPotShift>(a, flav, accumulation, vdwType, computeEnergy, cfg, stream);
        case c_flavEwaldForceSwitch:
            return launchTunedFlavour<c_flavEwaldForceSwitch>(a, flav, accumulation, vdwType, computeEnergy, cfg, stream);
        case c_flavEwaldTabPotShift:
            if (sizeof(float) * flav.ewaldTableSize > c_maxLdsEwaldTableBytes)
            {
                return launchNonbondedRuntimeFlavour(a, flav, accumulation, vdwType, computeEnergy, cfg, stream);
            }
            return launchTunedFlavour<c_flavEwaldTabPotShift>(a, flav, accumulation, vdwType, computeEnergy, cfg, stream);
        case c_flavEwaldPotSwitch:
        case c_flavEwaldTwin:
        case c_flavEwaldLjPme: return launchNonbondedFlavouredB(a, flav, accumulation, vdwType, computeEnergy, cfg, stream);
        case c_flavRuntime: return launchNonbondedRuntimeFlavour(a, flav, accumulation, vdwType, computeEnergy, cfg, stream);
        default: return hipErrorInvalidValue;
    }
}

} // namespace nbnxm_amd
