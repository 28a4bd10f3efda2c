// Instantiations of the tuned list walk for three more flavour classes
// compiled with their flavour fixed: analytical Ewald + LJ potential switch,
// analytical Ewald + LJ potential shift with a twin cut-off, and analytical
// Ewald + LJ-PME.  A translation unit of its own so that these instantiations
// build in parallel with the others.
#include "nbnxm_kernel_body.h"

namespace nbnxm_amd
{

hipError_t launchNonbondedFlavouredB(const NbKernelArgs& a,
                                     const NbFlavourParams& flav,
                                     int accumulation,
                                     int vdwType,
                                     bool computeEnergy,
                                     const NbLaunchConfig& cfg,
                                     hipStream_t stream)
{
    switch (flav.flavourClass)
    {
        case c_flavEwaldPotSwitch:
            return launchTunedFlavour<c_flavEwaldPotSwitch>(a, flav, accumulation, vdwType, computeEnergy, cfg, stream);
        case c_flavEwaldTwin: return launchTunedFlavour<c_flavEwaldTwin>(a, flav, accumulation, vdwType, computeEnergy, cfg, stream);
        case c_flavEwaldLjPme: return launchTunedFlavour<c_flavEwaldLjPme>(a, flav, accumulation, vdwType, computeEnergy, cfg, stream);
        default: return hipErrorInvalidValue;
    }
}

} // namespace nbnxm_amd
