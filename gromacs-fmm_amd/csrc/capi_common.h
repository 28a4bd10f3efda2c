// Error reporting and small helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "nbnxm_amd.h"

namespace nbnxm_amd
{

void setError(const std::string& msg);

// c_one4PiEps0 in kJ mol^-1 nm e^-2, derived from the same CODATA constants as
// api/legacy/include/gromacs/math/units.h:88-110.
inline double one4PiEps0()
{
    const double epsilon0Si = 8.8541878128e-12;
    const double electronCharge = 1.602176634e-19;
    const double avogadro = 6.02214076e23;
    const double epsilon0 = (epsilon0Si * 1e-9 * 1e3) / (electronCharge * electronCharge * avogadro);
    return 1.0 / (4.0 * M_PI * epsilon0);
}
int fail(int code, const std::string& msg);
int hipFail(hipError_t e, const char* what);

// The HIP stream of a GPU task (for callers inside the library that order
// their own work against it).
hipStream_t gpuTaskStream(const nbnxm_gpu_t* nb);

// Per-system parameter blocks of the tuned and FP64 kernels
// (NbKernelArgs::sysParamsA/B/D) from the interaction constants.
void systemParamBlocks(const nbnxm_interaction_const_t& ic, float4* pa, float2* pb, double pd[8]);

// A single-system grid and list that live on the device (the output of a
// device-input search build), for a GPU task to adopt without a host trip.
// Cluster, group and mask indices are the single-system ones.
struct DeviceListView
{
    int numAtoms = 0;
    int numSlots = 0;                            // grid slots, 64 per cell
    const int* atomIndex = nullptr;              // slot -> atom, -1 for fillers
    const float4* xw = nullptr;                  // wrapped x, y, z per atom
    const nbnxm_sci_t* sci = nullptr;
    const nbnxm_sci_t* hostSci = nullptr;        // host copy of sci, or null (lists derived on the device)
    int nsci = 0;
    const nbnxm_cj_packed_t* cjPacked = nullptr;
    int ncjPacked = 0;
    const nbnxm_excl_t* excl = nullptr;          // excl[0] is all ones
    int nexcl = 0;
    hipStream_t stream = nullptr;                // the stream that wrote them
};

// A GPU task adopts a device list and builds its grid-order atom data on the
// device: xq from the wrapped coordinates and per-atom charges, types (with
// the dummy type for fillers) and per-type combination parameters, the
// slot/atom maps of the buffer ops, the shift vectors of `box`.  Per-atom
// inputs are device arrays; typeComb may be null for the type-table kernel.
// Replaces gpu_init_atomdata + gpu_upload_shiftvec + gpu_init_pairlist +
// gpu_init_x_to_nbat_x of a task created by gpu_init.
int gpuTaskAdoptDeviceList(nbnxm_gpu_t* nb,
                           const DeviceListView& v,
                           const float* charges,
                           const int* types,
                           const float2* typeComb,
                           int dummyType,
                           const double box[3]);

#define NBNXM_HIP_CHECK(call)                          \
    do                                                 \
    {                                                  \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess)                          \
        {                                              \
            return ::nbnxm_amd::hipFail(e_, #call);    \
        }                                              \
    } while (0)

// Device buffer that grows on demand and frees itself.
template<typename T>
struct DeviceBuffer
{
    T* ptr = nullptr;
    size_t capacity = 0;
    size_t size = 0;

    hipError_t resize(size_t n)
    {
        if (n > capacity)
        {
            if (ptr)
            {
                hipError_t e = hipFree(ptr);
                ptr = nullptr;
                capacity = 0;
                if (e != hipSuccess)
                {
                    return e;
                }
            }
            size_t alloc = n + n / 5 + 16; // over-allocate like the reference's reallocateDeviceBuffer
            hipError_t e = hipMalloc(reinterpret_cast<void**>(&ptr), alloc * sizeof(T));
            if (e != hipSuccess)
            {
                return e;
            }
            capacity = alloc;
        }
        size = n;
        return hipSuccess;
    }
    void release()
    {
        if (ptr)
        {
            (void)hipFree(ptr);
        }
        ptr = nullptr;
        capacity = 0;
        size = 0;
    }
    ~DeviceBuffer() { release(); }
};

} // namespace nbnxm_amd
