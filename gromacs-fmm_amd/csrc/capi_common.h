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
