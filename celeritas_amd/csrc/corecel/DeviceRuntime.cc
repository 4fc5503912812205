//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/DeviceRuntime.cc
//---------------------------------------------------------------------------//
#include "DeviceRuntime.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
bool device_available()
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess)
    {
        (void)hipGetLastError();
        return false;
    }
    return count > 0;
}

void require_device()
{
    int count = 0;
    hipError_t err = hipGetDeviceCount(&count);
    if (err != hipSuccess || count == 0)
    {
        (void)hipGetLastError();
        throw RuntimeError(
            "celeritas_amd: no HIP device is available; the transport loop "
            "runs only on gfx950 (there is no CPU fallback)");
    }
}

//---------------------------------------------------------------------------//
DeviceArena::~DeviceArena()
{
    if (base_)
        (void)hipFree(base_);
}

std::size_t DeviceArena::reserve(std::size_t bytes)
{
    constexpr std::size_t align = 256;
    std::size_t const offset = total_;
    total_ += (bytes + align - 1) / align * align;
    return offset;
}

void DeviceArena::allocate()
{
    if (total_ > 0)
    {
        CELER_HIP_CALL(hipMalloc(&base_, total_));
        CELER_HIP_CALL(hipMemset(base_, 0, total_));
    }
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
