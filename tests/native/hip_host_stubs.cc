// Host stand-ins for the HIP runtime calls the library's host code makes,
// so its params and per-track bodies run on the CPU: device memory is
// host memory, copies are memcpy, and streams are synchronous.
#include <cstdlib>
#include <cstring>

#include <hip/hip_runtime_api.h>

extern "C" {
hipError_t hipMalloc(void** p, size_t n)
{
    *p = std::calloc(n ? n : 1, 1);
    return hipSuccess;
}
hipError_t hipFree(void* p)
{
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void* d, void const* s, size_t n, hipMemcpyKind)
{
    std::memcpy(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, void const* s, size_t n, hipMemcpyKind, hipStream_t)
{
    std::memcpy(d, s, n);
    return hipSuccess;
}
hipError_t hipMemset(void* d, int v, size_t n)
{
    std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t)
{
    std::memset(d, v, n);
    return hipSuccess;
}
char const* hipGetErrorString(hipError_t)
{
    return "host stub";
}
hipError_t hipGetLastError()
{
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t)
{
    return hipSuccess;
}
hipError_t hipDeviceSynchronize()
{
    return hipSuccess;
}
}

