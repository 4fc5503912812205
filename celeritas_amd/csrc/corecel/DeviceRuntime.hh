//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/DeviceRuntime.hh
//! HIP error checking, device memory arenas and typed device vectors.
//---------------------------------------------------------------------------//
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! Error for invalid user input or configuration (CELER_VALIDATE)
class RuntimeError : public std::runtime_error
{
  public:
    using std::runtime_error::runtime_error;
};

//! Throw a RuntimeError with a streamed message if a condition fails
#define CELER_VALIDATE(COND, MSG)                                   \
    do                                                              \
    {                                                               \
        if (!(COND))                                                \
        {                                                           \
            std::ostringstream celer_msg_;                          \
            celer_msg_ MSG;                                         \
            throw ::celeritas_amd::RuntimeError(celer_msg_.str());  \
        }                                                           \
    } while (0)

//! Check a HIP runtime call and throw on failure
#define CELER_HIP_CALL(EXPR)                                                 \
    do                                                                       \
    {                                                                        \
        hipError_t celer_err_ = (EXPR);                                      \
        if (celer_err_ != hipSuccess)                                        \
        {                                                                    \
            throw ::celeritas_amd::RuntimeError(                             \
                std::string("HIP error in ") + #EXPR + ": "                  \
                + hipGetErrorString(celer_err_) + " (" __FILE__ ":"          \
                + std::to_string(__LINE__) + ")");                           \
        }                                                                    \
    } while (0)

//---------------------------------------------------------------------------//
//! Whether a HIP device is usable (throws nothing)
bool device_available();

//! Require a device; throws a RuntimeError naming why it is missing
void require_device();

//---------------------------------------------------------------------------//
/*!
 * Bump allocator over one device allocation, 256-byte aligned slices.
 *
 * Used to lay out all per-slot arrays of a state contiguously in HBM.
 */
class DeviceArena
{
  public:
    DeviceArena() = default;
    ~DeviceArena();
    DeviceArena(DeviceArena const&) = delete;
    DeviceArena& operator=(DeviceArena const&) = delete;

    //! Plan a slice of the given size; returns offset
    std::size_t reserve(std::size_t bytes);
    //! Allocate all planned slices
    void allocate();
    //! Get a device pointer for a planned offset
    template<class T>
    T* get(std::size_t offset) const
    {
        return reinterpret_cast<T*>(static_cast<char*>(base_) + offset);
    }
    std::size_t size() const { return total_; }
    void* data() const { return base_; }

  private:
    void* base_{nullptr};
    std::size_t total_{0};
};

//---------------------------------------------------------------------------//
/*!
 * Owning device array of trivially copyable values.
 */
template<class T>
class DeviceVector
{
  public:
    DeviceVector() = default;
    explicit DeviceVector(std::size_t n) { this->resize(n); }
    ~DeviceVector()
    {
        if (data_)
            (void)hipFree(data_);
    }
    DeviceVector(DeviceVector const&) = delete;
    DeviceVector& operator=(DeviceVector const&) = delete;
    DeviceVector(DeviceVector&& o) noexcept : data_(o.data_), size_(o.size_)
    {
        o.data_ = nullptr;
        o.size_ = 0;
    }
    DeviceVector& operator=(DeviceVector&& o) noexcept
    {
        std::swap(data_, o.data_);
        std::swap(size_, o.size_);
        return *this;
    }

    void resize(std::size_t n)
    {
        if (data_)
            CELER_HIP_CALL(hipFree(data_));
        data_ = nullptr;
        size_ = n;
        if (n > 0)
            CELER_HIP_CALL(hipMalloc(&data_, n * sizeof(T)));
    }
    void copy_from_host(T const* src, std::size_t n)
    {
        if (n > size_)
            this->resize(n);
        if (n > 0)
            CELER_HIP_CALL(
                hipMemcpy(data_, src, n * sizeof(T), hipMemcpyHostToDevice));
    }
    void copy_from_host(std::vector<T> const& v)
    {
        this->copy_from_host(v.data(), v.size());
    }
    std::vector<T> copy_to_host() const
    {
        std::vector<T> result(size_);
        if (size_ > 0)
            CELER_HIP_CALL(hipMemcpy(
                result.data(), data_, size_ * sizeof(T), hipMemcpyDeviceToHost));
        return result;
    }
    T* data() const { return data_; }
    std::size_t size() const { return size_; }

  private:
    T* data_{nullptr};
    std::size_t size_{0};
};

//! Upload a host vector into a new device vector
template<class T>
DeviceVector<T> to_device(std::vector<T> const& v)
{
    DeviceVector<T> result;
    result.copy_from_host(v);
    return result;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
