// Error plumbing shared by the C ABI implementation files.  No exception
// crosses the ABI: every entry point runs its body through pbrt_guard(), which
// converts exceptions and HIP failures into a status code plus a thread-local
// message readable through pbrt_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>

#include "../../include/pbrt_c.h"

namespace pbrt {

struct PbrtError : std::runtime_error {
    int32_t status;
    PbrtError(int32_t s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

void set_last_error(const std::string& msg);

[[noreturn]] inline void fail(int32_t status, const std::string& msg) { throw PbrtError(status, msg); }

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) fail(PBRT_ERR_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
        fail(PBRT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}

#define PBRT_HIP(call) ::pbrt::hip_check((call), #call)
#define PBRT_REQUIRE(cond, msg)                                     \
    do {                                                            \
        if (!(cond)) ::pbrt::fail(PBRT_ERR_INVALID_ARG, (msg));     \
    } while (0)

template <class F>
int32_t guard(F&& body) {
    try {
        body();
        set_last_error("");
        return PBRT_OK;
    } catch (const PbrtError& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return PBRT_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return PBRT_ERR_INTERNAL;
    } catch (...) {
        set_last_error("unknown error");
        return PBRT_ERR_INTERNAL;
    }
}

// Device buffer owned by a C ABI object.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void upload(const void* src, size_t n) {
        release();
        if (n == 0) return;
        PBRT_HIP(hipMalloc(&ptr, n));
        bytes = n;
        PBRT_HIP(hipMemcpy(ptr, src, n, hipMemcpyHostToDevice));
    }
    void allocate(size_t n) {
        release();
        if (n == 0) return;
        PBRT_HIP(hipMalloc(&ptr, n));
        bytes = n;
    }
    void swap(DeviceBuffer& other) {
        std::swap(ptr, other.ptr);
        std::swap(bytes, other.bytes);
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

}  // namespace pbrt
