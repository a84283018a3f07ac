// Shared host-side plumbing of libsitk_hip: status codes, the per-thread error
// message, the process-wide backend (device + stream), device allocations and
// the optional per-kernel event profiler.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/sitk_hip.h"

namespace sitkhip {

// A failure carried out of the C++ layer to the C ABI: a status code (one of
// sitkhip_status) and a message, the way CudaError carries its variant and
// text (crates/sitk/src/cuda/error.rs:11-108).
struct Error {
    int status;
    std::string message;
};

[[noreturn]] void fail(int status, const std::string &message);
void set_last_error(const std::string &message);

// Throws Error{SITKHIP_ERR_DRIVER} naming the call and the HIP error string.
void check_hip(hipError_t e, const char *what, const char *file, int line);
#define HIP_OK(call) ::sitkhip::check_hip((call), #call, __FILE__, __LINE__)

// backend() (crates/sitk/src/cuda/backend.rs:154-160): one device, one stream,
// initialised once per process. Throws SITKHIP_ERR_NO_DEVICE without a GPU.
struct Backend {
    int device = -1;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    std::string name;
};
Backend &backend();
// Looks up the device address of the static scratch of the allocation-free
// smoothing entry point on `device` (current); called once per device when
// the backend is initialised there (gaussian.hip).
void gaussian_scratch_setup(int device);

// Device memory through a caching allocator: a freed block is kept in a
// per-size-class free list and handed to the next request of that class
// (hipMalloc and hipFree cost tens of microseconds and hipFree waits for the
// device). Every operation of this library is ordered on its one stream, so a
// block freed after an enqueued use and reused by a later enqueued operation
// is never touched by both at once (work recorded to be enqueued later, as a
// PrepBatch does, runs with device_defer_frees on until it is enqueued).
// sitkhip_empty_cache() returns the cached blocks to the runtime.
void *device_alloc(size_t bytes);
void device_free(void *ptr, size_t bytes);
void device_empty_cache();
// While on, freed blocks are held back instead of being reused (a use that
// is recorded but not yet enqueued is not ordered by the stream); switching
// it off, once every such use is enqueued, hands them to the free lists.
void device_defer_frees(bool on);

// Pinned, device-mapped host slots of `bytes` (<= 256) for small per-metric
// result blocks, from a pool of pinned pages: returns the host view and sets
// *dev to the device view.
void *pinned_slot_alloc(size_t bytes, void **dev);
void pinned_slot_free(void *host);

// RAII device allocation (crates/sitk/src/cuda/buffer.rs:12-149). A zero-length
// request allocates one element, so a handle is never null
// (crates/sitk/src/cuda/ops/resident.rs:544-546).
template <typename T>
struct DeviceBuffer {
    T *ptr = nullptr;
    size_t len = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { alloc(n); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : ptr(o.ptr), len(o.len) { o.ptr = nullptr; o.len = 0; }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; len = o.len; o.ptr = nullptr; o.len = 0; }
        return *this;
    }
    ~DeviceBuffer() { release(); }
    void alloc(size_t n) {
        release();
        ptr = static_cast<T *>(device_alloc((n == 0 ? 1 : n) * sizeof(T)));
        len = n;
    }
    void release() {
        if (ptr) { device_free(ptr, (len == 0 ? 1 : len) * sizeof(T)); ptr = nullptr; }
        len = 0;
    }
    size_t bytes() const { return len * sizeof(T); }
    void upload(const T *host, size_t n) {
        HIP_OK(hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, backend().stream));
    }
    void download(T *host, size_t n) const {
        HIP_OK(hipMemcpyAsync(host, ptr, n * sizeof(T), hipMemcpyDeviceToHost, backend().stream));
    }
};

void sync();

// ---------------------------------------------------------------- profiler
// When enabled, launches made through Launch bracket themselves with events on
// the library stream; durations are folded per kernel name at query time.
bool profiling_enabled();
void profile_record(const char *name, hipEvent_t start, hipEvent_t stop);
hipEvent_t profile_event();

struct LaunchScope {
    const char *name;
    hipEvent_t start = nullptr;
    explicit LaunchScope(const char *n) : name(n) {
        if (profiling_enabled()) {
            start = profile_event();
            HIP_OK(hipEventRecord(start, backend().stream));
        }
    }
    ~LaunchScope() noexcept(false) {
        if (start) {
            hipEvent_t stop = profile_event();
            HIP_OK(hipEventRecord(stop, backend().stream));
            profile_record(name, start, stop);
        }
    }
};

// Enqueues `kernel` on the library stream, profiled under `name`. With the
// profiler on, the two events ride on the dispatch packet itself
// (hipExtLaunchKernelGGL), so the recorded span is the kernel's execution as a
// kernel trace reports it, without the marker packets LaunchScope puts around
// a launch (about 3 us more per launch on gfx950).
template <typename F, typename... Args>
void launch_kernel(const char *name, F kernel, dim3 grid, dim3 block, size_t shm, Args... args) {
    if (profiling_enabled()) {
        hipEvent_t a = profile_event(), b = profile_event();
        hipExtLaunchKernelGGL(kernel, grid, block, static_cast<std::uint32_t>(shm), backend().stream, a, b,
                              0u, args...);
        profile_record(name, a, b);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shm, backend().stream, args...);
    }
}

// Checks the launch itself (bad configuration) right after it is enqueued.
void check_launch(const char *name);

// Number of workgroups for a grid-stride kernel over `n` items, capped.
inline unsigned grid_for(size_t n, unsigned block, unsigned cap) {
    size_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<unsigned>(g);
}

// Refuses a sampling fraction that the draw cannot turn into a count: the
// reference casts ceil(1 / p) (regular) and voxels * p (random) to usize, which
// saturates in Rust but is undefined in C++ for NaN, infinities and negatives.
// Only strategies 1 (regular) and 2 (random) read the fraction.
inline void check_sampling_fraction(int strategy, double p, size_t voxels) {
    if (strategy != 1 && strategy != 2) return;
    const double inv = 1.0 / p;
    if (!(p > 0.0) || !(inv < 9.0e18) || !(static_cast<double>(voxels) * p < 9.0e18))
        fail(SITKHIP_ERR_INVALID_ARGUMENT,
             "sampling percentage must be a finite number above 0 (got " + std::to_string(p) + ")");
}

}  // namespace sitkhip
