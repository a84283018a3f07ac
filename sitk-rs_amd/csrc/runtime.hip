// Backend, errors, profiler, device images and masks of libsitk_hip.
//
// Mirrors the reference's device runtime (crates/sitk/src/cuda/backend.rs,
// buffer.rs, image.rs, mask.rs, error.rs): one device and stream per process,
// float32 device volumes with an on-device upload cast, binary masks, and named
// refusals checked on the host before any launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

#include "common.hpp"
#include "image.hpp"

namespace sitkhip {

// ------------------------------------------------------------------ errors
static thread_local std::string g_last_error;

void set_last_error(const std::string &m) { g_last_error = m; }

[[noreturn]] void fail(int status, const std::string &message) { throw Error{status, message}; }

void check_hip(hipError_t e, const char *what, const char *file, int line) {
    if (e == hipSuccess) return;
    // A failed call also sets the runtime's last-error slot; clear it so the
    // next launch check reports only its own launch.
    (void)hipGetLastError();
    char buf[512];
    std::snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", static_cast<int>(e),
                  hipGetErrorString(e), what, file, line);
    fail(SITKHIP_ERR_DRIVER, buf);
}

void check_launch(const char *name) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fail(SITKHIP_ERR_DRIVER,
             std::string("launch of ") + name + " failed: " + hipGetErrorString(e));
    }
}

// ------------------------------------------------------ caching allocator
// Freed blocks go to a free list per size class and serve the next request of
// that class. Blocks below kSlabMax are carved from 256 MiB slabs (one
// hipMalloc per slab instead of one per block: a batch of registrations asks
// for thousands of small volumes); larger ones are allocated one by one.
// empty_cache returns the free large blocks and every slab none of whose
// blocks is in use.
namespace {
std::mutex g_alloc_mu;
std::unordered_map<size_t, std::vector<void *>> g_free_blocks;

constexpr size_t kSlabBytes = static_cast<size_t>(256) << 20;
constexpr size_t kSlabMax = static_cast<size_t>(16) << 20;

struct Slab {
    char *base;
    size_t used;
    size_t live;  // carved blocks not in a free list
};
std::map<char *, Slab> g_slabs;  // by base address
Slab *g_open_slab = nullptr;     // the slab new blocks are carved from

Slab *slab_of(void *ptr) {
    if (g_slabs.empty()) return nullptr;
    auto it = g_slabs.upper_bound(static_cast<char *>(ptr));
    if (it == g_slabs.begin()) return nullptr;
    --it;
    Slab &sl = it->second;
    return static_cast<char *>(ptr) < sl.base + kSlabBytes ? &sl : nullptr;
}

// Size classes: 8 steps per power of two (at most 12.5% rounding), >= 256 B.
size_t size_class(size_t bytes) {
    if (bytes <= 256) return 256;
    int lg = 63 - __builtin_clzll(static_cast<unsigned long long>(bytes));
    const size_t step = static_cast<size_t>(1) << (lg >= 3 ? lg - 3 : 0);
    return (bytes + step - 1) / step * step;
}

// A block of class c from the open slab (a new slab when it is full); nullptr
// when the device has no room for a slab. Called with g_alloc_mu held.
void *carve(size_t c) {
    c = (c + 255) & ~static_cast<size_t>(255);  // blocks start 256-byte aligned, as hipMalloc's do
    if (!g_open_slab || g_open_slab->used + c > kSlabBytes) {
        void *base = nullptr;
        if (hipMalloc(&base, kSlabBytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        char *b = static_cast<char *>(base);
        g_open_slab = &(g_slabs[b] = Slab{b, 0, 0});
    }
    void *p = g_open_slab->base + g_open_slab->used;
    g_open_slab->used += c;
    ++g_open_slab->live;
    return p;
}
}  // namespace

void *device_alloc(size_t bytes) {
    const size_t c = size_class(bytes);
    {
        std::lock_guard<std::mutex> lk(g_alloc_mu);
        auto it = g_free_blocks.find(c);
        if (it != g_free_blocks.end() && !it->second.empty()) {
            void *p = it->second.back();
            it->second.pop_back();
            if (Slab *sl = slab_of(p)) ++sl->live;
            return p;
        }
        if (c < kSlabMax) {
            if (void *p = carve(c)) return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, c);
    if (e != hipSuccess) {
        // out of memory with blocks cached: give them back and try once more
        (void)hipGetLastError();
        device_empty_cache();
        HIP_OK(hipMalloc(&p, c));
    }
    return p;
}

static bool g_defer_frees = false;
static std::vector<std::pair<void *, size_t>> g_deferred;  // (block, size class)

void device_free(void *ptr, size_t bytes) {
    if (!ptr) return;
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    if (g_defer_frees) {
        g_deferred.emplace_back(ptr, size_class(bytes));
        return;
    }
    g_free_blocks[size_class(bytes)].push_back(ptr);
    if (Slab *sl = slab_of(ptr)) --sl->live;
}

void device_defer_frees(bool on) {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    g_defer_frees = on;
    if (on) return;
    for (auto &d : g_deferred) {
        g_free_blocks[d.second].push_back(d.first);
        if (Slab *sl = slab_of(d.first)) --sl->live;
    }
    g_deferred.clear();
}

void device_empty_cache() {
    std::vector<void *> release;
    {
        std::lock_guard<std::mutex> lk(g_alloc_mu);
        // free blocks of idle slabs are dropped with their slab; free large
        // blocks are released one by one; free blocks of busy slabs stay
        for (auto &kv : g_free_blocks) {
            std::vector<void *> keep;
            for (void *p : kv.second) {
                Slab *sl = slab_of(p);
                if (!sl) release.push_back(p);
                else if (sl->live != 0) keep.push_back(p);
            }
            kv.second.swap(keep);
        }
        for (auto it = g_slabs.begin(); it != g_slabs.end();) {
            if (it->second.live == 0) {
                if (g_open_slab == &it->second) g_open_slab = nullptr;
                release.push_back(it->second.base);
                it = g_slabs.erase(it);
            } else {
                ++it;
            }
        }
    }
    if (release.empty()) return;
    (void)hipStreamSynchronize(backend().stream);
    for (void *p : release) (void)hipFree(p);
}

namespace {
constexpr size_t kSlotBytes = 256;
constexpr size_t kSlotsPerPage = 4096;
std::mutex g_pin_mu;
std::vector<char *> g_free_slots;                       // host addresses
std::unordered_map<char *, char *> g_slot_dev;          // host slot -> device view
}  // namespace

void *pinned_slot_alloc(size_t bytes, void **dev) {
    if (bytes > kSlotBytes) fail(SITKHIP_ERR_INVALID_ARGUMENT, "pinned slot too large");
    std::lock_guard<std::mutex> lk(g_pin_mu);
    if (g_free_slots.empty()) {
        char *page = nullptr, *dpage = nullptr;
        HIP_OK(hipHostMalloc(reinterpret_cast<void **>(&page), kSlotBytes * kSlotsPerPage,
                             hipHostMallocMapped | hipHostMallocCoherent));
        HIP_OK(hipHostGetDevicePointer(reinterpret_cast<void **>(&dpage), page, 0));
        for (size_t i = kSlotsPerPage; i-- > 0;) {
            g_free_slots.push_back(page + i * kSlotBytes);
            g_slot_dev[page + i * kSlotBytes] = dpage + i * kSlotBytes;
        }
    }
    char *h = g_free_slots.back();
    g_free_slots.pop_back();
    std::memset(h, 0, kSlotBytes);
    *dev = g_slot_dev[h];
    return h;
}

void pinned_slot_free(void *host) {
    if (!host) return;
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_free_slots.push_back(static_cast<char *>(host));
}

// ----------------------------------------------------------------- backend
static std::mutex g_backend_mu;
static Backend g_backend;
static bool g_backend_ready = false;

static void init_backend_locked(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        fail(SITKHIP_ERR_NO_DEVICE, std::string("no usable HIP device: ") +
                                        (e != hipSuccess ? hipGetErrorString(e) : "count is 0"));
    }
    if (device < 0 || device >= count) {
        fail(SITKHIP_ERR_NO_DEVICE, "device index " + std::to_string(device) + " out of range (" +
                                        std::to_string(count) + " devices)");
    }
    HIP_OK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, device));
    g_backend.device = device;
    g_backend.cu_count = prop.multiProcessorCount;
    g_backend.name = prop.gcnArchName;
    // Default-priority, non-blocking stream: every operation of the library is
    // enqueued here and synchronised before the call returns.
    HIP_OK(hipStreamCreateWithFlags(&g_backend.stream, hipStreamNonBlocking));
    gaussian_scratch_setup(device);
    g_backend_ready = true;
}

Backend &backend() {
    std::lock_guard<std::mutex> lk(g_backend_mu);
    if (!g_backend_ready) init_backend_locked(0);
    HIP_OK(hipSetDevice(g_backend.device));
    return g_backend;
}

void sync() { HIP_OK(hipStreamSynchronize(backend().stream)); }

// ---------------------------------------------------------------- profiler
static std::atomic<bool> g_prof_on{false};
static std::mutex g_prof_mu;
struct PendingEvent {
    std::string name;
    hipEvent_t start, stop;
};
static std::vector<PendingEvent> g_prof_pending;
static std::vector<hipEvent_t> g_prof_free;
struct ProfTotal {
    uint64_t count = 0;
    double ms = 0.0;
};
static std::map<std::string, ProfTotal> g_prof_totals;

bool profiling_enabled() { return g_prof_on.load(std::memory_order_relaxed); }

hipEvent_t profile_event() {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof_free.empty()) {
        hipEvent_t e = g_prof_free.back();
        g_prof_free.pop_back();
        return e;
    }
    hipEvent_t e;
    HIP_OK(hipEventCreate(&e));
    return e;
}

void profile_record(const char *name, hipEvent_t start, hipEvent_t stop) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_pending.push_back({name, start, stop});
}

static void profile_fold_locked() {
    for (auto &p : g_prof_pending) {
        HIP_OK(hipEventSynchronize(p.stop));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, p.start, p.stop));
        auto &t = g_prof_totals[p.name];
        t.count += 1;
        t.ms += ms;
        g_prof_free.push_back(p.start);
        g_prof_free.push_back(p.stop);
    }
    g_prof_pending.clear();
}

// ------------------------------------------------------------------ images
static std::atomic<uint64_t> g_next_id{1};
uint64_t next_id() { return g_next_id.fetch_add(1); }

size_t geometry_len(const sitkhip_geometry &g) { return g.size[0] * g.size[1] * g.size[2]; }

bool same_geometry(const sitkhip_geometry &a, const sitkhip_geometry &b) {
    return std::memcmp(&a, &b, sizeof(a)) == 0;
}

void validate_geometry(const sitkhip_geometry &g, const char *what) {
    for (int d = 0; d < 3; ++d) {
        if (g.size[d] == 0)
            fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, std::string(what) + ": zero size on an axis");
        if (!(g.spacing[d] > 0.0))
            fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, std::string(what) + ": spacing must be > 0");
    }
}

// (float)(double)x for each scalar pixel type (crates/sitk/src/cuda/image.rs:93-106).
template <typename T>
__global__ void cast_to_f32(const T *__restrict__ src, float *__restrict__ dst, size_t n) {
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = static_cast<float>(static_cast<double>(src[i]));
}

__global__ void fill_f32(float *__restrict__ dst, float v, size_t n) {
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = v;
}

__global__ void mask_from_f32(const float *__restrict__ src, uint8_t *__restrict__ dst, size_t n) {
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i] != 0.0f ? 1 : 0;
}

__global__ void mask_normalize(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, size_t n) {
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i] != 0 ? 1 : 0;
}

__global__ void mask_and(const uint8_t *__restrict__ a, const uint8_t *__restrict__ b,
                         uint8_t *__restrict__ dst, size_t n) {
    size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = (a[i] != 0 && b[i] != 0) ? 1 : 0;
}

static size_t pixel_bytes(int t) {
    switch (t) {
        case SITKHIP_UINT8: case SITKHIP_INT8: return 1;
        case SITKHIP_UINT16: case SITKHIP_INT16: return 2;
        case SITKHIP_UINT32: case SITKHIP_INT32: case SITKHIP_FLOAT32: return 4;
        case SITKHIP_UINT64: case SITKHIP_INT64: case SITKHIP_FLOAT64: return 8;
        default: return 0;
    }
}

template <typename T>
static void launch_cast(const void *src, float *dst, size_t n) {
    LaunchScope scope("cast_to_f32");
    hipLaunchKernelGGL(cast_to_f32<T>, dim3(grid_for(n, 256, 4096)), dim3(256), 0,
                       backend().stream, static_cast<const T *>(src), dst, n);
    check_launch("cast_to_f32");
}

Image *image_upload(int pixel_type, const void *data, size_t len, const sitkhip_geometry &g) {
    validate_geometry(g, "image upload");
    size_t n = geometry_len(g);
    if (len != n)
        fail(SITKHIP_ERR_INVALID_ARGUMENT, "upload length " + std::to_string(len) +
                                               " does not match the geometry's voxel count " +
                                               std::to_string(n));
    size_t pb = pixel_bytes(pixel_type);
    if (pb == 0)
        fail(SITKHIP_ERR_UNSUPPORTED_PIXEL_TYPE,
             "pixel type " + std::to_string(pixel_type) + " is not a scalar pixel type");
    if (!data) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null upload data");
    auto img = new Image();
    img->geom = g;
    img->id = next_id();
    img->buf.alloc(n);
    if (pixel_type == SITKHIP_FLOAT32) {
        img->buf.upload(static_cast<const float *>(data), n);
    } else {
        DeviceBuffer<uint8_t> staging(n * pb);
        staging.upload(static_cast<const uint8_t *>(data), n * pb);
        switch (pixel_type) {
            case SITKHIP_UINT8: launch_cast<uint8_t>(staging.ptr, img->buf.ptr, n); break;
            case SITKHIP_INT8: launch_cast<int8_t>(staging.ptr, img->buf.ptr, n); break;
            case SITKHIP_UINT16: launch_cast<uint16_t>(staging.ptr, img->buf.ptr, n); break;
            case SITKHIP_INT16: launch_cast<int16_t>(staging.ptr, img->buf.ptr, n); break;
            case SITKHIP_UINT32: launch_cast<uint32_t>(staging.ptr, img->buf.ptr, n); break;
            case SITKHIP_INT32: launch_cast<int32_t>(staging.ptr, img->buf.ptr, n); break;
            case SITKHIP_UINT64: launch_cast<uint64_t>(staging.ptr, img->buf.ptr, n); break;
            case SITKHIP_INT64: launch_cast<int64_t>(staging.ptr, img->buf.ptr, n); break;
            case SITKHIP_FLOAT64: launch_cast<double>(staging.ptr, img->buf.ptr, n); break;
        }
        sync();  // staging is freed at scope exit
    }
    sync();
    return img;
}

Image *image_filled(const sitkhip_geometry &g, float value) {
    validate_geometry(g, "filled image");
    size_t n = geometry_len(g);
    auto img = new Image();
    img->geom = g;
    img->id = next_id();
    img->buf.alloc(n);
    {
        LaunchScope scope("fill_f32");
        hipLaunchKernelGGL(fill_f32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, backend().stream,
                           img->buf.ptr, value, n);
        check_launch("fill_f32");
    }
    sync();
    return img;
}

Image *image_like(const sitkhip_geometry &g) {
    auto img = new Image();
    img->geom = g;
    img->id = next_id();
    img->buf.alloc(geometry_len(g));
    return img;
}

Mask *mask_upload(const uint8_t *data, size_t len, const sitkhip_geometry &g) {
    validate_geometry(g, "mask upload");
    size_t n = geometry_len(g);
    if (len != n) fail(SITKHIP_ERR_INVALID_ARGUMENT, "mask length does not match its geometry");
    if (!data) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null mask data");
    auto m = new Mask();
    m->geom = g;
    m->buf.alloc(n);
    DeviceBuffer<uint8_t> staging(n);
    staging.upload(data, n);
    {
        LaunchScope scope("mask_normalize");
        hipLaunchKernelGGL(mask_normalize, dim3(grid_for(n, 256, 4096)), dim3(256), 0,
                           backend().stream, staging.ptr, m->buf.ptr, n);
        check_launch("mask_normalize");
    }
    sync();
    return m;
}

Mask *mask_from_image(const Image &img) {
    size_t n = img.len();
    auto m = new Mask();
    m->geom = img.geom;
    m->buf.alloc(n);
    {
        LaunchScope scope("mask_from_f32");
        hipLaunchKernelGGL(mask_from_f32, dim3(grid_for(n, 256, 4096)), dim3(256), 0,
                           backend().stream, img.buf.ptr, m->buf.ptr, n);
        check_launch("mask_from_f32");
    }
    sync();
    return m;
}

Mask *mask_intersect(const Mask &a, const Mask &b) {
    if (!same_geometry(a.geom, b.geom))
        fail(SITKHIP_ERR_UNSUPPORTED_GEOMETRY, "mask intersect: the masks' geometries differ");
    size_t n = a.len();
    auto m = new Mask();
    m->geom = a.geom;
    m->buf.alloc(n);
    {
        LaunchScope scope("mask_and");
        hipLaunchKernelGGL(mask_and, dim3(grid_for(n, 256, 4096)), dim3(256), 0, backend().stream,
                           a.buf.ptr, b.buf.ptr, m->buf.ptr, n);
        check_launch("mask_and");
    }
    sync();
    return m;
}

}  // namespace sitkhip

// =================================================================== C ABI
using namespace sitkhip;

template <typename F>
static int guarded(F &&f) {
    try {
        f();
        return SITKHIP_OK;
    } catch (const Error &e) {
        set_last_error(e.message);
        return e.status;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return SITKHIP_ERR_DRIVER;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return SITKHIP_ERR_DRIVER;
    }
}

namespace sitkhip {
int run_guarded(const std::function<void()> &f) { return guarded(f); }
}  // namespace sitkhip

extern "C" {

const char *sitkhip_status_name(int s) {
    switch (s) {
        case SITKHIP_OK: return "Ok";
        case SITKHIP_ERR_NO_DEVICE: return "NoDevice";
        case SITKHIP_ERR_DRIVER: return "Driver";
        case SITKHIP_ERR_UNSUPPORTED_PIXEL_TYPE: return "UnsupportedPixelType";
        case SITKHIP_ERR_UNSUPPORTED_GEOMETRY: return "UnsupportedGeometry";
        case SITKHIP_ERR_DEGENERATE_INPUT: return "DegenerateInput";
        case SITKHIP_ERR_MASKED_EXPLICIT_POINTS: return "MaskedExplicitPoints";
        case SITKHIP_ERR_SAMPLE_INDEX_OUT_OF_GRID: return "SampleIndexOutOfGrid";
        case SITKHIP_ERR_POINT_MAP_STAGE_COUNT: return "PointMapStageCount";
        case SITKHIP_ERR_HISTOGRAM_SHAPE: return "HistogramShape";
        case SITKHIP_ERR_NOT_SCALAR: return "NotScalar";
        case SITKHIP_ERR_INVALID_ARGUMENT: return "InvalidArgument";
        case SITKHIP_ERR_TOO_FEW_BINS: return "TooFewHistogramBins";
        case SITKHIP_ERR_CONSTANT_INTENSITY: return "ConstantIntensity";
        case SITKHIP_ERR_TOO_MANY_BINS: return "TooManyHistogramBins";
        case SITKHIP_ERR_NOT_BUILT: return "NotBuilt";
        case SITKHIP_ERR_MASK_SIZE_MISMATCH: return "MaskSizeMismatch";
        default: return "Unknown";
    }
}

const char *sitkhip_last_error(void) { return g_last_error.c_str(); }

int sitkhip_backend_init(int device) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_backend_mu);
        if (g_backend_ready) {
            if (g_backend.device != device)
                fail(SITKHIP_ERR_INVALID_ARGUMENT,
                     "backend already initialised on device " + std::to_string(g_backend.device));
            return;
        }
        init_backend_locked(device);
    });
}

int sitkhip_backend_stream(void **stream) {
    return guarded([&] {
        if (!stream) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null stream out-pointer");
        *stream = static_cast<void *>(backend().stream);
    });
}

int sitkhip_device_name(char *buf, size_t len) {
    return guarded([&] {
        if (!buf || len == 0) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null buffer");
        const std::string &n = backend().name;
        std::snprintf(buf, len, "%s", n.c_str());
    });
}

int sitkhip_image_upload(int pixel_type, const void *data, size_t len,
                         const sitkhip_geometry *geometry, sitkhip_image *out) {
    return guarded([&] {
        if (!geometry || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null geometry or out-pointer");
        *out = reinterpret_cast<sitkhip_image>(image_upload(pixel_type, data, len, *geometry));
    });
}

int sitkhip_image_filled(const sitkhip_geometry *geometry, float value, sitkhip_image *out) {
    return guarded([&] {
        if (!geometry || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null geometry or out-pointer");
        *out = reinterpret_cast<sitkhip_image>(image_filled(*geometry, value));
    });
}

int sitkhip_image_to_host(sitkhip_image h, float *out, size_t len) {
    return guarded([&] {
        auto img = reinterpret_cast<Image *>(h);
        if (!img || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null image or output");
        if (len != img->len())
            fail(SITKHIP_ERR_INVALID_ARGUMENT, "to_host length does not match the voxel count");
        img->buf.download(out, len);
        sync();
    });
}

int sitkhip_image_geometry(sitkhip_image h, sitkhip_geometry *out) {
    return guarded([&] {
        auto img = reinterpret_cast<Image *>(h);
        if (!img || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null image or output");
        *out = img->geom;
    });
}

size_t sitkhip_image_len(sitkhip_image h) {
    auto img = reinterpret_cast<Image *>(h);
    return img ? img->len() : 0;
}

uint64_t sitkhip_image_id(sitkhip_image h) {
    auto img = reinterpret_cast<Image *>(h);
    return img ? img->id : 0;
}

const float *sitkhip_image_device_ptr(sitkhip_image h) {
    auto img = reinterpret_cast<Image *>(h);
    return img ? img->buf.ptr : nullptr;
}

void sitkhip_image_free(sitkhip_image h) { delete reinterpret_cast<Image *>(h); }

int sitkhip_mask_upload(const uint8_t *data, size_t len, const sitkhip_geometry *geometry,
                        sitkhip_mask *out) {
    return guarded([&] {
        if (!geometry || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null geometry or out-pointer");
        *out = reinterpret_cast<sitkhip_mask>(mask_upload(data, len, *geometry));
    });
}

int sitkhip_mask_from_image(sitkhip_image h, sitkhip_mask *out) {
    return guarded([&] {
        auto img = reinterpret_cast<Image *>(h);
        if (!img || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null image or out-pointer");
        *out = reinterpret_cast<sitkhip_mask>(mask_from_image(*img));
    });
}

int sitkhip_mask_intersect(sitkhip_mask a, sitkhip_mask b, sitkhip_mask *out) {
    return guarded([&] {
        auto ma = reinterpret_cast<Mask *>(a);
        auto mb = reinterpret_cast<Mask *>(b);
        if (!ma || !mb || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null mask or out-pointer");
        *out = reinterpret_cast<sitkhip_mask>(mask_intersect(*ma, *mb));
    });
}

int sitkhip_mask_to_host(sitkhip_mask h, uint8_t *out, size_t len) {
    return guarded([&] {
        auto m = reinterpret_cast<Mask *>(h);
        if (!m || !out) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null mask or output");
        if (len != m->len()) fail(SITKHIP_ERR_INVALID_ARGUMENT, "to_host length mismatch");
        m->buf.download(out, len);
        sync();
    });
}

void sitkhip_mask_free(sitkhip_mask h) { delete reinterpret_cast<Mask *>(h); }

int sitkhip_profile_enable(int on) {
    return guarded([&] { g_prof_on.store(on != 0); });
}

int sitkhip_profile_reset(void) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        profile_fold_locked();
        g_prof_totals.clear();
    });
}

int sitkhip_profile_query(const char *prefix, uint64_t *count, double *total_ms) {
    int matched = 0;
    int rc = guarded([&] {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        profile_fold_locked();
        uint64_t c = 0;
        double ms = 0.0;
        std::string p = prefix ? prefix : "";
        for (auto &kv : g_prof_totals) {
            if (kv.first.compare(0, p.size(), p) == 0) {
                c += kv.second.count;
                ms += kv.second.ms;
                ++matched;
            }
        }
        if (count) *count = c;
        if (total_ms) *total_ms = ms;
    });
    return rc == SITKHIP_OK ? matched : -rc;
}

int sitkhip_empty_cache(void) {
    return run_guarded([&] { device_empty_cache(); });
}

int sitkhip_profile_dump(char *buf, size_t len) {
    return guarded([&] {
        if (!buf || len == 0) fail(SITKHIP_ERR_INVALID_ARGUMENT, "null buffer");
        std::lock_guard<std::mutex> lk(g_prof_mu);
        profile_fold_locked();
        std::string s;
        for (auto &kv : g_prof_totals) {
            char line[256];
            std::snprintf(line, sizeof(line), "%s %llu %.6f\n", kv.first.c_str(),
                          static_cast<unsigned long long>(kv.second.count), kv.second.ms);
            s += line;
        }
        std::snprintf(buf, len, "%s", s.c_str());
    });
}

}  // extern "C"
