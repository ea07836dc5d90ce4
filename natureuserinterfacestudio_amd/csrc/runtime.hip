// Runtime layer of libnuif.so: device selection, allocation accounting,
// streams, events, graphs and fills.  Replaces the reference's process-wide
// OpenCL context + single in-order queue (CLU/NuiOpenCLGlobal.cpp:75-220),
// its allocation registry (CLU/NuiGPUMemManager.cpp) and its memset kernels
// (CLU/NuiOpenCLFoundationUtils.cpp:8-150, K/foundation.cl:3-32).  There are no
// singletons beyond the chosen device: every session owns its streams.

#include "nuif_internal.h"

#include <atomic>
#include <mutex>
#include <unordered_map>

namespace nuif {

namespace {
thread_local std::string t_last_error;
thread_local int t_device_set = -1;
std::atomic<int> g_device{-1};
std::atomic<int> g_generation{0};
hipStream_t g_default_stream = nullptr;
std::mutex g_mem_mu;
std::unordered_map<void*, size_t> g_allocs;
size_t g_bytes_in_use = 0;

void ensure_thread_device() {
    int d = g_device.load();
    if (d >= 0 && t_device_set != d) {
        (void)hipSetDevice(d);
        t_device_set = d;
    }
}
}  // namespace

void set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    t_last_error = buf;
}

int hip_fail(hipError_t e, const char* what) {
    set_error("%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
    return e == hipErrorOutOfMemory ? NUIF_ERR_OOM : NUIF_ERR_HIP;
}

bool initialised() {
    if (g_device.load() < 0) return false;
    ensure_thread_device();
    return true;
}

hipStream_t stream_of(nuif_stream_t s) { return s ? reinterpret_cast<hipStream_t>(s) : g_default_stream; }

// 1-D fill kernels (graph-capture safe replacement of hipMemsetAsync)
__global__ void fill_u8_kernel(uint8_t* __restrict__ p, uint8_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    uint32_t v4 = 0x01010101u * v;
    size_t n4 = n / 4;
    uint32_t* p4 = reinterpret_cast<uint32_t*>(p);
    for (size_t k = i; k < n4; k += stride) p4[k] = v4;
    for (size_t k = n4 * 4 + i; k < n; k += stride) p[k] = v;
}
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < n; k += stride) p[k] = v;
}

}  // namespace nuif

using namespace nuif;

extern "C" {

const char* nuif_version(void) { return "nuif 0.1 (gfx950)"; }

const char* nuif_last_error(void) { return t_last_error.c_str(); }

int nuif_device_count(int* count) {
    if (!count) return NUIF_ERR_INVALID;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return NUIF_OK;
}

int nuif_init(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_error("no HIP device available");
        return NUIF_ERR_NODEV;
    }
    if (device < 0 || device >= n) {
        set_error("device %d out of range (count %d)", device, n);
        return NUIF_ERR_INVALID;
    }
    if (g_device.load() == device) return NUIF_OK;
    NUIF_HIP(hipSetDevice(device));
    t_device_set = device;
    hipStream_t s = nullptr;
    NUIF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    g_default_stream = s;
    g_device.store(device);
    g_generation.fetch_add(1);
    // the allocation-free ICP entry point's scratch address on this device
    void* p = nullptr;
    return static_scratch(&p);
}

int nuif_device_info(char* name, int name_len, int* cu_count, size_t* hbm_bytes, int* clock_khz) {
    NUIF_REQUIRE_INIT();
    hipDeviceProp_t p;
    NUIF_HIP(hipGetDeviceProperties(&p, g_device.load()));
    if (name && name_len > 0) snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
    if (clock_khz) *clock_khz = p.clockRate;
    return NUIF_OK;
}

int nuif_malloc(void** dptr, size_t bytes) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(dptr != nullptr, "nuif_malloc: null out pointer");
    *dptr = nullptr;
    if (bytes == 0) bytes = 256;
    void* p = nullptr;
    NUIF_HIP(hipMalloc(&p, bytes));
    // The fill runs on the null stream, which the library's non-blocking streams
    // do not wait for: finish it before the memory is handed out.
    hipError_t e = hipMemsetAsync(p, 0, bytes, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return hip_fail(e, "hipMemset");
    }
    {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        g_allocs[p] = bytes;
        g_bytes_in_use += bytes;
    }
    *dptr = p;
    return NUIF_OK;
}

int nuif_free(void* dptr) {
    NUIF_REQUIRE_INIT();
    if (!dptr) return NUIF_OK;
    {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        auto it = g_allocs.find(dptr);
        NUIF_CHECK_ARG(it != g_allocs.end(), "nuif_free: pointer %p was not allocated by nuif_malloc", dptr);
        g_bytes_in_use -= it->second;
        g_allocs.erase(it);
    }
    NUIF_HIP(hipFree(dptr));
    return NUIF_OK;
}

int nuif_mem_in_use(size_t* bytes) {
    NUIF_CHECK_ARG(bytes != nullptr, "nuif_mem_in_use: null out pointer");
    std::lock_guard<std::mutex> lk(g_mem_mu);
    *bytes = g_bytes_in_use;
    return NUIF_OK;
}

int nuif_host_alloc(void** hptr, size_t bytes) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(hptr != nullptr, "nuif_host_alloc: null out pointer");
    NUIF_HIP(hipHostMalloc(hptr, bytes ? bytes : 64, hipHostMallocDefault));
    return NUIF_OK;
}

int nuif_host_free(void* hptr) {
    NUIF_REQUIRE_INIT();
    if (hptr) NUIF_HIP(hipHostFree(hptr));
    return NUIF_OK;
}

int nuif_memcpy_h2d(void* dst, const void* src, size_t bytes, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    if (!bytes) return NUIF_OK;
    NUIF_CHECK_ARG(dst && src, "nuif_memcpy_h2d: null pointer");
    NUIF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_of(s)));
    return NUIF_OK;
}

int nuif_memcpy_d2h(void* dst, const void* src, size_t bytes, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    if (!bytes) return NUIF_OK;
    NUIF_CHECK_ARG(dst && src, "nuif_memcpy_d2h: null pointer");
    NUIF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_of(s)));
    return NUIF_OK;
}

int nuif_memcpy_d2d(void* dst, const void* src, size_t bytes, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    if (!bytes) return NUIF_OK;
    NUIF_CHECK_ARG(dst && src, "nuif_memcpy_d2d: null pointer");
    NUIF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_of(s)));
    return NUIF_OK;
}

int nuif_fill_u8(void* dst, uint8_t value, size_t bytes, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    if (!bytes) return NUIF_OK;
    NUIF_CHECK_ARG(dst != nullptr, "nuif_fill_u8: null pointer");
    NUIF_CHECK_ARG(((uintptr_t)dst & 3) == 0, "nuif_fill_u8: pointer must be 4-byte aligned");
    unsigned blocks = std::min<unsigned>(div_up(bytes / 4 + 1, 256), 4096u);
    fill_u8_kernel<<<blocks, 256, 0, stream_of(s)>>>((uint8_t*)dst, value, bytes);
    NUIF_LAUNCHED("fill_u8_kernel");
    return NUIF_OK;
}

int nuif_fill_u32(void* dst, uint32_t value, size_t count, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    if (!count) return NUIF_OK;
    NUIF_CHECK_ARG(dst != nullptr, "nuif_fill_u32: null pointer");
    unsigned blocks = std::min<unsigned>(div_up(count, 256), 4096u);
    fill_u32_kernel<<<blocks, 256, 0, stream_of(s)>>>((uint32_t*)dst, value, count);
    NUIF_LAUNCHED("fill_u32_kernel");
    return NUIF_OK;
}

int nuif_stream_create(nuif_stream_t* s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(s != nullptr, "nuif_stream_create: null out pointer");
    hipStream_t h = nullptr;
    NUIF_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    *s = reinterpret_cast<nuif_stream_t>(h);
    return NUIF_OK;
}

int nuif_stream_destroy(nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    if (s) NUIF_HIP(hipStreamDestroy(reinterpret_cast<hipStream_t>(s)));
    return NUIF_OK;
}

int nuif_stream_synchronize(nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_HIP(hipStreamSynchronize(stream_of(s)));
    return NUIF_OK;
}

int nuif_device_synchronize(void) {
    NUIF_REQUIRE_INIT();
    NUIF_HIP(hipDeviceSynchronize());
    return NUIF_OK;
}

int nuif_event_create(nuif_event_t* e) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(e != nullptr, "nuif_event_create: null out pointer");
    hipEvent_t h = nullptr;
    NUIF_HIP(hipEventCreate(&h));
    *e = reinterpret_cast<nuif_event_t>(h);
    return NUIF_OK;
}

int nuif_event_destroy(nuif_event_t e) {
    NUIF_REQUIRE_INIT();
    if (e) NUIF_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(e)));
    return NUIF_OK;
}

int nuif_event_record(nuif_event_t e, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(e != nullptr, "nuif_event_record: null event");
    NUIF_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(e), stream_of(s)));
    return NUIF_OK;
}

int nuif_event_synchronize(nuif_event_t e) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(e != nullptr, "nuif_event_synchronize: null event");
    NUIF_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(e)));
    return NUIF_OK;
}

int nuif_event_elapsed_ms(nuif_event_t start, nuif_event_t stop, float* ms) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(start && stop && ms, "nuif_event_elapsed_ms: null argument");
    NUIF_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop)));
    return NUIF_OK;
}

int nuif_stream_wait_event(nuif_stream_t s, nuif_event_t e) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(e != nullptr, "nuif_stream_wait_event: null event");
    NUIF_HIP(hipStreamWaitEvent(stream_of(s), reinterpret_cast<hipEvent_t>(e), 0));
    return NUIF_OK;
}

int nuif_graph_begin(nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(s != nullptr, "nuif_graph_begin: capture needs an explicit stream");
    NUIF_HIP(hipStreamBeginCapture(reinterpret_cast<hipStream_t>(s), hipStreamCaptureModeThreadLocal));
    return NUIF_OK;
}

int nuif_graph_end(nuif_stream_t s, nuif_graph_t* g) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(s && g, "nuif_graph_end: null argument");
    hipGraph_t graph = nullptr;
    NUIF_HIP(hipStreamEndCapture(reinterpret_cast<hipStream_t>(s), &graph));
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *g = reinterpret_cast<nuif_graph_t>(exec);
    return NUIF_OK;
}

int nuif_graph_launch(nuif_graph_t g, nuif_stream_t s) {
    NUIF_REQUIRE_INIT();
    NUIF_CHECK_ARG(g != nullptr, "nuif_graph_launch: null graph");
    NUIF_HIP(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(g), stream_of(s)));
    return NUIF_OK;
}

int nuif_graph_destroy(nuif_graph_t g) {
    NUIF_REQUIRE_INIT();
    if (g) NUIF_HIP(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(g)));
    return NUIF_OK;
}

}  // extern "C"
