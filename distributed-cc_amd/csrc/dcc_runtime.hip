// Device, memory, stream and event entry points of libdcc (see include/dcc.h).
#include "dcc_common.h"

#include <cstring>

namespace dcc {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

}  // namespace dcc

namespace {

template <typename T>
__global__ void fill_kernel(T *__restrict__ dst, int64_t n, T value) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = value;
}

}  // namespace

extern "C" {

int dcc_abi_version(void) { return 1; }

const char *dcc_last_error(void) { return dcc::g_last_error.c_str(); }

int dcc_device_count(int *count) {
    DCC_CHECK(count != nullptr, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    (void)hipGetLastError();
    *count = n;
    return 0;
}

int dcc_set_device(int device, char *arch_name, int64_t arch_len) {
    DCC_HIP(hipSetDevice(device));
    if (arch_name != nullptr && arch_len > 0) {
        hipDeviceProp_t prop;
        DCC_HIP(hipGetDeviceProperties(&prop, device));
        std::strncpy(arch_name, prop.gcnArchName, (size_t)arch_len - 1);
        arch_name[arch_len - 1] = '\0';
    }
    return 0;
}

int dcc_malloc(void **ptr, int64_t bytes) {
    DCC_CHECK(ptr != nullptr && bytes >= 0, "bad arguments");
    *ptr = nullptr;
    if (bytes == 0) return 0;
    DCC_HIP(hipMalloc(ptr, (size_t)bytes));
    return 0;
}

int dcc_free(void *ptr) {
    if (ptr != nullptr) DCC_HIP(hipFree(ptr));
    return 0;
}

int dcc_memcpy_htod(void *dst, const void *src, int64_t bytes, dcc_stream_t stream) {
    if (bytes <= 0) return 0;
    DCC_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, dcc::as_stream(stream)));
    DCC_HIP(hipStreamSynchronize(dcc::as_stream(stream)));
    return 0;
}

int dcc_memcpy_dtoh(void *dst, const void *src, int64_t bytes, dcc_stream_t stream) {
    if (bytes <= 0) return 0;
    DCC_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, dcc::as_stream(stream)));
    DCC_HIP(hipStreamSynchronize(dcc::as_stream(stream)));
    return 0;
}

int dcc_memcpy_dtod(void *dst, const void *src, int64_t bytes, dcc_stream_t stream) {
    if (bytes <= 0) return 0;
    DCC_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, dcc::as_stream(stream)));
    return 0;
}

int dcc_fill_f64(double *dst, int64_t n, double value, dcc_stream_t stream) {
    if (n <= 0) return 0;
    DCC_CHECK(dst != nullptr, "null dst");
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, dcc::as_stream(stream), dst, n,
                       value);
    DCC_LAUNCH_CHECK();
    return 0;
}

int dcc_fill_f32(float *dst, int64_t n, float value, dcc_stream_t stream) {
    if (n <= 0) return 0;
    DCC_CHECK(dst != nullptr, "null dst");
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, dcc::as_stream(stream), dst, n,
                       value);
    DCC_LAUNCH_CHECK();
    return 0;
}

int dcc_mem_get_info(int64_t *free_bytes, int64_t *total_bytes) {
    DCC_CHECK(free_bytes != nullptr && total_bytes != nullptr, "null pointer");
    size_t f = 0, t = 0;
    DCC_HIP(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return 0;
}

int dcc_stream_create(dcc_stream_t *stream) {
    DCC_CHECK(stream != nullptr, "null stream");
    hipStream_t s;
    DCC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = reinterpret_cast<dcc_stream_t>(s);
    return 0;
}

int dcc_stream_destroy(dcc_stream_t stream) {
    if (stream != nullptr) DCC_HIP(hipStreamDestroy(dcc::as_stream(stream)));
    return 0;
}

int dcc_stream_synchronize(dcc_stream_t stream) {
    DCC_HIP(hipStreamSynchronize(dcc::as_stream(stream)));
    return 0;
}

int dcc_device_synchronize(void) {
    DCC_HIP(hipDeviceSynchronize());
    return 0;
}

int dcc_event_create(dcc_event_t *event) {
    DCC_CHECK(event != nullptr, "null event");
    hipEvent_t e;
    DCC_HIP(hipEventCreate(&e));
    *event = reinterpret_cast<dcc_event_t>(e);
    return 0;
}

int dcc_event_destroy(dcc_event_t event) {
    if (event != nullptr) DCC_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
    return 0;
}

int dcc_event_record(dcc_event_t event, dcc_stream_t stream) {
    DCC_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(event), dcc::as_stream(stream)));
    return 0;
}

int dcc_event_elapsed_ms(float *ms, dcc_event_t start, dcc_event_t stop) {
    DCC_CHECK(ms != nullptr, "null ms");
    DCC_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
    DCC_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop)));
    return 0;
}

}  // extern "C"
