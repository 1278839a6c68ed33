// bh_runtime.hip -- device selection, memory, stream and event plumbing.
#include <stdarg.h>
#include <stdio.h>

#include "bh_common.hpp"

namespace bh {

static thread_local char g_err[512] = "";
static thread_local hipStream_t g_user_stream = nullptr;
static hipStream_t g_own_stream = nullptr;
static int g_device = -1;

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? BH_ERR_ALLOC : BH_ERR_HIP;
}

static int ensure_init() {
    if (g_device >= 0) return BH_OK;
    return bh_init(0);
}

hipStream_t stream() {
    if (g_user_stream) return g_user_stream;
    return g_own_stream;
}

}  // namespace bh

using namespace bh;

extern "C" {

int bh_init(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        set_error("no HIP device visible (%s)", hipGetErrorString(e));
        return BH_ERR_NODEVICE;
    }
    if (device < 0 || device >= count) {
        set_error("device %d out of range (%d visible)", device, count);
        return -1;
    }
    hipDeviceProp_t prop;
    BH_HIP(hipGetDeviceProperties(&prop, device));
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) {
        set_error("device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);
        return BH_ERR_NODEVICE;
    }
    BH_HIP(hipSetDevice(device));
    if (g_own_stream == nullptr || g_device != device) {
        // Default priority only: the library never changes stream priorities.
        BH_HIP(hipStreamCreateWithFlags(&g_own_stream, hipStreamNonBlocking));
    }
    g_device = device;
    return BH_OK;
}

int bh_device_count(int *count) {
    BH_ARG(count != nullptr, 1);
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) *count = 0;
    return BH_OK;
}

const char *bh_last_error(void) { return g_err; }

int bh_malloc(void **dptr, size_t bytes) {
    BH_ARG(dptr != nullptr, 1);
    int st = ensure_init();
    if (st) return st;
    *dptr = nullptr;
    if (bytes == 0) return BH_OK;
    BH_HIP(hipMalloc(dptr, bytes));
    return BH_OK;
}

int bh_free(void *dptr) {
    if (dptr == nullptr) return BH_OK;
    BH_HIP(hipFree(dptr));
    return BH_OK;
}

int bh_memcpy_htod(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return BH_OK;
    int st = ensure_init();
    if (st) return st;
    BH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    BH_HIP(hipStreamSynchronize(stream()));
    return BH_OK;
}

int bh_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return BH_OK;
    int st = ensure_init();
    if (st) return st;
    BH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    BH_HIP(hipStreamSynchronize(stream()));
    return BH_OK;
}

int bh_memcpy_dtod(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return BH_OK;
    int st = ensure_init();
    if (st) return st;
    BH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
    return BH_OK;
}

int bh_memset(void *dptr, int value, size_t bytes) {
    if (bytes == 0) return BH_OK;
    int st = ensure_init();
    if (st) return st;
    BH_HIP(hipMemsetAsync(dptr, value, bytes, stream()));
    return BH_OK;
}

int bh_sync(void) {
    int st = ensure_init();
    if (st) return st;
    BH_HIP(hipStreamSynchronize(stream()));
    return BH_OK;
}

int bh_set_stream(void *s) {
    g_user_stream = (hipStream_t)s;
    return BH_OK;
}

void *bh_get_stream(void) { return (void *)stream(); }

int bh_event_create(void **ev) {
    BH_ARG(ev != nullptr, 1);
    int st = ensure_init();
    if (st) return st;
    hipEvent_t e;
    BH_HIP(hipEventCreate(&e));
    *ev = (void *)e;
    return BH_OK;
}

int bh_event_record(void *ev) {
    BH_HIP(hipEventRecord((hipEvent_t)ev, stream()));
    return BH_OK;
}

int bh_event_elapsed_ms(void *start, void *stop, float *ms) {
    BH_HIP(hipEventSynchronize((hipEvent_t)stop));
    BH_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return BH_OK;
}

int bh_event_destroy(void *ev) {
    if (ev) BH_HIP(hipEventDestroy((hipEvent_t)ev));
    return BH_OK;
}

}  // extern "C"
