// bh_runtime.hip -- device selection, memory, stream and event plumbing.
#include <stdarg.h>
#include <stdio.h>

#include "bh_common.hpp"

namespace bh {

static thread_local char g_err[512] = "";
static thread_local hipStream_t g_user_stream = nullptr;
static hipStream_t g_own_stream = nullptr;
static int g_device = -1;
// Library-owned device memory (allocated by bh_init, grown only when a call
// needs more than any earlier one): the entry points take their info words,
// flags and temporaries from here instead of allocating per call.  Calls are
// stream-ordered on the library stream; calls that overlap on different
// streams must not share the library (LAPACK-like reentrancy: one handle per
// task).
static i64 *g_ctrl = nullptr;
static void *g_scratch = nullptr;
static size_t g_scratch_bytes = 0;
constexpr size_t kCtrlBytes = 256;
constexpr size_t kScratchInitial = 8u << 20;

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

int ctrl_words(i64 **p) {
    int st = ensure_init();
    if (st) return st;
    *p = g_ctrl;
    return BH_OK;
}

int scratch_buffer(size_t bytes, void **p) {
    int st = ensure_init();
    if (st) return st;
    if (bytes > g_scratch_bytes) {
        // grow: earlier work on the stream may still read the old buffer
        BH_HIP(hipStreamSynchronize(stream()));
        if (g_scratch) BH_HIP(hipFree(g_scratch));
        g_scratch = nullptr;
        g_scratch_bytes = 0;
        const size_t nb = bytes + bytes / 2;
        BH_HIP(hipMalloc(&g_scratch, nb));
        g_scratch_bytes = nb;
    }
    *p = g_scratch;
    return BH_OK;
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
        // library-owned control words and scratch on this device
        g_ctrl = nullptr;
        g_scratch = nullptr;
        g_scratch_bytes = 0;
        BH_HIP(hipMalloc((void **)&g_ctrl, kCtrlBytes));
        BH_HIP(hipMalloc(&g_scratch, kScratchInitial));
        g_scratch_bytes = kScratchInitial;
    }
    g_device = device;
    // the no-allocation entry point's scratch address on this device; a
    // failure is recorded there and reported by that entry point only
    (void)noalloc_setup(device);
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
