// HIP runtime plumbing behind the thin C ABI (memory, streams, events, graphs).
#include <stdio.h>
#include <string.h>

#include "ust_common.h"

static thread_local char g_err[512];

void ust_set_hip_error(hipError_t e, const char *file, int line) {
    snprintf(g_err, sizeof(g_err), "%s (%s:%d)", hipGetErrorString(e), file, line);
}

void ust_set_error_text(const char *text) { snprintf(g_err, sizeof(g_err), "%s", text); }

#define HIPCALL(x)                                         \
    do {                                                   \
        hipError_t _e = (x);                               \
        if (_e != hipSuccess) {                            \
            ust_set_hip_error(_e, __FILE__, __LINE__);     \
            return UST_EDEVICE;                            \
        }                                                  \
    } while (0)

extern "C" {

const char *ust_dev_error_string(void) { return g_err; }

int ust_dev_set(int device) {
    HIPCALL(hipSetDevice(device));
    return UST_OK;
}

int ust_dev_malloc(void **p, size_t bytes) {
    *p = NULL;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        ust_set_hip_error(e, __FILE__, __LINE__);
        *p = NULL;
        return UST_ENOMEM;
    }
    return UST_OK;
}

int ust_dev_free(void *p) {
    if (p) HIPCALL(hipFree(p));
    return UST_OK;
}

int ust_host_malloc_pinned(void **p, size_t bytes) {
    HIPCALL(hipHostMalloc(p, bytes == 0 ? 256 : bytes, hipHostMallocDefault));
    return UST_OK;
}

int ust_host_free_pinned(void *p) {
    if (p) HIPCALL(hipHostFree(p));
    return UST_OK;
}

int ust_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return UST_OK;
    HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ust_stream(stream)));
    HIPCALL(hipStreamSynchronize(ust_stream(stream)));
    return UST_OK;
}

int ust_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return UST_OK;
    HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ust_stream(stream)));
    HIPCALL(hipStreamSynchronize(ust_stream(stream)));
    return UST_OK;
}

int ust_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return UST_OK;
    HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ust_stream(stream)));
    return UST_OK;
}

int ust_stream_create(void **stream) {
    hipStream_t s;
    HIPCALL(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return UST_OK;
}

int ust_stream_destroy(void *stream) {
    if (stream) HIPCALL(hipStreamDestroy(ust_stream(stream)));
    return UST_OK;
}

int ust_stream_sync(void *stream) {
    HIPCALL(hipStreamSynchronize(ust_stream(stream)));
    return UST_OK;
}

int ust_event_create(void **ev) {
    hipEvent_t e;
    HIPCALL(hipEventCreate(&e));
    *ev = (void *)e;
    return UST_OK;
}

int ust_event_destroy(void *ev) {
    if (ev) HIPCALL(hipEventDestroy((hipEvent_t)ev));
    return UST_OK;
}

int ust_event_record(void *ev, void *stream) {
    HIPCALL(hipEventRecord((hipEvent_t)ev, ust_stream(stream)));
    return UST_OK;
}

int ust_event_elapsed_ms(void *a, void *b, float *ms) {
    HIPCALL(hipEventSynchronize((hipEvent_t)b));
    HIPCALL(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return UST_OK;
}

// Stream capture of a fixed launch sequence into an instantiated HIP graph.  Only kernel launches are
// captured by the callers (no memset nodes), so replays see exactly the captured arguments.
int ust_graph_begin(void *stream) {
    HIPCALL(hipStreamBeginCapture(ust_stream(stream), hipStreamCaptureModeThreadLocal));
    return UST_OK;
}

int ust_graph_end(void *stream, void **exec) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    *exec = nullptr;
    HIPCALL(hipStreamEndCapture(ust_stream(stream), &g));
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        ust_set_hip_error(e, __FILE__, __LINE__);
        return UST_EDEVICE;
    }
    *exec = (void *)ge;
    return UST_OK;
}

int ust_graph_abort(void *stream) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ust_stream(stream), &g);
    if (g) hipGraphDestroy(g);
    (void)hipGetLastError();
    return e == hipSuccess ? UST_OK : UST_EDEVICE;
}

int ust_graph_launch(void *exec, void *stream) {
    HIPCALL(hipGraphLaunch((hipGraphExec_t)exec, ust_stream(stream)));
    return UST_OK;
}

int ust_graph_destroy(void *exec) {
    if (exec) HIPCALL(hipGraphExecDestroy((hipGraphExec_t)exec));
    return UST_OK;
}

}  // extern "C"
