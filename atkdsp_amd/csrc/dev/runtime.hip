// HIP runtime wrappers for the C host layer: device/pinned allocation,
// default-priority streams, copies and events. Every failure maps to
// ATKDSP_E_FFT (-4) and nothing is printed (the ABI does no I/O).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"

static char g_devname[128] = "";

extern "C" int atkgpu_noalloc_setup(void);

extern "C" int atkgpu_init(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return -4;
    // the allocation-free entry point's scratch address on this device
    if (atkgpu_noalloc_setup() != 0) return -4;
    if (!g_devname[0]) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess)
            snprintf(g_devname, sizeof g_devname, "%s %s", prop.name, prop.gcnArchName);
    }
    return 0;
}

extern "C" const char *atkgpu_device_name(void) { return g_devname; }

extern "C" void *atkgpu_dmalloc(size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 256;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();            // the caller handles NULL; no error is left for a launch check
        return nullptr;
    }
    return p;
}

extern "C" void atkgpu_dfree(void *p) {
    if (p) (void)hipFree(p);
}

extern "C" void *atkgpu_hmalloc(size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 256;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

extern "C" void atkgpu_hfree(void *p) {
    if (p) (void)hipHostFree(p);
}

extern "C" int atkgpu_stream_create(void **s) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return -4;
    *s = (void *)st;
    return 0;
}

extern "C" void atkgpu_stream_destroy(void *s) {
    if (s) (void)hipStreamDestroy((hipStream_t)s);
}

extern "C" int atkgpu_sync(void *s) {
    return hipStreamSynchronize((hipStream_t)s) == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_h2d(void *s, void *dst, const void *src, size_t bytes) {
    if (!bytes) return 0;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)s) == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_d2h(void *s, void *dst, const void *src, size_t bytes) {
    if (!bytes) return 0;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s) == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_d2d(void *s, void *dst, const void *src, size_t bytes) {
    if (!bytes) return 0;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)s) == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_d2h_2d(void *s, void *dst, size_t dpitch, const void *src, size_t spitch,
                             size_t width, size_t height) {
    if (!width || !height) return 0;
    return hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToHost,
                            (hipStream_t)s) == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_d2d_2d(void *s, void *dst, size_t dpitch, const void *src, size_t spitch,
                             size_t width, size_t height) {
    if (!width || !height) return 0;
    return hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToDevice,
                            (hipStream_t)s) == hipSuccess ? 0 : -4;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// A fill kernel rather than hipMemsetAsync, so that callers may capture it in a graph.
extern "C" int atkgpu_fill_u32(void *s, void *dst, uint32_t v, size_t count) {
    if (!count) return 0;
    size_t blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)s,
                       (uint32_t *)dst, v, count);
    return LAUNCH_OK();
}

extern "C" int atkgpu_event_create(void **e) {
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) return -4;
    *e = (void *)ev;
    return 0;
}

// an event for ordering only (no timestamps: cheaper to record and wait on)
extern "C" int atkgpu_event_create_order(void **e) {
    hipEvent_t ev;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return -4;
    *e = (void *)ev;
    return 0;
}

// work submitted to s after this call waits for e's last record
extern "C" int atkgpu_stream_wait_event(void *s, void *e) {
    return hipStreamWaitEvent((hipStream_t)s, (hipEvent_t)e, 0) == hipSuccess ? 0 : -4;
}

extern "C" void atkgpu_event_destroy(void *e) {
    if (e) (void)hipEventDestroy((hipEvent_t)e);
}

extern "C" int atkgpu_event_record(void *e, void *s) {
    return hipEventRecord((hipEvent_t)e, (hipStream_t)s) == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_event_elapsed_ms(void *e0, void *e1, float *ms) {
    if (hipEventSynchronize((hipEvent_t)e1) != hipSuccess) return -4;
    return hipEventElapsedTime(ms, (hipEvent_t)e0, (hipEvent_t)e1) == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_event_query(void *e) {
    const hipError_t r = hipEventQuery((hipEvent_t)e);
    if (r == hipSuccess) return 1;
    if (r == hipErrorNotReady) {
        (void)hipGetLastError();            // "not yet" is no error for a later launch check
        return 0;
    }
    return -4;
}

extern "C" int atkgpu_event_sync(void *e) {
    return hipEventSynchronize((hipEvent_t)e) == hipSuccess ? 0 : -4;
}

// a synchronous device-to-host copy on the null stream (reads of results
// already complete on their own stream)
extern "C" int atkgpu_d2h_sync(void *dst, const void *src, size_t bytes) {
    if (!bytes) return 0;
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -4;
}

extern "C" int atkgpu_d2h_2d_sync(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                                  size_t height) {
    if (!width || !height) return 0;
    return hipMemcpy2D(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -4;
}
