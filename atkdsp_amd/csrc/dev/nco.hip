// The phase-continuous complex mixer (reference src/nco.c:33-98):
//   out[i] = in[i] * exp(j * (phase + i*step)),   phase += n*step, wrapped.
// The phase is carried in FP64 and is the only FP64 on the path: each
// sample's phase is phase + step*i in FP64 (the reference's exact tail form,
// src/nco.c:87-93, used for every sample here, so any chunking gives the same
// values), reduced to [-pi, pi] in FP64, then rotated in FP32 (sincos of the
// reduced angle). The carried phase is wrapped with fmod exactly as the
// reference's atk_wrap (src/internal.h:36-41); contraction is off so the
// device's FP64 arithmetic is the host's, operation for operation.
#include <hip/hip_runtime.h>
#include "atkgpu_dev.h"
#include "dev_common.hpp"

namespace {

constexpr double NCO_PI = 3.14159265358979323846;

__device__ __forceinline__ double nco_wrap(double ph) {
#pragma clang fp contract(off)
    ph = fmod(ph, 2.0 * NCO_PI);
    if (ph >= NCO_PI) ph -= 2.0 * NCO_PI;
    if (ph < -NCO_PI) ph += 2.0 * NCO_PI;
    return ph;
}

__device__ __forceinline__ float2 nco_rotate(float2 x, double ph0, double step, long long i) {
#pragma clang fp contract(off)
    double ph = ph0 + step * (double)i;
    ph -= 2.0 * NCO_PI * rint(ph * (1.0 / (2.0 * NCO_PI)));     // into [-pi, pi]
    float s, c;
    sincosf((float)ph, &s, &c);
    return make_float2(x.x * c - x.y * s, x.x * s + x.y * c);
}

}  // namespace

// drop-in path: samples base .. base+n-1 of a call whose phase is ph0
__global__ void __launch_bounds__(256) k_nco_mix(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                 long long n, long long base, double ph0, double step) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        out[i] = nco_rotate(in[i], ph0, step, base + i);
}

// device-resident path: the phase lives in st[par] = {phase, step}; the
// next phase goes to st[par ^ 1] (and to log[k] when log is set), so no
// block reads a phase another block is writing
__global__ void __launch_bounds__(256) k_nco_run(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                 long long n, const double *__restrict__ st_in,
                                                 double *__restrict__ st_out, double *__restrict__ logp) {
#pragma clang fp contract(off)
    const double ph0 = st_in[0], step = st_in[1];
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        out[i] = nco_rotate(in[i], ph0, step, i);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const double nx = nco_wrap(ph0 + step * (double)n);
        st_out[0] = nx;
        st_out[1] = step;
        if (logp) *logp = nx;
    }
}

static unsigned nco_grid(long long n) {
    long long b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b < 1 ? 1 : b);
}

extern "C" int atkgpu_nco_mix(void *s, const void *in, void *out, size_t n, size_t base, double ph0, double step) {
    if (!n) return 0;
    hipLaunchKernelGGL(k_nco_mix, dim3(nco_grid((long long)n)), dim3(256), 0, (hipStream_t)s, (const float2 *)in,
                       (float2 *)out, (long long)n, (long long)base, ph0, step);
    return LAUNCH_OK();
}

extern "C" int atkgpu_nco_run_dev(void *s, const void *in, void *out, size_t n, const double *st_in, double *st_out,
                              double *logp) {
    hipLaunchKernelGGL(k_nco_run, dim3(nco_grid((long long)n)), dim3(256), 0, (hipStream_t)s, (const float2 *)in,
                       (float2 *)out, (long long)n, st_in, st_out, logp);
    return LAUNCH_OK();
}
