// Shared helpers for the ephip HIP library: status codes, the last-error string, argument checks
// and small device utilities. Everything here is internal; the public C-ABI is include/ephip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../../include/ephip.h"

namespace ephip {

void set_error(const char *fmt, ...);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace ephip

#define EPHIP_TRY_HIP(expr)                                                                    \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            ephip::set_error("%s failed: %s", #expr, hipGetErrorString(e_));                 \
            return EPHIP_ERR_HIP;                                                              \
        }                                                                                      \
    } while (0)

#define EPHIP_REQUIRE(cond, ...)                                                               \
    do {                                                                                       \
        if (!(cond)) {                                                                         \
            ephip::set_error(__VA_ARGS__);                                                     \
            return EPHIP_ERR_ARG;                                                              \
        }                                                                                      \
    } while (0)

// Launch check: a bad launch configuration surfaces here, asynchronous faults at the next sync.
#define EPHIP_LAUNCHED()                                                                       \
    do {                                                                                       \
        hipError_t e_ = hipGetLastError();                                                     \
        if (e_ != hipSuccess) {                                                                \
            ephip::set_error("kernel launch failed: %s", hipGetErrorString(e_));              \
            return EPHIP_ERR_HIP;                                                              \
        }                                                                                      \
    } while (0)

namespace ephip {

// Complex value with interleaved (re, im), the layout of complex64/complex128 device arrays.
template <class T>
struct cplx {
    T re, im;
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Make this lane's LDS writes visible to the other lanes of its own wavefront and keep the
// compiler from moving LDS accesses across the point. No cross-wave synchronisation.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// exp(2 pi i t) for an FP64 argument: reduce t to [-1/2, 1/2] in FP64, then evaluate sin/cos
// of pi * 2t in the target precision.
__device__ __forceinline__ void cis2pi(double t, float &c, float &s) {
    t -= rint(t);
    sincospif(2.0f * static_cast<float>(t), &s, &c);
}
__device__ __forceinline__ void cis2pi(double t, double &c, double &s) {
    t -= rint(t);
    sincospi(2.0 * t, &s, &c);
}

}  // namespace ephip
