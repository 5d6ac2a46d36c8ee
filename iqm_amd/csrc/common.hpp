// Shared definitions of the HIP library: error handling, complex helpers,
// launch constants.  Every kernel in this library is written for gfx950
// (wave64, 160 KB LDS per CU); no other target is built.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "iqm/iqm.h"

namespace iqm {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define IQM_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::iqm::Error(IQM_ERR_DEVICE,                                          \
                               std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define IQM_CHECK(cond, code, msg)                                                      \
    do {                                                                                \
        if (!(cond)) throw ::iqm::Error((code), (msg));                                 \
    } while (0)

__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) {
    return make_float2(a.x + b.x, a.y + b.y);
}
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) {
    return make_float2(a.x - b.x, a.y - b.y);
}
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) {
    return make_float2(a.x * s, a.y * s);
}
/* multiply by s*i, s = +1 or -1 */
template <int S>
__host__ __device__ __forceinline__ float2 cmul_si(float2 a) {
    return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

constexpr int kBlock = 256;          // default workgroup size (4 wavefronts)
constexpr int kBlockFft = 512;       // workgroup size of the LDS line-transform kernels

/* Workgroup barrier that orders LDS only: waits for this wave's LDS
 * operations, then joins the barrier.  Global stores issued before it stay in
 * flight (a __syncthreads() would wait for them).  Use only where the
 * barrier protects LDS data. */
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}
/* Streaming (non-temporal) 8-byte accesses for data touched once. */
typedef float v2f_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void nt_store(float2* p, float2 v) {
    v2f_t t = {v.x, v.y};
    __builtin_nontemporal_store(t, (v2f_t*)p);
}
__device__ __forceinline__ float2 nt_load(const float2* p) {
    const v2f_t t = __builtin_nontemporal_load((const v2f_t*)p);
    return make_float2(t.x, t.y);
}

/* base + i elements with the byte offset formed in 32 bits (i * sizeof(T)
 * < 4 GiB): a workgroup-uniform base then stays in SGPRs and the access
 * needs no 64-bit address arithmetic per lane. */
template <class T>
__device__ __forceinline__ T* at32(T* base, unsigned i) {
    return (T*)((char*)base + i * (unsigned)sizeof(T));
}

/* Device-side launch gate: a kernel given a gate returns at once when the
 * flag reads 0 (a decision an earlier kernel of the stream wrote), so a
 * captured launch sequence can skip work without a host round trip.  A null
 * gate means always run. */
__device__ __forceinline__ bool gate_closed(const int* gate) { return gate != nullptr && *gate == 0; }

constexpr int kMaxLdsBytes = 65536;  // per-workgroup LDS budget used by every kernel

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace iqm
