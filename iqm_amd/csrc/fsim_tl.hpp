// Host entry points of the two-level FSIM column and row passes (fsim_tl.hip).
#pragma once

#include "ctx.hpp"

namespace iqm {

struct TlColArgs {
    const float2* A;      /* row spectra of the chunk, [pair][2][N] */
    float2* T;            /* response planes of the chunk, [pair][32][N] */
    const float* bank;    /* 16 filter planes */
    const float2* tw;     /* exp(-2 pi i j / hd) */
    int wd, npairs, bx, xcds;
    size_t lds_budget;
    const int* gate;      /* nullable device flag: 0 skips the launch's work */
};

struct TlRowArgs {
    const float2* T;
    float* an_e;
    float* keys;
    unsigned* hist;
    const float2* tw;     /* exp(-2 pi i j / wd) */
    int hd, npairs, rb;
    float invN;
    size_t lds_budget;
    const int* gate;
};

struct TlGradArgs {
    const float* yiq;
    float2* A;
    const float2* tw;     /* exp(-2 pi i j / wd) */
    int hd, npairs, rb;
    size_t lds_budget;
    const int* gate;
};

/* F12 (fsim_sum_filter_responses.glsl:33-83) for one pixel and orientation
 * from the 4 scales' responses before the inverse transform's 1/N: with
 * eo = invN * eo_u, the energy normaliser |sum eo| + 1e-4 is invN * (|sum eo_u|
 * + 1e-4 * N), so the unit vector (me, mo) is the same and An, E scale by
 * invN and the median key |eo_0|^2 by invN^2.  One reciprocal replaces the two
 * divisions; square roots are the hardware's (1 ulp). */
__device__ __forceinline__ void fsim_response_epilogue(const float2 (&eo)[4], float invN, float epsN, float& amp,
                                                       float& energy, float& key) {
    float a = __builtin_amdgcn_sqrtf(eo[0].x * eo[0].x + eo[0].y * eo[0].y), sr = eo[0].x, si = eo[0].y;
#pragma unroll
    for (int s = 1; s < 4; ++s) {
        a += __builtin_amdgcn_sqrtf(eo[s].x * eo[s].x + eo[s].y * eo[s].y);
        sr += eo[s].x;
        si += eo[s].y;
    }
    const float rx = __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(sr * sr + si * si) + epsN);
    const float me = sr * rx, mo = si * rx;
    /* sum_s (eo_s . m - |eo_s x m|) with the dot products summed as
     * (sum_s eo_s) . m */
    float d = fabsf(eo[0].x * mo - eo[0].y * me);
#pragma unroll
    for (int s = 1; s < 4; ++s) d += fabsf(eo[s].x * mo - eo[s].y * me);
    const float e = (sr * me + si * mo) - d;
    amp = a * invN;
    energy = e * invN;
    key = (eo[0].x * eo[0].x + eo[0].y * eo[0].y) * (invN * invN);
}

bool tl_supported(int n);
void tl_rowfft(Ctx& c, int wd, const TlGradArgs& a);
void tl_colpass(Ctx& c, int hd, const TlColArgs& a);
void tl_rowpass(Ctx& c, int wd, const TlRowArgs& a);

}  // namespace iqm
