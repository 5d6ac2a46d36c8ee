// Line FFTs in LDS: the project's replacement of VkFFT
// (reference call sites src/gpu/fsim/fsim.cpp:370-437 and :527-550).
//
// A workgroup transforms `nlines` lines held in LDS with a Stockham autosort
// schedule: stage j of radix r reads element i + t*(n/r) of the source
// buffer, applies the twiddle exp(S*2*pi*i*t*k/(r*p)) with k = i mod p
// (p = product of the earlier radices), runs a radix-r DFT and writes
// element (i/p)*r*p + k + t*p of the destination buffer.  Buffers ping-pong
// between stages and the result ends in natural order.
//
// Radices 2, 3, 4, 5, 7, 8, 11 and 13 have compile-time butterflies.  A
// length with a larger prime factor goes through Bluestein's chirp-z
// identity X_k = c_k * sum_j (x_j c_j) conj(c_{k-j}), c_j = exp(S*pi*i*j^2/n),
// evaluated as a circular convolution of length m >= 2n-1 (the smallest such
// m the line transforms handle directly).
// Twiddles and chirps are computed in double on the host and stored as float.
#pragma once

#include "common.hpp"

#include <vector>

namespace iqm {

constexpr int kMaxStages = 16;
constexpr int kMaxLdsTwiddles = 4096;

/* Kernel-argument view of a line plan (passed by value). */
struct LinePlan {
    int n;                 /* transform length                                  */
    int nst;               /* direct plan: number of stages                     */
    int rad[kMaxStages];   /* direct plan: radices in application order         */
    const float2* tw;      /* direct plan: exp(-2 pi i j / n), j < n            */
    int blue;              /* 1: Bluestein                                      */
    int m;                 /* Bluestein convolution length (power of two)       */
    int mst;               /* stages of the length-m plan                       */
    int mrad[kMaxStages];
    const float2* mtw;     /* exp(-2 pi i j / m), j < m                         */
    const float2* chirp;   /* exp(-pi i (j^2 mod 2n) / n), j < n                */
    const float2* gf;      /* FFT_m of the forward convolution kernel           */
    const float2* gi;      /* FFT_m of the inverse convolution kernel           */
    int lds_len;           /* elements a line occupies in LDS (n or m, or the
                              two-level exchange area when that is longer)      */
    int tn1, tn2;          /* runtime two-level split n = tn1 * tn2 (0: none)   */
    int tm1, tm2;          /* the same for the Bluestein length m               */
};

/* Sizes of the register DFTs the runtime two-level transform dispatches to
 * (fft_line.hpp); a length n = n1 * n2 with both in this set is transformed
 * with one LDS exchange instead of one LDS pass per radix stage. */
#define IQM_TL2_SIZES(X)                                                                                    \
    X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(18) X(20) X(21) X(22) \
    X(24) X(25) X(26) X(27) X(28) X(30) X(32)

inline bool tl2_size(int m) {
    switch (m) {
#define IQM_TL2_CASE(k) case k:
        IQM_TL2_SIZES(IQM_TL2_CASE)
#undef IQM_TL2_CASE
        return true;
        default: return false;
    }
}

/* Most balanced split n = n1 * n2 with both factors in IQM_TL2_SIZES
 * (n1 <= n2); false when there is none. */
inline bool tl2_split(int n, int& n1, int& n2) {
    n1 = n2 = 0;
    for (int a = 2; a * a <= n; ++a)
        if (n % a == 0 && tl2_size(a) && tl2_size(n / a)) {
            n1 = a;
            n2 = n / a;
        }
    return n1 > 0;
}

/* ---- small-radix DFTs (y_t = sum_q x_q exp(S*2*pi*i*q*t/R)) -------------- */

template <int R>
struct OddTab;
template <>
struct OddTab<3> {
    static constexpr float C[3] = {1.0f, -0.5f, -0.5f};
    static constexpr float S[3] = {0.0f, 0.86602540378443864676f, -0.86602540378443864676f};
};
template <>
struct OddTab<5> {
    static constexpr float C[5] = {1.0f, 0.30901699437494742410f, -0.80901699437494742410f,
                                   -0.80901699437494742410f, 0.30901699437494742410f};
    static constexpr float S[5] = {0.0f, 0.95105651629515357212f, 0.58778525229247312917f,
                                   -0.58778525229247312917f, -0.95105651629515357212f};
};
template <>
struct OddTab<7> {
    static constexpr float C[7] = {1.0f, 0.62348980185873353053f, -0.22252093395631440429f,
                                   -0.90096886790241912624f, -0.90096886790241912624f,
                                   -0.22252093395631440429f, 0.62348980185873353053f};
    static constexpr float S[7] = {0.0f, 0.78183148246802980871f, 0.97492791218182360702f,
                                   0.43388373911755812048f, -0.43388373911755812048f,
                                   -0.97492791218182360702f, -0.78183148246802980871f};
};
template <>
struct OddTab<11> {
    static constexpr float C[11] = {1.0f, 0.84125353283118116886f, 0.41541501300188642553f,
                                    -0.14231483827328514044f, -0.65486073394528506406f,
                                    -0.95949297361449738989f, -0.95949297361449738989f,
                                    -0.65486073394528506406f, -0.14231483827328514044f,
                                    0.41541501300188642553f, 0.84125353283118116886f};
    static constexpr float S[11] = {0.0f, 0.54064081745559758211f, 0.90963199535451837141f,
                                    0.98982144188093273238f, 0.75574957435425828377f,
                                    0.28173255684142969771f, -0.28173255684142969771f,
                                    -0.75574957435425828377f, -0.98982144188093273238f,
                                    -0.90963199535451837141f, -0.54064081745559758211f};
};
template <>
struct OddTab<13> {
    static constexpr float C[13] = {1.0f, 0.88545602565320989590f, 0.56806474673115580251f,
                                    0.12053668025532305335f, -0.35460488704253562597f,
                                    -0.74851074817110109863f, -0.97094181742605202716f,
                                    -0.97094181742605202716f, -0.74851074817110109863f,
                                    -0.35460488704253562597f, 0.12053668025532305335f,
                                    0.56806474673115580251f, 0.88545602565320989590f};
    static constexpr float S[13] = {0.0f, 0.46472317204376854566f, 0.82298386589365639458f,
                                    0.99270887409805399280f, 0.93501624268541482344f,
                                    0.66312265824079520238f, 0.23931566428755776715f,
                                    -0.23931566428755776715f, -0.66312265824079520238f,
                                    -0.93501624268541482344f, -0.99270887409805399280f,
                                    -0.82298386589365639458f, -0.46472317204376854566f};
};

/* Odd prime radix: pairs q and R-q share the cosine terms. */
template <int R, int S>
__host__ __device__ __forceinline__ void dft_odd(float2* x) {
    constexpr int H = (R - 1) / 2;
    float2 a[H + 1], b[H + 1];
    float2 y0 = x[0];
#pragma unroll
    for (int q = 1; q <= H; ++q) {
        a[q] = cadd(x[q], x[R - q]);
        b[q] = csub(x[q], x[R - q]);
        y0 = cadd(y0, a[q]);
    }
    float2 out[R];
    out[0] = y0;
#pragma unroll
    for (int t = 1; t <= H; ++t) {
        float2 P = x[0];
        float2 Q = make_float2(0.f, 0.f);
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            const float c = OddTab<R>::C[(q * t) % R];
            const float s = OddTab<R>::S[(q * t) % R];
            P.x += a[q].x * c;
            P.y += a[q].y * c;
            Q.x += b[q].x * s;
            Q.y += b[q].y * s;
        }
        const float2 iq = cmul_si<S>(Q);
        out[t] = cadd(P, iq);
        out[R - t] = csub(P, iq);
    }
#pragma unroll
    for (int t = 0; t < R; ++t) x[t] = out[t];
}

template <int S>
__host__ __device__ __forceinline__ void dft2(float2* x) {
    float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
}

template <int S>
__host__ __device__ __forceinline__ void dft4(float2* x) {
    float2 a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
    float2 a2 = cadd(x[1], x[3]), a3 = cmul_si<S>(csub(x[1], x[3]));
    x[0] = cadd(a0, a2);
    x[2] = csub(a0, a2);
    x[1] = cadd(a1, a3);
    x[3] = csub(a1, a3);
}

template <int S>
__host__ __device__ __forceinline__ void dft8(float2* x) {
    /* radix-4 DFTs of the even and odd samples, then a radix-2 combine with
     * w8^k = exp(S*2*pi*i*k/8) */
    float2 e[4] = {x[0], x[2], x[4], x[6]};
    float2 o[4] = {x[1], x[3], x[5], x[7]};
    dft4<S>(e);
    dft4<S>(o);
    const float h = 0.70710678118654752440f;
    const float2 w1 = make_float2(h, S * h), w3 = make_float2(-h, S * h);
    o[1] = cmul(o[1], w1);
    o[2] = cmul_si<S>(o[2]);
    o[3] = cmul(o[3], w3);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        x[k] = cadd(e[k], o[k]);
        x[k + 4] = csub(e[k], o[k]);
    }
}

template <int R, int S>
__host__ __device__ __forceinline__ void dft_r(float2* x) {
    if constexpr (R == 2) dft2<S>(x);
    else if constexpr (R == 4) dft4<S>(x);
    else if constexpr (R == 8) dft8<S>(x);
    else dft_odd<R, S>(x);
}

/* q = a / d for 0 <= a < 2^22 via a float reciprocal and one correction. */
__host__ __device__ __forceinline__ int fast_div(int a, int d, float inv) {
    int q = (int)((float)a * inv);
    const int r = a - q * d;
    if (r >= d) ++q;
    else if (r < 0) --q;
    return q;
}

/* One Stockham stage over nlines lines (line l at src + l*ld); tw holds
 * exp(-2 pi i j / n). */
template <int R, int S>
__host__ __device__ __forceinline__ void stockham_stage(const float2* src, float2* dst, int ld,
                                                        int nlines, int n, int p,
                                                        const float2* tw, int tid, int nthr) {
    const int T = n / R;
    const int total = nlines * T;
    const int tws = n / (R * p);
    const float invT = 1.0f / (float)T, invp = 1.0f / (float)p;
    for (int idx = tid; idx < total; idx += nthr) {
        const int line = fast_div(idx, T, invT);
        const int i = idx - line * T;
        const int k = i - fast_div(i, p, invp) * p;
        const float2* s = src + line * ld;
        float2 x[R];
#pragma unroll
        for (int t = 0; t < R; ++t) x[t] = s[i + t * T];
        /* k = 0 multiplies by tw[0] = (1, -0), which leaves a finite x[t]
         * unchanged; no branch on the per-lane k */
#pragma unroll
        for (int t = 1; t < R; ++t) {
            float2 w = tw[t * k * tws];
            if (S > 0) w.y = -w.y;
            x[t] = cmul(x[t], w);
        }
        dft_r<R, S>(x);
        float2* d = dst + line * ld + (i - k) * R + k;
#pragma unroll
        for (int t = 0; t < R; ++t) d[t * p] = x[t];
    }
}

template <int S>
__host__ __device__ __forceinline__ void stockham_dispatch(int r, const float2* src, float2* dst,
                                                           int ld, int nlines, int n, int p,
                                                           const float2* tw, int tid, int nthr) {
    switch (r) {
        case 2: stockham_stage<2, S>(src, dst, ld, nlines, n, p, tw, tid, nthr); break;
        case 3: stockham_stage<3, S>(src, dst, ld, nlines, n, p, tw, tid, nthr); break;
        case 4: stockham_stage<4, S>(src, dst, ld, nlines, n, p, tw, tid, nthr); break;
        case 5: stockham_stage<5, S>(src, dst, ld, nlines, n, p, tw, tid, nthr); break;
        case 7: stockham_stage<7, S>(src, dst, ld, nlines, n, p, tw, tid, nthr); break;
        case 8: stockham_stage<8, S>(src, dst, ld, nlines, n, p, tw, tid, nthr); break;
        case 11: stockham_stage<11, S>(src, dst, ld, nlines, n, p, tw, tid, nthr); break;
        case 13: stockham_stage<13, S>(src, dst, ld, nlines, n, p, tw, tid, nthr); break;
        default: break;
    }
}

/* Direct mixed-radix transform; returns the buffer holding the result.
 * Ends with a workgroup barrier. */
template <int S>
__device__ __forceinline__ float2* lds_fft_direct(float2* buf0, float2* buf1, int ld, int nlines,
                                                  int n, int nst, const int* rad,
                                                  const float2* tw) {
    float2* src = buf0;
    float2* dst = buf1;
    int p = 1;
    for (int st = 0; st < nst; ++st) {
        const int r = rad[st];
        stockham_dispatch<S>(r, src, dst, ld, nlines, n, p, tw, threadIdx.x, blockDim.x);
        lds_barrier();
        float2* t = src;
        src = dst;
        dst = t;
        p *= r;
    }
    return src;
}

/* Number of float2 twiddles a kernel stages in LDS for this plan (0 when the
 * plan reads them from global memory). */
__host__ __device__ __forceinline__ int lds_twiddles(const LinePlan& P) {
    return (!P.blue && P.n <= kMaxLdsTwiddles) ? P.n : 0;
}

/* Copies the direct plan's twiddles to LDS (when it fits) and returns the
 * table the transforms should read.  The caller must barrier before use. */
__device__ __forceinline__ const float2* stage_twiddles(const LinePlan& P, float2* tw_lds) {
    if (lds_twiddles(P) == 0) return P.blue ? P.mtw : P.tw;
    for (int i = threadIdx.x; i < P.n; i += blockDim.x) tw_lds[i] = P.tw[i];
    return tw_lds;
}

/* Host-side planner: radices for n (largest power-of-two radices first,
 * then the odd primes); false when n has a prime factor above 13. */
inline bool plan_radices(int n, std::vector<int>& rad) {
    rad.clear();
    if (n < 1) return false;
    int r = n;
    while (r % 8 == 0) { rad.push_back(8); r /= 8; }
    while (r % 4 == 0) { rad.push_back(4); r /= 4; }
    while (r % 2 == 0) { rad.push_back(2); r /= 2; }
    for (int p : {3, 5, 7, 11, 13})
        while (r % p == 0) { rad.push_back(p); r /= p; }
    return r == 1 && (int)rad.size() <= kMaxStages;
}

/* LDS row pitch (in float2) for a line occupying len elements: odd, so that
 * the transposed accesses of the column kernels spread over the LDS banks. */
__host__ __device__ __forceinline__ int lds_pitch(int len) { return (len % 2 == 0) ? len + 1 : len; }

}  // namespace iqm
