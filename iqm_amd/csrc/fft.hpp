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
// Radices 2, 3, 4, 5, 7, 8, 11 and 13 have compile-time butterflies, and the
// run-time two-level transform (fft_line.hpp) has register DFTs of every size
// up to 32, primes 17..31 included.  A length n = q * p with a larger prime p
// goes through Rader's identity (fft_line.hpp) when p - 1 has a two-level
// split and q is 1 or a register-DFT size; any other length through
// Bluestein's chirp-z identity X_k = c_k * sum_j (x_j c_j) conj(c_{k-j}),
// c_j = exp(S*pi*i*j^2/n), evaluated as a circular convolution of length
// m >= 2n-1 (the smallest such m the line transforms handle directly).
// Twiddles and chirps are computed in double on the host and stored as float.
#pragma once

#include "common.hpp"

#include <algorithm>
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
    int m;                 /* convolution length: Bluestein's m, Rader's p - 1  */
    int mst;               /* Bluestein: stages of the length-m plan            */
    int mrad[kMaxStages];
    const float2* mtw;     /* exp(-2 pi i j / m), j < m                         */
    const float2* chirp;   /* Bluestein: exp(-pi i (j^2 mod 2n) / n), j < n     */
    const float2* gf;      /* FFT_m of the forward convolution kernel           */
    const float2* gi;      /* FFT_m of the inverse convolution kernel           */
    int lds_len;           /* elements a line occupies in LDS (n or m, or the
                              two-level exchange area when that is longer, or
                              Rader's q sub-lines)                              */
    int tn1, tn2;          /* runtime two-level split n = tn1 * tn2 (0: none)   */
    int tm1, tm2;          /* the same for the convolution length m             */
    int rader;             /* 1: Rader, n = rq * rp with rp prime               */
    int rq, rp;
    int rpitch;            /* Rader: LDS pitch of a sub-line of rp points       */
    const int* rin;        /* Rader: slot of sub-line input j (rp entries)      */
    const int* rout;       /* Rader: g^b mod rp, b < rp - 1 (g a primitive root) */
};

/* Sizes of the register DFTs the runtime two-level transform dispatches to
 * (fft_line.hpp); a length n = n1 * n2 with both in this set is transformed
 * with one LDS exchange instead of one LDS pass per radix stage. */
#define IQM_TL2_SIZES(X)                                                                                    \
    X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20) \
    X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32)

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

/* ---- compile-time sine and cosine of 2*pi*e/M -------------------------- */

constexpr double kPiD = 3.14159265358979323846;
constexpr double kHalfSqrt2 = 0.70710678118654752440;

constexpr double cx_sin_poly(double x) { /* |x| <= pi/8 */
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i < 12; ++i) {
        term *= -x2 / (double)((2 * i) * (2 * i + 1));
        sum += term;
    }
    return sum;
}
constexpr double cx_cos_poly(double x) {
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i < 12; ++i) {
        term *= -x2 / (double)((2 * i - 1) * (2 * i));
        sum += term;
    }
    return sum;
}
/* cos and sin of 2*pi*e/M: nearest octant q*pi/4 (exact sine/cosine) plus a
 * remainder r = 2*pi*(8e - qM)/(8M), |r| <= pi/8, from an exact integer. */
constexpr double cx_cos2pi(long e, long M) {
    e %= M;
    if (e < 0) e += M;
    const long q = (8 * e + M / 2) / M;
    const double r = 2.0 * kPiD * (double)(8 * e - q * M) / (double)(8 * M);
    const double cq[8] = {1, kHalfSqrt2, 0, -kHalfSqrt2, -1, -kHalfSqrt2, 0, kHalfSqrt2};
    const double sq[8] = {0, kHalfSqrt2, 1, kHalfSqrt2, 0, -kHalfSqrt2, -1, -kHalfSqrt2};
    return cq[q % 8] * cx_cos_poly(r) - sq[q % 8] * cx_sin_poly(r);
}
constexpr double cx_sin2pi(long e, long M) {
    e %= M;
    if (e < 0) e += M;
    const long q = (8 * e + M / 2) / M;
    const double r = 2.0 * kPiD * (double)(8 * e - q * M) / (double)(8 * M);
    const double cq[8] = {1, kHalfSqrt2, 0, -kHalfSqrt2, -1, -kHalfSqrt2, 0, kHalfSqrt2};
    const double sq[8] = {0, kHalfSqrt2, 1, kHalfSqrt2, 0, -kHalfSqrt2, -1, -kHalfSqrt2};
    return sq[q % 8] * cx_cos_poly(r) + cq[q % 8] * cx_sin_poly(r);
}

template <int M>
struct RegTw {
    float c[M];
    float s[M];
    constexpr RegTw() : c(), s() {
        for (int j = 0; j < M; ++j) {
            c[j] = (float)cx_cos2pi(j, M);
            s[j] = (float)cx_sin2pi(j, M);
        }
    }
};
template <int M>
struct RegTwTab {
    static constexpr RegTw<M> t{};
};

/* ---- Rader plans (host) ---------------------------------------------------- */

/* Largest prime factor of n >= 2. */
inline int largest_prime_factor(int n) {
    int best = 1;
    for (int f = 2; (long)f * f <= n; ++f)
        while (n % f == 0) {
            best = f;
            n /= f;
        }
    return n > 1 ? std::max(best, n) : best;
}

/* Rader split of n: n = q * p with p its largest prime factor, p > 13, q = 1
 * or a register-DFT size, and p - 1 = m1 * m2 a two-level split; false when
 * n has no such split. */
inline bool rader_split(int n, int& q, int& p, int& m1, int& m2) {
    q = p = m1 = m2 = 0;
    if (n < 17) return false;
    const int pf = largest_prime_factor(n);
    if (pf <= 13) return false;
    const int qq = n / pf;
    if (qq != 1 && !tl2_size(qq)) return false;
    if (!tl2_split(pf - 1, m1, m2)) return false;
    q = qq;
    p = pf;
    return true;
}

/* LDS pitch of a Rader sub-line: the p - 1 convolution points or the
 * two-level exchange areas of p - 1 (m1 rows of m2 + 1 for the forward
 * transform, m2 rows of m1 + 1 for the inverse, which takes the split the
 * other way round), whichever is longest, then the slots of r_0 (pitch - 1)
 * and X_0 (pitch - 2); odd, so that sub-lines read in lockstep spread over
 * the banks. */
inline int rader_pitch(int p, int m1, int m2) {
    const int r = std::max({p - 1, m1 * (m2 + 1), m2 * (m1 + 1)}) + 2;
    return (r % 2 == 0) ? r + 1 : r;
}

/* Smallest primitive root of the prime p. */
inline int primitive_root(int p) {
    std::vector<int> fs;
    int r = p - 1;
    for (int f = 2; f * f <= r; ++f)
        if (r % f == 0) {
            fs.push_back(f);
            while (r % f == 0) r /= f;
        }
    if (r > 1) fs.push_back(r);
    auto powmod = [&](long b, int e) {
        long acc = 1;
        for (b %= p; e; e >>= 1, b = b * b % p)
            if (e & 1) acc = acc * b % p;
        return acc;
    };
    for (int g = 2; g < p; ++g) {
        bool ok = true;
        for (int f : fs) ok = ok && powmod(g, (p - 1) / f) != 1;
        if (ok) return g;
    }
    return 1; /* p = 2 */
}

/* Index tables of Rader's identity for the prime p with primitive root g:
 * rout[b] = g^b mod p (b < p - 1); rin[j] = (-log_g j) mod (p - 1), the
 * position of input j in the convolution sequence u_a = x[g^-a] (j >= 1),
 * and rin[0] = pitch - 1, the slot that keeps x_0. */
inline void rader_tables(int p, int pitch, std::vector<int>& rin, std::vector<int>& rout) {
    const int g = primitive_root(p);
    rout.assign(p - 1, 0);
    rin.assign(p, 0);
    long v = 1;
    for (int b = 0; b < p - 1; ++b) {
        rout[b] = (int)v;
        rin[v] = (p - 1 - b) % (p - 1);
        v = v * g % p;
    }
    rin[0] = pitch - 1;
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

/* cos and sin of 2*pi*e/R for the odd prime radices: the literal tables
 * above up to 13, the compile-time table for 17..31 (register DFTs of the
 * run-time two-level transform only) */
template <int R>
__host__ __device__ __forceinline__ constexpr float odd_cos(int e) {
    if constexpr (R <= 13) return OddTab<R>::C[e];
    else return RegTwTab<R>::t.c[e];
}
template <int R>
__host__ __device__ __forceinline__ constexpr float odd_sin(int e) {
    if constexpr (R <= 13) return OddTab<R>::S[e];
    else return RegTwTab<R>::t.s[e];
}

/* Odd prime radix: pairs q and R-q share the cosine terms.  Output t goes
 * to emit(t, y_t) as soon as it is formed (t = 0, then the pairs t, R - t),
 * so that a caller that stores the outputs keeps only the H sums and
 * differences live, not R outputs besides them. */
template <int R, int S, class Emit>
__host__ __device__ __forceinline__ void dft_odd_emit(const float2* x, const Emit& emit) {
    constexpr int H = (R - 1) / 2;
    float2 a[H + 1], b[H + 1];
    float2 y0 = x[0];
#pragma unroll
    for (int q = 1; q <= H; ++q) {
        a[q] = cadd(x[q], x[R - q]);
        b[q] = csub(x[q], x[R - q]);
        y0 = cadd(y0, a[q]);
    }
    const float2 x0 = x[0];
    emit(0, y0);
#pragma unroll
    for (int t = 1; t <= H; ++t) {
        float2 P = x0;
        float2 Q = make_float2(0.f, 0.f);
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            const float c = odd_cos<R>((q * t) % R);
            const float s = odd_sin<R>((q * t) % R);
            P.x += a[q].x * c;
            P.y += a[q].y * c;
            Q.x += b[q].x * s;
            Q.y += b[q].y * s;
        }
        const float2 iq = cmul_si<S>(Q);
        emit(t, cadd(P, iq));
        emit(R - t, csub(P, iq));
    }
}
template <int R, int S>
__host__ __device__ __forceinline__ void dft_odd(float2* x) {
    float2 out[R];
    dft_odd_emit<R, S>(x, [&](int t, float2 v) { out[t] = v; });
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
