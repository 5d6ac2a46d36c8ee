// Line transforms in LDS for any length: the runtime two-level transform
// (one LDS exchange, register DFTs picked by a uniform switch on the split),
// the Stockham schedule of fft.hpp for lengths without such a split, and
// Bluestein's chirp-z identity around either for lengths with a prime factor
// above 13.
#pragma once

#include "fft_reg.hpp"

namespace iqm {

/* Element accessors of a line held in memory: value i, and store to i. */
struct PtrLoad {
    const float2* p;
    __host__ __device__ float2 operator()(int i) const { return p[i]; }
};
struct PtrStore {
    float2* p;
    __host__ __device__ void operator()(int i, float2 v) const { p[i] = v; }
};

/* Step 1 of the runtime two-level transform for one item (line, j2) of a
 * line of n = N1 * n2 points: x[j1] = in(n2*j1 + j2), DFT_N1, twiddles
 * W_n^{S j2 k1} (tw holds exp(-2 pi i j / n), j < n), stored to the line's
 * exchange area at z[k1 * (n2 + 1) + j2]. */
template <int N1, int S, class Load>
__host__ __device__ __forceinline__ void tl2_step1_f(const Load& in, int j2, int n2, const float2* __restrict__ tw,
                                                     float2* __restrict__ z) {
    float2 x[N1];
#pragma unroll
    for (int j1 = 0; j1 < N1; ++j1) x[j1] = in(n2 * j1 + j2);
    RegDft<N1, S>::run(x);
    const int p2 = n2 + 1;
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
        float2 v = x[k1];
        /* j2 = 0 multiplies by tw[0] = (1, -0), which leaves a finite v
         * unchanged; no branch on the run-time j2 */
        if (k1 != 0) {
            float2 w = tw[j2 * k1];
            if (S > 0) w.y = -w.y;
            v = cmul(v, w);
        }
        z[k1 * p2 + j2] = v;
    }
}
template <int N1, int S>
__host__ __device__ __forceinline__ void tl2_step1(const float2* __restrict__ line, int j2, int n2,
                                                   const float2* __restrict__ tw, float2* __restrict__ z) {
    tl2_step1_f<N1, S>(PtrLoad{line}, j2, n2, tw, z);
}

/* Step 2 for one item (line, k1): DFT_N2 of z[k1 * (N2 + 1) + j2], j2 < N2,
 * given to out(k1 + n1 * k2, value). */
template <int N2, int S, class Store>
__host__ __device__ __forceinline__ void tl2_step2_f(const float2* __restrict__ z, int k1, int n1, const Store& out) {
    float2 y[N2];
    const int p2 = N2 + 1;
#pragma unroll
    for (int j2 = 0; j2 < N2; ++j2) y[j2] = z[k1 * p2 + j2];
    RegDft<N2, S>::run(y);
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) out(k1 + n1 * k2, y[k2]);
}
template <int N2, int S>
__host__ __device__ __forceinline__ void tl2_step2(const float2* __restrict__ z, int k1, int n1, float2* __restrict__ out) {
    tl2_step2_f<N2, S>(z, k1, n1, PtrStore{out});
}

/* One item of step 1 / step 2 with the DFT size picked at run time (the
 * switch is uniform across a workgroup). */
template <int S, class Load>
__host__ __device__ __forceinline__ void tl2_item1_f(int n1, const Load& in, int j2, int n2, const float2* tw,
                                                     float2* z) {
    switch (n1) {
#define IQM_TL2_STEP1(k) \
    case k: tl2_step1_f<k, S>(in, j2, n2, tw, z); break;
        IQM_TL2_SIZES(IQM_TL2_STEP1)
#undef IQM_TL2_STEP1
        default: break;
    }
}
template <int S>
__host__ __device__ __forceinline__ void tl2_item1(int n1, const float2* line, int j2, int n2, const float2* tw,
                                                   float2* z) {
    tl2_item1_f<S>(n1, PtrLoad{line}, j2, n2, tw, z);
}
template <int S, class Store>
__host__ __device__ __forceinline__ void tl2_item2_f(int n2, const float2* z, int k1, int n1, const Store& out) {
    switch (n2) {
#define IQM_TL2_STEP2(k) \
    case k: tl2_step2_f<k, S>(z, k1, n1, out); break;
        IQM_TL2_SIZES(IQM_TL2_STEP2)
#undef IQM_TL2_STEP2
        default: break;
    }
}
template <int S>
__host__ __device__ __forceinline__ void tl2_item2(int n2, const float2* z, int k1, int n1, float2* out) {
    tl2_item2_f<S>(n2, z, k1, n1, PtrStore{out});
}

/* Two-level transform of nlines lines of n = n1 * n2 points (both factors in
 * IQM_TL2_SIZES) held in `in` (line l at l * ld), through the exchange area
 * `z` (same line pitch, >= n1 * (n2 + 1) elements per line); the result
 * replaces the input in `in`, which is returned.  The caller's barrier makes
 * `in` visible; ends with a barrier. */
template <int S>
__device__ __forceinline__ float2* tl2_lines(float2* in, float2* z, int ld, int nlines, int n1, int n2, const float2* tw) {
    const float inv2 = 1.0f / (float)n2, inv1 = 1.0f / (float)n1;
    for (int it = threadIdx.x; it < nlines * n2; it += blockDim.x) {
        const int l = fast_div(it, n2, inv2), j2 = it - l * n2;
        tl2_item1<S>(n1, in + l * ld, j2, n2, tw, z + l * ld);
    }
    lds_barrier();
    for (int it = threadIdx.x; it < nlines * n1; it += blockDim.x) {
        const int l = fast_div(it, n1, inv1), k1 = it - l * n1;
        tl2_item2<S>(n2, z + l * ld, k1, n1, in + l * ld);
    }
    lds_barrier();
    return in;
}

/* Direct transform of a length-n line set with plan radices or split. */
template <int S>
__device__ __forceinline__ float2* line_direct(float2* b0, float2* b1, int ld, int nlines, int n, int s1, int s2,
                                               int nst, const int* rad, const float2* tw) {
    if (s1 > 0) return tl2_lines<S>(b0, b1, ld, nlines, s1, s2, tw);
    return lds_fft_direct<S>(b0, b1, ld, nlines, n, nst, rad, tw);
}

/* Full line transform of nlines lines whose first n elements hold the input;
 * lines are ld >= P.lds_len apart.  Returns the buffer holding the result
 * (first n elements of each line).  Starts and ends at a barrier. */
template <int S>
__device__ __forceinline__ float2* line_fft(float2* b0, float2* b1, int ld, int nlines, const LinePlan& P,
                            const float2* tw) {
    if (!P.blue) return line_direct<S>(b0, b1, ld, nlines, P.n, P.tn1, P.tn2, P.nst, P.rad, tw);
    const int n = P.n, m = P.m;
    const float im_ = 1.0f / (float)m, in_ = 1.0f / (float)n;
    for (int idx = threadIdx.x; idx < nlines * m; idx += blockDim.x) {
        const int l = fast_div(idx, m, im_), j = idx - l * m;
        float2 v = make_float2(0.f, 0.f);
        if (j < n) {
            float2 c = P.chirp[j];
            if (S > 0) c.y = -c.y;
            v = cmul(b0[l * ld + j], c);
        }
        b0[l * ld + j] = v;
    }
    lds_barrier();
    float2* r = line_direct<-1>(b0, b1, ld, nlines, m, P.tm1, P.tm2, P.mst, P.mrad, P.mtw);
    float2* other = (r == b0) ? b1 : b0;
    const float2* g = (S < 0) ? P.gf : P.gi;
    for (int idx = threadIdx.x; idx < nlines * m; idx += blockDim.x) {
        const int l = fast_div(idx, m, im_), j = idx - l * m;
        r[l * ld + j] = cmul(r[l * ld + j], g[j]);
    }
    lds_barrier();
    float2* q = line_direct<1>(r, other, ld, nlines, m, P.tm1, P.tm2, P.mst, P.mrad, P.mtw);
    const float inv_m = 1.0f / (float)m;
    for (int idx = threadIdx.x; idx < nlines * n; idx += blockDim.x) {
        const int l = fast_div(idx, n, in_), k = idx - l * n;
        float2 c = P.chirp[k];
        if (S > 0) c.y = -c.y;
        q[l * ld + k] = cscale(cmul(q[l * ld + k], c), inv_m);
    }
    lds_barrier();
    return q;
}

/* Line transform of nlines lines whose input element j of line l is
 * load(l)(j) and whose output element k is given to store(l)(k, value):
 * the runtime two-level transform reads its step-1 inputs and writes its
 * step-2 outputs through them (no input or output pass over LDS), Stockham
 * and Bluestein lines materialise the input in b0 first (Bluestein folds
 * the load into its first chirp product and the store into its last).
 * LineFast: work items take the line index fastest, so that lines which
 * are columns of a row-major plane are read and written coalesced; else the
 * element index is fastest (lines that are rows).  b0 and b1 are line
 * buffers of pitch ld (b0 unused by the two-level transform, whose exchange
 * area is b1).  Starts after the caller's barrier and ends with a barrier. */
template <int S, bool LineFast, class MakeLoad, class MakeStore>
__device__ __forceinline__ void line_fft_io(const MakeLoad& load, const MakeStore& store, float2* b0, float2* b1,
                                            int ld, int nlines, const LinePlan& P, const float2* tw) {
    const int n = P.n;
    const float inl = 1.0f / (float)nlines;
    /* work item it of a pass over nlines lines of len elements -> (line, element) */
    auto split = [&](int it, int len, float ilen, int& l, int& e) {
        if (LineFast) {
            e = fast_div(it, nlines, inl);
            l = it - e * nlines;
        } else {
            l = fast_div(it, len, ilen);
            e = it - l * len;
        }
    };
    if (!P.blue && P.tn1 > 0) {
        const int n1 = P.tn1, n2 = P.tn2;
        const float i1 = 1.0f / (float)n1, i2 = 1.0f / (float)n2;
        for (int it = threadIdx.x; it < nlines * n2; it += blockDim.x) {
            int l, j2;
            split(it, n2, i2, l, j2);
            tl2_item1_f<S>(n1, load(l), j2, n2, tw, b1 + l * ld);
        }
        lds_barrier();
        for (int it = threadIdx.x; it < nlines * n1; it += blockDim.x) {
            int l, k1;
            split(it, n1, i1, l, k1);
            tl2_item2_f<S>(n2, b1 + l * ld, k1, n1, store(l));
        }
        lds_barrier();
        return;
    }
    const float in_ = 1.0f / (float)n;
    if (!P.blue) {
        for (int idx = threadIdx.x; idx < nlines * n; idx += blockDim.x) {
            int l, j;
            split(idx, n, in_, l, j);
            b0[l * ld + j] = load(l)(j);
        }
        lds_barrier();
        const float2* r = lds_fft_direct<S>(b0, b1, ld, nlines, n, P.nst, P.rad, tw);
        for (int idx = threadIdx.x; idx < nlines * n; idx += blockDim.x) {
            int l, k;
            split(idx, n, in_, l, k);
            store(l)(k, r[l * ld + k]);
        }
        lds_barrier();
        return;
    }
    const int m = P.m;
    const float im_ = 1.0f / (float)m;
    for (int idx = threadIdx.x; idx < nlines * m; idx += blockDim.x) {
        int l, j;
        split(idx, m, im_, l, j);
        float2 v = make_float2(0.f, 0.f);
        if (j < n) {
            float2 c = P.chirp[j];
            if (S > 0) c.y = -c.y;
            v = cmul(load(l)(j), c);
        }
        b0[l * ld + j] = v;
    }
    lds_barrier();
    float2* r = line_direct<-1>(b0, b1, ld, nlines, m, P.tm1, P.tm2, P.mst, P.mrad, P.mtw);
    float2* other = (r == b0) ? b1 : b0;
    const float2* g = (S < 0) ? P.gf : P.gi;
    for (int idx = threadIdx.x; idx < nlines * m; idx += blockDim.x) {
        const int l = fast_div(idx, m, im_), j = idx - l * m;
        r[l * ld + j] = cmul(r[l * ld + j], g[j]);
    }
    lds_barrier();
    const float2* q = line_direct<1>(r, other, ld, nlines, m, P.tm1, P.tm2, P.mst, P.mrad, P.mtw);
    const float inv_m = 1.0f / (float)m;
    for (int idx = threadIdx.x; idx < nlines * n; idx += blockDim.x) {
        int l, k;
        split(idx, n, in_, l, k);
        float2 c = P.chirp[k];
        if (S > 0) c.y = -c.y;
        store(l)(k, cscale(cmul(q[l * ld + k], c), inv_m));
    }
    lds_barrier();
}

}  // namespace iqm
