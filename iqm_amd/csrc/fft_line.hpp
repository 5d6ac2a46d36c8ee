// Line transforms in LDS for any length: the runtime two-level transform
// (one LDS exchange, register DFTs picked by a uniform switch on the split),
// the Stockham schedule of fft.hpp for lengths without such a split, Rader's
// identity for lengths q * p with a prime p whose p - 1 has a two-level
// split, and Bluestein's chirp-z identity around either direct transform for
// the remaining lengths.
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
 * W_n^{S j2 k1} (tw holds exp(-2 pi i j / n), j < n), stored to z[k1 * zs]. */
template <int N1, int S, class Load>
__host__ __device__ __forceinline__ void tl2_step1_g(const Load& in, int j2, int n2, const float2* __restrict__ tw,
                                                     float2* __restrict__ z, int zs) {
    float2 x[N1];
#pragma unroll
    for (int j1 = 0; j1 < N1; ++j1) x[j1] = in(n2 * j1 + j2);
    RegDft<N1, S>::run_emit(x, [&](int k1, float2 v) {
        /* j2 = 0 multiplies by tw[0] = (1, -0), which leaves a finite v
         * unchanged; no branch on the run-time j2 */
        if (k1 != 0) {
            float2 w = tw[j2 * k1];
            if (S > 0) w.y = -w.y;
            v = cmul(v, w);
        }
        z[k1 * zs] = v;
    });
}
/* ... to the line's exchange area at z[k1 * (n2 + 1) + j2] */
template <int N1, int S, class Load>
__host__ __device__ __forceinline__ void tl2_step1_f(const Load& in, int j2, int n2, const float2* __restrict__ tw,
                                                     float2* __restrict__ z) {
    tl2_step1_g<N1, S>(in, j2, n2, tw, z + j2, n2 + 1);
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
    RegDft<N2, S>::run_emit(y, [&](int k2, float2 v) { out(k1 + n1 * k2, v); });
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

/* ---- Rader's identity for a prime factor ---------------------------------
 * A line of n = q * p points (p prime, q = 1 or a register-DFT size) is split
 * like the two-level transform: step 1 takes DFT_q over x[p*j1 + j2] and the
 * twiddles W_n^{S j2 k1}, giving q sub-lines r (k1 < q) of p points whose
 * DFTs are the outputs X[k1 + q*k2].  A sub-line's DFT follows from Rader's
 * identity: with g a primitive root of p and w = exp(S*2*pi*i/p),
 *   R[g^b] = r_0 + sum_{a < p-1} u_a v_{b-a},  u_a = r[g^-a],  v_c = w^(g^c),
 *   R[0]   = r_0 + sum_a u_a,
 * a cyclic convolution of length p - 1, evaluated as
 * IFFT(FFT(u) * V) with V = FFT(v) / (p - 1) (the plan's gf / gi) by the
 * run-time two-level transform of p - 1 = m1 * m2 (S = -1 forward, +1
 * inverse).  Step 1 stores r_j straight to its convolution position rin[j]
 * (rin[0] is the slot pitch - 1 that keeps r_0).  The inverse transform
 * takes the split the other way round (m2 * m1): its step-1 item j2' = k1
 * needs U[k1 + m1 * k2] * V[k1 + m1 * k2] for k2 < m2, exactly the outputs
 * of the forward transform's step-2 item k1, so one phase (conv_mid) runs
 * forward step 2, the product with V and inverse step 1 in registers, and
 * keeps R[0] = r_0 + U[0] in slot pitch - 2.  Sub-line (l, k1) of an
 * nlines-line set lives at (k1 * nlines + l) * pitch of the buffer
 * (pitch = P.rpitch >= the p - 1 points and both exchange areas, plus the
 * two slots). */
template <int S, class Load>
__host__ __device__ __forceinline__ void rader_item1_f(int q, const Load& in, int j2, int p, const float2* tw,
                                                       float2* z, int zs) {
    switch (q) {
        case 1: tl2_step1_g<1, S>(in, j2, p, tw, z, zs); break;
#define IQM_RADER_STEP1(k) \
    case k: tl2_step1_g<k, S>(in, j2, p, tw, z, zs); break;
        IQM_TL2_SIZES(IQM_RADER_STEP1)
#undef IQM_RADER_STEP1
        default: break;
    }
}

/* The middle phase of a convolution by two-level transforms of length
 * m = m1 * m2 (Rader's p - 1, Bluestein's m) for one item k1 < m1: forward
 * step 2 (DFT_M2 of exchange row k1 of z, the forward transform's exchange
 * area), giving U[k1 + m1 * k2]; with a slot (Rader: item 0 keeps
 * R[0] = r_0 + U[0] in slot[0], r_0 being slot[1]); the products with the
 * kernel spectrum V; then step 1 of the inverse transform with the split the
 * other way round (m2 * m1), whose item j2' = k1 needs exactly these values:
 * DFT_M2 with S = +1 and the twiddles W_m^{+k1 k1'} (tw = exp(-2 pi i j / m)),
 * stored to the inverse exchange area zo[k1' * (m1 + 1) + k1], k1' < m2. */
template <int M2>
__host__ __device__ __forceinline__ void conv_mid(const float2* __restrict__ z, int k1, int m1,
                                                  const float2* __restrict__ V, const float2* __restrict__ tw,
                                                  float2* __restrict__ zo, float2* slot) {
    float2 y[M2];
#pragma unroll
    for (int j2 = 0; j2 < M2; ++j2) y[j2] = z[k1 * (M2 + 1) + j2];
    RegDft<M2, -1>::run(y);
    if (slot != nullptr && k1 == 0) slot[0] = cadd(y[0], slot[1]);
#pragma unroll
    for (int k2 = 0; k2 < M2; ++k2) y[k2] = cmul(y[k2], V[k1 + m1 * k2]);
    RegDft<M2, 1>::run(y);
    const int p2 = m1 + 1;
#pragma unroll
    for (int kk = 0; kk < M2; ++kk) {
        float2 v = y[kk];
        /* k1 = 0 multiplies by tw[0] = (1, -0): a finite v is unchanged */
        if (kk != 0) {
            float2 w = tw[k1 * kk];
            w.y = -w.y;
            v = cmul(v, w);
        }
        zo[kk * p2 + k1] = v;
    }
}
__host__ __device__ __forceinline__ void conv_mid_item(int m2, const float2* z, int k1, int m1, const float2* V,
                                                       const float2* tw, float2* zo, float2* slot) {
    switch (m2) {
#define IQM_CONV_MID(k) \
    case k: conv_mid<k>(z, k1, m1, V, tw, zo, slot); break;
        IQM_TL2_SIZES(IQM_CONV_MID)
#undef IQM_CONV_MID
        default: break;
    }
}

/* Bluestein's first and last steps folded into the convolution's loads and
 * stores: input j < n times the chirp c_j (zero for n <= j < m), and output
 * k < n times c_k / m (outputs k >= n dropped); S > 0 takes the conjugate
 * chirp. */
template <int S, class Load>
struct ChirpLoad {
    Load in;
    const float2* chirp;
    int n;
    __host__ __device__ float2 operator()(int j) const {
        if (j >= n) return make_float2(0.f, 0.f);
        float2 c = chirp[j];
        if (S > 0) c.y = -c.y;
        return cmul(in(j), c);
    }
};
template <int S, class Store>
struct ChirpStore {
    Store out;
    const float2* chirp;
    int n;
    float inv_m;
    __host__ __device__ void operator()(int k, float2 v) const {
        if (k >= n) return;
        float2 c = chirp[k];
        if (S > 0) c.y = -c.y;
        out(k, cscale(cmul(v, c), inv_m));
    }
};

/* Output b of the inverse convolution transform of sub-line s = (line, k1):
 * R[g^b] = c_b + r_0 is output k1 + q * g^b of the line; b = 0 also gives
 * output k1 (R[0]). */
template <class Store>
struct RaderOutStore {
    Store out;
    const float2* s;
    const int* rout;
    int k1, q, rp;
    __host__ __device__ void operator()(int b, float2 v) const {
        out(k1 + q * rout[b], cadd(v, s[rp - 1]));
        if (b == 0) out(k1, s[rp - 2]);
    }
};

/* Rader transform of nlines lines held in b0 (line l at l * ld); the result
 * is returned in b0.  The sub-lines and exchange areas of the convolution
 * transforms alternate between b1 and b0.  Starts after the caller's
 * barrier, ends with one. */
template <int S>
__device__ __forceinline__ float2* line_rader(float2* b0, float2* b1, int ld, int nlines, const LinePlan& P,
                                              const float2* tw) {
    const int n = P.n, q = P.rq, p = P.rp, rp = P.rpitch, m1 = P.tm1, m2 = P.tm2;
    const int nl2 = nlines * q;
    const float ip = 1.0f / (float)p, i1 = 1.0f / (float)m1, i2 = 1.0f / (float)m2;
    for (int it = threadIdx.x; it < nlines * p; it += blockDim.x) {
        const int l = fast_div(it, p, ip), j2 = it - l * p;
        rader_item1_f<S>(q, PtrLoad{b0 + l * ld}, j2, p, tw, b1 + l * rp + P.rin[j2], nlines * rp);
    }
    lds_barrier();
    for (int it = threadIdx.x; it < nl2 * m2; it += blockDim.x) {
        const int l2 = fast_div(it, m2, i2), j2 = it - l2 * m2;
        tl2_item1<-1>(m1, b1 + l2 * rp, j2, m2, P.mtw, b0 + l2 * rp);
    }
    lds_barrier();
    const float2* V = (S < 0) ? P.gf : P.gi;
    for (int it = threadIdx.x; it < nl2 * m1; it += blockDim.x) {
        const int l2 = fast_div(it, m1, i1), k1 = it - l2 * m1;
        conv_mid_item(m2, b0 + l2 * rp, k1, m1, V, P.mtw, b1 + l2 * rp, b1 + l2 * rp + rp - 2);
    }
    lds_barrier();
    /* inverse step 2 (split m2 * m1), each output stored straight to its
     * place in its line in b0 (line l at l * ld), free since the middle
     * phase; the sub-line's slots stay in b1 */
    const float inl = 1.0f / (float)nlines;
    for (int it = threadIdx.x; it < nl2 * m2; it += blockDim.x) {
        const int l2 = fast_div(it, m2, i2), kk = it - l2 * m2;
        const int r1 = fast_div(l2, nlines, inl), l = l2 - r1 * nlines;
        tl2_item2_f<1>(m1, b1 + l2 * rp, kk, m2,
                       RaderOutStore<PtrStore>{PtrStore{b0 + l * ld}, b1 + l2 * rp, P.rout, r1, q, rp});
    }
    lds_barrier();
    return b0;
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
    if (P.rader) return line_rader<S>(b0, b1, ld, nlines, P, tw);
    if (!P.blue) return line_direct<S>(b0, b1, ld, nlines, P.n, P.tn1, P.tn2, P.nst, P.rad, tw);
    const int n = P.n, m = P.m;
    if (P.tm1 > 0) {
        /* three phases: forward step 1 with the chirp on its loads (b0 ->
         * b1), forward step 2 + kernel product + inverse step 1 (b1 -> b0),
         * inverse step 2 with the chirp and 1/m on its stores (b0 -> b1) */
        const int m1 = P.tm1, m2 = P.tm2;
        const float i1 = 1.0f / (float)m1, i2 = 1.0f / (float)m2;
        for (int it = threadIdx.x; it < nlines * m2; it += blockDim.x) {
            const int l = fast_div(it, m2, i2), j2 = it - l * m2;
            tl2_item1_f<-1>(m1, ChirpLoad<S, PtrLoad>{PtrLoad{b0 + l * ld}, P.chirp, n}, j2, m2, P.mtw, b1 + l * ld);
        }
        lds_barrier();
        const float2* g = (S < 0) ? P.gf : P.gi;
        for (int it = threadIdx.x; it < nlines * m1; it += blockDim.x) {
            const int l = fast_div(it, m1, i1), k1 = it - l * m1;
            conv_mid_item(m2, b1 + l * ld, k1, m1, g, P.mtw, b0 + l * ld, nullptr);
        }
        lds_barrier();
        const float inv_m = 1.0f / (float)m;
        for (int it = threadIdx.x; it < nlines * m2; it += blockDim.x) {
            const int l = fast_div(it, m2, i2), kk = it - l * m2;
            tl2_item2_f<1>(m1, b0 + l * ld, kk, m2, ChirpStore<S, PtrStore>{PtrStore{b1 + l * ld}, P.chirp, n, inv_m});
        }
        lds_barrier();
        return b1;
    }
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
    if (P.rader) {
        /* sub-line l2 = k1 * nlines + l; with LineFast the sub-line index is
         * fastest, so the final stores of neighbouring lines stay together */
        const int q = P.rq, p = P.rp, rp = P.rpitch, m1 = P.tm1, m2 = P.tm2;
        const int nl2 = nlines * q;
        const float ip = 1.0f / (float)p, i1 = 1.0f / (float)m1, i2 = 1.0f / (float)m2, inl2 = 1.0f / (float)nl2;
        auto split2 = [&](int it, int len, float ilen, int& l2, int& e) {
            if (LineFast) {
                e = fast_div(it, nl2, inl2);
                l2 = it - e * nl2;
            } else {
                l2 = fast_div(it, len, ilen);
                e = it - l2 * len;
            }
        };
        for (int it = threadIdx.x; it < nlines * p; it += blockDim.x) {
            int l, j2;
            split(it, p, ip, l, j2);
            rader_item1_f<S>(q, load(l), j2, p, tw, b1 + l * rp + P.rin[j2], nlines * rp);
        }
        lds_barrier();
        for (int it = threadIdx.x; it < nl2 * m2; it += blockDim.x) {
            int l2, j2;
            split2(it, m2, i2, l2, j2);
            tl2_item1<-1>(m1, b1 + l2 * rp, j2, m2, P.mtw, b0 + l2 * rp);
        }
        lds_barrier();
        const float2* V = (S < 0) ? P.gf : P.gi;
        for (int it = threadIdx.x; it < nl2 * m1; it += blockDim.x) {
            int l2, k1;
            split2(it, m1, i1, l2, k1);
            conv_mid_item(m2, b0 + l2 * rp, k1, m1, V, P.mtw, b1 + l2 * rp, b1 + l2 * rp + rp - 2);
        }
        lds_barrier();
        /* inverse step 2 (split m2 * m1) */
        using St = decltype(store(0));
        for (int it = threadIdx.x; it < nl2 * m2; it += blockDim.x) {
            int l2, kk;
            split2(it, m2, i2, l2, kk);
            const int r1 = fast_div(l2, nlines, inl), l = l2 - r1 * nlines;
            tl2_item2_f<1>(m1, b1 + l2 * rp, kk, m2, RaderOutStore<St>{store(l), b1 + l2 * rp, P.rout, r1, q, rp});
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
    if (P.tm1 > 0) {
        /* Bluestein in three phases, as in line_fft */
        const int m1 = P.tm1, m2 = P.tm2;
        const float i1 = 1.0f / (float)m1, i2 = 1.0f / (float)m2;
        using Ld = decltype(load(0));
        using St = decltype(store(0));
        for (int it = threadIdx.x; it < nlines * m2; it += blockDim.x) {
            int l, j2;
            split(it, m2, i2, l, j2);
            tl2_item1_f<-1>(m1, ChirpLoad<S, Ld>{load(l), P.chirp, n}, j2, m2, P.mtw, b1 + l * ld);
        }
        lds_barrier();
        const float2* g = (S < 0) ? P.gf : P.gi;
        for (int it = threadIdx.x; it < nlines * m1; it += blockDim.x) {
            int l, k1;
            split(it, m1, i1, l, k1);
            conv_mid_item(m2, b1 + l * ld, k1, m1, g, P.mtw, b0 + l * ld, nullptr);
        }
        lds_barrier();
        const float inv_m = 1.0f / (float)m;
        for (int it = threadIdx.x; it < nlines * m2; it += blockDim.x) {
            int l, kk;
            split(it, m2, i2, l, kk);
            tl2_item2_f<1>(m1, b0 + l * ld, kk, m2, ChirpStore<S, St>{store(l), P.chirp, n, inv_m});
        }
        lds_barrier();
        return;
    }
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
