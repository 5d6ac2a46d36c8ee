// Register-resident DFTs and the two-level (four-step) line transform.
//
// RegDft<M, S> computes an M-point DFT (y_k = sum_j x_j exp(S*2*pi*i*j*k/M))
// entirely in registers, fully unrolled, by splitting M = R * Q:
//   y[k1 + R*k2] = sum_{j2} W_Q^{j2 k2} W_M^{j2 k1} sum_{j1} x[Q*j1 + j2] W_R^{j1 k1},
// with the radix-R butterflies of fft.hpp and compile-time twiddles; the
// primes 17..31 are single butterflies (sizes up to 32 need no split by them).
//
// A line of n = N1 * N2 points is then transformed with one LDS exchange:
//   step 1, item j2 < N2:  z[k1][j2] = W_n^{j2 k1} * DFT_N1(x[N2*j1 + j2])[k1]
//   step 2, item k1 < N1:  X[k1 + N1*k2] = DFT_N2(z[k1][.])[k2]
// Both steps read and write with unit stride across neighbouring items, so
// the global loads of step 1 and the stores of step 2 are coalesced.
#pragma once

#include "fft.hpp"

namespace iqm {

/* radix used to split M: a native butterfly size, largest power of two first */
constexpr bool is_native_radix(int m) {
    return m == 2 || m == 3 || m == 4 || m == 5 || m == 7 || m == 8 || m == 11 || m == 13 || m == 17 || m == 19 ||
           m == 23 || m == 29 || m == 31;
}
constexpr int split_radix(int M) {
    if (M % 8 == 0 && M != 8) return 8;
    if (M % 4 == 0 && M != 4) return 4;
    if (M % 2 == 0 && M != 2) return 2;
    for (int p : {3, 5, 7, 11, 13})
        if (M % p == 0 && M != p) return p;
    return M;
}
constexpr bool reg_dft_ok(int M) {
    if (M == 1 || is_native_radix(M)) return true;
    const int R = split_radix(M);
    return R != M && is_native_radix(R) && reg_dft_ok(M / R);
}

template <int M, int S>
struct RegDft {
    static_assert(reg_dft_ok(M), "register DFT size must factor into the native radices");
    /* y_k to emit(k, y_k), k < M; x is clobbered.  The odd primes above 8
     * give each output away as it is formed (dft_odd_emit). */
    template <class Emit>
    __host__ __device__ __forceinline__ static void run_emit(float2* x, const Emit& emit) {
        if constexpr (M > 8 && is_native_radix(M)) {
            dft_odd_emit<M, S>(x, emit);
        } else {
            run(x);
#pragma unroll
            for (int k = 0; k < M; ++k) emit(k, x[k]);
        }
    }
    __host__ __device__ __forceinline__ static void run(float2* x) {
        if constexpr (M == 1) {
            return;
        } else if constexpr (is_native_radix(M)) {
            dft_r<M, S>(x);
        } else {
            constexpr int R = split_radix(M);
            constexpr int Q = M / R;
            float2 y[M]; /* y[k1*Q + j2] */
#pragma unroll
            for (int j2 = 0; j2 < Q; ++j2) {
                float2 t[R];
#pragma unroll
                for (int j1 = 0; j1 < R; ++j1) t[j1] = x[Q * j1 + j2];
                dft_r<R, S>(t);
#pragma unroll
                for (int k1 = 0; k1 < R; ++k1) {
                    const int e = (j2 * k1) % M;
                    if (e == 0) {
                        y[k1 * Q + j2] = t[k1];
                    } else {
                        const float c = RegTwTab<M>::t.c[e];
                        const float s = (float)S * RegTwTab<M>::t.s[e];
                        y[k1 * Q + j2] = cmul(t[k1], make_float2(c, s));
                    }
                }
            }
#pragma unroll
            for (int k1 = 0; k1 < R; ++k1) {
                RegDft<Q, S>::run(&y[k1 * Q]);
#pragma unroll
                for (int k2 = 0; k2 < Q; ++k2) x[k1 + R * k2] = y[k1 * Q + k2];
            }
        }
    }
};

/* Two-level split of a line length that the specialised kernels handle. */
struct TwoLevel {
    int n1, n2;
};

/* LDS layout of the exchange buffer of one line: z[k1 * (N2 + 1) + j2]. */
template <int N1, int N2>
struct Exch {
    static constexpr int pitch = N2 + 1;
    static constexpr int line = N1 * (N2 + 1) + ((N1 * (N2 + 1)) % 2 == 0 ? 1 : 0);
};

/* Step 1 of the two-level transform for one item (line, j2): x holds the N1
 * inputs x[N2*j1 + j2]; applies DFT_N1 and the twiddles W_n^{S j2 k1} (tw is
 * exp(-2 pi i j / n), j < n) and stores to z (the line's exchange area). */
template <int N1, int N2, int S>
__host__ __device__ __forceinline__ void two_level_step1(float2* x, int j2, const float2* tw, float2* z) {
    RegDft<N1, S>::run(x);
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
        z[k1 * Exch<N1, N2>::pitch + j2] = v;
    }
}

/* Step 2 for one item (line, k1): loads z[k1][.] and leaves X[k1 + N1*k2]
 * in y[k2]. */
template <int N1, int N2, int S>
__host__ __device__ __forceinline__ void two_level_step2(const float2* z, int k1, float2* y) {
#pragma unroll
    for (int j2 = 0; j2 < N2; ++j2) y[j2] = z[k1 * Exch<N1, N2>::pitch + j2];
    RegDft<N2, S>::run(y);
}

}  // namespace iqm
