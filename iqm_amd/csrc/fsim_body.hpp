// Device bodies of the FSIM stages and the host helpers that size their
// launches.  Shared by fsim.hip (same-size kernels and the launch sequence),
// fsim_gen.hip (the generic line-transform passes) and fsim_mix.hip (mixed
// batches), so that each stage's arithmetic is written once.
#pragma once

#include "ctx.hpp"
#include "fft_line.hpp"
#include "fsim_tl.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace iqm {

/* ---- hot path kernels -------------------------------------------------- */

constexpr int kHistBins = 1024; /* level-1 radix-select bins: key bits 30..21 */

/* F2: fsim_downsample.glsl:38-54.  A workgroup produces `cols` downscaled
 * pixels of `prb` consecutive rows for both images: the bytes of the
 * full-resolution rows they cover are staged in LDS for both images at once,
 * as 16-byte aligned chunks (an aligned 16-byte block that holds a byte of the
 * image lies in that byte's page, so these loads never leave mapped memory),
 * with each thread's loads issued before its LDS stores.  Each thread then
 * sums the R, G and B bytes of its F x F box as integers (exact; zero outside
 * the image) and applies the YIQ matrix once, divided by F^2 -- the same
 * value as the shader's per-pixel sum in exact arithmetic.  yiq planes per
 * pair: Y_in, Y_ref, I_in, Q_in, I_ref, Q_ref.  Workgroup (0, 0) of each pair
 * also clears the pair's median histograms. */
constexpr int kPrepMaxCols = 256;
constexpr int kPrepMaxRows = 64;

/* 16-byte chunks staged per full-resolution row for a segment of seg bytes
 * (one spare chunk keeps the box sums' whole-dword reads inside the row) */
__host__ __device__ __forceinline__ int prep_row_chunks(int seg) { return seg / 16 + 3; }

inline size_t prep_lds_bytes(int F, int ch, int cols, int prb) {
    return (size_t)2 * prb * F * prep_row_chunks(cols * F * ch) * 16;
}
/* output columns per prep workgroup: the widest power of two whose two
 * staged images fit kMaxLdsBytes */
inline int prep_cols(int F, int ch, int prb) {
    int c = kPrepMaxCols;
    while (c > 1 && prep_lds_bytes(F, ch, c, prb) > (size_t)kMaxLdsBytes) c >>= 1;
    return c;
}

/* Box sums of the F x F box (F = FT known at compile time, window not
 * clipped horizontally) over the nk staged rows starting at row index k0r of
 * image block rbase, from byte xoff of each row.  Each row's F*CH bytes are
 * read as whole dwords and realigned with v_alignbyte; bytes are added in
 * 16-bit lanes (even bytes, odd bytes), which hold up to 257 rows of 255.
 * CH = 3: byte p of the window is channel p % 3; CH = 4: p % 4 (alpha
 * skipped); CH = 1: all bytes are gray. */
template <int FT, int CH>
__device__ __forceinline__ void prep_box_swar(const uint8_t* __restrict__ stage, const int* rowoff, int rbase,
                                              int k0r, int nk, int xoff, int& sr, int& sg, int& sb) {
    constexpr int L = FT * CH;
    constexpr int DW = (L + 3) / 4;
    unsigned e[DW], o[DW];
#pragma unroll
    for (int j = 0; j < DW; ++j) e[j] = o[j] = 0u;
#pragma unroll
    for (int kk = 0; kk < FT; ++kk) {
        if (kk < nk) {
            const int sb0 = rowoff[rbase + k0r + kk] + xoff;
            const unsigned* b = (const unsigned*)(stage + (sb0 & ~3));
            const unsigned sh = (unsigned)(sb0 & 3);
            unsigned d[DW + 1];
#pragma unroll
            for (int j = 0; j < DW; ++j) d[j] = b[j];
            d[DW] = ((int)sh + L > 4 * DW) ? b[DW] : 0u;
#pragma unroll
            for (int j = 0; j < DW; ++j) {
                unsigned w = __builtin_amdgcn_alignbyte(d[j + 1], d[j], sh);
                if (j == DW - 1 && (L % 4) != 0) w &= (1u << (8 * (L % 4))) - 1u;
                e[j] += w & 0x00FF00FFu;
                o[j] += (w >> 8) & 0x00FF00FFu;
            }
        }
    }
    int acc[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < DW; ++j) {
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const int p = 4 * j + bb;
            if (p < L) {
                const unsigned v = (bb == 0) ? (e[j] & 0xFFFFu) : (bb == 1) ? (o[j] & 0xFFFFu)
                                 : (bb == 2) ? (e[j] >> 16) : (o[j] >> 16);
                acc[CH == 1 ? 0 : p % CH] += (int)v;
            }
        }
    }
    sr = acc[0];
    sg = CH == 1 ? acc[0] : acc[1];
    sb = CH == 1 ? acc[0] : acc[2];
}

template <int FT>
__device__ __forceinline__ void fsim_prep_body(const uint8_t* __restrict__ img, const uint8_t* __restrict__ ref,
                                               long pair_stride, int w, int h, int ch, int F, int wd, int hd,
                                               int cols, int prb, float* __restrict__ yiq,
                                               unsigned* __restrict__ hist, int bx, int by, int pair) {
    extern __shared__ uint4 stage4[];
    __shared__ int rowoff[2 * kPrepMaxRows];
    const uint8_t* stage = (const uint8_t*)stage4;
    const int tid = threadIdx.x;
    if (bx == 0 && by == 0) {
        unsigned* hp = hist + (long)pair * 8 * kHistBins;
        for (int i = tid; i < 8 * kHistBins; i += blockDim.x) hp[i] = 0u;
    }
    const int y0 = by * prb;
    const int nout = min(prb, hd - y0);
    const int xd0 = bx * cols;
    const int ncols = min(cols, wd - xd0);
    /* full-resolution window of this block, clipped to the image */
    const int c_lo = xd0 * F - F / 2, c_hi = (xd0 + ncols) * F - F / 2;
    const int r_lo = y0 * F - F / 2, r_hi = (y0 + nout) * F - F / 2;
    const int cc0 = max(c_lo, 0), cc1 = min(c_hi, w);
    const int rr0 = max(r_lo, 0), rr1 = min(r_hi, h);
    const int seg = max(cc1 - cc0, 0) * ch; /* bytes of each staged row */
    const int nrows = max(rr1 - rr0, 0);
    const int rc = prep_row_chunks(seg);     /* chunks per staged row */
    const uint8_t* src0 = img + (long)pair * pair_stride;
    const uint8_t* src1 = ref + (long)pair * pair_stride;
    const int per_im = nrows * rc;
    /* staged rows: image 0 rows, then image 1 rows; (row, chunk) flattened
     * over the workgroup, 8 loads per thread in flight */
    const int total = 2 * per_im;
    for (int base = 0; base < total; base += 8 * (int)blockDim.x) {
        uint4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = base + u * blockDim.x + tid;
            v[u] = make_uint4(0u, 0u, 0u, 0u);
            if (i < total) {
                const int rr = i / rc, k = i - rr * rc;
                const int im = rr >= nrows, r = rr - im * nrows;
                const uintptr_t a0 = (uintptr_t)((im ? src1 : src0) + ((long)(rr0 + r) * w + cc0) * ch);
                const uintptr_t a = (a0 & ~(uintptr_t)15) + (uintptr_t)k * 16;
                if (a <= ((a0 + seg - 1) & ~(uintptr_t)15)) v[u] = *(const uint4*)a;
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = base + u * blockDim.x + tid;
            if (i < total) stage4[i] = v[u];
        }
    }
    if (tid < 2 * nrows) {
        const int im = tid >= nrows, r = tid - im * nrows;
        const uintptr_t a0 = (uintptr_t)((im ? src1 : src0) + ((long)(rr0 + r) * w + cc0) * ch);
        rowoff[tid] = (im * per_im + r * rc) * 16 + (int)(a0 & 15);
    }
    __syncthreads();
    const long N = (long)wd * hd;
    float* base = yiq + (long)pair * 6 * N;
    const float inv = 1.0f / ((float)F * (float)F);
    for (int t = tid; t < ncols * nout; t += blockDim.x) {
        const int oy = t / ncols, xc = t - oy * ncols;
        const int x = xd0 + xc, y = y0 + oy;
        const int xs = x * F - F / 2, ys = y * F - F / 2;
        const int j0 = max(xs, cc0), j1 = min(xs + F, cc1);
        const int k0 = max(ys, rr0), k1 = min(ys + F, rr1);
        const int nb = (j1 - j0) * ch;
        const long idx = (long)y * wd + x;
        const bool full = FT > 0 && j0 == xs && j1 == xs + F;
        for (int im = 0; im < 2; ++im) {
            int sr = 0, sg = 0, sb = 0;
            if constexpr (FT > 0) {
                if (full) {
                    const int xoff = (j0 - cc0) * ch;
                    if (ch == 3) prep_box_swar<FT, 3>(stage, rowoff, im * nrows, k0 - rr0, k1 - k0, xoff, sr, sg, sb);
                    else if (ch == 4) prep_box_swar<FT, 4>(stage, rowoff, im * nrows, k0 - rr0, k1 - k0, xoff, sr, sg, sb);
                    else prep_box_swar<FT, 1>(stage, rowoff, im * nrows, k0 - rr0, k1 - k0, xoff, sr, sg, sb);
                }
            }
            if (!full) {
                for (int k = k0; k < k1; ++k) {
                    const uint8_t* p = stage + rowoff[im * nrows + (k - rr0)] + (j0 - cc0) * ch;
                    if (ch == 1) {
                        for (int b = 0; b < nb; ++b) sr += p[b];
                    } else {
                        for (int b = 0; b < nb; b += ch) {
                            sr += p[b];
                            sg += p[b + 1];
                            sb += p[b + 2];
                        }
                    }
                }
                if (ch == 1) sg = sb = sr;
            }
            const float R = (float)sr, G = (float)sg, B = (float)sb;
            base[(im ? 1 : 0) * N + idx] = (R * 0.299f + G * 0.587f + B * 0.114f) * inv;
            base[(im ? 4 : 2) * N + idx] = (R * 0.596f - G * 0.274f - B * 0.322f) * inv;
            base[(im ? 5 : 3) * N + idx] = (R * 0.211f - G * 0.523f + B * 0.312f) * inv;
        }
    }
}

/* F2 for boxes of F >= kPrep2MinF on inputs whose rows are a multiple of
 * 16 bytes apart (so every row of a box has the same alignment): a
 * workgroup produces kPrep2Cols downscaled pixels of one row.  Phase 1: each
 * thread takes aligned 16-byte chunks of the box's column range and adds the
 * chunk over the F rows in registers, bytes in 16-bit lanes (even bytes, odd
 * bytes), then writes the 16 column sums to LDS as u16.  Phase 2: each
 * thread adds the column sums of its pixel's F x C bytes per channel.  The
 * integer sums, and so the planes, are those of fsim_prep_body; LDS holds one
 * row of sums instead of F rows of bytes. */
constexpr int kPrep2Cols = 256;
constexpr int kPrep2MinF = 4; /* measured: F = 8 507 -> 301 us, F = 4 99 -> 93 us per 30-32 pairs */
__host__ __device__ __forceinline__ int prep2_chunks(int F, int ch, int cols) { return (cols * F * ch) / 16 + 2; }
inline size_t prep2_lds_bytes(int F, int ch, int cols) {
    return (size_t)2 * prep2_chunks(F, ch, cols) * 16 * sizeof(uint16_t);
}

template <int FT>
__device__ __forceinline__ void fsim_prep2_body(const uint8_t* __restrict__ img, const uint8_t* __restrict__ ref,
                                                long pair_stride, int w, int h, int ch, int wd, int hd, int cols,
                                                float* __restrict__ yiq, unsigned* __restrict__ hist, int bx,
                                                int y, int pair) {
    extern __shared__ uint4 p2_smem[];
    uint16_t* cs = (uint16_t*)p2_smem; /* [image][chunk * 16 + byte] */
    constexpr int F = FT;
    const int tid = threadIdx.x;
    if (bx == 0 && y == 0) {
        unsigned* hp = hist + (long)pair * 8 * kHistBins;
        for (int i = tid; i < 8 * kHistBins; i += blockDim.x) hp[i] = 0u;
    }
    const int xd0 = bx * cols;
    const int ncols = min(cols, wd - xd0);
    const int cc0 = max(xd0 * F - F / 2, 0), cc1 = min((xd0 + ncols) * F - F / 2, w);
    const int ys = y * F - F / 2;
    const int k0 = max(ys, 0), nk = min(ys + F, h) - k0;
    const int seg = max(cc1 - cc0, 0) * ch;
    const uint8_t* src[2] = {img + (long)pair * pair_stride, ref + (long)pair * pair_stride};
    /* the box rows share their alignment d */
    const uintptr_t a00 = (uintptr_t)(src[0] + ((long)k0 * w + cc0) * ch);
    const uintptr_t a01 = (uintptr_t)(src[1] + ((long)k0 * w + cc0) * ch);
    const int d0 = (int)(a00 & 15), d1 = (int)(a01 & 15);
    const int K0 = (d0 + seg + 15) / 16, K1 = (d1 + seg + 15) / 16;
    const int KS = prep2_chunks(F, ch, cols); /* chunk slots per image */
    const long pitch = (long)w * ch;
    for (int t = tid; t < K0 + K1; t += blockDim.x) {
        const int im = t >= K0;
        const int k = t - im * K0;
        const uint4* p = (const uint4*)(((im ? a01 : a00) & ~(uintptr_t)15) + (uintptr_t)k * 16);
        uint4 v[FT];
#pragma unroll
        for (int r = 0; r < FT; ++r) {
            typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
            /* streaming loads: each image byte is read once */
            const u32x4_t t = r < nk ? __builtin_nontemporal_load((const u32x4_t*)((const uint8_t*)p + r * pitch))
                                     : u32x4_t{0u, 0u, 0u, 0u};
            v[r] = make_uint4(t.x, t.y, t.z, t.w);
        }
        unsigned e[4] = {0u, 0u, 0u, 0u}, o[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int r = 0; r < FT; ++r) {
            const unsigned q[4] = {v[r].x, v[r].y, v[r].z, v[r].w};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                e[i] += q[i] & 0x00FF00FFu;
                o[i] += (q[i] >> 8) & 0x00FF00FFu;
            }
        }
        /* u16 sums of bytes 4i .. 4i+3: (e lo, o lo, e hi, o hi) */
        uint4 lo, hi;
        lo.x = (e[0] & 0xFFFFu) | (o[0] << 16);
        lo.y = (e[0] >> 16) | (o[0] & 0xFFFF0000u);
        lo.z = (e[1] & 0xFFFFu) | (o[1] << 16);
        lo.w = (e[1] >> 16) | (o[1] & 0xFFFF0000u);
        hi.x = (e[2] & 0xFFFFu) | (o[2] << 16);
        hi.y = (e[2] >> 16) | (o[2] & 0xFFFF0000u);
        hi.z = (e[3] & 0xFFFFu) | (o[3] << 16);
        hi.w = (e[3] >> 16) | (o[3] & 0xFFFF0000u);
        uint4* dst = p2_smem + ((long)im * KS + k) * 2;
        dst[0] = lo;
        dst[1] = hi;
    }
    __syncthreads();
    const long N = (long)wd * hd;
    float* base = yiq + (long)pair * 6 * N;
    const float inv = 1.0f / ((float)F * (float)F);
    for (int xc = tid; xc < ncols; xc += blockDim.x) {
        const int x = xd0 + xc;
        const int xs = x * F - F / 2;
        const int j0 = max(xs, cc0), j1 = min(xs + F, cc1);
        const long idx = (long)y * wd + x;
        for (int im = 0; im < 2; ++im) {
            const uint16_t* c = cs + (long)im * KS * 16 + (im ? d1 : d0) + (j0 - cc0) * ch;
            int sr = 0, sg = 0, sb = 0;
            if (ch == 1) {
                for (int j = 0; j < j1 - j0; ++j) sr += c[j];
                sg = sb = sr;
            } else {
                for (int j = 0; j < j1 - j0; ++j) {
                    sr += c[j * ch];
                    sg += c[j * ch + 1];
                    sb += c[j * ch + 2];
                }
            }
            const float R = (float)sr, G = (float)sg, B = (float)sb;
            base[(im ? 1 : 0) * N + idx] = (R * 0.299f + G * 0.587f + B * 0.114f) * inv;
            base[(im ? 4 : 2) * N + idx] = (R * 0.596f - G * 0.274f - B * 0.322f) * inv;
            base[(im ? 5 : 3) * N + idx] = (R * 0.211f - G * 0.523f + B * 0.312f) * inv;
        }
    }
}


/* FT = F for the common factors 1..8 (compile-time box width), 0 otherwise */

/* Copies element idx < total from load(idx) to store(idx, v), the block's
 * threads taking idx = base + tid + q * blockDim in rounds of K per thread:
 * all K loads of a round are in flight before its first store. */
template <int K, class Load, class Store>
__device__ __forceinline__ void staged_copy(int total, const Load& load, const Store& store) {
    for (int base = 0; base < total; base += K * (int)blockDim.x) {
        float2 v[K];
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int idx = base + threadIdx.x + q * blockDim.x;
            v[q] = idx < total ? load(idx) : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int idx = base + threadIdx.x + q * blockDim.x;
            if (idx < total) store(idx, v[q]);
        }
    }
}

/* Row half of F5 on the packed rows Y_in + i*Y_ref of `rb` rows; the two
 * real-input row spectra are split with X(k) = (Z(k) + conj Z(-k)) / 2,
 * Y(k) = (Z(k) - conj Z(-k)) / (2i). */
__device__ __forceinline__ void fsim_rowfft_body(const float* __restrict__ yiq, float2* __restrict__ A, int wd,
                                                 int hd, int rb, const LinePlan& P, int bx, int pair) {
    extern __shared__ float2 lds[];
    const int ld = lds_pitch(P.lds_len);
    float2* b0 = lds;
    float2* b1 = lds + rb * ld;
    const float2* tw = stage_twiddles(P, lds + 2 * rb * ld);
    const long N = (long)wd * hd;
    const int y0 = bx * rb;
    const int nr = min(rb, hd - y0);
    const float* Yin = yiq + (long)pair * 6 * N;
    const float* Yref = Yin + N;
    staged_copy<8>(nr * wd, [&](int idx) {
        const int r = idx / wd, x = idx - r * wd;
        const long g = (long)(y0 + r) * wd + x;
        return make_float2(Yin[g], Yref[g]);
    }, [&](int idx, float2 v) {
        const int r = idx / wd, x = idx - r * wd;
        b0[r * ld + x] = v;
    });
    __syncthreads();
    float2* R = line_fft<-1>(b0, b1, ld, nr, P, tw);
    float2* Ain = A + (long)pair * 2 * N;
    float2* Aref = Ain + N;
    for (int idx = threadIdx.x; idx < nr * wd; idx += blockDim.x) {
        const int r = idx / wd, k = idx - r * wd;
        const int m = (k == 0) ? 0 : wd - k;
        const float2 zk = R[r * ld + k], zm = R[r * ld + m];
        const long g = (long)(y0 + r) * wd + k;
        Ain[g] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        Aref[g] = make_float2(0.5f * (zk.y + zm.y), 0.5f * (zm.x - zk.x));
    }
}


/* Column half of F5, then F9 and the column half of F11.  A workgroup owns
 * `bx` adjacent columns of one image of one pair: the forward column FFT
 * gives the spectrum; each of the 16 filter products H_z * F is inverse
 * transformed along the columns and written to plane img*16 + z.
 * The flat grid is laid out (pair, img, padded column block) with the
 * column-block count padded to a multiple of 8, so column block cb always
 * runs on XCD cb mod 8 and each XCD's L2 holds only its slice of the bank. */
/* element y of column line c: spectrum times filter; output y to the plane */
/* LDS line buffers of the column pass (the spectrum stays in global
 * memory): the two-level exchange area, or the two work buffers of a
 * Stockham, Rader or Bluestein transform */
__host__ __device__ __forceinline__ int colpass_bufs(const LinePlan& P) { return (!P.blue && P.tn1 > 0) ? 1 : 2; }

/* element y of column line c of a plane (row pitch wd), and its store */
struct ColLoad {
    const float2* a; /* the plane at the column */
    int wd;
    __device__ float2 operator()(int y) const { return *at32(a, (unsigned)(y * wd)); }
};
struct ColStore {
    float2* a;
    int wd;
    __device__ void operator()(int y, float2 v) const { *at32(a, (unsigned)(y * wd)) = v; }
};
struct ColFilterLoad {
    const float2* fs; /* the column's spectrum (global, in place of its input) */
    const float* h;   /* the filter plane at the column */
    int wd;
    __device__ float2 operator()(int y) const {
        const unsigned o = (unsigned)(y * wd);
        return cscale(*at32(fs, o), *at32(h, o));
    }
};
struct ColStreamStore {
    float2* d; /* the response plane at the column */
    int wd;
    __device__ void operator()(int y, float2 v) const { nt_store(at32(d, (unsigned)(y * wd)), v); }
};

__device__ __forceinline__ void fsim_colpass_body(float2* __restrict__ A, float2* __restrict__ Tout,
                                                  const float* __restrict__ bank, int wd, int hd, int bx,
                                                  int nbx_pad, const LinePlan& P, int blk) {
    extern __shared__ float2 lds[];
    const int cb = blk % nbx_pad;
    const int rest = blk / nbx_pad;
    const int img = rest & 1, pair = rest >> 1;
    const int x0 = cb * bx;
    if (x0 >= wd) return;
    const int ld = lds_pitch(P.lds_len);
    const int nbuf = colpass_bufs(P);
    float2* b0 = lds;
    float2* b1 = lds + (nbuf - 1) * bx * ld; /* one buffer: the exchange area only */
    const float2* tw = stage_twiddles(P, lds + nbuf * bx * ld);
    const long N = (long)wd * hd;
    const int nc = min(bx, wd - x0);
    /* plane offsets y * wd + c (< N) in 32 bits from workgroup-uniform bases */
    float2* sp = A + ((long)pair * 2 + img) * N + x0;
    __syncthreads(); /* twiddles staged */
    /* forward column transform in place: every transform reads all its
     * inputs in its first phase and stores in its last, and no other
     * workgroup reads these columns */
    line_fft_io<-1, true>([&](int c) { return ColLoad{sp + c, wd}; }, [&](int c) { return ColStore{sp + c, wd}; },
                          b0, b1, ld, nc, P, tw);
    __syncthreads(); /* the spectrum's stores visible to the whole workgroup */
    for (int z = 0; z < 16; ++z) {
        const float* hz = bank + (long)z * N + x0;
        float2* dst = Tout + ((long)pair * 32 + img * 16 + z) * N + x0;
        /* the inverse transform reads H_z * F straight from the spectrum and
         * the bank (requested together), and streams its outputs to the row
         * pass (out of the L2) */
        line_fft_io<1, true>([&](int c) { return ColFilterLoad{sp + c, hz + c, wd}; },
                             [&](int c) { return ColStreamStore{dst + c, wd}; }, b0, b1, ld, nc, P, tw);
    }
}


/* element j of one row of a response plane (zeros for rows past the image),
 * read once */
struct RowLoad {
    const float2* p;
    bool ok;
    __device__ float2 operator()(int j) const { return ok ? nt_load(p + j) : make_float2(0.f, 0.f); }
};

/* Row half of F11 (normalised by 1/N) fused with F12
 * (fsim_sum_filter_responses.glsl:33-83) and the F13 packing
 * (fsim_pack_for_median.glsl:33-44) for `rb` rows of one image, one
 * orientation after the other; T planes img*16 + 4*o + s.  Writes per image
 * the plane sum_o An (summed in the order o = 0..3, as the PC step
 * (fsim_phase_congruency.glsl:55-60) adds them) and E_o, plus the keys and
 * their level-1 histogram per orientation. */
__host__ __device__ __forceinline__ size_t rowpass_lds_bytes(const LinePlan& P, int rb, int wd) {
    const size_t bufbytes = (size_t)4 * rb * lds_pitch(P.lds_len) * sizeof(float2);
    return 2 * bufbytes + (size_t)lds_twiddles(P) * sizeof(float2) + kHistBins * sizeof(unsigned) +
           sizeof(float) * (size_t)rb * wd;
}

__device__ __forceinline__ void fsim_rowpass_body(const float2* __restrict__ Tin, float* __restrict__ an_e,
                                                  float* __restrict__ keys, unsigned* __restrict__ hist, int wd,
                                                  int hd, int rb, const LinePlan& P, float invN, int bx, int img,
                                                  int pair) {
    extern __shared__ float2 lds[];
    const int ld = lds_pitch(P.lds_len);
    const int nl = 4 * rb;
    float2* b0 = lds;
    float2* b1 = lds + nl * ld;
    const int ntw = lds_twiddles(P);
    const float2* tw = stage_twiddles(P, lds + 2 * nl * ld);
    /* the histogram has its own LDS: cleared once here, then by each thread
     * for the bins it flushes, so no orientation spends a phase on it */
    unsigned* hl = (unsigned*)(lds + 2 * nl * ld + ntw);
    float* sa = (float*)(hl + kHistBins); /* rb * wd running sums of An */
    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) hl[i] = 0u;
    const long N = (long)wd * hd;
    const int y0 = bx * rb;
    const int nr = min(rb, hd - y0);
    float* ap = an_e + ((long)pair * 2 + img) * kAnePlanes * N;
    const float epsN = 0.0001f / invN;
    /* when the 4 * rb lines fit kRowPrefetch elements per thread, the next
     * orientation's lines are loaded into registers while the current one is
     * transformed (the transform's barriers order LDS only) */
    constexpr int kRowPrefetch = 8;
    const int tot = nl * wd;
    /* two-level and Rader lines read their first-phase inputs straight from
     * the planes (no copy into LDS first); Stockham and Bluestein lines copy
     * the tile into b0 */
    const bool direct = !P.blue && (P.rader || P.tn1 > 0);
    const bool pf = !direct && tot <= kRowPrefetch * (int)blockDim.x;
    float2 pre[kRowPrefetch];
    const float iwd = 1.0f / (float)wd, irb = 1.0f / (float)rb;
    /* row offsets (y0 + r) * wd + x (< N) in 32 bits from per-plane bases */
    auto fetch = [&](int o) {
#pragma unroll
        for (int q = 0; q < kRowPrefetch; ++q) {
            const int idx = threadIdx.x + q * blockDim.x;
            pre[q] = make_float2(0.f, 0.f);
            if (idx < tot) {
                const int l = fast_div(idx, wd, iwd), x = idx - l * wd;
                const int s = fast_div(l, rb, irb), r = l - s * rb;
                if (r < nr)
                    pre[q] = nt_load(at32(Tin + ((long)pair * 32 + (img * 4 + o) * 4 + s) * N,
                                          (unsigned)((y0 + r) * wd + x)));
            }
        }
    };
    if (pf) fetch(0);
    for (int o = 0; o < 4; ++o) {
        const int io = img * 4 + o;
        float2* R = b0;
        if (direct) {
            lds_barrier(); /* also orders the twiddles staged before the loop */
            line_fft_io<1, false>(
                [&](int l) {
                    const int s = fast_div(l, rb, irb), r = l - s * rb;
                    return RowLoad{Tin + ((long)pair * 32 + io * 4 + s) * N + (long)(y0 + min(r, nr - 1)) * wd,
                                   r < nr};
                },
                [&](int l) { return PtrStore{b0 + l * ld}; }, b0, b1, ld, nl, P, tw);
        } else {
            if (pf) {
#pragma unroll
                for (int q = 0; q < kRowPrefetch; ++q) {
                    const int idx = threadIdx.x + q * blockDim.x;
                    if (idx < tot) {
                        const int l = fast_div(idx, wd, iwd), x = idx - l * wd;
                        b0[l * ld + x] = pre[q];
                    }
                }
                if (o < 3) fetch(o + 1);
            } else {
                /* rounds of kRowPrefetch elements per thread, all loads of a
                 * round in flight before its LDS stores */
                for (int base = 0; base < tot; base += kRowPrefetch * (int)blockDim.x) {
                    float2 v[kRowPrefetch];
#pragma unroll
                    for (int q = 0; q < kRowPrefetch; ++q) {
                        const int idx = base + threadIdx.x + q * blockDim.x;
                        v[q] = make_float2(0.f, 0.f);
                        if (idx < tot) {
                            const int l = fast_div(idx, wd, iwd), x = idx - l * wd;
                            const int s = fast_div(l, rb, irb), r = l - s * rb;
                            if (r < nr)
                                v[q] = nt_load(
                                    at32(Tin + ((long)pair * 32 + io * 4 + s) * N, (unsigned)((y0 + r) * wd + x)));
                        }
                    }
#pragma unroll
                    for (int q = 0; q < kRowPrefetch; ++q) {
                        const int idx = base + threadIdx.x + q * blockDim.x;
                        if (idx < tot) {
                            const int l = fast_div(idx, wd, iwd), x = idx - l * wd;
                            b0[l * ld + x] = v[q];
                        }
                    }
                }
            }
            lds_barrier(); /* also orders the twiddles staged before the loop */
            R = line_fft<1>(b0, b1, ld, nl, P, tw);
        }
        float* enp = ap + (long)(1 + o) * N;
        float* kp = keys + ((long)pair * 8 + io) * N;
        const int sstride = rb * ld; /* lines of one scale */
        for (int idx = threadIdx.x; idx < nr * wd; idx += blockDim.x) {
            const int r = fast_div(idx, wd, iwd), x = idx - r * wd;
            const float2* er = R + r * ld + x;
            float2 eo[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) eo[s] = er[s * sstride];
            float amp, energy, key;
            fsim_response_epilogue(eo, invN, epsN, amp, energy, key);
            const unsigned g = (unsigned)(y0 * wd + idx); /* < N */
            /* sum over o, in the order o = 0..3 (each idx stays with one thread) */
            if (o == 0) sa[idx] = amp;
            else if (o < 3) sa[idx] += amp;
            else *at32(ap, g) = sa[idx] + amp;
            *at32(enp, g) = energy;
            *at32(kp, g) = key;
            atomicAdd(&hl[(__float_as_uint(key) >> 21) & 1023u], 1u);
        }
        lds_barrier(); /* histogram complete; line buffers free for the next orientation */
        unsigned* hg = hist + ((long)pair * 8 + io) * kHistBins;
        for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) {
            const unsigned c = hl[i];
            if (c) atomicAdd(&hg[i], c);
            hl[i] = 0u;
        }
        /* the next orientation's first barrier orders these before its
         * epilogue; its loads go to b0 only */
    }
}


/* Finds the bucket holding the element of rank k (0-based, ascending) in
 * counts[0..nbins); results in *bucket / *rest (shared).  All threads of
 * the block must call it. */
__device__ void rank_bucket(const unsigned* counts, int nbins, unsigned k, int* bucket,
                            unsigned* rest, unsigned* wsum) {
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int per = (nbins + nthr - 1) / nthr;
    const int b0 = tid * per;
    unsigned local = 0;
    for (int b = b0; b < b0 + per && b < nbins; ++b) local += counts[b];
    /* wave64 inclusive scan */
    const int lane = tid & 63, wid = tid >> 6;
    unsigned inc = local;
    for (int d = 1; d < 64; d <<= 1) {
        unsigned t = __shfl_up(inc, d, 64);
        if (lane >= d) inc += t;
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    if (tid == 0) {
        unsigned acc = 0;
        const int nw = (nthr + 63) / 64;
        for (int i = 0; i < nw; ++i) {
            unsigned t = wsum[i];
            wsum[i] = acc;
            acc += t;
        }
    }
    __syncthreads();
    const unsigned excl = wsum[wid] + inc - local;
    if (local > 0 && k >= excl && k < excl + local) {
        unsigned cum = excl;
        for (int b = b0; b < b0 + per && b < nbins; ++b) {
            const unsigned c = counts[b];
            if (k < cum + c) {
                *bucket = b;
                *rest = k - cum;
                break;
            }
            cum += c;
        }
    }
    __syncthreads();
}

using NoiseArgs = FsimNoiseArgs;

/* F13 + F15 thresholds: exact element N/2 of the ascending keys
 * (fsim_noise_power.cpp:198-199) by a 3-level radix select on the float bit
 * patterns (keys are >= 0, so the bits order like the values and bit 31 is
 * clear): level 1 uses the histogram built by fsim_rowpass (bits 30..21),
 * level 2 bits 20..10, level 3 bits 9..0.  The level-2 pass also copies the
 * keys of the level-1 bucket (about 5-8% of them on natural images) into
 * LDS, so level 3 reads them from there; when more than kSelectCap keys fall
 * in that bucket, level 3 reads the keys from memory again. */
constexpr int kSelectCap = 16384; /* keys of the level-1 bucket kept in LDS */
constexpr int kSelectLoads = 8;    /* 16-byte key loads in flight per thread */

__device__ __forceinline__ void fsim_select_body(const float* __restrict__ keys, const unsigned* __restrict__ hist,
                                                 float* __restrict__ thr, double* __restrict__ stats, long N,
                                                 const NoiseArgs& na, int io, int pair) {
    extern __shared__ unsigned cand[]; /* kSelectCap keys of bucket b1 */
    __shared__ unsigned h[2048];
    __shared__ unsigned wsum[16];
    __shared__ int bucket;
    __shared__ unsigned rest;
    __shared__ unsigned ncand;
    const long slot = (long)pair * 8 + io;
    const unsigned* kb = (const unsigned*)(keys + slot * N);
    const unsigned* hg = hist + slot * kHistBins;
    const unsigned k = (unsigned)(N / 2);
    const int lane = threadIdx.x & 63;

    for (int i = threadIdx.x; i < kHistBins; i += blockDim.x) h[i] = hg[i];
    if (threadIdx.x == 0) { bucket = 0; rest = 0; ncand = 0u; }
    __syncthreads();
    rank_bucket(h, kHistBins, k, &bucket, &rest, wsum);
    const unsigned b1 = (unsigned)bucket, k1 = rest;

    for (int i = threadIdx.x; i < 2048; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    /* level 2 + compaction: every lane runs the same number of iterations so
     * the wave-wide ballot sees all lanes */
    auto level2 = [&](unsigned u, bool valid) {
        const bool m = valid && ((u >> 21) & 1023u) == b1;
        if (m) atomicAdd(&h[(u >> 10) & 2047u], 1u);
        const unsigned long long mask = __ballot(m);
        if (mask) {
            const int leader = __ffsll((long long)mask) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(&ncand, (unsigned)__popcll(mask));
            base = __shfl(base, leader, 64);
            if (m) {
                const unsigned at = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
                if (at < (unsigned)kSelectCap) cand[at] = u;
            }
        }
    };
    /* the same for a round of 4 * kSelectLoads keys per lane: the lane counts
     * its keys of bucket b1 (adding them to the level-2 histogram), one
     * wave-wide prefix sum and one reservation place them, then the lane
     * writes them in order */
    auto level2_round = [&](const uint4 (&q)[kSelectLoads], const bool (&ok)[kSelectLoads]) {
        unsigned cnt = 0;
#pragma unroll
        for (int v = 0; v < kSelectLoads; ++v) {
            const unsigned e[4] = {q[v].x, q[v].y, q[v].z, q[v].w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool m = ok[v] && ((e[j] >> 21) & 1023u) == b1;
                if (m) atomicAdd(&h[(e[j] >> 10) & 2047u], 1u);
                cnt += m ? 1u : 0u;
            }
        }
        unsigned inc = cnt; /* inclusive scan over the wave */
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned t = __shfl_up(inc, d, 64);
            if (lane >= d) inc += t;
        }
        const unsigned total = __shfl(inc, 63, 64);
        if (total == 0u) return;
        unsigned base = 0;
        if (lane == 63) base = atomicAdd(&ncand, total);
        unsigned at = __shfl(base, 63, 64) + inc - cnt;
        if (cnt == 0u) return;
#pragma unroll
        for (int v = 0; v < kSelectLoads; ++v) {
            const unsigned e[4] = {q[v].x, q[v].y, q[v].z, q[v].w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (ok[v] && ((e[j] >> 21) & 1023u) == b1) {
                    if (at < (unsigned)kSelectCap) cand[at] = e[j];
                    ++at;
                }
            }
        }
    };
    /* keys are read as uint4 when every plane starts 16-byte aligned */
    const bool vec = (N % 4) == 0;
    const long stride = blockDim.x;
    if (vec) {
        /* kSelectLoads loads in flight per thread */
        const uint4* k4 = (const uint4*)kb;
        const long n4 = N / 4;
        for (long i0 = 0; i0 < n4; i0 += kSelectLoads * stride) {
            uint4 q[kSelectLoads];
            bool ok[kSelectLoads];
#pragma unroll
            for (int u = 0; u < kSelectLoads; ++u) {
                const long i = i0 + u * stride + threadIdx.x;
                ok[u] = i < n4;
                q[u] = ok[u] ? k4[i] : make_uint4(0u, 0u, 0u, 0u);
            }
            level2_round(q, ok);
        }
    } else {
        for (long i0 = 0; i0 < N; i0 += stride) {
            const long i = i0 + threadIdx.x;
            const bool ok = i < N;
            level2(ok ? kb[i] : 0u, ok);
        }
    }
    __syncthreads();
    rank_bucket(h, 2048, k1, &bucket, &rest, wsum);
    const unsigned b2 = (unsigned)bucket, k2 = rest;

    for (int i = threadIdx.x; i < 1024; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    const unsigned pre = (b1 << 11) | b2;
    const unsigned nc = ncand;
    if (nc <= (unsigned)kSelectCap) {
        for (unsigned i = threadIdx.x; i < nc; i += blockDim.x) {
            const unsigned u = cand[i];
            if (((u >> 10) & 0x1FFFFFu) == pre) atomicAdd(&h[u & 1023u], 1u);
        }
    } else if (vec) {
        const uint4* k4 = (const uint4*)kb;
        for (long i = threadIdx.x; i < N / 4; i += blockDim.x) {
            const uint4 q = k4[i];
            const unsigned u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (((u[j] >> 10) & 0x1FFFFFu) == pre) atomicAdd(&h[u[j] & 1023u], 1u);
        }
    } else {
        for (long i = threadIdx.x; i < N; i += blockDim.x) {
            const unsigned u = kb[i];
            if (((u >> 10) & 0x1FFFFFu) == pre) atomicAdd(&h[u & 1023u], 1u);
        }
    }
    __syncthreads();
    rank_bucket(h, 1024, k2, &bucket, &rest, wsum);
    if (threadIdx.x == 0) {
        const unsigned bits = (b1 << 21) | (b2 << 10) | (unsigned)bucket;
        const double median = (double)__uint_as_float(bits);
        const int o = io & 3;
        /* fsim_noise_power.cpp:199-201 */
        const double mean = -median / log(0.5);
        const double np = mean / na.em_n[o];
        /* fsim_phase_congruency.glsl:45-53 */
        const double pi = 3.141592653589;
        const double e2 = 2.0 * np * na.est[2 * o] + 4.0 * np * na.est[2 * o + 1];
        const double tau = sqrt(e2 / 2.0);
        const double ene = tau * sqrt(pi / 2.0);
        const double sig = sqrt((2.0 - pi / 2.0) * (tau * tau));
        const double T = (ene + 2.0 * sig) / 1.7;
        thr[slot] = (float)T;
        double* st = stats + slot * 4;
        st[0] = median;
        st[1] = np;
        st[2] = T;
        st[3] = 0.0;
    }
}


__device__ __forceinline__ double wave_sum(double v) {
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

/* F17 (fsim_final_multiply.cpp:294-300): fixed-order sum of the block
 * partials; fsim = sum sim / sum pcm, fsimc = sum simc / sum pcm. */
__device__ __forceinline__ void fsim_finalize_body(const double* __restrict__ partial, int nblk,
                                                   double* __restrict__ res, int pair) {
    /* strided per-thread sums, butterfly sums within each wave, then the
     * waves in order: a fixed order for a given nblk and block size */
    __shared__ double sh[3][kBlock / 64];
    double s[3] = {0, 0, 0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        const double* p = partial + ((long)pair * nblk + b) * 4;
        for (int j = 0; j < 3; ++j) s[j] += p[j];
    }
    for (int j = 0; j < 3; ++j) s[j] = wave_sum(s[j]);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int j = 0; j < 3; ++j) sh[j][wid] = s[j];
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[3] = {0, 0, 0};
        for (int w = 0; w < (int)(blockDim.x / 64); ++w)
            for (int j = 0; j < 3; ++j) t[j] += sh[j][w];
        double* r = res + (long)pair * 8;
        r[0] = t[0];
        r[1] = t[1];
        r[2] = t[2];
        r[3] = t[1] / t[0];
        r[4] = t[2] / t[0];
        r[5] = r[6] = r[7] = 0.0;
    }
}

/* x^y for 0 <= x <= 1 as exp2(y log2 x) on the hardware log and exp (about 1
 * ulp each, as a shader's pow); a denormal x is scaled into the normal range
 * first */
__device__ __forceinline__ float pow_unit(float x, float y) {
    const bool tiny = x < 1.17549435e-38f;
    const float l = __builtin_amdgcn_logf(tiny ? x * 18446744073709551616.0f : x) - (tiny ? 64.0f : 0.0f);
    return __builtin_amdgcn_exp2f(y * l);
}

/* F3 (fsim_gradientmap.glsl:38-56) + F15 (fsim_phase_congruency.glsl:40-63)
 * + F16 (fsim_final_multiply.glsl:26-57) + per-block FP64 partials of F17.
 * A workgroup covers a kPcTx x kPcTy tile, each thread kPcRows pixels of one
 * column, 4 rows apart; the Y planes of both images are staged in LDS with a
 * one-pixel halo for the Scharr stencil. */
constexpr int kPcTx = 64, kPcTy = 4, kPcRows = kPcTx * kPcTy / kBlock;
static_assert(kPcRows * kBlock == kPcTx * kPcTy && kBlock / kPcTx == 4, "4 thread rows per tile");

__device__ __forceinline__ void fsim_pcfinal_body(const float* __restrict__ yiq, float* __restrict__ grad,
                                                  const float* __restrict__ an_e, const float* __restrict__ thr,
                                                  double* __restrict__ partial, int wd, int hd, float* simc_map,
                                                  float* pc_map, float* pcm_map, float* sim_map,
                                                  const float* __restrict__ pc_src, int bx, int by, int gx,
                                                  int gy, int pair) {
    constexpr int HX = kPcTx + 2, HY = kPcTy + 2;
    __shared__ float ys[2][HY][HX];
    __shared__ double red[3][kBlock / 64];
    const long N = (long)wd * hd;
    const int x0 = bx * kPcTx, y0 = by * kPcTy;
    const float* yq = yiq + (long)pair * 6 * N;
    const int tx = threadIdx.x % kPcTx, ty0 = threadIdx.x / kPcTx;
    const int x = x0 + tx;
    /* the halo loads are issued first, then every pixel's own operands; only
     * the halo values are awaited before the LDS writes and the barrier, so
     * the other loads stay in flight across it */
    constexpr int kHalo = 2 * HY * HX, kStage = (kHalo + kBlock - 1) / kBlock;
    float sv[kStage];
#pragma unroll
    for (int st = 0; st < kStage; ++st) {
        const int i = threadIdx.x + st * kBlock;
        const int im = i / (HY * HX), rem = i - im * HY * HX;
        const int ly = rem / HX, lx = rem - ly * HX;
        const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
        sv[st] = (i < kHalo && gx >= 0 && gx < wd && gy >= 0 && gy < hd)
                     ? *at32(yq, (unsigned)(im * (int)N + gy * wd + gx))
                     : 0.f;
    }
    float a_n[kPcRows][2], a_e[kPcRows][2][4], iq[kPcRows][4];
#pragma unroll
    for (int r = 0; r < kPcRows; ++r) {
        const int y = y0 + ty0 + 4 * r;
        if (x < wd && y < hd) {
            /* plane offsets (below 6 N) in 32 bits from per-pair bases */
            const unsigned idx = (unsigned)(y * wd + x), n1 = (unsigned)N;
            for (int im = 0; im < 2; ++im) {
                if (pc_src) {
                    a_n[r][im] = *at32(pc_src + (long)pair * 2 * N, im * n1 + idx);
                    for (int o = 0; o < 4; ++o) a_e[r][im][o] = 0.f;
                } else {
                    /* plane 0: sum_o An (added in the order o = 0..3 by the row pass) */
                    const float* ae = an_e + ((long)pair * 2 + im) * kAnePlanes * N;
                    a_n[r][im] = __builtin_nontemporal_load(at32(ae, idx));
                    for (int o = 0; o < 4; ++o) a_e[r][im][o] = __builtin_nontemporal_load(at32(ae, (1 + o) * n1 + idx));
                }
            }
            /* streaming loads for the planes read once (the Y halo above is
             * shared by neighbouring tiles through the cache) */
            for (int c = 0; c < 4; ++c) iq[r][c] = __builtin_nontemporal_load(at32(yq, (2 + c) * n1 + idx));
        }
    }
#pragma unroll
    for (int st = 0; st < kStage; ++st) {
        const int i = threadIdx.x + st * kBlock;
        if (i < kHalo) (&ys[0][0][0])[i] = sv[st];
    }
    lds_barrier();
    double v0 = 0, v1 = 0, v2 = 0;
#pragma unroll
    for (int r = 0; r < kPcRows; ++r) {
        const int ty = ty0 + 4 * r, y = y0 + ty;
        if (!(x < wd && y < hd)) continue;
        const long idx = (long)y * wd + x;
        /* Scharr pair as a correlation, x offset j outer, y offset k inner;
         * the staged zeros outside the image are the shader's skipped taps */
        float gm[2];
        for (int im = 0; im < 2; ++im) {
            float vs = 0.f, hs = 0.f;
            for (int j = -1; j < 2; ++j) {
                for (int k = -1; k < 2; ++k) {
                    const float v = ys[im][ty + 1 + k][tx + 1 + j];
                    vs += v * ((float)(-j) * (k == 0 ? 10.f : 3.f));  /* verticalArray   */
                    hs += v * ((float)(-k) * (j == 0 ? 10.f : 3.f));  /* horizontalArray */
                }
            }
            const float gv = vs / 16.0f, gh = hs / 16.0f;
            gm[im] = __builtin_amdgcn_sqrtf(gv * gv + gh * gh);
        }
        if (grad) {
            grad[(long)pair * 2 * N + idx] = gm[0];
            grad[(long)pair * 2 * N + N + idx] = gm[1];
        }
        float pc[2];
        for (int im = 0; im < 2 && pc_src; ++im) pc[im] = a_n[r][im];
        for (int im = 0; im < 2 && !pc_src; ++im) {
            const float* T = thr + (long)pair * 8 + im * 4;
            float et = 0.f;
            for (int o = 0; o < 4; ++o) et += fmaxf(a_e[r][im][o] - T[o], 0.0f);
            pc[im] = et * __builtin_amdgcn_rcpf(a_n[r][im]); /* no epsilon: NaN where sum An = 0 */
        }
        const float T1 = 0.85f, T2 = 160.0f, T3 = 200.0f, T4 = 200.0f, lambda = 0.03f;
        const float p1 = pc[0], p2 = pc[1];
        const float pcm = (p1 < p2) ? p2 : p1;
        /* the ratios with the hardware reciprocal (1 ulp), as a shader's divide */
        const float pcsim = (2.0f * p1 * p2 + T1) * __builtin_amdgcn_rcpf(p1 * p1 + p2 * p2 + T1);
        const float g1 = gm[0], g2 = gm[1];
        const float gsim = (2.0f * g1 * g2 + T2) * __builtin_amdgcn_rcpf(g1 * g1 + g2 * g2 + T2);
        const float sim = gsim * pcsim * pcm;
        const float i1 = iq[r][0], q1 = iq[r][1];
        const float i2 = iq[r][2], q2 = iq[r][3];
        const float isim = (2.0f * i1 * i2 + T3) * __builtin_amdgcn_rcpf(i1 * i1 + i2 * i2 + T3);
        const float qsim = (2.0f * q1 * q2 + T4) * __builtin_amdgcn_rcpf(q1 * q1 + q2 * q2 + T4);
        /* |isim qsim| <= 1 */
        const float simc = gsim * pcsim * pow_unit(fabsf(isim * qsim), lambda) * pcm;
        v0 += (double)pcm;
        v1 += (double)sim;
        v2 += (double)simc;
        const long mo = (long)pair * N + idx;
        if (simc_map) simc_map[mo] = simc;
        if (pcm_map) pcm_map[mo] = pcm;
        if (sim_map) sim_map[mo] = sim;
        if (pc_map) {
            pc_map[(long)pair * 2 * N + idx] = p1;
            pc_map[(long)pair * 2 * N + N + idx] = p2;
        }
    }
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    v2 = wave_sum(v2);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wid] = v0;
        red[1][wid] = v1;
        red[2][wid] = v2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s3[3] = {0, 0, 0};
        for (int w = 0; w < (int)(blockDim.x / 64); ++w)
            for (int j = 0; j < 3; ++j) s3[j] += red[j][w];
        const int nblk = gx * gy;
        double* p = partial + ((long)pair * nblk + by * gx + bx) * 4;
        p[0] = s3[0];
        p[1] = s3[1];
        p[2] = s3[2];
        p[3] = 0.0;
    }
}



/* [fsim, fsimc] of each pair from the n x 8 result array */


/* ---- generic line-transform passes (fsim_gen.hip), launched by fsim.hip -- */

__global__ __launch_bounds__(kBlock) void fsim_rowfft_kernel(const float* __restrict__ yiq,
                                                             float2* __restrict__ A, int wd, int hd,
                                                             int rb, LinePlan P, const int* __restrict__ gate);
__global__ __launch_bounds__(kBlockFft) void fsim_colpass_kernel(float2* __restrict__ A,
                                                              float2* __restrict__ Tout,
                                                              const float* __restrict__ bank,
                                                              int wd, int hd, int bx, int nbx_pad,
                                                              LinePlan P, const int* __restrict__ gate);
__global__ __launch_bounds__(kBlockFft) void fsim_rowpass_kernel(const float2* __restrict__ Tin,
                                                              float* __restrict__ an_e,
                                                              float* __restrict__ keys,
                                                              unsigned* __restrict__ hist,
                                                              int wd, int hd, int rb, LinePlan P,
                                                              float invN, const int* __restrict__ gate);

/* ---- host helpers ------------------------------------------------------- */

/* the kernels address a pair's planes (at most 6 float or 1 float2 plane
 * apart from their base) with 32-bit byte offsets */
inline bool fsim_planes_fit32(long N) { return N > 0 && N * 8 <= 0xFFFFFFFFL && N * 6 * 4 <= 0xFFFFFFFFL; }

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct FsimRegions {
    size_t yiq, grad, A, T, an_e, keys, hist, thr, stats, partial, res, total;
};
inline FsimRegions fsim_regions(long N, int n, int chunk, int nblk) {
    FsimRegions r{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += align256(bytes);
        return o;
    };
    r.yiq = take(sizeof(float) * 6 * N * n);
    r.grad = take(sizeof(float) * 2 * N * n);
    r.A = take(sizeof(float2) * 2 * N * n);
    r.T = take(sizeof(float2) * 32 * N * chunk);
    r.an_e = take(sizeof(float) * 2 * kAnePlanes * N * n);
    r.keys = take(sizeof(float) * 8 * N * n);
    r.hist = take(sizeof(unsigned) * 8 * kHistBins * n);
    r.thr = take(sizeof(float) * 8 * n);
    r.stats = take(sizeof(double) * 8 * 4 * n);
    r.partial = take(sizeof(double) * 4 * (size_t)nblk * n);
    r.res = take(sizeof(double) * 8 * n);
    r.total = off;
    return r;
}

inline int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return (v && *v) ? std::atoi(v) : dflt;
}

/* Tile and chunk choices of one FSIM launch sequence. */
struct FsimTiles {
    int grad_rb, col_bx, row_rb, chunk;
};

inline FsimTiles pick_tiles(const Ctx& c, const LinePlan& Px, const LinePlan& Py, long N, const FsimTune& tune) {
    FsimTiles t;
    /* 64 KiB per workgroup keeps two workgroups per CU; when that holds
     * fewer than 4 lines (or the cap), long (Bluestein) lines take the
     * device's whole LDS for more work per barrier */
    /* The generic transforms (two-level with a run-time split, Stockham,
     * Bluestein) fill a workgroup's threads only with enough lines: without
     * a measured tune they take up to 32 columns, 8 x 4 rows and 32 rows per
     * workgroup, in up to the whole LDS (mixed 256-pair batch: 39.6 ms
     * against 49.3 ms with 16 / 2 x 4 / 8 in 64 KiB; the row-pass cap
     * 4 -> 8 took its row pass from 8.9 to 8.0 ms) */
    const bool measured = tune.us_per_pair != 0; /* tuned or explicitly set tiles */
    const size_t pref = (size_t)env_int("IQM_FSIM_LDS_PREF", (int)c.lds_per_block());
    auto fit = [&](const LinePlan& P, int bufs, int cap, int group) {
        const int l = lines_that_fit(P, bufs, std::min<size_t>(c.lds_per_block(), pref), group * cap) / group;
        if (l >= std::min(cap, 4)) return l;
        return std::max(l, lines_that_fit(P, bufs, c.lds_per_block(), group * cap) / group);
    };
    /* row FFT (256 threads): at most 64 KiB, so that two or more workgroups
     * share a CU (mixed batch: 0.50 -> 0.33 ms against the whole LDS) */
    {
        const int cap = env_int("IQM_FSIM_GRB", measured ? tune.fft_rb : 32);
        const int l = lines_that_fit(Px, 2, std::min<size_t>({c.lds_per_block(), pref, 65536}), cap);
        t.grad_rb = l > 0 ? l : lines_that_fit(Px, 2, c.lds_per_block(), cap);
    }
    t.col_bx = fit(Py, colpass_bufs(Py), env_int("IQM_FSIM_BX", measured ? tune.col_bx : 64), 1);
    /* row pass: 4 scales per row, plus the histogram and the An sums */
    auto fit_row = [&](int cap) {
        auto best = [&](size_t budget) {
            int b = 0;
            for (int r = 1; r <= cap; ++r)
                if (rowpass_lds_bytes(Px, r, Px.n) <= budget) b = r;
            return b;
        };
        const int l = best(std::min<size_t>(c.lds_per_block(), pref));
        if (l >= std::min(cap, 4)) return l;
        return std::max(l, best(c.lds_per_block()));
    };
    t.row_rb = fit_row(env_int("IQM_FSIM_RB", measured ? tune.row_rb : 8));
    /* pairs per column/row-pass launch: one launch over the whole batch
     * measured fastest on MI355X (larger grids beat cache residency of the
     * response planes); the cap bounds the response-plane workspace */
    const long chunk_mb = env_int("IQM_FSIM_CHUNK_MB", 16384);
    t.chunk = (int)std::max<long>(1, (chunk_mb << 20) / (long)(sizeof(float2) * 32 * N));
    IQM_CHECK(t.grad_rb > 0 && t.col_bx > 0 && t.row_rb > 0, IQM_ERR_UNSUPPORTED,
              "FSIM grid too large for the LDS line transforms");
    return t;
}

}  // namespace iqm
