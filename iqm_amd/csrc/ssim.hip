// Fused SSIM on MI355X.
//
// Replaces the reference's three dispatches (shaders/ssim/ssim_lumapack.glsl,
// ssim_gaussinput.glsl, ssim.glsl) and the host-side MSSIM loop
// (src/gpu/ssim/ssim.cpp:386-401) with one LDS-tiled kernel plus a small
// fixed-order reduction:
//   1. a (TX+2R) x (TY+2R) tile of both images is read once from HBM and
//      turned into luma (0.299 R + 0.581 G + 0.114 B on UNORM values) minus a
//      per-tile shift (the luma of the tile centre), which keeps the
//      E[x^2] - mu^2 form of the variance free of cancellation;
//   2. a horizontal 2R+1-tap pass writes the five weighted sums x, y, x^2,
//      y^2, xy per row to LDS;
//   3. a vertical pass finishes the 2-D sums, renormalises them by the
//      in-bounds weight mass (the Gaussian window is separable, so the
//      in-bounds set is a rectangle and its mass is Gx(x) * Gy(y)),
//      evaluates the SSIM formula and stores the map;
//   4. pixels inside the MSSIM crop [R, W-R] x [R, H-R] are summed into one
//      FP64 partial per workgroup; ssim_finalize_kernel adds the partials in
//      a fixed order.
#include "ctx.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

namespace iqm {

constexpr int kMaxRadius = 15;

struct SsimWeights {
    float g[2 * kMaxRadius + 1]; /* g[d + R] = exp(-d^2 / (2 sigma^2)) */
};

__device__ __forceinline__ float luma_at(const uint8_t* p, int ch) {
    float r, g, b;
    if (ch == 1) {
        r = g = b = (float)p[0] / 255.0f;
    } else {
        r = (float)p[0] / 255.0f;
        g = (float)p[1] / 255.0f;
        b = (float)p[2] / 255.0f;
    }
    return 0.299f * r + 0.581f * g + 0.114f * b;
}

template <int R, int TX, int TY>
__device__ __forceinline__ void ssim_body(const uint8_t* __restrict__ img, const uint8_t* __restrict__ ref,
                                          long pair_stride, int w, int h, int ch, const SsimWeights& wt,
                                          float c1, float c2, float* __restrict__ map,
                                          double* __restrict__ partial, int bx, int by, int gx, int gy,
                                          int pair) {
    constexpr int LW = TX + 2 * R, LH = TY + 2 * R;
    __shared__ float lx[LH][LW];
    __shared__ float ly[LH][LW];
    __shared__ float hp[5][LH][TX];
    __shared__ double red[kBlock / 64];

    const uint8_t* A = img + (long)pair * pair_stride;
    const uint8_t* B = ref + (long)pair * pair_stride;
    const int x0 = bx * TX, y0 = by * TY;
    const int tid = threadIdx.x;

    /* the bytes of the tile's rows (clipped to the image) of both images are
     * staged as 16-byte aligned chunks in the space of hp, with every load of
     * a thread issued before its LDS stores; an aligned 16-byte block holding
     * a byte of the image lies in that byte's page */
    const int c0 = max(x0 - R, 0), cend = min(x0 + TX + R, w);
    const int r0 = max(y0 - R, 0), rend = min(y0 + TY + R, h);
    const int seg = (cend - c0) * ch, nrows = rend - r0;
    const int rc = seg / 16 + 2;
    static_assert(2 * LH * ((LW * 4) / 16 + 2) * 16 <= (int)sizeof(hp), "SSIM staging does not fit");
    uint4* stage = (uint4*)&hp[0][0][0];
    const uint8_t* sb = (const uint8_t*)stage;
    {
        const int total = 2 * nrows * rc;
        for (int base = 0; base < total; base += 8 * (int)blockDim.x) {
            uint4 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = base + u * blockDim.x + tid;
                v[u] = make_uint4(0u, 0u, 0u, 0u);
                if (i < total) {
                    const int rr = i / rc, k = i - rr * rc;
                    const int im = rr >= nrows, r = rr - im * nrows;
                    const uintptr_t a0 = (uintptr_t)((im ? B : A) + ((long)(r0 + r) * w + c0) * ch);
                    const uintptr_t a = (a0 & ~(uintptr_t)15) + (uintptr_t)k * 16;
                    if (a <= ((a0 + seg - 1) & ~(uintptr_t)15)) v[u] = *(const uint4*)a;
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = base + u * blockDim.x + tid;
                if (i < total) stage[i] = v[u];
            }
        }
    }
    __syncthreads();
    /* byte address in the staging of pixel (gx, gy) of image im */
    auto staged = [&](int im, int gx, int gy) {
        const uintptr_t a0 = (uintptr_t)((im ? B : A) + ((long)gy * w + c0) * ch);
        return sb + ((im * nrows + (gy - r0)) * rc) * 16 + (int)(a0 & 15) + (gx - c0) * ch;
    };
    const int cx = min(x0 + TX / 2, w - 1), cy = min(y0 + TY / 2, h - 1);
    const float sx = luma_at(staged(0, cx, cy), ch);
    const float sy = luma_at(staged(1, cx, cy), ch);

    for (int i = tid; i < LH * LW; i += blockDim.x) {
        const int ry = i / LW, rx = i - ry * LW;
        const int gx = x0 - R + rx, gy = y0 - R + ry;
        float a = 0.f, b = 0.f;
        if (gx >= 0 && gx < w && gy >= 0 && gy < h) {
            a = luma_at(staged(0, gx, gy), ch) - sx;
            b = luma_at(staged(1, gx, gy), ch) - sy;
        }
        lx[ry][rx] = a;
        ly[ry][rx] = b;
    }
    __syncthreads(); /* hp is written next */

    for (int i = tid; i < LH * TX; i += blockDim.x) {
        const int ry = i / TX, c = i - ry * TX;
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
#pragma unroll
        for (int d = 0; d <= 2 * R; ++d) {
            const float g = wt.g[d];
            const float a = lx[ry][c + d], b = ly[ry][c + d];
            const float ga = g * a, gb = g * b;
            s0 += ga;
            s1 += gb;
            s2 += ga * a;
            s3 += gb * b;
            s4 += ga * b;
        }
        hp[0][ry][c] = s0;
        hp[1][ry][c] = s1;
        hp[2][ry][c] = s2;
        hp[3][ry][c] = s3;
        hp[4][ry][c] = s4;
    }
    __syncthreads();

    double crop = 0.0;
    /* each thread finishes RPT adjacent rows of one column, reading every
     * input row of hp once (the sums keep the tap order d = 0..2R) */
    constexpr int ROWS_PER_PASS = kBlock / TX;
    constexpr int RPT = TY / ROWS_PER_PASS;
    static_assert(TY % ROWS_PER_PASS == 0, "rows per thread");
    const int c = tid % TX;
    const int rb = (tid / TX) * RPT;
    const int x = x0 + c;
    float gxs = 0.f;
#pragma unroll
    for (int d = -R; d <= R; ++d)
        if (x + d >= 0 && x + d < w) gxs += wt.g[d + R];
    float m[RPT][5];
#pragma unroll
    for (int q = 0; q < RPT; ++q)
#pragma unroll
        for (int k = 0; k < 5; ++k) m[q][k] = 0.f;
#pragma unroll
    for (int j = 0; j < RPT + 2 * R; ++j) {
        float hv[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) hv[k] = hp[k][rb + j][c];
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int d = j - q;
            if (d >= 0 && d <= 2 * R) {
                const float g = wt.g[d];
#pragma unroll
                for (int k = 0; k < 5; ++k) m[q][k] += g * hv[k];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int y = y0 + rb + q;
        if (x >= w || y >= h) continue;
        float gys = 0.f;
#pragma unroll
        for (int d = 0; d <= 2 * R; ++d) {
            const int yy = y + d - R;
            if (yy >= 0 && yy < h) gys += wt.g[d];
        }
        const float inv = 1.0f / (gxs * gys);
        const float ma = m[q][0] * inv, mb = m[q][1] * inv;
        const float va = m[q][2] * inv - ma * ma;
        const float vb = m[q][3] * inv - mb * mb;
        const float cov = m[q][4] * inv - ma * mb;
        const float mi = ma + sx, mr = mb + sy;
        const float v = ((2.0f * mi * mr + c1) * (2.0f * cov + c2)) /
                        ((mi * mi + mr * mr + c1) * (va + vb + c2));
        if (map) map[(long)pair * w * h + (long)y * w + x] = v;
        if (x >= R && x <= w - R && y >= R && y <= h - R) crop += (double)v;
    }
    for (int d = 32; d > 0; d >>= 1) crop += __shfl_xor(crop, d, 64);
    if ((tid & 63) == 0) red[tid >> 6] = crop;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int i = 0; i < kBlock / 64; ++i) s += red[i];
        partial[((long)pair * gy + by) * gx + bx] = s;
    }
}

template <int R, int TX, int TY>
__global__ __launch_bounds__(kBlock) void ssim_kernel(const uint8_t* __restrict__ img,
                                                      const uint8_t* __restrict__ ref,
                                                      long pair_stride, int w, int h, int ch,
                                                      SsimWeights wt, float c1, float c2,
                                                      float* __restrict__ map,
                                                      double* __restrict__ partial) {
    ssim_body<R, TX, TY>(img, ref, pair_stride, w, h, ch, wt, c1, c2, map, partial, blockIdx.x, blockIdx.y,
                         gridDim.x, gridDim.y, blockIdx.z);
}

/* ---- row-walker SSIM (window radius R <= kRwMaxRadius) -------------------
 *
 * A workgroup of kRwThreads threads produces a kRwTX x kRwTY tile; thread t
 * owns output columns 2t and 2t+1.  It walks down the kRwTY + 2R window rows
 * once:
 *   1. the bytes of both images' window rows arrive in LDS in chunks of
 *      K = 2R+1 rows, loaded straight into LDS (global_load_lds_dwordx4) one
 *      chunk ahead of their use;
 *   2. per window row, the shifted lumas a = x - sx and b = y - sy of the
 *      row's kRwTX + 2R columns give u = a + b, v = a - b, u^2 and v^2 in a
 *      double-buffered LDS row;
 *   3. each thread forms the four horizontal weighted sums of its two
 *      columns from that row and keeps them in a ring of the last K rows held
 *      in registers (slot = row mod K, so every ring index is a compile-time
 *      constant);
 *   4. once K rows are in the ring, the vertical sums give one output row.
 * SSIM needs only sigma_x^2 + sigma_y^2 and sigma_xy, and with
 * Var(u) = E[u^2] - E[u]^2 (likewise v) these are (Var u + Var v) / 2 and
 * (Var u - Var v) / 4, so four moments replace the five of ssim_body.  The
 * vertical pass reads no LDS and recomputes no halo rows; the tile's halo
 * costs (kRwTY + 2R) / kRwTY in rows.  Same shift and renormalisation as
 * ssim_body. */
constexpr int kRwThreads = 128;
constexpr int kRwTX = 2 * kRwThreads;
constexpr int kRwTY = 64;
constexpr int kRwMaxRadius = 6;

__host__ __device__ __forceinline__ int rw_row_chunks(int R, int ch) { return ((kRwTX + 2 * R) * ch) / 16 + 2; }
/* chunk buffers of the staged rows: two (the next chunk lands while this one
 * is used) for gray input; one for RGB and RGBA, whose stage rows are three
 * and four times longer -- twice the workgroups then share a CU and cover
 * each other's chunk loads (mixed 256-pair batch SSIM 3.8 -> 2.4 ms) */
__host__ __device__ __forceinline__ int rw_stage_bufs(int ch) { return ch >= 3 ? 1 : 2; }
/* dynamic LDS: the chunk buffers of 2 images x K rows, then the rows of u,
 * v, u^2, v^2, double-buffered */
static size_t rw_lds_bytes(int R, int ch) {
    return (size_t)rw_stage_bufs(ch) * 2 * (2 * R + 1) * rw_row_chunks(R, ch) * 16 +
           (size_t)2 * 4 * (kRwTX + 2 * R) * sizeof(float);
}

__device__ __forceinline__ float luma_rgb(float r, float g, float b) {
    return (0.299f * r + 0.581f * g + 0.114f * b) * (1.0f / 255.0f);
}
__device__ __forceinline__ float luma_rw(const uint8_t* p, int ch) {
    float r, g, b;
    if (ch == 1) {
        r = g = b = (float)p[0];
    } else {
        r = (float)p[0];
        g = (float)p[1];
        b = (float)p[2];
    }
    return luma_rgb(r, g, b);
}

template <int R>
__device__ __forceinline__ void ssim_rw_body(const uint8_t* __restrict__ img, const uint8_t* __restrict__ ref,
                                             long pair_stride, int w, int h, int ch, const SsimWeights& wt,
                                             float c1, float c2, float* __restrict__ map,
                                             double* __restrict__ partial, int bx, int by, int gx, int gy,
                                             int pair, int ty) {
    constexpr int K = 2 * R + 1;
    constexpr int LW = kRwTX + 2 * R; /* window columns (even) */
    extern __shared__ uint4 rw_smem[];
    __shared__ double red[kRwThreads / 64];

    const uint8_t* A = img + (long)pair * pair_stride;
    const uint8_t* B = ref + (long)pair * pair_stride;
    const int tid = threadIdx.x, wv = tid >> 6;
    const int x0 = bx * kRwTX, y0 = by * ty;
    const int c0 = max(x0 - R, 0), cend = min(x0 + kRwTX + R, w);
    const int seg = (cend - c0) * ch;
    const int rc = rw_row_chunks(R, ch);
    const int slots = 2 * K * rc;
    const int nbuf = rw_stage_bufs(ch);
    float* lum = (float*)(rw_smem + nbuf * slots); /* [buf][u, v, u^2, v^2][LW] */
    const int nwin = ty + 2 * R;
    const int nch = (nwin + K - 1) / K;

    /* shift: luma of the tile centre (read by every lane from one address) */
    const int cx = min(x0 + kRwTX / 2, w - 1), cy = min(y0 + ty / 2, h - 1);
    const float sx = luma_rw(A + ((long)cy * w + cx) * ch, ch);
    const float sy = luma_rw(B + ((long)cy * w + cx) * ch, ch);

    /* chunk c -> stage buffer: slot (im * K + r) * rc + k holds the aligned
     * 16-byte block k of window row c*K + r of image im; lane l of a wave
     * writes slot (wave base + l) */
    auto issue = [&](int c, uint4* st) {
        for (int base = 0; base < slots; base += kRwThreads) {
            const int i = base + tid;
            if (i < slots) {
                const int rr = i / rc, k = i - rr * rc;
                const int im = rr >= K, r = rr - im * K;
                const int wr = c * K + r, gyy = y0 - R + wr;
                if (wr < nwin && gyy >= 0 && gyy < h) {
                    const uintptr_t a0 = (uintptr_t)((im ? B : A) + ((long)gyy * w + c0) * ch);
                    const uintptr_t alast = (a0 + seg - 1) & ~(uintptr_t)15;
                    uintptr_t a = (a0 & ~(uintptr_t)15) + (uintptr_t)k * 16;
                    if (a > alast) a = alast; /* spare slot: a block of this row, never read */
                    __builtin_amdgcn_global_load_lds((const void*)a,
                                                     (__attribute__((address_space(3))) void*)(st + base + wv * 64),
                                                     16, 0, 0);
                }
            }
        }
    };

    /* reciprocal in-bounds weight mass per column, and of the full window */
    float rgx[2];
#pragma unroll
    for (int col = 0; col < 2; ++col) {
        const int x = x0 + 2 * tid + col;
        float sgx = 0.f;
#pragma unroll
        for (int d = -R; d <= R; ++d)
            if (x + d >= 0 && x + d < w) sgx += wt.g[d + R];
        rgx[col] = 1.0f / sgx; /* infinite only for columns past the image, which are not stored */
    }
    float gfull = 0.f;
#pragma unroll
    for (int d = 0; d <= 2 * R; ++d) gfull += wt.g[d];
    const float rgfull = 1.0f / gfull;
    /* reciprocal in-bounds weight mass of the first R and the last R image
     * rows (entry k: row k, entry R + k: row h - R + k), read by the rows
     * that need it; interior rows use rgfull */
    __shared__ float redge[2 * kRwMaxRadius];
    if (tid < 2 * R) {
        const int yk = tid < R ? tid : h - 2 * R + tid;
        float gys = 0.f;
#pragma unroll
        for (int d = 0; d <= 2 * R; ++d) {
            const int yy = yk + d - R;
            if (yy >= 0 && yy < h) gys += wt.g[d];
        }
        redge[tid] = 1.0f / gys; /* only rows inside the image are read */
    }

    float ring[K][2][4];
    double crop = 0.0;

    issue(0, rw_smem);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
        if (nbuf == 2 && c + 1 < nch) issue(c + 1, rw_smem + ((c + 1) & 1) * slots);
        const uint8_t* sb = (const uint8_t*)(rw_smem + (nbuf == 2 ? (c & 1) * slots : 0));
#pragma unroll
        for (int p = 0; p < K; ++p) {
            const int i = c * K + p; /* window row */
            if (i < nwin) {
                /* 2. luma of window row i */
                const int gyy = y0 - R + i;
                const bool rowin = gyy >= 0 && gyy < h;
                float* lq = lum + (i & 1) * 4 * LW; /* u, v, u^2, v^2 rows */
                int off[2] = {0, 0};
                if (rowin) {
                    off[0] = (p * rc) * 16 + (int)(((uintptr_t)(A + ((long)gyy * w + c0) * ch)) & 15);
                    off[1] = ((K + p) * rc) * 16 + (int)(((uintptr_t)(B + ((long)gyy * w + c0) * ch)) & 15);
                }
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int j = q == 0 ? 2 * tid : kRwTX + 2 * tid;
                    if (q == 0 || tid < R) {
                        float uv[2], vv[2];
                        const int gx0 = x0 - R + j;
                        if (ch >= 3 && rowin && gx0 >= 0 && gx0 + 1 < w) {
                            /* both pixels inside: their 6 or 8 bytes come
                             * from 3 aligned dwords, picked by alignbyte; the
                             * third dword may reach past the row's staged
                             * blocks into the next slot or the luma rows
                             * after the stage buffers (inside the LDS
                             * allocation), and its bytes past the pixels are
                             * never used */
                            const int bo = (gx0 - c0) * ch;
                            float lm[2][2];
#pragma unroll
                            for (int im = 0; im < 2; ++im) {
                                const int ob = off[im] + bo;
                                const unsigned* wv = (const unsigned*)(sb + (ob & ~3));
                                const unsigned w0 = wv[0], w1 = wv[1], w2 = wv[2];
                                const unsigned d0 = __builtin_amdgcn_alignbyte(w1, w0, (unsigned)(ob & 3));
                                const unsigned d1 = __builtin_amdgcn_alignbyte(w2, w1, (unsigned)(ob & 3));
                                const unsigned p1 = ch == 3 ? ((d0 >> 24) | (d1 << 8)) : d1;
                                lm[im][0] = luma_rgb((float)(d0 & 0xffu), (float)((d0 >> 8) & 0xffu),
                                                     (float)((d0 >> 16) & 0xffu));
                                lm[im][1] = luma_rgb((float)(p1 & 0xffu), (float)((p1 >> 8) & 0xffu),
                                                     (float)((p1 >> 16) & 0xffu));
                            }
#pragma unroll
                            for (int e = 0; e < 2; ++e) {
                                const float a = lm[0][e] - sx, b = lm[1][e] - sy;
                                uv[e] = a + b;
                                vv[e] = a - b;
                            }
                        } else {
#pragma unroll
                            for (int e = 0; e < 2; ++e) {
                                const int gxx = gx0 + e;
                                uv[e] = vv[e] = 0.f;
                                if (rowin && gxx >= 0 && gxx < w) {
                                    const int bo = (gxx - c0) * ch;
                                    const float a = luma_rw(sb + off[0] + bo, ch) - sx;
                                    const float b = luma_rw(sb + off[1] + bo, ch) - sy;
                                    uv[e] = a + b;
                                    vv[e] = a - b;
                                }
                            }
                        }
                        *(float2*)(lq + j) = make_float2(uv[0], uv[1]);
                        *(float2*)(lq + LW + j) = make_float2(vv[0], vv[1]);
                        *(float2*)(lq + 2 * LW + j) = make_float2(uv[0] * uv[0], uv[1] * uv[1]);
                        *(float2*)(lq + 3 * LW + j) = make_float2(vv[0] * vv[0], vv[1] * vv[1]);
                    }
                }
                lds_barrier();
                /* 3. horizontal sums of columns 2t, 2t+1 into ring slot p */
                float m[2][4];
#pragma unroll
                for (int col = 0; col < 2; ++col)
#pragma unroll
                    for (int k = 0; k < 4; ++k) m[col][k] = 0.f;
#pragma unroll
                for (int k2 = 0; k2 < 2 * R + 2; k2 += 2) {
                    float2 q2[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) q2[k] = *(const float2*)(lq + k * LW + 2 * tid + k2);
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int kk = k2 + e;
#pragma unroll
                        for (int col = 0; col < 2; ++col) {
                            const int d = kk - col;
                            if (d >= 0 && d <= 2 * R) {
#pragma unroll
                                for (int k = 0; k < 4; ++k) m[col][k] += wt.g[d] * (e ? q2[k].y : q2[k].x);
                            }
                        }
                    }
                }
#pragma unroll
                for (int col = 0; col < 2; ++col)
#pragma unroll
                    for (int k = 0; k < 4; ++k) ring[p][col][k] = m[col][k];
                /* 4. output row i - 2R */
                if (i >= 2 * R) {
                    const int y = y0 + i - 2 * R;
                    if (y < h) {
                        float rgy = rgfull;
                        if (y < R) rgy = redge[y];
                        else if (y + R >= h) rgy = redge[y - h + 2 * R];
                        float out[2];
#pragma unroll
                        for (int col = 0; col < 2; ++col) {
                            float sm[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                            for (int d = 0; d <= 2 * R; ++d) {
                                const int slot = (p + 1 + d) % K;
#pragma unroll
                                for (int k = 0; k < 4; ++k) sm[k] += wt.g[d] * ring[slot][col][k];
                            }
                            const float inv = rgx[col] * rgy;
                            const float mu = sm[0] * inv, mv = sm[1] * inv;
                            const float vu = sm[2] * inv - mu * mu; /* Var u = var_x + var_y + 2 cov */
                            const float vv = sm[3] * inv - mv * mv; /* Var v = var_x + var_y - 2 cov */
                            const float mi = 0.5f * (mu + mv) + sx, mr = 0.5f * (mu - mv) + sy;
                            const float num = (2.0f * mi * mr + c1) * (0.5f * (vu - vv) + c2);
                            const float den = (mi * mi + mr * mr + c1) * (0.5f * (vu + vv) + c2);
                            out[col] = num * __builtin_amdgcn_rcpf(den);
                        }
                        const int x = x0 + 2 * tid;
                        float* mrow = map ? map + (long)pair * w * h + (long)y * w : nullptr;
#pragma unroll
                        for (int col = 0; col < 2; ++col) {
                            if (x + col < w) {
                                if (mrow) mrow[x + col] = out[col];
                                /* adding +0 outside the crop leaves the sum unchanged */
                                const bool incrop = x + col >= R && x + col <= w - R && y >= R && y <= h - R;
                                crop += (double)(incrop ? out[col] : 0.0f);
                            }
                        }
                    }
                }
            }
        }
        /* one buffer: chunk c + 1 replaces chunk c, whose last reads (the luma
         * of its last row) every thread finished before that row's barrier */
        if (nbuf == 1 && c + 1 < nch) issue(c + 1, rw_smem);
        /* chunk c + 1 has landed in LDS for every wave */
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }
    for (int d = 32; d > 0; d >>= 1) crop += __shfl_xor(crop, d, 64);
    if ((tid & 63) == 0) red[tid >> 6] = crop;
    __syncthreads();
    if (tid == 0) {
        double sum = 0.0;
        for (int i = 0; i < kRwThreads / 64; ++i) sum += red[i];
        partial[((long)pair * gy + by) * gx + bx] = sum;
    }
}

template <int R>
__global__ __launch_bounds__(kRwThreads) void ssim_rw_kernel(const uint8_t* __restrict__ img,
                                                             const uint8_t* __restrict__ ref, long pair_stride,
                                                             int w, int h, int ch, SsimWeights wt, float c1,
                                                             float c2, float* __restrict__ map,
                                                             double* __restrict__ partial, int ty) {
    ssim_rw_body<R>(img, ref, pair_stride, w, h, ch, wt, c1, c2, map, partial, blockIdx.x, blockIdx.y, gridDim.x,
                    gridDim.y, blockIdx.z, ty);
}

__device__ __forceinline__ void ssim_finalize_body(const double* __restrict__ partial, int nblk, double count,
                                                   double* __restrict__ out, int pair) {
    __shared__ double sh[kBlock];
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += partial[(long)pair * nblk + b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[pair] = sh[0] / count;
}

__global__ __launch_bounds__(kBlock) void ssim_finalize_kernel(const double* __restrict__ partial,
                                                               int nblk, double count,
                                                               double* __restrict__ out) {
    ssim_finalize_body(partial, nblk, count, out, blockIdx.x);
}

static int env_int_ssim(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return (v && *v) ? std::atoi(v) : dflt;
}

/* output tile of the kernel used for radius R */
static void ssim_tile(int R, int& TX, int& TY) {
    if (R <= kRwMaxRadius) {
        TX = kRwTX;
        TY = kRwTY;
    } else {
        TX = R <= 7 ? 64 : 32;
        TY = R <= 7 ? 16 : 8;
    }
}

/* rows of the row walker's output tile (IQM_SSIM_TY overrides, at least 8) */
static int ssim_tile_rows(int R) {
    int TX, TY;
    ssim_tile(R, TX, TY);
    if (R <= kRwMaxRadius) TY = std::max(8, env_int_ssim("IQM_SSIM_TY", TY));
    return TY;
}

static dim3 ssim_grid(int w, int h, int ks, int n) {
    const int R = (ks - 1) / 2;
    int TX, TY;
    ssim_tile(R, TX, TY);
    TY = ssim_tile_rows(R);
    return dim3(ceil_div(w, TX), ceil_div(h, TY), n);
}

/* Launches of one same-size SSIM batch: the map/partials kernel (one FP64
 * partial per workgroup at partial[(pair * gy + by) * gx + bx], every one
 * written) and the fixed-order finalize into out[0..n).  `timed` routes the
 * launches through the context's stage timer (hipEvents); the
 * allocation-free entry point launches without it. */
void Ctx::ssim_enqueue(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch, long pair_stride, int n,
                       int ks, float k1, float k2, float sigma, float* d_maps, double* partial, double* out,
                       bool timed) {
    const int R = (ks - 1) / 2;
    SsimWeights wt{};
    const double sg = (double)sigma;
    for (int d = -R; d <= R; ++d) wt.g[d + R] = (float)std::exp(-(double)(d * d) / (2.0 * sg * sg));
    const dim3 grid = ssim_grid(w, h, ks, n);
    const int nblk = grid.x * grid.y;
    const float c1 = k1 * k1, c2 = k2 * k2;
    auto L = [&](const char* tag, auto kernel, dim3 g, dim3 b, size_t lds, auto... args) {
        if (timed) {
            launch(tag, kernel, g, b, lds, args...);
        } else {
            hipLaunchKernelGGL(kernel, g, b, lds, stream_, args...);
            IQM_HIP(hipGetLastError());
        }
    };
    if (R <= kRwMaxRadius) {
        const size_t lds = rw_lds_bytes(R, ch);
        const int ty = ssim_tile_rows(R);
        switch (R) {
#define IQM_SSIM_RW_CASE(r)                                                                                         \
    case r:                                                                                                         \
        L("ssim", ssim_rw_kernel<r>, grid, dim3(kRwThreads), lds, d_img, d_ref, pair_stride, w, h, ch, wt, c1, c2, \
          d_maps, partial, ty);                                                                                     \
        break;
            IQM_SSIM_RW_CASE(1) IQM_SSIM_RW_CASE(2) IQM_SSIM_RW_CASE(3)
            IQM_SSIM_RW_CASE(4) IQM_SSIM_RW_CASE(5) IQM_SSIM_RW_CASE(6)
#undef IQM_SSIM_RW_CASE
            default: break;
        }
    } else switch (R) {
#define IQM_SSIM_CASE(r, tx, ty)                                                                                \
    case r:                                                                                                     \
        L("ssim", ssim_kernel<r, tx, ty>, grid, dim3(kBlock), 0, d_img, d_ref, pair_stride, w, h, ch, wt, c1, c2, \
          d_maps, partial);                                                                                     \
        break;
        IQM_SSIM_CASE(7, 64, 16) IQM_SSIM_CASE(8, 32, 8) IQM_SSIM_CASE(9, 32, 8) IQM_SSIM_CASE(10, 32, 8)
        IQM_SSIM_CASE(11, 32, 8) IQM_SSIM_CASE(12, 32, 8) IQM_SSIM_CASE(13, 32, 8) IQM_SSIM_CASE(14, 32, 8)
        IQM_SSIM_CASE(15, 32, 8)
#undef IQM_SSIM_CASE
        default: break;
    }
    const double count = (double)((long)(w - 2 * R + 1) * (long)(h - 2 * R + 1));
    L("ssim_finalize", ssim_finalize_kernel, dim3(n), dim3(kBlock), 0, (const double*)partial, nblk, count, out);
}

static void ssim_check_args(int w, int h, int ks) {
    IQM_CHECK(ks >= 3 && ks <= 2 * kMaxRadius + 1 && (ks % 2) == 1, IQM_ERR_INVALID,
              "kernel_size must be odd and in [3, 31]");
    const int R = (ks - 1) / 2;
    /* ssim.cpp:390-400 needs a non-empty crop */
    IQM_CHECK(w - R >= R && h - R >= R, IQM_ERR_TOO_SMALL, "image smaller than the SSIM window");
}

double* Ctx::ssim_batch(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch,
                        long pair_stride, int n, int ks, float k1, float k2, float sigma,
                        float* d_maps, double* mssim_host) {
    ssim_check_args(w, h, ks);
    const dim3 grid = ssim_grid(w, h, ks, n);
    const size_t nblk = (size_t)grid.x * grid.y;
    double* partial = (double*)workspace(WS_SSIM, sizeof(double) * (nblk * n + n));
    double* out = partial + nblk * n;
    ssim_enqueue(d_img, d_ref, w, h, ch, pair_stride, n, ks, k1, k2, sigma, d_maps, partial, out, true);
    if (mssim_host) {
        IQM_HIP(hipMemcpyAsync(mssim_host, out, sizeof(double) * n, hipMemcpyDeviceToHost, stream_));
        IQM_HIP(hipStreamSynchronize(stream_));
    }
    return out;
}

/* ---- allocation-free SSIM batch ------------------------------------------
 * The FP64 workgroup partials of iqm_ssim_batch_dev_noalloc live in this
 * file-scope device array instead of a context workspace.  In host code its
 * name is only a stand-in: the device address of each GPU's copy is taken
 * once per device with hipGetSymbolAddress (ssim_noalloc_setup, called when a
 * context is created) and only that address is handed to kernels.  Every call
 * on one GPU shares that GPU's copy, so calls that overlap on one GPU (two
 * contexts, two streams) race; calls on one context are ordered by its
 * stream.  Each workgroup writes its own partial, so nothing needs clearing
 * between calls. */
static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxNoAllocDevices] = {};
}  // namespace

void ssim_noalloc_setup(int device) {
    IQM_CHECK(device >= 0 && device < kMaxNoAllocDevices, IQM_ERR_UNSUPPORTED,
              "device ordinal beyond the allocation-free scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    /* the caller has made `device` current */
    IQM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    IQM_CHECK(p != nullptr, IQM_ERR_DEVICE, "hipGetSymbolAddress returned no address for the SSIM scratch");
    g_scratch_ptr[device] = (double*)p;
}

void Ctx::ssim_batch_noalloc(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch, long pair_stride,
                             int n, int ks, float k1, float k2, float sigma, float* d_maps, double* d_mssim) {
    ssim_check_args(w, h, ks);
    const dim3 grid = ssim_grid(w, h, ks, n);
    const size_t need = (size_t)grid.x * grid.y * (size_t)n;
    IQM_CHECK(need <= (size_t)kNoAllocPartials, IQM_ERR_INVALID,
              "batch needs " + std::to_string(need) + " workgroup partials; the allocation-free SSIM holds at most " +
                  std::to_string(kNoAllocPartials) + " (fewer pairs per call)");
    double* partial = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        partial = g_scratch_ptr[device_];
    }
    IQM_CHECK(partial != nullptr, IQM_ERR_DEVICE, "allocation-free SSIM scratch not set up on this device");
    ssim_enqueue(d_img, d_ref, w, h, ch, pair_stride, n, ks, k1, k2, sigma, d_maps, partial, d_mssim, false);
}

/* ---- mixed batches ------------------------------------------------------ */

__device__ __forceinline__ int ssim_mix_find(const int* __restrict__ st, int n, int b) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (st[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int R, int TX, int TY>
__global__ __launch_bounds__(kBlock) void mix_ssim_kernel(const MixJob* __restrict__ jobs, const int* __restrict__ st,
                                                          int n, SsimWeights wt, float c1, float c2) {
    const int j = ssim_mix_find(st, n, blockIdx.x);
    const MixJob& J = jobs[j];
    const int local = blockIdx.x - st[j];
    ssim_body<R, TX, TY>(J.img, J.ref, 0, J.w, J.h, J.ch, wt, c1, c2, J.ssim_map, J.ssim_partial,
                         local % J.ssim_gx, local / J.ssim_gx, J.ssim_gx, J.ssim_gy, 0);
}

template <int R>
__global__ __launch_bounds__(kRwThreads) void mix_ssim_rw_kernel(const MixJob* __restrict__ jobs,
                                                                 const int* __restrict__ st, int n, SsimWeights wt,
                                                                 float c1, float c2) {
    const int j = ssim_mix_find(st, n, blockIdx.x);
    const MixJob& J = jobs[j];
    const int local = blockIdx.x - st[j];
    ssim_rw_body<R>(J.img, J.ref, 0, J.w, J.h, J.ch, wt, c1, c2, J.ssim_map, J.ssim_partial, local % J.ssim_gx,
                    local / J.ssim_gx, J.ssim_gx, J.ssim_gy, 0, kRwTY);
}

/* finalize, then the pair's MSSIM into its row of the batch result */
__global__ __launch_bounds__(kBlock) void mix_ssim_finalize_kernel(const MixJob* __restrict__ jobs,
                                                                   const int* __restrict__ st, int n) {
    const int j = ssim_mix_find(st, n, blockIdx.x);
    const MixJob& J = jobs[j];
    ssim_finalize_body(J.ssim_partial, J.ssim_gx * J.ssim_gy, J.ssim_count, J.ssim_out, 0);
    if (threadIdx.x == 0) J.out[0] = J.ssim_out[0];
}


void Ctx::mix_ssim_fill(const PairJob& pj, MixJob& mj, char* base, size_t& off, MixCounts& cnt, int* blk,
                        int* cls, const SsimArgs& sp) {
    const int R = (sp.ks - 1) / 2;
    int TX, TY;
    ssim_tile(R, TX, TY);
    mj.ssim_gx = ceil_div(pj.w, TX);
    mj.ssim_gy = ceil_div(pj.h, TY);
    mj.ssim_count = (double)((long)(pj.w - 2 * R + 1) * (long)(pj.h - 2 * R + 1));
    mj.ssim_map = pj.ssim_map;
    const int nb = mj.ssim_gx * mj.ssim_gy;
    char* b = base ? base + off : nullptr;
    mj.ssim_partial = (double*)b;
    mj.ssim_out = (double*)(b + sizeof(double) * nb);
    off += (sizeof(double) * (nb + 1) + 255) & ~(size_t)255;
    blk[MIX_SSIM] = nb;
    blk[MIX_SSIMFIN] = 1;
    cls[MIX_SSIM] = cls[MIX_SSIMFIN] = 0;
    cnt.blocks[MIX_SSIM][0] += nb;
    cnt.blocks[MIX_SSIMFIN][0] += 1;
    /* the row walker's stage rows are sized for the widest input of the batch */
    if (R <= kRwMaxRadius) cnt.lds[MIX_SSIM][0] = std::max(cnt.lds[MIX_SSIM][0], rw_lds_bytes(R, pj.ch));
}

template <int R>
static void launch_ssim_mix(Ctx& c, int nblocks, size_t lds, const MixJob* jobs, const int* st, int n,
                            const SsimWeights& wt, float c1, float c2) {
    if constexpr (R <= kRwMaxRadius)
        c.launch("ssim", mix_ssim_rw_kernel<R>, dim3(nblocks), dim3(kRwThreads), lds, jobs, st, n, wt, c1, c2);
    else if constexpr (R <= 7)
        c.launch("ssim", mix_ssim_kernel<R, 64, 16>, dim3(nblocks), dim3(kBlock), 0, jobs, st, n, wt, c1, c2);
    else
        c.launch("ssim", mix_ssim_kernel<R, 32, 8>, dim3(nblocks), dim3(kBlock), 0, jobs, st, n, wt, c1, c2);
}

void Ctx::mix_ssim_launch(const MixJob* d_jobs, const int* d_starts, int n, const MixCounts& cnt,
                          const SsimArgs& sp) {
    if (cnt.blocks[MIX_SSIM][0] == 0) return;
    const int R = (sp.ks - 1) / 2;
    SsimWeights wt{};
    const double sg = (double)sp.sigma;
    for (int d = -R; d <= R; ++d) wt.g[d + R] = (float)std::exp(-(double)(d * d) / (2.0 * sg * sg));
    const float c1 = sp.k1 * sp.k1, c2 = sp.k2 * sp.k2;
    const int* st = d_starts + (size_t)MIX_SSIM * 2 * (n + 1);
    const int nb = cnt.blocks[MIX_SSIM][0];
    switch (R) {
#define IQM_SSIM_MIX_CASE(r) \
    case r: launch_ssim_mix<r>(*this, nb, cnt.lds[MIX_SSIM][0], d_jobs, st, n, wt, c1, c2); break;
        IQM_SSIM_MIX_CASE(1) IQM_SSIM_MIX_CASE(2) IQM_SSIM_MIX_CASE(3) IQM_SSIM_MIX_CASE(4) IQM_SSIM_MIX_CASE(5)
        IQM_SSIM_MIX_CASE(6) IQM_SSIM_MIX_CASE(7) IQM_SSIM_MIX_CASE(8) IQM_SSIM_MIX_CASE(9) IQM_SSIM_MIX_CASE(10)
        IQM_SSIM_MIX_CASE(11) IQM_SSIM_MIX_CASE(12) IQM_SSIM_MIX_CASE(13) IQM_SSIM_MIX_CASE(14) IQM_SSIM_MIX_CASE(15)
#undef IQM_SSIM_MIX_CASE
        default: IQM_CHECK(false, IQM_ERR_INVALID, "kernel_size must be odd and in [3, 31]");
    }
    launch("ssim_finalize", mix_ssim_finalize_kernel, dim3(cnt.blocks[MIX_SSIMFIN][0]), dim3(kBlock), 0, d_jobs,
           d_starts + (size_t)MIX_SSIMFIN * 2 * (n + 1), n);
}

}  // namespace iqm

