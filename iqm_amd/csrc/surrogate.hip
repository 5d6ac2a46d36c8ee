// Learned phase-congruency surrogate (see surrogate.hpp).
//
// Layers 2-5 (32 -> 32 channels, 3x3) run as implicit GEMMs on the gfx950
// BF16 matrix cores (v_mfma_f32_16x16x32_bf16): a wavefront computes 16
// output pixels x 16 output channels per instruction chain, one instruction
// per tap covering the tap's 32 input channels (K = 32).  The input gradient
// of these layers reuses the same kernel with the flipped, transposed weights
// and the ReLU mask of the layer below.  Layers 1 and 6 (1 <-> 32 channels)
// and the weight gradients use FP32 FMA with register tiling; all
// accumulation is FP32, weight-gradient partials are summed in a fixed order.
#include "surrogate.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

namespace iqm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

/* FP32 -> BF16 bits, round to nearest even (one v_cvt_pk_bf16_f32 on gfx950) */
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
/* two values to BF16 in one conversion, lo in the low half */
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}
/* ReLU on BF16 bits: a BF16 value is negative exactly when its bits are as a
 * signed 16-bit integer, so max(bits, 0) is the ReLU (-0 becomes +0) */
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t v) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(i16x2, v), (i16x2){0, 0}));
}
__device__ __forceinline__ uint16_t relu_bf16(uint16_t v) { return (int16_t)v < 0 ? (uint16_t)0 : v; }

/* ---- forward ----------------------------------------------------------- */

/* An input value as the BF16 pair (hi, lo) = (bf16(v), bf16(v - hi)), hi in
 * the low half: 16 significant bits (v - hi is exact), and its products with
 * BF16 weights are exact in FP32. */
__device__ __forceinline__ uint32_t x_pair(float v) {
    const float hi = bf2f(f2bf(v));
    return pack_bf16(v, v - hi);
}

/* Layer-1 A fragment (weights; M = output channel co, K = 8 g .. 8 g + 7):
 * k = 8 dy + 2 dx + part for tap (dy, dx) and part (hi, lo) of the input
 * pair, dy = g; k % 8 >= 6 and g = 3 are zeros. */
__device__ __forceinline__ bf16x8 sur_first_a(const float* __restrict__ params, int co, int g) {
    bf16x8 a = {};
    if (g < 3)
#pragma unroll
        for (int e = 0; e < 6; ++e) a[e] = (__bf16)params[SurLayout::w1 + co * 9 + 3 * g + e / 2];
    return a;
}

/* Layer 1: 1 -> 32 channels, ReLU, on the matrix cores.  Image j of the
 * input is read at x + (j/2)*pstride + (j%2)*istride, times `scale` (so
 * FSIM's Y planes can be read in place).  Each input value enters as its
 * BF16 pair x_pair and the weights in BF16, so every product is exact and
 * the sums, from the bias, are FP32 inside one instruction: K = tap row dy
 * (lane group) x tap column dx x (hi, lo), 18 of the 32.  A wavefront
 * computes 16 pixels of a row x 32 channels (one instruction per 16
 * channels); the fused kernel builds the same operands, so its layer 1 has
 * the same bits. */
__global__ __launch_bounds__(kBlock) void sur_conv_first(const float* __restrict__ x, long pstride,
                                                         long istride, float scale,
                                                         const float* __restrict__ params,
                                                         uint16_t* __restrict__ out, int nimg, int h,
                                                         int w) {
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, g = lane >> 4;
    const int tw = (w + 15) / 16;
    const long tile = (long)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (tile >= (long)nimg * h * tw) return;
    const int j = (int)(tile / ((long)h * tw));
    const int rem = (int)(tile - (long)j * h * tw);
    const int y = rem / tw, px = (rem - y * tw) * 16 + li;
    const float* xi = x + (long)(j >> 1) * pstride + (long)(j & 1) * istride;
    /* B fragment: lane (pixel li, group g) holds tap row g, columns px-1..px+1 */
    uint32_t v[3] = {0u, 0u, 0u};
    const int yy = y + g - 1;
    if (g < 3 && yy >= 0 && yy < h)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            const int xc = px + dx - 1;
            if (xc >= 0 && xc < w) v[dx] = x_pair(xi[(long)yy * w + xc] * scale);
        }
    const bf16x8 F = __builtin_bit_cast(bf16x8, make_uint4(v[0], v[1], v[2], 0u));
    uint16_t* o = out + (((long)j * h + y) * w + px) * 32;
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
        const bf16x8 A = sur_first_a(params, 16 * hf + li, g);
        f32x4 c;
#pragma unroll
        for (int e = 0; e < 4; ++e) c[e] = params[SurLayout::b1 + 16 * hf + 4 * g + e];
        /* D layout: lane holds channels 16 hf + 4g..4g+3 of pixel li */
        const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, F, c, 0, 0, 0);
        if (px < w)
            *(uint2*)(o + 16 * hf + 4 * g) =
                make_uint2(relu_bf16x2(pack_bf16(d[0], d[1])), relu_bf16x2(pack_bf16(d[2], d[3])));
    }
}

/* Layers 2-5 (and their input gradient): 3x3, 32 -> 32, NHWC BF16, MFMA.
 * A 16 x 16 pixel tile per workgroup; wavefront v owns tile rows 4v..4v+3.
 * Weights arrive as [tap][co][ci] BF16 (9 x 32 x 32), so each lane loads its
 * B fragments (k = ci, n = co) straight from global memory, where the 18 KiB
 * of a layer stay cached.  FWD: out = relu(conv(in, W) + b).  !FWD: out =
 * conv(in, W) * (mask > 0) (the input gradient through the ReLU of the layer
 * below, with W the flipped, transposed weights). */
template <bool FWD>
__global__ __launch_bounds__(kBlock) void sur_conv_mid(const uint16_t* __restrict__ in,
                                                       const uint16_t* __restrict__ wgt,
                                                       const float* __restrict__ bias,
                                                       const uint16_t* __restrict__ mask,
                                                       uint16_t* __restrict__ out, int h, int w) {
    __shared__ __align__(16) uint16_t tin[18 * 18 * 32];
    __shared__ __align__(16) uint16_t tout[16 * 16 * 32];
    const int img = blockIdx.z, x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    const long HW = (long)h * w;
    const uint16_t* src = in + (long)img * HW * 32;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, g = lane >> 4;
    bf16x8 B[9][2];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) B[t][nb] = *(const bf16x8*)&wgt[(t * 32 + nb * 16 + li) * 32 + 8 * g];
    /* input tile with a one-pixel zero halo, 16-byte chunks of 8 channels */
    for (int i = threadIdx.x; i < 18 * 18 * 4; i += blockDim.x) {
        const int px = i >> 2, part = i & 3;
        const int ly = px / 18, lx = px - ly * 18;
        const int gy = y0 - 1 + ly, gx = x0 - 1 + lx;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (gy >= 0 && gy < h && gx >= 0 && gx < w) v = ((const uint4*)(src + ((long)gy * w + gx) * 32))[part];
        ((uint4*)tin)[i] = v;
    }
    float b0 = 0.f, b1 = 0.f;
    if (FWD) {
        b0 = bias[li];
        b1 = bias[16 + li];
    }
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
        const int r = wave * 4 + rr;
        /* the forward sums start from the bias (the accumulator's initial
         * value); ReLU is taken on the BF16 result */
        f32x4 acc0 = {b0, b0, b0, b0}, acc1 = {b1, b1, b1, b1};
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int dy = t / 3, dx = t % 3;
            const bf16x8 A = *(const bf16x8*)&tin[((r + dy) * 18 + li + dx) * 32 + 8 * g];
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[t][0], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[t][1], acc1, 0, 0, 0);
        }
        /* D layout: lane holds pixel rows 4g..4g+3 of column li */
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int px = 4 * g + e;
            uint16_t v0 = f2bf(acc0[e]), v1 = f2bf(acc1[e]);
            if (FWD) {
                v0 = relu_bf16(v0);
                v1 = relu_bf16(v1);
            }
            tout[(r * 16 + px) * 32 + li] = v0;
            tout[(r * 16 + px) * 32 + 16 + li] = v1;
        }
    }
    __syncthreads();
    uint16_t* dst = out + (long)img * HW * 32;
    const uint16_t* msk = mask ? mask + (long)img * HW * 32 : nullptr;
    for (int i = threadIdx.x; i < 16 * 16 * 4; i += blockDim.x) {
        const int px = i >> 2, part = i & 3;
        const int gy = y0 + (px >> 4), gx = x0 + (px & 15);
        if (gy >= h || gx >= w) continue;
        const long o = ((long)gy * w + gx) * 32;
        uint4 v = ((const uint4*)tout)[i];
        if (!FWD) {
            /* keep a gradient where the activation below is > 0 (BF16 bits
             * 0x0001..0x7fff) */
            const uint4 m = ((const uint4*)(msk + o))[part];
            uint32_t* vv = (uint32_t*)&v;
            const uint32_t* mm = (const uint32_t*)&m;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t lo = mm[k] & 0xffffu, hi = mm[k] >> 16;
                const uint32_t keep = ((lo != 0u && lo < 0x8000u) ? 0x0000ffffu : 0u) |
                                      ((hi != 0u && hi < 0x8000u) ? 0xffff0000u : 0u);
                vv[k] &= keep;
            }
        }
        ((uint4*)(dst + o))[part] = v;
    }
}

/* Layer 6: 32 -> 1 channel, FP32 accumulation, no activation. */
__global__ __launch_bounds__(kBlock) void sur_conv_last(const uint16_t* __restrict__ a5,
                                                        const float* __restrict__ params,
                                                        float* __restrict__ out, int nimg, int h, int w) {
    __shared__ float wsh[32 * 9 + 1];
    for (int i = threadIdx.x; i < 32 * 9 + 1; i += blockDim.x) {
        const float v = params[SurLayout::w6 + i];
        wsh[i] = i < 32 * 9 ? bf2f(f2bf(v)) : v;
    }
    __syncthreads();
    const long HW = (long)h * w;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nimg * HW) return;
    const int j = (int)(p / HW);
    const long q = p - (long)j * HW;
    const int y = (int)(q / w), x = (int)(q - (long)y * w);
    float s = wsh[32 * 9];
    for (int t = 0; t < 9; ++t) {
        const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
        if (yy < 0 || yy >= h || xx < 0 || xx >= w) continue;
        const uint4* src = (const uint4*)(a5 + ((long)j * HW + (long)yy * w + xx) * 32);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint4 v = src[k];
            const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int ci = 8 * k + 2 * m;
                s += wsh[ci * 9 + t] * bf2f((uint16_t)(u[m] & 0xffffu));
                s += wsh[(ci + 1) * 9 + t] * bf2f((uint16_t)(u[m] >> 16));
            }
        }
    }
    out[p] = s;
}

/* All six layers of the forward pass in one kernel, for inference.  A
 * workgroup streams strips of kFuSW output columns of one image down their
 * rows (a persistent grid: each workgroup takes an even share of all strip
 * rows, as segments with six halo rows on either side); every step s ends
 * with one barrier.  The 3x3 layers 2-5 run on the
 * matrix cores: at step s layer l reads one row q = s - 2l + 3 of the layer
 * below from LDS and applies the three tap rows of its kernel to it at once:
 * row q is tap row 0 of output row q + 1 (whose sums start here, from the
 * bias), tap row 1 of output row q and tap row 2 of output row q - 1, which
 * is then complete and is written to the layer's ring.  So every row is read
 * from LDS once, three open output rows live in registers, and a layer's ring
 * holds two rows (the one being written and the one the next layer reads).
 * An output's taps are summed in the order t = 0..8 (three per step, dx
 * ascending), so the bits are those of sur_conv_mid.  A layer row is five
 * 16-pixel tiles.  Every wave works on the matrix cores: waves k and k + 4
 * share a SIMD and split the row of layer k + 2 (fu_has), and wave k + 4
 * also does a quarter of the rest, on the matrix cores too: the input row
 * s + 2 (into a 4-row ring, each value as its BF16 pair x_pair), layer-1
 * units of row s with the operands of sur_conv_first (so the same bits),
 * and output tile k of layer 6, whose instructions give the three tap rows'
 * sums of a pixel (three open output rows, like the 3x3 layers; its sums in
 * another order than sur_conv_last's).  The FP32 conversions and stores are
 * spread between the matrix-core instructions: a SIMD runs about two of them
 * per matrix-core instruction at no cost, while a wave of FP32 work beside a
 * matrix-core wave slows both.  Output row s - 2(l-1) of layer l completes at
 * step s; steps run from -2 to h + 9 (layer 6 completes row h-1).  Rows
 * outside the image read as zeros: rows -2, -1 and h of every layer are
 * written as zeros.  The halo shrinks by one column per layer (layer l
 * covers kFuSW + 2(6-l) columns; the five tiles of a row cover more, and
 * those extra columns only feed extra columns), and columns outside the
 * image are zeros, the zero padding of the layer-by-layer network. */
constexpr int kFuSW = 72;      /* output columns per strip */
constexpr int kFuPitch = 96;   /* ring row pitch (pixels): layer 1 computes six tiles */
constexpr int kFuXPitch = 112; /* input ring row pitch (BF16 pairs); rows 0-3 the ring, row 4 zeros */
constexpr int kFuXCols = 98;   /* input columns x0 - 6 .. x0 + 91 */
constexpr int kFuThreads = 512;
constexpr size_t kFuRingBytes = (size_t)5 * 2 * kFuPitch * 32 * 2;
constexpr size_t kFuXBytes = 2304; /* input rows, then the rings */
constexpr size_t kFuLds = kFuXBytes + kFuRingBytes;
static_assert(5 * kFuXPitch * 4 <= (int)kFuXBytes && kFuXBytes % 256 == 0, "input ring, rings on 256-byte boundaries");
static_assert(kFuPitch % 4 == 0 && kFuXCols <= kFuXPitch, "ring pitches");
static_assert((kFuSW + 8 + 15) / 16 == 5 && (kFuSW + 2 + 15) / 16 == 5, "five tiles per row of layers 2-5");
static_assert((kFuSW + 10 + 15) / 16 == 6 && kFuPitch >= 6 * 16 && kFuXCols >= 6 * 16 + 2, "six tiles per layer-1 row");

/* Element offset of channel c of pixel p within a ring row.  A pixel's 64
 * bytes hold its four 16-byte chunks of 8 channels in the order
 * chunk ^ ((p >> 1) & 3), so the 8 lanes of one 128-byte LDS access that read
 * one chunk of 8 consecutive pixels (a matrix-core fragment) fall into 8
 * different 16-byte bank groups.  Moving 16 pixels keeps the order. */
__device__ __forceinline__ int fu_pix(int p, int c) {
    return p * 32 + ((((c >> 3) ^ (p >> 1)) & 3) << 3) + (c & 7);
}
/* element offset of ring row `slot` (0, 1) of layer 1..5 */
__device__ __forceinline__ int fu_row(int layer, int slot) { return ((layer - 1) * 2 + slot) * kFuPitch * 32; }

/* Tiles of a layer row per wave: wave k (EXTRA false) has tiles 0 and 1
 * (both channel halves) and half 0 of tile 2, wave k + 4 (EXTRA) half 1 of
 * tile 2 and tiles 3 and 4; local tile lt is tile lt + fu_t0(EXTRA). */
__host__ __device__ constexpr int fu_t0(bool extra) { return extra ? 2 : 0; }
__host__ __device__ constexpr bool fu_has(bool extra, int lt, int hf) {
    return extra ? (lt != 0 || hf == 1) : (lt != 2 || hf == 0);
}

/* One wave of the fused kernel: its three tiles of layer k + 2; EXTRA
 * (waves 4-7) adds the input row, a share of layer 1 and a quarter of
 * layer 6. */
template <bool EXTRA>
__device__ __forceinline__ void fu_wave(uint32_t* __restrict__ xr, uint16_t* __restrict__ ring,
                                        const float* __restrict__ x, long pstride, long istride,
                                        float* __restrict__ out, float scale, const float* __restrict__ params,
                                        const uint16_t* __restrict__ wbf, int k, int lane, int tid, int nimg,
                                        int himg, int w) {
    constexpr int T0 = fu_t0(EXTRA);
    const int ml = 2 + k;
    const int li = lane & 15, g = lane >> 4;
    /* B fragments (k = ci, n = co) of both channel halves */
    const uint16_t* wl = wbf + (ml - 2) * 9216;
    bf16x8 B[2][9];
#pragma unroll
    for (int hf = 0; hf < 2; ++hf)
#pragma unroll
        for (int t = 0; t < 9; ++t) B[hf][t] = *(const bf16x8*)&wl[(t * 32 + hf * 16 + li) * 32 + 8 * g];
    /* biases of the channels this lane's outputs hold: 16 hf + 4 g + e */
    f32x4 b4[2];
#pragma unroll
    for (int hf = 0; hf < 2; ++hf)
#pragma unroll
        for (int e = 0; e < 4; ++e) b4[hf][e] = params[SurLayout::bl(ml) + 16 * hf + 4 * g + e];
    /* lane offsets: fragment of pixel li + dx (chunk g), outputs of pixel li */
    int rd[3], wr[2];
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) rd[dx] = fu_pix(li + dx, 8 * g);
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) wr[hf] = fu_pix(li, 16 * hf + 4 * g);
    /* layer 1: channel half k & 1 of tiles (k >> 1) + 2i, i < 3 */
    const int hf1 = k & 1, t1 = k >> 1;
    bf16x8 A1 = {};
    f32x4 c1 = {0.f, 0.f, 0.f, 0.f};
    if (EXTRA) {
        A1 = sur_first_a(params, 16 * hf1 + li, g);
#pragma unroll
        for (int e = 0; e < 4; ++e) c1[e] = params[SurLayout::b1 + 16 * hf1 + 4 * g + e];
    }
    /* layer 6 (not EXTRA): output tile k on the matrix cores, one instruction
     * per tap column dx: A6[dx] has M = tap row dy (rows 0-2 of 16) and
     * K = the 32 channels, so D[dy][p] = sum over dx and channels of the
     * dy-th tap row's products for pixel p, and lanes of group 0 hold the
     * three tap rows' sums of their pixel */
    bf16x8 A6[3];
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) {
        A6[dx] = bf16x8{};
        if (!EXTRA && li < 3)
#pragma unroll
            for (int e = 0; e < 8; ++e) A6[dx][e] = (__bf16)params[SurLayout::w6 + (8 * g + e) * 9 + 3 * li + dx];
    }
    const float b6 = params[SurLayout::b6];
    /* every set-up load has arrived before the segments (an empty asm that
     * reads the values makes the compiler wait for them here), so the loop
     * waits only for its own */
    auto ready = [](const auto& v) { asm volatile("" ::"v"(v)); };
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
        ready(b4[hf]);
#pragma unroll
        for (int t = 0; t < 9; ++t) ready(B[hf][t]);
    }
    ready(A1);
    ready(c1);
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) ready(A6[dx]);
    ready(b6);
    const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
    const uint16_t* in = ring + fu_row(ml - 1, 0);
    uint16_t* outr = ring + fu_row(ml, 0);
    uint16_t* ring1 = ring + fu_row(1, 0);
    const uint16_t* ring5 = ring + fu_row(5, 0);

    /* One segment: output rows olo .. ohi - 1 of the h rows of a strip at
     * column x0 that start at xi (input) and oj (output).  Rows outside
     * 0 .. h - 1 are zeros; the rows computed for the segment reach six rows
     * beyond the stored ones on either side (where the strip has them), the
     * depth of the network, so zeros that stand in for real rows never reach
     * a stored output. */
    auto segment = [&](const float* __restrict__ xi, float* __restrict__ oj, int x0, int h, int olo, int ohi) {
    /* column of ring position p: x0 - (6 - ml) + p */
    const int gx0 = x0 - (6 - ml);
    const bool edge = gx0 + 16 * T0 < 0 || gx0 + 16 * (T0 + 3) > w;
    unsigned keep = 0; /* bit lt: this lane's pixel of local tile lt is a column of the image */
#pragma unroll
    for (int lt = 0; lt < 3; ++lt) {
        const int gx = gx0 + 16 * (T0 + lt) + li;
        if (gx >= 0 && gx < w) keep |= 1u << lt;
    }

    /* ---- FP32 share (EXTRA): thread vt = 64 k + lane */
    const int vt = k * 64 + lane;
    /* input: thread vt < kFuXCols owns column x0 - 6 + vt; row s + 2 is
     * loaded when step s starts and stored when it ends (rows and columns
     * outside the image as zeros), by buffer loads whose row offset is a
     * scalar and whose column offset is a fixed lane register */
    const int xgx = x0 - 6 + vt;
    const bool xown = EXTRA && vt < kFuXCols;
    const bool xcol = xown && xgx >= 0 && xgx < w;
    const __amdgpu_buffer_rsrc_t xrs =
        __builtin_amdgcn_make_buffer_rsrc((void*)xi, (short)0, (int)(h * w * 4), 0x00020000);
    const int xoff = min(max(xgx, 0), w - 1) * 4;
    unsigned keep1 = 0; /* bit i: this lane's pixel of layer-1 tile t1 + 2i is a column of the image */
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int gx = x0 - 5 + 16 * (t1 + 2 * i) + li;
        if (gx >= 0 && gx < w) keep1 |= 1u << i;
    }
    const int gx6 = x0 + 16 * k + li;
    float a6[3] = {0.f, 0.f, 0.f};
    /* wave 0 also has output tile 4 (columns 64 .. kFuSW - 1 of the strip) */
    float a6b[3] = {0.f, 0.f, 0.f};

    f32x4 acc[3][2][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int hf = 0; hf < 2; ++hf)
#pragma unroll
            for (int lt = 0; lt < 3; ++lt) acc[a][hf][lt] = z4;

    auto step = [&](auto kk, int s) {
        constexpr int kn = decltype(kk)::value; /* slot of the row started now */
        constexpr int km = (kn + 2) % 3;        /* slot of the row started last step */
        constexpr int kc = (kn + 1) % 3;        /* slot of the row completed now */
        /* ---- FP32 share, reads first: input row s + 2 to LDS (row s + 3
         * issued), the layer-5 row q6 = s - 9 of layer 6 and the layer-1
         * fragments of row s: K = tap row g (lane group) x column dx x
         * (hi, lo) from the input ring; group 3 reads the ring's zero row */
        const int q6 = s - 9;
        const uint16_t* r5 = ring5 + (q6 & 1) * kFuPitch * 32 + k * 16 * 32;
        const uint32_t* xrow = xr + (g < 3 ? ((s + g - 1) & 3) : 4) * kFuXPitch + li;
        uint16_t* r1 = ring1 + (s & 1) * kFuPitch * 32;
        const bool live1 = s >= 0 && s < h;
        bf16x8 F6[3];
        f32x4 d6 = z4;
        bf16x8 F1[3];
        f32x4 d1[3];
        auto l6_read = [&](int dx) { F6[dx] = *(const bf16x8*)(r5 + rd[dx]); };
        auto l6_mfma = [&](int dx) { d6 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A6[dx], F6[dx], d6, 0, 0, 0); };
        auto l1_read = [&](int i) {
            const uint32_t* xp = xrow + 16 * (t1 + 2 * i);
            F1[i] = __builtin_bit_cast(bf16x8, make_uint4(xp[0], xp[1], xp[2], 0u));
        };
        float xv = 0.f;
        if (EXTRA && k < 2) /* the waves holding input columns */
            xv = __uint_as_float(
                __builtin_amdgcn_raw_buffer_load_b32(xrs, xoff, min(max(s + 2, 0), h - 1) * w * 4, 0));
        /* layer-1 unit i: sums from the bias; row s: rows -2, -1 and h..
         * as zeros, columns outside the image as zeros */
        auto l1_mfma = [&](int i) { d1[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, F1[i], c1, 0, 0, 0); };
        auto l1_store = [&](int i) {
            const bool kp = live1 && ((keep1 >> i) & 1u);
            *(uint2*)(r1 + wr[hf1] + (t1 + 2 * i) * 16 * 32) = make_uint2(kp ? relu_bf16x2(pack_bf16(d1[i][0], d1[i][1])) : 0u,
                                                kp ? relu_bf16x2(pack_bf16(d1[i][2], d1[i][3])) : 0u);
        };

        const int q = s - 2 * ml + 3;
        if (q >= -1 && q <= h) {
            /* row o = q - 1 completes */
            const int o = q - 1;
            const uint16_t* row = in + (q & 1) * kFuPitch * 32;
            uint16_t* orow = outr + (o & 1) * kFuPitch * 32;
            /* D layout: lane holds channels 16 hf + 4g..4g+3 of pixel li of
             * a tile: BF16, ReLU on the bits, one 8-byte store */
            auto store = [&](int lt) {
#pragma unroll
                for (int hf = 0; hf < 2; ++hf)
                    if (fu_has(EXTRA, lt, hf)) {
                        const f32x4 a = acc[kc][hf][lt];
                        *(uint2*)(orow + wr[hf] + (T0 + lt) * 16 * 32) =
                            make_uint2(relu_bf16x2(pack_bf16(a[0], a[1])), relu_bf16x2(pack_bf16(a[2], a[3])));
                    }
            };
            /* fragment of tap column i / 3, local tile i % 3; the next one
             * is read while the matrix cores work on the current one */
            auto frag = [&](int i) { return *(const bf16x8*)(row + rd[i / 3] + (T0 + i % 3) * 16 * 32); };
            bf16x8 Fn = frag(0);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                const int dx = i / 3, lt = i % 3;
                const bf16x8 F = Fn;
                if (i + 1 < 9) Fn = frag(i + 1);
                /* weights as the first operand: D = W^T X^T, so a lane's
                 * four outputs are 4 channels of one pixel; taps dx (new
                 * row), 3 + dx (open row), 6 + dx (completing row) */
#pragma unroll
                for (int hf = 0; hf < 2; ++hf)
                    if (fu_has(EXTRA, lt, hf)) {
                        acc[kn][hf][lt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            B[hf][dx], F, dx == 0 ? b4[hf] : acc[kn][hf][lt], 0, 0, 0);
                        acc[km][hf][lt] =
                            __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[hf][3 + dx], F, acc[km][hf][lt], 0, 0, 0);
                        acc[kc][hf][lt] =
                            __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[hf][6 + dx], F, acc[kc][hf][lt], 0, 0, 0);
                    }
                /* between the matrix-core instructions: the conversions
                 * and stores of tile lt - 1 of row o (complete), and
                 * layer 1 (EXTRA waves) or layer 6 (the others): reads,
                 * instructions and stores three fragments apart */
                if (dx == 2 && lt > 0) store(lt - 1);
                if (EXTRA) {
                    if (i <= 2)
                        l1_read(i);
                    else if (i <= 5)
                        l1_mfma(i - 3);
                    else
                        l1_store(i - 6);
                } else {
                    if (i <= 2)
                        l6_read(i);
                    else if (i <= 5)
                        l6_mfma(i - 3);
                }
                __builtin_amdgcn_sched_group_barrier(0x100, 4, 0); /* the reads */
#pragma unroll
                for (int m = 0; m < 6; ++m) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0); /* one MFMA */
                    __builtin_amdgcn_sched_group_barrier(0x002, 2, 0); /* two VALU */
                }
                __builtin_amdgcn_sched_group_barrier(0x200, 3, 0); /* the stores */
                __builtin_amdgcn_sched_barrier(0);
            }
            store(2);
            /* rows -2, -1 as zeros (the next layer reads row -1 as padding
             * and reads neither slot at this step); columns outside the
             * image as zeros */
            if (o < 0) {
#pragma unroll
                for (int lt = 0; lt < 3; ++lt)
#pragma unroll
                    for (int hf = 0; hf < 2; ++hf)
                        if (fu_has(EXTRA, lt, hf))
                            *(uint2*)(orow + wr[hf] + (T0 + lt) * 16 * 32) = make_uint2(0u, 0u);
            } else if (edge) {
#pragma unroll
                for (int lt = 0; lt < 3; ++lt)
                    if (!((keep >> lt) & 1u))
#pragma unroll
                        for (int hf = 0; hf < 2; ++hf)
                            if (fu_has(EXTRA, lt, hf))
                                *(uint2*)(orow + wr[hf] + (T0 + lt) * 16 * 32) = make_uint2(0u, 0u);
            }
        } else {
            if (q == h + 1) {
                /* row h as zeros: the next layer reads it as padding (its
                 * slot still holds row h - 2) */
                uint16_t* orow = outr + (h & 1) * kFuPitch * 32;
#pragma unroll
                for (int lt = 0; lt < 3; ++lt)
#pragma unroll
                    for (int hf = 0; hf < 2; ++hf)
                        if (fu_has(EXTRA, lt, hf))
                            *(uint2*)(orow + wr[hf] + (T0 + lt) * 16 * 32) = make_uint2(0u, 0u);
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                if (EXTRA) {
                    l1_read(i);
                    l1_mfma(i);
                    l1_store(i);
                } else {
                    l6_read(i);
                    l6_mfma(i);
                }
            }
        }
        if (!EXTRA) {
            /* layer 6: row q6 is tap row 0 of output row q6 + 1 (started),
             * 1 of row q6 and 2 of row q6 - 1 (complete) */
            a6[kn] = d6[0];
            a6[km] += d6[1];
            const float v = a6[kc] + d6[2];
            const int o6 = q6 - 1;
            if (g == 0 && o6 >= olo && o6 < ohi && gx6 < w) oj[(long)o6 * w + gx6] = v + b6;
            if (k == 0) {
                const uint16_t* r5b = ring5 + (q6 & 1) * kFuPitch * 32 + 4 * 16 * 32;
                f32x4 e = z4;
#pragma unroll
                for (int dx = 0; dx < 3; ++dx)
                    e = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A6[dx], *(const bf16x8*)(r5b + rd[dx]), e, 0, 0, 0);
                a6b[kn] = e[0];
                a6b[km] += e[1];
                const float vb = a6b[kc] + e[2];
                if (g == 0 && li < kFuSW - 64 && o6 >= olo && o6 < ohi && gx6 + 64 < w)
                    oj[(long)o6 * w + gx6 + 64] = vb + b6;
            }
        } else if (k < 2) {
            /* input row s + 2, read by layer 1 from the next step on */
            const int r = s + 2;
            const float v = (xcol && r >= 0 && r < h) ? xv * scale : 0.f;
            if (xown) xr[(r & 3) * kFuXPitch + vt] = x_pair(v);
        }
        lds_barrier();
    };
    /* the rings start as zeros */
    for (int i = tid; i < (int)(kFuLds / 16); i += kFuThreads) ((uint4*)xr)[i] = make_uint4(0, 0, 0, 0);
    __syncthreads();
    for (int u = 0; u < h + 12; u += 3) {
        step(std::integral_constant<int, 0>{}, u - 2);
        step(std::integral_constant<int, 1>{}, u - 1);
        step(std::integral_constant<int, 2>{}, u);
    }
    };

    /* this workgroup's rows of the strips (image j, strip sx) laid end to
     * end, strip after strip: an even share of all of them, so every
     * workgroup ends at about the same time */
    const int S = (w + kFuSW - 1) / kFuSW;
    const long R = (long)nimg * S * himg;
    const long L = (R + gridDim.x - 1) / gridDim.x;
    long g0 = (long)blockIdx.x * L;
    const long g1 = min(R, g0 + L);
    while (g0 < g1) {
        const long ks = g0 / himg;
        const int r0 = (int)(g0 - ks * himg);
        const int r1 = (int)min((long)himg, r0 + (g1 - g0));
        const int j = (int)(ks / S), sx = (int)(ks - (long)j * S);
        const int top = max(0, r0 - 6), bot = min(himg, r1 + 6);
        segment(x + (long)(j >> 1) * pstride + (long)(j & 1) * istride + (long)top * w,
                out + ((long)j * himg + top) * w, sx * kFuSW, bot - top, r0 - top, r1 - top);
        g0 += r1 - r0;
    }
}

/* A persistent grid: each workgroup takes an even share of the rows of all
 * strips (fu_wave). */
__global__ __launch_bounds__(kFuThreads) void sur_fused_kernel(const float* __restrict__ x, long pstride,
                                                               long istride, float scale,
                                                               const float* __restrict__ params,
                                                               const uint16_t* __restrict__ wbf,
                                                               float* __restrict__ out, int nimg, int h, int w) {
    extern __shared__ __align__(16) uint8_t fu_smem[];
    uint32_t* xr = (uint32_t*)fu_smem;                 /* [5][kFuXPitch] */
    uint16_t* ring = (uint16_t*)(fu_smem + kFuXBytes); /* [layer 1..5][2][kFuPitch][32] */
    const int tid = threadIdx.x;
    /* both kinds of wave run the same steps, one barrier each */
    if (tid < 256)
        fu_wave<false>(xr, ring, x, pstride, istride, out, scale, params, wbf, tid >> 6, tid & 63, tid, nimg, h, w);
    else
        fu_wave<true>(xr, ring, x, pstride, istride, out, scale, params, wbf, (tid >> 6) - 4, tid & 63, tid, nimg, h,
                      w);
}

/* ---- backward ---------------------------------------------------------- */

/* dpred = 2 w (pred - target) / P and per-block FP64 partial sums of the
 * weighted squared error.  eps < 0: w = 1 (mean squared error); eps >= 0:
 * w = max(target, target of the FSIM pair partner) + eps, where images 2i
 * and 2i + 1 (pd pixels each) are a pair: FSIM weights a pixel's similarity
 * by the larger of the two phase congruencies, so errors where both are
 * small move the score little. */
__global__ __launch_bounds__(kBlock) void sur_loss(const float* __restrict__ pred,
                                                   const float* __restrict__ target,
                                                   float* __restrict__ dpred, double* __restrict__ part,
                                                   long P, long pd, float eps) {
    __shared__ double sh[kBlock];
    double acc = 0;
    const float sc = 2.0f / (float)P;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
        const float d = pred[p] - target[p];
        float wt = 1.0f;
        if (eps >= 0.0f) {
            const long j = p / pd;
            wt = fmaxf(target[p], target[p + ((j ^ 1) - j) * pd]) + eps;
        }
        dpred[p] = sc * wt * d;
        acc += (double)wt * (double)d * (double)d;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void sur_loss_final(const double* part, int n, long P, double* loss) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0;
        for (int i = 0; i < n; ++i) s += part[i];
        *loss = s / (double)P;
    }
}

/* Input gradient of layer 6 through the ReLU of layer 5:
 * dz5[p][ci] = (a5[p][ci] > 0) * sum_tap dpred[p - tap] * w6[ci][tap]. */
__global__ __launch_bounds__(kBlock) void sur_dlast(const float* __restrict__ dpred,
                                                    const float* __restrict__ params,
                                                    const uint16_t* __restrict__ a5,
                                                    uint16_t* __restrict__ dz, int nimg, int h, int w) {
    __shared__ float wsh[32 * 9];
    for (int i = threadIdx.x; i < 32 * 9; i += blockDim.x) wsh[i] = bf2f(f2bf(params[SurLayout::w6 + i]));
    __syncthreads();
    const long HW = (long)h * w;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nimg * HW) return;
    const int j = (int)(p / HW);
    const long q = p - (long)j * HW;
    const int y = (int)(q / w), x = (int)(q - (long)y * w);
    float d[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        /* output pixel o = p - offset(t) contributes w[t] to input pixel p */
        const int yy = y - (t / 3 - 1), xx = x - (t % 3 - 1);
        d[t] = (yy >= 0 && yy < h && xx >= 0 && xx < w) ? dpred[(long)j * HW + (long)yy * w + xx] : 0.f;
    }
    const uint16_t* a = a5 + p * 32;
    uint16_t* o = dz + p * 32;
    for (int ci = 0; ci < 32; ++ci) {
        float s = 0.f;
#pragma unroll
        for (int t = 0; t < 9; ++t) s += d[t] * wsh[ci * 9 + t];
        o[ci] = bf2f(a[ci]) > 0.f ? f2bf(s) : (uint16_t)0;
    }
}

/* Weight gradient of a 32 -> 32 layer: dW[co][ci][tap] = sum_p dz[p][co] *
 * a[p + offset(tap)][ci] and db[co] = sum_p dz[p][co], on the matrix cores.
 * A workgroup of 4 waves walks 16 x 16 tiles (grid-stride).  The sum runs
 * over pixels, so the operands need pixels contiguous: each tile is staged
 * in LDS channel-major, dz as [co][256 pixels] and the activation halo as
 * three copies [dx][ci][18 rows][16 columns] shifted by the tap column, so
 * that every fragment (8 consecutive pixels of one channel) is one aligned
 * 16-byte read.  One instruction per 32 pixels (two tile rows), tap and
 * 16 x 16 (co, ci) block: D[co][ci] += dz^T a_tap.  Wave w owns taps w,
 * w + 4, w + 8; waves 1 and 2 also sum db (one instruction with a B of
 * ones per co half).  Products of BF16 values are exact in FP32; the sums
 * run in a fixed order (instruction, 32-pixel block, tile), so the partial
 * rows are deterministic. */
constexpr int kDwTaLds = 3 * 32 * 18 * 16;       /* elements of the shifted halo copies */
constexpr size_t kDwLds = (size_t)(32 * 256 + kDwTaLds) * 2;
__global__ __launch_bounds__(256) void sur_dw_mid(const uint16_t* __restrict__ dz,
                                                  const uint16_t* __restrict__ a,
                                                  float* __restrict__ part, int nimg, int h, int w) {
    extern __shared__ __align__(16) uint16_t dw_smem[];
    uint16_t* tdz = dw_smem;            /* [32 co][256 px] */
    uint16_t* ta = dw_smem + 32 * 256;  /* [3 dx][32 ci][18][16] */
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int nt = wave == 0 ? 3 : 2; /* taps wave, wave + 4 (, wave + 8) */
    f32x4 acc[3][2][2];
    f32x4 accb = {0.f, 0.f, 0.f, 0.f};
    const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int d = 0; d < 2; ++d) acc[i][c][d] = z4;
    bf16x8 ones;
#pragma unroll
    for (int e = 0; e < 8; ++e) ones[e] = (__bf16)1.0f;
    const long HW = (long)h * w;
    const int tx = (w + 15) / 16, ty = (h + 15) / 16;
    const long ntiles = (long)nimg * tx * ty;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int img = (int)(tile / (tx * ty));
        const int rem = (int)(tile - (long)img * tx * ty);
        const int y0 = (rem / tx) * 16, x0 = (rem % tx) * 16;
        const uint16_t* dzi = dz + (long)img * HW * 32;
        const uint16_t* ai = a + (long)img * HW * 32;
        __syncthreads();
        /* dz: pixel pairs (p, p + 1) x 8-channel chunks; one 4-byte store
         * per channel holds the pair */
        for (int i = tid; i < 128 * 4; i += blockDim.x) {
            const int pp = i >> 2, ch = i & 3, p = 2 * pp;
            const int gy = y0 + (p >> 4), gx = x0 + (p & 15);
            uint4 v0 = make_uint4(0, 0, 0, 0), v1 = make_uint4(0, 0, 0, 0);
            if (gy < h && gx < w) v0 = ((const uint4*)(dzi + ((long)gy * w + gx) * 32))[ch];
            if (gy < h && gx + 1 < w) v1 = ((const uint4*)(dzi + ((long)gy * w + gx + 1) * 32))[ch];
            const uint32_t u0[4] = {v0.x, v0.y, v0.z, v0.w}, u1[4] = {v1.x, v1.y, v1.z, v1.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int co = 8 * ch + 2 * k;
                *(uint32_t*)&tdz[co * 256 + p] = (u0[k] & 0xffffu) | (u1[k] << 16);
                *(uint32_t*)&tdz[(co + 1) * 256 + p] = (u0[k] >> 16) | (u1[k] & 0xffff0000u);
            }
        }
        /* activation halo: copy dx holds column c at c + dx - 1 of the tile's
         * columns; column pairs (c, c + 1) x rows x 8-channel chunks */
        for (int i = tid; i < 3 * 18 * 8 * 4; i += blockDim.x) {
            const int ch = i & 3, cp = (i >> 2) & 7, rest = i >> 5;
            const int r = rest % 18, dx = rest / 18;
            const int c = 2 * cp;
            const int gy = y0 - 1 + r, gx = x0 - 1 + c + dx;
            uint4 v0 = make_uint4(0, 0, 0, 0), v1 = make_uint4(0, 0, 0, 0);
            if (gy >= 0 && gy < h) {
                if (gx >= 0 && gx < w) v0 = ((const uint4*)(ai + ((long)gy * w + gx) * 32))[ch];
                if (gx + 1 >= 0 && gx + 1 < w) v1 = ((const uint4*)(ai + ((long)gy * w + gx + 1) * 32))[ch];
            }
            const uint32_t u0[4] = {v0.x, v0.y, v0.z, v0.w}, u1[4] = {v1.x, v1.y, v1.z, v1.w};
            uint16_t* base = ta + (dx * 32) * 18 * 16 + r * 16 + c;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int ci = 8 * ch + 2 * k;
                *(uint32_t*)&base[ci * 18 * 16] = (u0[k] & 0xffffu) | (u1[k] << 16);
                *(uint32_t*)&base[(ci + 1) * 18 * 16] = (u0[k] >> 16) | (u1[k] & 0xffff0000u);
            }
        }
        __syncthreads();
        /* 32 pixels per step: tile rows 2kb, 2kb + 1; lane group g holds
         * pixels 32 kb + 8g .. + 7 (row 2kb + g / 2, columns 8 (g & 1) ..) */
        for (int kb = 0; kb < 8; ++kb) {
            bf16x8 A[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) A[c] = *(const bf16x8*)&tdz[(16 * c + li) * 256 + 32 * kb + 8 * g];
            const int ly = 2 * kb + (g >> 1), lx = 8 * (g & 1);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                if (i < nt) {
                    const int t = wave + 4 * i, dy = t / 3, dx = t - 3 * dy;
#pragma unroll
                    for (int d = 0; d < 2; ++d) {
                        const bf16x8 Bf = *(const bf16x8*)&ta[((dx * 32) + 16 * d + li) * 18 * 16 + (ly + dy) * 16 + lx];
#pragma unroll
                        for (int c = 0; c < 2; ++c)
                            acc[i][c][d] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[c], Bf, acc[i][c][d], 0, 0, 0);
                    }
                }
            }
            if (wave == 1 || wave == 2)
                accb = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[wave - 1], ones, accb, 0, 0, 0);
        }
    }
    /* partial row: [32 co][32 ci][9 taps] weights then [32] biases; D layout:
     * lane holds D[4g + e][li] of its 16 x 16 block */
    float* pr = part + (long)blockIdx.x * (32 * 32 * 9 + 32);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (i < nt) {
            const int t = wave + 4 * i;
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int d = 0; d < 2; ++d)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        pr[((16 * c + 4 * g + e) * 32 + 16 * d + li) * 9 + t] = acc[i][c][d][e];
        }
    }
    if ((wave == 1 || wave == 2) && li == 0)
#pragma unroll
        for (int e = 0; e < 4; ++e) pr[32 * 32 * 9 + 16 * (wave - 1) + 4 * g + e] = accb[e];
}

/* Weight gradients of layer 1 (1 -> 32) and layer 6 (32 -> 1) and their
 * biases.  Thread (lane q, channel c): c = tid % 32, pixels p = q, q + 8, ...
 * of the workgroup's grid-stride share; 9 + 9 FP32 accumulators per thread,
 * folded over the 8 lanes in LDS into one partial row per workgroup. */
__global__ __launch_bounds__(kBlock) void sur_dw_edge(const float* __restrict__ x, long pstride,
                                                      long istride, float scale,
                                                      const uint16_t* __restrict__ dz1,
                                                      const uint16_t* __restrict__ a5,
                                                      const float* __restrict__ dpred,
                                                      float* __restrict__ part1,
                                                      float* __restrict__ part6, int nimg, int h,
                                                      int w) {
    __shared__ float red[8][32][20];
    const int c = threadIdx.x & 31, q = threadIdx.x >> 5;
    const long HW = (long)h * w, P = (long)nimg * HW;
    float g1[9], g6[9], gb1 = 0.f, gb6 = 0.f;
#pragma unroll
    for (int t = 0; t < 9; ++t) g1[t] = g6[t] = 0.f;
    for (long p = (long)blockIdx.x * 8 + q; p < P; p += (long)gridDim.x * 8) {
        const int j = (int)(p / HW);
        const long r = p - (long)j * HW;
        const int y = (int)(r / w), xx = (int)(r - (long)y * w);
        const float* xi = x + (long)(j >> 1) * pstride + (long)(j & 1) * istride;
        const float d1 = bf2f(dz1[p * 32 + c]);
        const float d6 = dpred[p];
        gb1 += d1;
        gb6 += d6;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int yy = y + t / 3 - 1, xc = xx + t % 3 - 1;
            if (yy < 0 || yy >= h || xc < 0 || xc >= w) continue;
            g1[t] += d1 * (xi[(long)yy * w + xc] * scale);
            g6[t] += d6 * bf2f(a5[((long)j * HW + (long)yy * w + xc) * 32 + c]);
        }
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        red[q][c][t] = g1[t];
        red[q][c][9 + t] = g6[t];
    }
    red[q][c][18] = gb1;
    red[q][c][19] = gb6;
    __syncthreads();
    if (q == 0) {
        float s[20];
#pragma unroll
        for (int k = 0; k < 20; ++k) s[k] = 0.f;
        for (int l = 0; l < 8; ++l)
#pragma unroll
            for (int k = 0; k < 20; ++k) s[k] += red[l][c][k];
        float* p1 = part1 + (long)blockIdx.x * (288 + 32);
        float* p6 = part6 + (long)blockIdx.x * (288 + 1);
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            p1[c * 9 + t] = s[t];
            p6[c * 9 + t] = s[9 + t];
        }
        p1[288 + c] = s[18];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        /* every channel lane saw the same dpred values: take channel 0's sum */
        float b6 = 0.f;
        for (int l = 0; l < 8; ++l) b6 += red[l][0][19];
        part6[(long)blockIdx.x * (288 + 1) + 288] = b6;
    }
}

/* grad[dst + k] = sum_b part[b * stride + k], fixed order */
__global__ __launch_bounds__(kBlock) void sur_reduce(const float* __restrict__ part, int nb,
                                                     int count, int stride, float* __restrict__ grad,
                                                     int dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    double s = 0;
    for (int b = 0; b < nb; ++b) s += (double)part[(long)b * stride + k];
    grad[dst + k] = (float)s;
}

/* Adam (Kingma & Ba) with FP32 master weights. */
__global__ __launch_bounds__(kBlock) void sur_adam(float* __restrict__ w, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, int n,
                                                   float lr, float b1, float b2, float eps, float c1,
                                                   float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

/* BF16 copies of the mid-layer weights in the matrix-core B layout
 * [tap][n][k]: forward W[tap][co][ci], and for the input gradient the flipped
 * transpose Wt[tap'][ci][co] = W[co][ci][8 - tap']. */
__global__ __launch_bounds__(kBlock) void sur_cast(const float* __restrict__ master,
                                                   uint16_t* __restrict__ wbf, uint16_t* __restrict__ wbt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4 * 9216) return;
    const int l = i / 9216, r = i - l * 9216;
    const int co = r / 288, r2 = r - co * 288, ci = r2 / 9, t = r2 - ci * 9;
    const uint16_t b = f2bf(master[SurLayout::wl(l + 2) + r]);
    wbf[l * 9216 + (t * 32 + co) * 32 + ci] = b;
    wbt[l * 9216 + ((8 - t) * 32 + ci) * 32 + co] = b;
}

/* ---- training data ----------------------------------------------------- */

__device__ __forceinline__ uint64_t sm64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float u01(uint64_t z) { return ((float)(sm64(z) >> 40) + 0.5f) * (1.0f / 16777216.0f); }

/* white N(0,1) noise planes (Box-Muller on counter-based uniforms) */
__global__ __launch_bounds__(kBlock) void sur_noise(float2* planes, long per, int n, uint64_t seed) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per * n) return;
    const uint64_t base = seed * 0x632BE59BD9B4E019ull + (uint64_t)i * 2;
    const float u1 = u01(base), u2 = u01(base + 1);
    planes[i] = make_float2(sqrtf(-2.f * logf(u1)) * cosf(6.2831853f * u2), 0.f);
}

/* multiply spectra by 1/|f| on the fftfreq grid (DC := 0): power ~ 1/f^2 */
__global__ __launch_bounds__(kBlock) void sur_pink(float2* planes, int s, int n) {
    const long per = (long)s * s;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per * n) return;
    const long q = i % per;
    const int y = (int)(q / s), x = (int)(q % s);
    const float fy = (float)(y < (s + 1) / 2 ? y : y - s) / s, fx = (float)(x < (s + 1) / 2 ? x : x - s) / s;
    const float f = sqrtf(fx * fx + fy * fy);
    planes[i] = f > 0.f ? cscale(planes[i], 1.0f / f) : make_float2(0.f, 0.f);
}

/* per patch: min-max rescale to [0,255]; half of the patches get a seeded
 * contrast change, additive noise and 16-level quantisation; u8 out */
__global__ __launch_bounds__(1024) void sur_to_u8(const float2* planes, uint8_t* img, int s, int n,
                                                  uint64_t seed) {
    __shared__ float smin[1024], smax[1024];
    const long per = (long)s * s;
    const int k = blockIdx.x;
    const float2* p = planes + (long)k * per;
    float lo = 3e38f, hi = -3e38f;
    for (long i = threadIdx.x; i < per; i += blockDim.x) {
        lo = fminf(lo, p[i].x);
        hi = fmaxf(hi, p[i].x);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + st]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + st]);
        }
        __syncthreads();
    }
    lo = smin[0];
    hi = smax[0];
    const float sc = hi > lo ? 255.f / (hi - lo) : 0.f;
    const uint64_t ks = seed * 0x9E3779B97F4A7C15ull + (uint64_t)k * 0xD1B54A32D192ED03ull;
    const bool distort = (sm64(ks) & 1ull) != 0;
    const float contrast = 0.7f + 0.6f * u01(ks + 11), sigma = 2.f + 18.f * u01(ks + 13);
    for (long i = threadIdx.x; i < per; i += blockDim.x) {
        float v = (p[i].x - lo) * sc;
        if (distort) {
            const uint64_t z = ks + 0x100000ull + (uint64_t)i * 2;
            const float nz = sqrtf(-2.f * logf(u01(z))) * cosf(6.2831853f * u01(z + 1));
            v = 127.5f + contrast * (v - 127.5f) + sigma * nz;
            v = 17.f * rintf(v / 17.f);
        }
        img[(long)k * per + i] = (uint8_t)fminf(fmaxf(rintf(v), 0.f), 255.f);
    }
}

__global__ __launch_bounds__(kBlock) void sur_gather_y(const float* yiq, long N, int npairs, float* x) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * N * npairs) return;
    const long j = i / N, q = i - j * N;
    x[i] = yiq[(j >> 1) * 6 * N + (j & 1) * N + q] * (1.0f / 255.0f);
}

__global__ __launch_bounds__(kBlock) void sur_gather_pc(const float* pc, long N, int npairs, float* t) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * N * npairs) return;
    const float v = pc[i];
    t[i] = isfinite(v) ? v : 0.f; /* flat patches: the reference's 0/0 */
}

/* Spot check: block (t, im, j) compares the network PC of image im of pair
 * j * every with the exact PC on one 32 x 32 tile at a seeded position and
 * writes the tile's max-abs difference (NaN exact values, the reference's
 * 0/0 on flat regions, are skipped). */
constexpr int kSpotTiles = 4, kSpotSize = 32;

__global__ __launch_bounds__(kBlock) void sur_spot_kernel(const float* __restrict__ pc_net,
                                                          const float* __restrict__ pc_exact, int wd,
                                                          int hd, int every, uint64_t seed,
                                                          float* __restrict__ err) {
    __shared__ float red[kBlock / 64];
    const int t = blockIdx.x, im = blockIdx.y, j = blockIdx.z;
    const long N = (long)wd * hd;
    const int tw = min(kSpotSize, wd), th = min(kSpotSize, hd);
    const uint64_t z = sm64(seed ^ ((uint64_t)j * 8 + (uint64_t)im * 4 + (uint64_t)t) * 0x9E3779B97F4A7C15ull);
    const int x0 = (int)((z & 0xffffffffull) % (uint64_t)(wd - tw + 1));
    const int y0 = (int)((z >> 32) % (uint64_t)(hd - th + 1));
    const float* a = pc_net + ((long)j * every * 2 + im) * N;
    const float* b = pc_exact + ((long)j * 2 + im) * N;
    float m = 0.f;
    for (int i = threadIdx.x; i < tw * th; i += blockDim.x) {
        const long q = (long)(y0 + i / tw) * wd + x0 + i % tw;
        const float d = fabsf(a[q] - b[q]);
        if (d == d) m = fmaxf(m, d);
    }
    for (int d = 32; d > 0; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.f;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) r = fmaxf(r, red[w]);
        err[((long)j * 2 + im) * kSpotTiles + t] = r;
    }
}

/* ---- host ---------------------------------------------------------------- */

Surrogate::Surrogate(Ctx& c, uint64_t seed) : c_(c) {
    const int n = SurLayout::total;
    IQM_HIP(hipMalloc(&m.master, sizeof(float) * n));
    IQM_HIP(hipMalloc(&m.grad, sizeof(float) * n));
    IQM_HIP(hipMalloc(&m.adam_m, sizeof(float) * n));
    IQM_HIP(hipMalloc(&m.adam_v, sizeof(float) * n));
    IQM_HIP(hipMalloc(&m.wbf, sizeof(uint16_t) * 4 * 9216));
    IQM_HIP(hipMalloc(&m.wbf_t, sizeof(uint16_t) * 4 * 9216));
    IQM_HIP(hipMalloc(&d_loss_, sizeof(double) * 1025));
    m.seed = seed;
    /* He-uniform initialisation from a SplitMix64 stream; zero biases */
    std::vector<float> p(n, 0.f);
    uint64_t st = seed ? seed : 1;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((double)(z >> 11) + 0.5) / 9007199254740992.0;
    };
    auto fill = [&](int off, int count, int fan_in) {
        const double lim = std::sqrt(6.0 / fan_in);
        for (int i = 0; i < count; ++i) p[off + i] = (float)((2.0 * next() - 1.0) * lim);
    };
    fill(SurLayout::w1, 32 * 9, 9);
    for (int l = 2; l <= 5; ++l) fill(SurLayout::wl(l), 9216, 288);
    fill(SurLayout::w6, 32 * 9, 288);
    set_params(p.data());
    IQM_HIP(hipMemsetAsync(m.adam_m, 0, sizeof(float) * n, c_.stream()));
    IQM_HIP(hipMemsetAsync(m.adam_v, 0, sizeof(float) * n, c_.stream()));
    IQM_HIP(hipMemsetAsync(m.grad, 0, sizeof(float) * n, c_.stream()));
    c_.sync();
}

Surrogate::~Surrogate() {
    (void)hipStreamSynchronize(c_.stream());
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second.exec);
    for (void* q : {(void*)m.master, (void*)m.grad, (void*)m.adam_m, (void*)m.adam_v, (void*)m.wbf,
                    (void*)m.wbf_t, (void*)d_loss_, (void*)acts_.x, (void*)acts_.pred,
                    (void*)acts_.dpred, (void*)acts_.dz[0], (void*)acts_.dz[1], (void*)acts_.partial,
                    (void*)spot_pc_, (void*)spot_err_, (void*)d_out_, (void*)d_flags_})
        if (q) (void)hipFree(q);
    for (auto* a : acts_.a)
        if (a) (void)hipFree(a);
}

void Surrogate::refresh_bf16() {
    c_.launch("sur_cast", sur_cast, dim3(ceil_div(4 * 9216, kBlock)), dim3(kBlock), 0,
              (const float*)m.master, m.wbf, m.wbf_t);
}

/* Host copies are ordered on the context's (non-blocking) stream. */
void Surrogate::set_params(const float* host) {
    IQM_HIP(hipMemcpyAsync(m.master, host, sizeof(float) * SurLayout::total, hipMemcpyHostToDevice, c_.stream()));
    refresh_bf16();
    c_.sync();
}
void Surrogate::get_params(float* host) const {
    IQM_HIP(hipMemcpyAsync(host, m.master, sizeof(float) * SurLayout::total, hipMemcpyDeviceToHost, c_.stream()));
    IQM_HIP(hipStreamSynchronize(c_.stream()));
}
void Surrogate::get_grad(float* host) const {
    IQM_HIP(hipMemcpyAsync(host, m.grad, sizeof(float) * SurLayout::total, hipMemcpyDeviceToHost, c_.stream()));
    IQM_HIP(hipStreamSynchronize(c_.stream()));
}

void Surrogate::ensure_acts(int nimg, int h, int w) {
    if (acts_.nimg >= nimg && acts_.h == h && acts_.w == w) return;
    c_.sync();
    for (void* q : {(void*)acts_.x, (void*)acts_.pred, (void*)acts_.dpred, (void*)acts_.dz[0],
                    (void*)acts_.dz[1], (void*)acts_.partial})
        if (q) (void)hipFree(q);
    for (auto*& a : acts_.a)
        if (a) {
            (void)hipFree(a);
            a = nullptr;
        }
    const long P = (long)nimg * h * w;
    acts_.nimg = nimg;
    acts_.h = h;
    acts_.w = w;
    IQM_HIP(hipMalloc(&acts_.x, sizeof(float) * P));
    IQM_HIP(hipMalloc(&acts_.pred, sizeof(float) * P));
    IQM_HIP(hipMalloc(&acts_.dpred, sizeof(float) * P));
    for (auto*& a : acts_.a) IQM_HIP(hipMalloc(&a, sizeof(uint16_t) * 32 * P));
    IQM_HIP(hipMalloc(&acts_.dz[0], sizeof(uint16_t) * 32 * P));
    IQM_HIP(hipMalloc(&acts_.dz[1], sizeof(uint16_t) * 32 * P));
    acts_.npart = 512;
    IQM_HIP(hipMalloc(&acts_.partial, sizeof(float) * (size_t)acts_.npart * (9216 + 32 + 2 * 320)));
}

static void launch_forward(Ctx& c, const SurModel& m, const float* x, long ps, long is, float scale,
                           uint16_t* const a[5], float* out, int nimg, int h, int w) {
    const long P = (long)nimg * h * w;
    const long tiles1 = (long)nimg * h * ceil_div(w, 16);
    c.launch("sur_conv1", sur_conv_first, dim3(ceil_div(tiles1, kBlock / 64)), dim3(kBlock), 0, x, ps, is, scale,
             (const float*)m.master, a[0], nimg, h, w);
    const dim3 g(ceil_div(w, 16), ceil_div(h, 16), nimg);
    for (int l = 2; l <= 5; ++l)
        c.launch("sur_conv_mid", sur_conv_mid<true>, g, dim3(kBlock), 0, (const uint16_t*)a[l - 2],
                 (const uint16_t*)(m.wbf + (l - 2) * 9216), (const float*)(m.master + SurLayout::bl(l)),
                 (const uint16_t*)nullptr, a[l - 1], h, w);
    c.launch("sur_conv6", sur_conv_last, dim3(ceil_div(P, kBlock)), dim3(kBlock), 0, (const uint16_t*)a[4],
             (const float*)m.master, out, nimg, h, w);
}

void Surrogate::forward(int nimg, int h, int w) {
    launch_forward(c_, m, acts_.x, 2L * h * w, (long)h * w, 1.0f, acts_.a, acts_.pred, nimg, h, w);
}

static void launch_fused(Ctx& c, const SurModel& m, const float* x, long ps, long is, float scale, float* out,
                         int nimg, int h, int w) {
    /* one workgroup per CU, each with an even share of the strip rows (at
     * least 48, so the six halo rows on either side stay a small part) */
    const long rows = (long)nimg * ceil_div(w, kFuSW) * h;
    const int groups = (int)std::max(1L, std::min((long)c.cus(), (rows + 47) / 48));
    c.launch("sur_fused", sur_fused_kernel, dim3(groups), dim3(kFuThreads), kFuLds, x, ps, is, scale,
             (const float*)m.master, (const uint16_t*)m.wbf, out, nimg, h, w);
}

void Surrogate::infer_strided(const float* d_x, long pstride, long istride, float scale, float* d_out,
                              int nimg, int h, int w) {
    launch_fused(c_, m, d_x, pstride, istride, scale, d_out, nimg, h, w);
}

void Surrogate::infer(const float* d_x, float* d_out, int nimg, int h, int w, bool fused) {
    if (fused) {
        launch_fused(c_, m, d_x, 2L * h * w, (long)h * w, 1.0f, d_out, nimg, h, w);
        return;
    }
    ensure_acts(nimg, h, w);
    launch_forward(c_, m, d_x, 2L * h * w, (long)h * w, 1.0f, acts_.a, d_out, nimg, h, w);
}

double Surrogate::backward(const float* d_target) {
    const int nimg = acts_.nimg, h = acts_.h, w = acts_.w;
    const long P = (long)nimg * h * w;
    const int lb = 1024;
    IQM_CHECK(loss_eps_ < 0.0f || nimg % 2 == 0, IQM_ERR_INVALID,
              "the FSIM-weighted loss takes images in pairs (an even count)");
    c_.launch("sur_loss", sur_loss, dim3(lb), dim3(kBlock), 0, (const float*)acts_.pred, d_target,
              acts_.dpred, d_loss_, P, (long)h * w, loss_eps_);
    c_.launch("sur_loss_final", sur_loss_final, dim3(1), dim3(64), 0, (const double*)d_loss_, lb, P,
              d_loss_ + 1024);
    /* layer 6 -> dz5 */
    c_.launch("sur_dlast", sur_dlast, dim3(ceil_div(P, kBlock)), dim3(kBlock), 0, (const float*)acts_.dpred,
              (const float*)m.master, (const uint16_t*)acts_.a[4], acts_.dz[0], nimg, h, w);
    const dim3 g(ceil_div(w, 16), ceil_div(h, 16), nimg);
    const int np = acts_.npart;
    float* part = acts_.partial;
    int cur = 0;
    for (int l = 5; l >= 2; --l) {
        /* dW_l from dz_l and a_{l-1}; then dz_{l-1} = conv(dz_l, Wt_l) * (a_{l-1} > 0) */
        c_.launch("sur_dw_mid", sur_dw_mid, dim3(np), dim3(256), kDwLds, (const uint16_t*)acts_.dz[cur],
                  (const uint16_t*)acts_.a[l - 2], part, nimg, h, w);
        c_.launch("sur_reduce", sur_reduce, dim3(ceil_div(9216 + 32, kBlock)), dim3(kBlock), 0,
                  (const float*)part, np, 9216 + 32, 9216 + 32, m.grad, SurLayout::wl(l));
        c_.launch("sur_conv_dx", sur_conv_mid<false>, g, dim3(kBlock), 0, (const uint16_t*)acts_.dz[cur],
                  (const uint16_t*)(m.wbf_t + (l - 2) * 9216), (const float*)nullptr,
                  (const uint16_t*)acts_.a[l - 2], acts_.dz[cur ^ 1], h, w);
        cur ^= 1;
    }
    /* layer 1 and layer 6 weights (dz1 = dz[cur]) */
    float* p1 = part;
    float* p6 = part + (size_t)np * 320;
    c_.launch("sur_dw_edge", sur_dw_edge, dim3(np), dim3(kBlock), 0, (const float*)acts_.x,
              2L * h * w, (long)h * w, 1.0f, (const uint16_t*)acts_.dz[cur], (const uint16_t*)acts_.a[4],
              (const float*)acts_.dpred, p1, p6, nimg, h, w);
    c_.launch("sur_reduce", sur_reduce, dim3(ceil_div(320, kBlock)), dim3(kBlock), 0, (const float*)p1, np,
              320, 320, m.grad, SurLayout::w1);
    c_.launch("sur_reduce", sur_reduce, dim3(ceil_div(289, kBlock)), dim3(kBlock), 0, (const float*)p6, np,
              289, 289, m.grad, SurLayout::w6);
    double loss = 0;
    IQM_HIP(hipMemcpyAsync(&loss, d_loss_ + 1024, sizeof(double), hipMemcpyDeviceToHost, c_.stream()));
    IQM_HIP(hipStreamSynchronize(c_.stream()));
    return loss;
}

double Surrogate::loss_and_grad(const float* d_x, const float* d_target, int nimg, int h, int w) {
    ensure_acts(nimg, h, w);
    IQM_HIP(hipMemcpyAsync(acts_.x, d_x, sizeof(float) * nimg * h * w, hipMemcpyDeviceToDevice, c_.stream()));
    forward(nimg, h, w);
    return backward(d_target);
}

double Surrogate::train_step(const float* d_x, const float* d_target, int nimg, int h, int w, float lr) {
    const double loss = loss_and_grad(d_x, d_target, nimg, h, w);
    ++m.step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - (float)std::pow(0.9, (double)m.step), c2 = 1.f - (float)std::pow(0.999, (double)m.step);
    c_.launch("sur_adam", sur_adam, dim3(ceil_div(SurLayout::total, kBlock)), dim3(kBlock), 0, m.master,
              (const float*)m.grad, m.adam_m, m.adam_v, SurLayout::total, lr, b1, b2, 1e-8f, c1, c2);
    refresh_bf16();
    m.data_cursor += (uint64_t)nimg;
    return loss;
}

void Surrogate::make_patches(int n, int size, uint64_t seed, float* d_x, float* d_pc, int scale) {
    IQM_CHECK(n > 0 && n % 2 == 0 && size >= 16 && scale >= 1, IQM_ERR_INVALID,
              "patch count must be even, size >= 16 and scale >= 1");
    /* the fields are made at scale * size and FSIM's own box downscale
     * brings them to size x size, as it does for large inputs */
    const int full = size * scale;
    {
        const int F = std::max(1, (int)std::round(full / 256.0));
        IQM_CHECK((int)std::round((float)full / (float)F) == size, IQM_ERR_INVALID,
                  "scale * size does not downscale to size in FSIM");
    }
    const long per = (long)full * full;  /* full-resolution pixels per patch */
    const long pd = (long)size * size;   /* network-grid pixels per patch */
    const int chunk = std::min(n, std::max(2, (256 / (scale * scale)) & ~1));
    float2* planes = nullptr;
    uint8_t* u8 = nullptr;
    float* pcs = nullptr;
    IQM_HIP(hipMalloc(&planes, sizeof(float2) * per * chunk));
    IQM_HIP(hipMalloc(&u8, per * chunk));
    IQM_HIP(hipMalloc(&pcs, sizeof(float) * 2 * pd * (chunk / 2)));
    struct Free {
        void* a; void* b; void* c;
        ~Free() { (void)hipFree(a); (void)hipFree(b); (void)hipFree(c); }
    } guard{planes, u8, pcs};
    for (int k0 = 0; k0 < n; k0 += chunk) {
        const int k = std::min(chunk, n - k0);
        const uint64_t s = seed + (uint64_t)k0 * 1315423911ull;
        c_.launch("sur_noise", sur_noise, dim3(ceil_div(per * k, kBlock)), dim3(kBlock), 0, planes, per, k, s);
        c_.fft2d_dev(planes, full, full, k, -1, false);
        c_.launch("sur_pink", sur_pink, dim3(ceil_div(per * k, kBlock)), dim3(kBlock), 0, planes, full, k);
        c_.fft2d_dev(planes, full, full, k, +1, true);
        c_.launch("sur_to_u8", sur_to_u8, dim3(k), dim3(1024), 0, (const float2*)planes, u8, full, k, s);
        /* exact PC targets: patches 2i and 2i+1 as one FSIM pair (gray) */
        FsimBatchOut out;
        out.d_pc = pcs;
        c_.fsim_batch(u8, u8 + per, full, full, 1, 2 * per, k / 2, out);
        Ctx::FsimLayout L = c_.fsim_layout(pd, k / 2, c_.fsim_ws_base());
        c_.launch("sur_gather_y", sur_gather_y, dim3(ceil_div(pd * k, kBlock)), dim3(kBlock), 0,
                  (const float*)L.yiq, pd, k / 2, d_x + (long)k0 * pd);
        c_.launch("sur_gather_pc", sur_gather_pc, dim3(ceil_div(pd * k, kBlock)), dim3(kBlock), 0,
                  (const float*)pcs, pd, k / 2, d_pc + (long)k0 * pd);
        c_.sync();
    }
}

/* Spot-check verdict: the largest tile error; the batch falls back to exact
 * PC when it is not within tol (NaN included). */
__global__ __launch_bounds__(kBlock) void sur_decide_kernel(const float* __restrict__ err, int count, float tol,
                                                            SpotFlags* __restrict__ flags) {
    __shared__ float red[kBlock / 64];
    float m = 0.f;
    for (int i = threadIdx.x; i < count; i += blockDim.x) m = fmaxf(m, err[i]);
    for (int d = 32; d > 0; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.f;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) r = fmaxf(r, red[w]);
        flags->err = r;
        flags->use_exact = (r <= tol) ? 0 : 1;
    }
}

/* result of the batch: the exact array when the verdict fell back, else the
 * network's */
__global__ __launch_bounds__(kBlock) void sur_select_kernel(const double* __restrict__ net,
                                                            const double* __restrict__ exact,
                                                            const SpotFlags* __restrict__ flags, int count,
                                                            double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = flags->use_exact ? exact[i] : net[i];
}

void Surrogate::ensure_batch_buffers(int n, int m, long N) {
    bool moved = false;
    if ((size_t)n > out_cap_) {
        c_.sync();
        if (d_out_) (void)hipFree(d_out_);
        d_out_ = nullptr;
        out_cap_ = 0;
        IQM_HIP(hipMalloc(&d_out_, sizeof(double) * 8 * n));
        out_cap_ = (size_t)n;
        moved = true;
    }
    if ((size_t)m * 2 * N > spot_cap_) {
        c_.sync();
        if (spot_pc_) (void)hipFree(spot_pc_);
        if (spot_err_) (void)hipFree(spot_err_);
        spot_pc_ = nullptr;
        spot_err_ = nullptr;
        spot_cap_ = 0;
        IQM_HIP(hipMalloc(&spot_pc_, sizeof(float) * 2 * N * m));
        IQM_HIP(hipMalloc(&spot_err_, sizeof(float) * 2 * kSpotTiles * m));
        spot_cap_ = (size_t)m * 2 * N;
        moved = true;
    }
    if (!d_flags_) {
        IQM_HIP(hipMalloc(&d_flags_, sizeof(SpotFlags)));
        moved = true;
    }
    if (moved) drop_graphs();
}

double* Surrogate::enqueue(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch, long pair_stride,
                           int n, int spot_every, float tol) {
    const int F = std::max(1, (int)std::round(std::min(w, h) / 256.0));
    const int wd = (int)std::round((float)w / (float)F), hd = (int)std::round((float)h / (float)F);
    const long N = (long)wd * hd;
    const int m = spot_every > 0 ? (n + spot_every - 1) / spot_every : 0;
    ensure_batch_buffers(n, m, N);
    const double* net = c_.fsim_batch_pcfn(d_img, d_ref, w, h, ch, pair_stride, n, FsimBatchOut{},
                                           [&](const float* yiq, long Np, int np, float* pc) {
                                               infer_strided(yiq, 6 * Np, Np, 1.0f / 255.0f, pc, 2 * np, hd, wd);
                                               pc_net_ = pc;
                                           });
    if (spot_every <= 0) return const_cast<double*>(net);
    /* exact PC of every spot_every-th pair, compared on seeded tiles */
    FsimBatchOut ex;
    ex.d_pc = spot_pc_;
    c_.fsim_batch(d_img, d_ref, w, h, ch, pair_stride * spot_every, m, ex);
    c_.launch("sur_spot", sur_spot_kernel, dim3(kSpotTiles, 2, m), dim3(kBlock), 0, (const float*)pc_net_,
              (const float*)spot_pc_, wd, hd, spot_every, (uint64_t)0x5EED5EEDull, spot_err_);
    c_.launch("sur_decide", sur_decide_kernel, dim3(1), dim3(kBlock), 0, (const float*)spot_err_,
              2 * kSpotTiles * m, tol, d_flags_);
    /* the whole batch exactly, each kernel gated by the verdict: it does its
     * work only when the batch falls back */
    c_.set_gate(&d_flags_->use_exact);
    const double* exact = nullptr;
    try {
        exact = c_.fsim_batch(d_img, d_ref, w, h, ch, pair_stride, n, FsimBatchOut{});
    } catch (...) {
        c_.set_gate(nullptr);
        throw;
    }
    c_.set_gate(nullptr);
    c_.launch("sur_select", sur_select_kernel, dim3(ceil_div(8 * n, kBlock)), dim3(kBlock), 0, net, exact,
              (const SpotFlags*)d_flags_, 8 * n, d_out_);
    return d_out_;
}

void Surrogate::drop_graphs() {
    if (!graphs_.empty()) (void)hipStreamSynchronize(c_.stream());
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second.exec);
    graphs_.clear();
}

double* Surrogate::enqueue_graph(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch,
                                 long pair_stride, int n, int spot_every, float tol) {
    if (!c_.graphs() || c_.timing()) return enqueue(d_img, d_ref, w, h, ch, pair_stride, n, spot_every, tol);
    if (graph_gen_ != c_.graph_generation()) {
        drop_graphs();
        graph_gen_ = c_.graph_generation();
    }
    uint32_t tolbits;
    std::memcpy(&tolbits, &tol, 4);
    const GraphKey key{d_img, d_ref, pair_stride, w, h, ch, n, spot_every, tolbits};
    auto it = graphs_.find(key);
    if (it != graphs_.end()) {
        it->second.last_use = ++graph_tick_;
        IQM_HIP(hipGraphLaunch(it->second.exec, c_.stream()));
        return it->second.res;
    }
    /* an eager run sizes every workspace and builds the per-size plans */
    double* res = enqueue(d_img, d_ref, w, h, ch, pair_stride, n, spot_every, tol);
    if (graph_gen_ != c_.graph_generation()) {
        drop_graphs();
        graph_gen_ = c_.graph_generation();
    }
    hipGraph_t g = nullptr;
    IQM_HIP(hipStreamBeginCapture(c_.stream(), hipStreamCaptureModeThreadLocal));
    c_.set_capturing(true);
    double* res2 = nullptr;
    try {
        res2 = enqueue(d_img, d_ref, w, h, ch, pair_stride, n, spot_every, tol);
    } catch (...) {
        c_.set_capturing(false);
        (void)hipStreamEndCapture(c_.stream(), &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    c_.set_capturing(false);
    IQM_HIP(hipStreamEndCapture(c_.stream(), &g));
    IQM_CHECK(res2 == res, IQM_ERR_DEVICE, "surrogate batch buffers moved during capture");
    GraphEntry e;
    e.res = res;
    e.last_use = ++graph_tick_;
    const hipError_t ie = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    IQM_HIP(ie);
    if (graphs_.size() >= 8) {
        auto victim = graphs_.begin();
        for (auto i = graphs_.begin(); i != graphs_.end(); ++i)
            if (i->second.last_use < victim->second.last_use) victim = i;
        (void)hipStreamSynchronize(c_.stream());
        (void)hipGraphExecDestroy(victim->second.exec);
        graphs_.erase(victim);
    }
    graphs_.emplace(key, e);
    return res; /* the capture ran nothing: the eager run above produced it */
}

void Surrogate::fsim_batch(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch, long pair_stride,
                           int n, int spot_every, float tol, double* fsim_out, double* fsimc_out, float* spot_err,
                           int* fell_back) {
    const int F = std::max(1, (int)std::round(std::min(w, h) / 256.0));
    const auto size = std::make_pair((int)std::round((float)w / (float)F), (int)std::round((float)h / (float)F));
    const bool known_bad = spot_every > 0 && remember_ && failed_.count(size);
    const double* res;
    if (known_bad) {
        /* the network failed this size's spot check before: exact PC */
        res = c_.fsim_batch_graph(d_img, d_ref, w, h, ch, pair_stride, n, FsimBatchOut{});
    } else {
        res = enqueue_graph(d_img, d_ref, w, h, ch, pair_stride, n, spot_every, tol);
    }
    char* host = (char*)c_.host_pinned(sizeof(double) * 8 * n + sizeof(SpotFlags));
    double* hres = (double*)host;
    SpotFlags* hflags = (SpotFlags*)(host + sizeof(double) * 8 * n);
    IQM_HIP(hipMemcpyAsync(hres, res, sizeof(double) * 8 * n, hipMemcpyDeviceToHost, c_.stream()));
    if (spot_every > 0 && !known_bad)
        IQM_HIP(hipMemcpyAsync(hflags, d_flags_, sizeof(SpotFlags), hipMemcpyDeviceToHost, c_.stream()));
    c_.sync();
    for (int i = 0; i < n; ++i) {
        if (fsim_out) fsim_out[i] = hres[(size_t)8 * i + 3];
        if (fsimc_out) fsimc_out[i] = hres[(size_t)8 * i + 4];
    }
    float err = 0.f;
    int back = 0;
    if (known_bad) {
        err = failed_[size];
        back = 1;
    } else if (spot_every > 0) {
        err = hflags->err;
        back = hflags->use_exact;
        if (back && remember_) failed_[size] = err;
    }
    if (spot_err) *spot_err = err;
    if (fell_back) *fell_back = back;
}

/* versioned little-endian blob: magic, version, step, seed, cursor, then the
 * FP32 master weights and both Adam moments */
std::vector<uint8_t> Surrogate::save() const {
    const int n = SurLayout::total;
    std::vector<uint8_t> blob(8 + 4 + 3 * 8 + 3 * 4 * (size_t)n);
    std::memcpy(blob.data(), "IQMSUR01", 8);
    const uint32_t ver = 1;
    std::memcpy(blob.data() + 8, &ver, 4);
    const uint64_t hdr[3] = {(uint64_t)m.step, m.seed, m.data_cursor};
    std::memcpy(blob.data() + 12, hdr, sizeof(hdr));
    uint8_t* p = blob.data() + 12 + sizeof(hdr);
    for (const float* src : {m.master, m.adam_m, m.adam_v}) {
        IQM_HIP(hipMemcpyAsync(p, src, sizeof(float) * n, hipMemcpyDeviceToHost, c_.stream()));
        p += sizeof(float) * n;
    }
    IQM_HIP(hipStreamSynchronize(c_.stream()));
    return blob;
}

void Surrogate::load(const std::vector<uint8_t>& blob) {
    const int n = SurLayout::total;
    IQM_CHECK(blob.size() == 8 + 4 + 3 * 8 + 3 * 4 * (size_t)n && std::memcmp(blob.data(), "IQMSUR01", 8) == 0,
              IQM_ERR_INVALID, "not a surrogate checkpoint of this version");
    uint64_t hdr[3];
    std::memcpy(hdr, blob.data() + 12, sizeof(hdr));
    m.step = (long)hdr[0];
    m.seed = hdr[1];
    m.data_cursor = hdr[2];
    const uint8_t* p = blob.data() + 12 + sizeof(hdr);
    for (float* dst : {m.master, m.adam_m, m.adam_v}) {
        IQM_HIP(hipMemcpyAsync(dst, p, sizeof(float) * n, hipMemcpyHostToDevice, c_.stream()));
        p += sizeof(float) * n;
    }
    refresh_bf16();
    c_.sync();
}

}  // namespace iqm
