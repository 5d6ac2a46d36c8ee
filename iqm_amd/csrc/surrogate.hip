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

/* Layer 1: 1 -> 32 channels, FP32 FMA, ReLU.  Image j of the input is read
 * at x + (j/2)*pstride + (j%2)*istride, times `scale` (so FSIM's Y planes can
 * be read in place). */
__global__ __launch_bounds__(kBlock) void sur_conv_first(const float* __restrict__ x, long pstride,
                                                         long istride, float scale,
                                                         const float* __restrict__ params,
                                                         uint16_t* __restrict__ out, int nimg, int h,
                                                         int w) {
    __shared__ float wsh[32 * 9 + 32];
    for (int i = threadIdx.x; i < 32 * 9 + 32; i += blockDim.x) {
        const float v = params[SurLayout::w1 + i];
        wsh[i] = i < 32 * 9 ? bf2f(f2bf(v)) : v; /* BF16 weights, FP32 bias */
    }
    __syncthreads();
    const long HW = (long)h * w;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nimg * HW) return;
    const int j = (int)(p / HW);
    const long q = p - (long)j * HW;
    const int y = (int)(q / w), xx = (int)(q - (long)y * w);
    const float* xi = x + (long)(j >> 1) * pstride + (long)(j & 1) * istride;
    float v[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int yy = y + t / 3 - 1, xc = xx + t % 3 - 1;
        v[t] = (yy >= 0 && yy < h && xc >= 0 && xc < w) ? xi[(long)yy * w + xc] * scale : 0.f;
    }
    uint32_t packed[16];
#pragma unroll
    for (int co = 0; co < 32; co += 2) {
        float s0 = wsh[32 * 9 + co], s1 = wsh[32 * 9 + co + 1];
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            s0 += wsh[co * 9 + t] * v[t];
            s1 += wsh[(co + 1) * 9 + t] * v[t];
        }
        packed[co / 2] = (uint32_t)f2bf(fmaxf(s0, 0.f)) | ((uint32_t)f2bf(fmaxf(s1, 0.f)) << 16);
    }
    uint4* o = (uint4*)(out + p * 32);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        o[k] = make_uint4(packed[4 * k], packed[4 * k + 1], packed[4 * k + 2], packed[4 * k + 3]);
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
 * workgroup streams a strip of kFuSW output columns of one image down its
 * rows.  At step s the matrix-core layer l (2..5) reads one new row q of
 * layer l-1 (the row that layer wrote at step s-1) and applies the three tap
 * rows of the 3x3 kernel to it at once: row q is tap row 0 of output row
 * q+1 (whose sums start here), tap row 1 of output row q and tap row 2 of
 * output row q-1, which is then complete and is written to the layer's ring.
 * So every row is read from LDS once, the three open output rows live in
 * registers, and a layer's ring holds two rows (the one being written and
 * the one the next layer reads).  An output's taps are still summed in the
 * order t = 0..8 (three per step, dx ascending), so the bits are those of
 * sur_conv_mid.  Layer 6 keeps its three open rows the same way; layer 1
 * (FP32, from a 4-row input ring) computes row s from input rows s-1..s+1 as
 * sur_conv_first does.  Output row s - 2(l-1) completes at step s; steps run
 * from -2 to h + 9 (layer 6 completes row h-1).  Rows outside the image read
 * as zeros: slots of negative rows are never written before they are read,
 * and rows >= h are written as zeros.  The halo shrinks by one column per
 * layer (layer l covers kFuSW + 2(6-l) columns), positions outside the image
 * are zeros (the zero padding of the layer-by-layer network).  Waves 2k and
 * 2k+1 own layer k+2, one 16-channel half each (its B fragments in
 * registers, all five pixel tiles, three accumulators per tile); the FP32
 * layers 1 and 6 are shared by all threads. */
constexpr int kFuSW = 64;                          /* output columns per strip */
constexpr int kFuPitch = 84;                       /* ring row pitch (pixels) */
constexpr int kFuXPitch = 80;                      /* input ring row pitch (floats) */
constexpr int kFuThreads = 512;
constexpr size_t kFuRingBytes = (size_t)5 * 2 * kFuPitch * 32 * 2;
constexpr size_t kFuXBytes = (size_t)4 * kFuXPitch * 4; /* input rows, then the rings */
constexpr size_t kFuLds = kFuXBytes + kFuRingBytes;
static_assert(kFuPitch % 4 == 0 && kFuXBytes % 256 == 0, "ring rows on 256-byte boundaries");

__global__ __launch_bounds__(kFuThreads) void sur_fused_kernel(const float* __restrict__ x, long pstride,
                                                               long istride, float scale,
                                                               const float* __restrict__ params,
                                                               const uint16_t* __restrict__ wbf,
                                                               float* __restrict__ out, int h, int w) {
    extern __shared__ __align__(16) uint8_t fu_smem[];
    __shared__ float b6s;
    float* xr = (float*)fu_smem;                                  /* [4][kFuXPitch] */
    uint16_t* ring = (uint16_t*)(fu_smem + kFuXBytes);            /* [layer 1..5][2][kFuPitch][32] */
    /* channel c of pixel p: a pixel's 64 bytes hold its four 16-byte chunks
     * of 8 channels in the order chunk ^ ((p >> 1) & 3), so the 8 lanes of
     * one 128-byte LDS access reading one chunk of 8 consecutive pixels (a
     * matrix-core fragment) fall into 8 different 16-byte bank groups */
    auto RS = [&](int layer, int row, int p, int c) {
        return ring + (((layer - 1) * 2 + (row & 1)) * kFuPitch + p) * 32 + ((((c >> 3) ^ (p >> 1)) & 3) << 3) +
               (c & 7);
    };
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * kFuSW;
    const int j = blockIdx.y;
    const float* xi = x + (long)(j >> 1) * pstride + (long)(j & 1) * istride;
    float* oj = out + (long)j * h * w;

    for (int i = tid; i < (int)(kFuLds / 16); i += blockDim.x) ((uint4*)fu_smem)[i] = make_uint4(0, 0, 0, 0);
    if (tid == 0) b6s = params[SurLayout::b6];
    /* this wave's layer, channel half and B fragments (k = ci, n = co) */
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int ml = 2 + (wave >> 1), nb = wave & 1;
    const uint16_t* wl = wbf + (ml - 2) * 9216;
    bf16x8 B[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) B[t] = *(const bf16x8*)&wl[(t * 32 + nb * 16 + li) * 32 + 8 * g];
    /* biases of the channels this lane's outputs hold: 16 nb + 4 g + e */
    f32x4 b4;
#pragma unroll
    for (int e = 0; e < 4; ++e) b4[e] = params[SurLayout::bl(ml) + 16 * nb + 4 * g + e];
    /* bit k: this lane's pixel of tile k lies inside the layer's columns and the image */
    unsigned inmask = 0;
    {
        const int lcols = kFuSW + 2 * (6 - ml);
        for (int k = 0; k < 5; ++k) {
            const int p = 16 * k + li, gx = x0 - (6 - ml) + p;
            if (p < lcols && gx >= 0 && gx < w) inmask |= 1u << k;
        }
    }
    /* layer 1: thread = (channel l1co, pixels 5 l1pg ..), its BF16-rounded weights */
    const int l1co = tid & 31, l1pg = tid >> 5;
    float w1r[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) w1r[t] = bf2f(f2bf(params[SurLayout::w1 + l1co * 9 + t]));
    const float b1r = params[SurLayout::b1 + l1co];
    unsigned l1mask = 0; /* bit k: pixel 5 l1pg + k is a column of the image */
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int gx = x0 - 5 + l1pg * 5 + k;
        if (gx >= 0 && gx < w) l1mask |= 1u << k;
    }
    /* layer 6: thread = (pixel tid / 8, channels 4 (tid % 8) ..); weights of
     * channel pairs packed as BF16 pairs for the two-term dot products */
    bf16x2 w6p[2][9];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int c = 4 * (tid & 7) + 2 * k;
            w6p[k][t] = bf16x2{(__bf16)params[SurLayout::w6 + c * 9 + t], (__bf16)params[SurLayout::w6 + (c + 1) * 9 + t]};
        }
    /* input rows enter two steps ahead: row s + 2 is stored at step s (from
     * a register loaded the step before) */
    const int xq = tid, xgx = x0 - 6 + tid;
    const bool xload = tid < kFuSW + 12 && xgx >= 0 && xgx < w;
    float xnext = (xload && 0 < h) ? xi[xgx] : 0.f;
    __syncthreads();

    /* open sums: [k] belongs to the output row started at step u = s + 2
     * with u % 3 == k */
    float a6[3] = {0.f, 0.f, 0.f};
    f32x4 acc[3][5];
    const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int c = 0; c < 5; ++c) acc[k][c] = z4;

    auto step = [&](auto kk, int s) {
        constexpr int kn = decltype(kk)::value;  /* slot of the row started now */
        constexpr int km = (kn + 2) % 3;         /* slot of the row started last step */
        constexpr int kc = (kn + 1) % 3;         /* slot of the row completed now */
        /* input row s + 2 to LDS; issue row s + 3 */
        if (tid < kFuSW + 12) {
            xr[((s + 2) & 3) * kFuXPitch + xq] = xnext * scale;
            xnext = (xload && s + 3 >= 0 && s + 3 < h) ? xi[(long)(s + 3) * w + xgx] : 0.f;
        }
        auto layer1 = [&]() {
            /* row s from input rows s-1..s+1: kFuSW + 10 columns x 32 channels */
            const int r = s;
            const int p0 = l1pg * 5;
            if (r < 0 || r > h || p0 >= kFuSW + 10) return;
            float xw[3][7];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int qq = 0; qq < 7; ++qq) xw[dy][qq] = xr[((r + dy - 1) & 3) * kFuXPitch + p0 + qq];
            const bool rin = r < h;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const int p = p0 + k;
                /* computed everywhere, kept where the column is real */
                float v = b1r;
#pragma unroll
                for (int t = 0; t < 9; ++t) v += w1r[t] * xw[t / 3][k + t % 3];
                v = (rin && ((l1mask >> k) & 1u)) ? fmaxf(v, 0.f) : 0.f;
                if (p < kFuSW + 10) *RS(1, r, p, l1co) = f2bf(v);
            }
        };
        auto layers_mid = [&]() {
            /* layer ml reads row q = s - 2 ml + 3 of layer ml - 1 and
             * completes row q - 1; five tiles, three open rows each */
            static_assert((kFuSW + 8 + 15) / 16 == 5 && (kFuSW + 2 + 15) / 16 == 5, "five tiles per layer row");
            const int q = s - 2 * ml + 3;
            /* q = h + 1 only writes row h as zeros: the next layer reads it
             * as padding and its slot still holds row h - 2 */
            if (q < -1 || q > h + 1) return;
            __builtin_amdgcn_s_setprio(1);
#pragma unroll
            for (int dx = 0; dx < 3 && q <= h; ++dx) {
                const uint16_t* row = RS(ml - 1, q, li + dx, 8 * g);
#pragma unroll
                for (int tile = 0; tile < 5; ++tile) {
                    /* tiles are 16 pixels apart, which keeps the chunk order */
                    const bf16x8 F = *(const bf16x8*)(row + tile * 16 * 32);
                    /* weights as the first operand: D = W^T X^T, so a lane's
                     * four outputs are 4 channels of one pixel; taps dx (new
                     * row), 3 + dx (open row), 6 + dx (completing row) */
                    acc[kn][tile] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[dx], F, dx == 0 ? b4 : acc[kn][tile], 0, 0, 0);
                    acc[km][tile] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[3 + dx], F, acc[km][tile], 0, 0, 0);
                    acc[kc][tile] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[6 + dx], F, acc[kc][tile], 0, 0, 0);
                }
            }
            __builtin_amdgcn_s_setprio(0);
            const int o = q - 1;
            if (o < 0) return;
            const bool live = o < h;
            /* D layout: lane holds channels 16 nb + 4g..4g+3 of pixel li of a
             * tile (the sums started from the bias): BF16, ReLU, one 8-byte
             * store */
#pragma unroll
            for (int tile = 0; tile < 5; ++tile) {
                const f32x4 a = acc[kc][tile];
                const bool keep = live && ((inmask >> tile) & 1u);
                const uint2 v = make_uint2(keep ? relu_bf16x2(pack_bf16(a[0], a[1])) : 0u,
                                           keep ? relu_bf16x2(pack_bf16(a[2], a[3])) : 0u);
                *(uint2*)RS(ml, o, tile * 16 + li, 16 * nb + 4 * g) = v;
            }
        };
        auto layer6 = [&]() {
            /* reads row q = s - 9 of layer 5, completes output row s - 10 */
            const int q = s - 9;
            if (q < -1 || q > h) return;
            const int p = tid >> 3, cg = tid & 7;
            float vn = 0.f, vm = a6[km], vc = a6[kc];
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const uint2 a = *(const uint2*)RS(5, q, p + dx, 4 * cg);
                const bf16x2 lo = __builtin_bit_cast(bf16x2, a.x), hi = __builtin_bit_cast(bf16x2, a.y);
                /* BF16 x BF16 products are exact in FP32 */
                vn = __builtin_amdgcn_fdot2_f32_bf16(lo, w6p[0][dx], vn, false);
                vn = __builtin_amdgcn_fdot2_f32_bf16(hi, w6p[1][dx], vn, false);
                vm = __builtin_amdgcn_fdot2_f32_bf16(lo, w6p[0][3 + dx], vm, false);
                vm = __builtin_amdgcn_fdot2_f32_bf16(hi, w6p[1][3 + dx], vm, false);
                vc = __builtin_amdgcn_fdot2_f32_bf16(lo, w6p[0][6 + dx], vc, false);
                vc = __builtin_amdgcn_fdot2_f32_bf16(hi, w6p[1][6 + dx], vc, false);
            }
            a6[kn] = vn;
            a6[km] = vm;
            vc += __shfl_xor(vc, 1, 64);
            vc += __shfl_xor(vc, 2, 64);
            vc += __shfl_xor(vc, 4, 64);
            const int o = q - 1, gx = x0 + p;
            if (cg == 0 && o >= 0 && o < h && gx < w) oj[(long)o * w + gx] = vc + b6s;
        };
        /* the layers of a step touch disjoint LDS rows, so their order is
         * free: waves 0-3 run the FP32 layers first, waves 4-7 the
         * matrix-core layer first, and each SIMD (waves k and k+4) overlaps
         * the two */
        if (wave < 4) {
            layer1();
            layer6();
            layers_mid();
        } else {
            layers_mid();
            layer1();
            layer6();
        }
        lds_barrier();
    };
    for (int u = 0; u < h + 12; u += 3) {
        step(std::integral_constant<int, 0>{}, u - 2);
        step(std::integral_constant<int, 1>{}, u - 1);
        step(std::integral_constant<int, 2>{}, u);
    }
}

/* All six layers of the forward pass in one kernel, for inference.  A
 * workgroup streams a strip of kFuSW output columns of one image down its
 * rows; every layer's activations stay in LDS in a ring of 4 rows.  Layer l
 * (1..6) works on row s - 2(l-1) at step s, so all layers advance together
 * with one barrier per step: layer l reads rows r-1..r+1 of layer l-1, which
 * were written in earlier steps, while layer l-1 writes row r+2 into the
 * fourth slot.  The halo shrinks by one column per layer (layer l covers
 * kFuSW + 2(6-l) columns), and positions outside the image are written as
 * zeros, which is the zero padding of the layer-by-layer network, so layers
 * 1-5 produce the same bits as sur_conv_first / sur_conv_mid.  Waves 2k and
 * 2k+1 own layer k+2 (B fragments of both 16-channel halves in registers for
 * the whole kernel, pixel tiles split between the two waves); the FP32 layers
 * 1 and 6 are shared by all threads. */
constexpr size_t kFuRingBytesV1 = (size_t)5 * 4 * kFuPitch * 32 * 2;
constexpr size_t kFuLdsV1 = (size_t)4 * kFuXPitch * 4 + kFuRingBytesV1;

__global__ __launch_bounds__(kFuThreads) void sur_fused_v1_kernel(const float* __restrict__ x, long pstride,
                                                               long istride, float scale,
                                                               const float* __restrict__ params,
                                                               const uint16_t* __restrict__ wbf,
                                                               float* __restrict__ out, int h, int w) {
    extern __shared__ __align__(16) uint8_t fu_smem[];
    __shared__ float b6s;
    float* xr = (float*)fu_smem;                                  /* [4][kFuXPitch] */
    uint16_t* ring = (uint16_t*)(fu_smem + 4 * kFuXPitch * 4);    /* [layer 1..5][4][kFuPitch][32] */
    auto R = [&](int layer, int row, int p) { return ring + (((layer - 1) * 4 + (row & 3)) * kFuPitch + p) * 32; };
    /* channel c of pixel p: a pixel's 64 bytes hold its four 16-byte chunks
     * of 8 channels in the order chunk ^ ((p >> 1) & 3), so the 8 lanes of
     * one 128-byte LDS access reading one chunk of 8 consecutive pixels (a
     * matrix-core A fragment) fall into 8 different 16-byte bank groups
     * (measured bank-conflict cycles: 39% of the kernel's cycles with the
     * plain layout, 3.5% with this one); ring rows start on 256-byte
     * boundaries */
    auto RS = [&](int layer, int row, int p, int c) {
        return R(layer, row, p) + ((((c >> 3) ^ (p >> 1)) & 3) << 3) + (c & 7);
    };
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * kFuSW;
    const int j = blockIdx.y;
    const float* xi = x + (long)(j >> 1) * pstride + (long)(j & 1) * istride;
    float* oj = out + (long)j * h * w;

    for (int i = tid; i < (int)(kFuLds / 16); i += blockDim.x) ((uint4*)fu_smem)[i] = make_uint4(0, 0, 0, 0);
    if (tid == 0) b6s = params[SurLayout::b6];
    /* this wave's layer and its B fragments (k = ci, n = co) and biases */
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, g = lane >> 4;
    const int ml = 2 + (wave >> 1), half = wave & 1;
    const uint16_t* wl = wbf + (ml - 2) * 9216;
    bf16x8 B[9][2];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) B[t][nb] = *(const bf16x8*)&wl[(t * 32 + nb * 16 + li) * 32 + 8 * g];
    /* biases of the channels this lane's outputs hold: 16 nb + 4 g + e */
    float bias[2][4];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int e = 0; e < 4; ++e) bias[nb][e] = params[SurLayout::bl(ml) + 16 * nb + 4 * g + e];
    /* layer 1: thread = (channel l1co, pixels 5 l1pg ..); layer 6: thread =
     * (pixel tid / 8, channels 4 (tid % 8) ..); their BF16-rounded weights */
    /* whether this lane's pixel of each of its mid-layer tiles lies inside the
     * layer's columns and the image (bit 0: tile 2h, 1: tile 2h+1, 2: tile 4) */
    unsigned inmask = 0;
    {
        const int wl_cols = kFuSW + 2 * (6 - ml);
        const int ps[3] = {2 * half * 16 + li, (2 * half + 1) * 16 + li, 64 + li};
        for (int k = 0; k < 3; ++k) {
            const int gx = x0 - (6 - ml) + ps[k];
            if (ps[k] < wl_cols && gx >= 0 && gx < w) inmask |= 1u << k;
        }
    }
    const int l1co = tid & 31, l1pg = tid >> 5;
    float w1r[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) w1r[t] = bf2f(f2bf(params[SurLayout::w1 + l1co * 9 + t]));
    const float b1r = params[SurLayout::b1 + l1co];
    /* layer 6 weights of channel pairs (4 cg, 4 cg + 1) and (4 cg + 2, 4 cg + 3),
     * packed as BF16 pairs for the two-term dot products */
    bf16x2 w6p[2][9];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int c = 4 * (tid & 7) + 2 * k;
            w6p[k][t] = bf16x2{(__bf16)params[SurLayout::w6 + c * 9 + t], (__bf16)params[SurLayout::w6 + (c + 1) * 9 + t]};
        }
    __syncthreads();
    /* input rows 0 and 1 (row -1 stays zero) */
    for (int i = tid; i < 2 * (kFuSW + 12); i += blockDim.x) {
        const int r = i / (kFuSW + 12), q = i - r * (kFuSW + 12);
        const int gx = x0 - 6 + q;
        xr[r * kFuXPitch + q] = (r < h && gx >= 0 && gx < w) ? xi[(long)r * w + gx] * scale : 0.f;
    }
    /* input rows are loaded one step before they are stored to LDS, so the
     * load's latency overlaps a step of work */
    const int xq = tid, xgx = x0 - 6 + tid;
    const bool xload = tid < kFuSW + 12 && xgx >= 0 && xgx < w;
    float xnext = (xload && 2 < h) ? xi[2L * w + xgx] : 0.f;
    __syncthreads();

    for (int s = 0; s < h + 10; ++s) {
        /* input row s + 2 (loaded last step); issue row s + 3 */
        if (tid < kFuSW + 12) {
            xr[((s + 2) & 3) * kFuXPitch + xq] = xnext * scale;
            xnext = (xload && s + 3 < h) ? xi[(long)(s + 3) * w + xgx] : 0.f;
        }
        auto layer1 = [&]() {
        /* layer 1, row s: kFuSW + 10 columns x 32 channels; thread = (channel,
         * 5 adjacent pixels) with the channel's weights in registers */
        {
            const int r = s;
            const int p0 = l1pg * 5;
            if (p0 < kFuSW + 10) {
                float xw[3][7];
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int q = 0; q < 7; ++q) xw[dy][q] = xr[((r + dy - 1) & 3) * kFuXPitch + p0 + q];
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    const int p = p0 + k;
                    const int gx = x0 - 5 + p;
                    float v = 0.f;
                    if (r < h && gx >= 0 && gx < w) {
                        v = b1r;
#pragma unroll
                        for (int t = 0; t < 9; ++t) v += w1r[t] * xw[t / 3][k + t % 3];
                        v = fmaxf(v, 0.f);
                    }
                    if (p < kFuSW + 10) *RS(1, r, p, l1co) = f2bf(v);
                }
            }
        }
        };
        auto layers_mid = [&]() {
        /* layers 2-5 on the matrix cores: a layer row is 5 pixel tiles x 2
         * channel halves; wave `half` of the layer takes tiles 2h and 2h+1
         * (both halves, one A load per two instructions) and half h of tile
         * 4, as five independent accumulator chains */
        {
            static_assert((kFuSW + 8 + 15) / 16 == 5 && (kFuSW + 2 + 15) / 16 == 5, "five tiles per layer row");
            const int r = s - 2 * (ml - 1);
            const int wl_cols = kFuSW + 2 * (6 - ml);
            if (r >= 0) {
                const bool live = r < h;
                const int ta = 2 * half, tb = 2 * half + 1;
                f32x4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = a0, b0 = a0, b1 = a0, c0 = a0;
                if (live) {
                    /* the matrix-core chain issues first when both waves of
                     * a SIMD are ready; the other wave's FP32 layers fill in */
                    __builtin_amdgcn_s_setprio(1);
#pragma unroll
                    for (int t = 0; t < 9; ++t) {
                        /* tiles are 16 pixels apart, which keeps the chunk order */
                        const uint16_t* row = RS(ml - 1, r + t / 3 - 1, li + t % 3, 8 * g);
                        const bf16x8 A = *(const bf16x8*)(row + ta * 16 * 32);
                        const bf16x8 Bv = *(const bf16x8*)(row + tb * 16 * 32);
                        const bf16x8 C = *(const bf16x8*)(row + 4 * 16 * 32);
                        /* weights as the first operand: D = W^T X^T, so a
                         * lane's four outputs are 4 channels of one pixel */
                        a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[t][0], A, a0, 0, 0, 0);
                        b0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[t][0], Bv, b0, 0, 0, 0);
                        c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(half ? B[t][1] : B[t][0], C, c0, 0, 0, 0);
                        a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[t][1], A, a1, 0, 0, 0);
                        b1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[t][1], Bv, b1, 0, 0, 0);
                    }
                    __builtin_amdgcn_s_setprio(0);
                }
                /* D layout: lane holds channels 16 nb + 4g..4g+3 of pixel li
                 * of a tile: bias, ReLU, BF16, one 8-byte store */
                auto put = [&](const f32x4& acc, int nb, int tile, unsigned bit) {
                    uint2 v = make_uint2(0u, 0u);
                    if (live && ((inmask >> bit) & 1u)) {
                        v.x = (uint32_t)f2bf(fmaxf(acc[0] + bias[nb][0], 0.f)) |
                              ((uint32_t)f2bf(fmaxf(acc[1] + bias[nb][1], 0.f)) << 16);
                        v.y = (uint32_t)f2bf(fmaxf(acc[2] + bias[nb][2], 0.f)) |
                              ((uint32_t)f2bf(fmaxf(acc[3] + bias[nb][3], 0.f)) << 16);
                    }
                    *(uint2*)RS(ml, r, tile * 16 + li, 16 * nb + 4 * g) = v;
                };
                put(a0, 0, ta, 0);
                put(a1, 1, ta, 0);
                put(b0, 0, tb, 1);
                put(b1, 1, tb, 1);
                put(c0, half, 4, 2);
            }
        }
        };
        auto layer6 = [&]() {
        /* layer 6, row s - 10: thread = (pixel, group of 4 channels) */
        {
            const int r = s - 10;
            const int p = tid >> 3, cg = tid & 7;
            float v = 0.f;
            if (r >= 0 && r < h) {
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                    const uint2 a = *(const uint2*)RS(5, r + t / 3 - 1, p + t % 3, 4 * cg);
                    /* BF16 x BF16 products are exact in FP32 */
                    v = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, a.x), w6p[0][t], v, false);
                    v = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, a.y), w6p[1][t], v, false);
                }
            }
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            v += __shfl_xor(v, 4, 64);
            const int gx = x0 + p;
            if (cg == 0 && r >= 0 && r < h && gx < w) oj[(long)r * w + gx] = v + b6s;
        }
        };
        /* the layers of a step touch disjoint ring rows, so their order is
         * free: waves 0-3 run the FP32 layers first, waves 4-7 the matrix-core
         * layer first, and each SIMD (waves k and k+4) overlaps the two */
        if (wave < 4) {
            layer1();
            layer6();
            layers_mid();
        } else {
            layers_mid();
            layer1();
            layer6();
        }
        lds_barrier();
    }
}

/* ---- backward ---------------------------------------------------------- */

/* dpred = 2 (pred - target) / P and per-block FP64 partial sums of the
 * squared error. */
__global__ __launch_bounds__(kBlock) void sur_loss(const float* __restrict__ pred,
                                                   const float* __restrict__ target,
                                                   float* __restrict__ dpred, double* __restrict__ part,
                                                   long P) {
    __shared__ double sh[kBlock];
    double acc = 0;
    const float sc = 2.0f / (float)P;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
        const float d = pred[p] - target[p];
        dpred[p] = sc * d;
        acc += (double)d * (double)d;
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
 * a[p + offset(tap)][ci] and db[co] = sum_p dz[p][co].  A workgroup of 288
 * threads walks 16 x 16 tiles (grid-stride); thread (tap, co block of 8, ci
 * block of 4) keeps 32 FP32 accumulators and writes one partial row. */
__global__ __launch_bounds__(288) void sur_dw_mid(const uint16_t* __restrict__ dz,
                                                  const uint16_t* __restrict__ a,
                                                  float* __restrict__ part, int nimg, int h, int w) {
    __shared__ __align__(16) uint16_t tdz[16 * 16 * 32];
    __shared__ __align__(16) uint16_t ta[18 * 18 * 32];
    const int tid = threadIdx.x;
    const int t = tid / 32, cb = (tid % 32) / 8, ib = tid % 8; /* tap, co block, ci block */
    const int dy = t / 3, dx = t % 3;
    float acc[8][4];
    for (int i = 0; i < 8; ++i)
        for (int k = 0; k < 4; ++k) acc[i][k] = 0.f;
    float db = 0.f;
    const long HW = (long)h * w;
    const int tx = (w + 15) / 16, ty = (h + 15) / 16;
    const long ntiles = (long)nimg * tx * ty;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int img = (int)(tile / (tx * ty));
        const int rem = (int)(tile - (long)img * tx * ty);
        const int y0 = (rem / tx) * 16, x0 = (rem % tx) * 16;
        __syncthreads();
        for (int i = tid; i < 16 * 16 * 4; i += blockDim.x) {
            const int px = i >> 2, part4 = i & 3;
            const int gy = y0 + (px >> 4), gx = x0 + (px & 15);
            uint4 v = make_uint4(0, 0, 0, 0);
            if (gy < h && gx < w) v = ((const uint4*)(dz + ((long)img * HW + (long)gy * w + gx) * 32))[part4];
            ((uint4*)tdz)[i] = v;
        }
        for (int i = tid; i < 18 * 18 * 4; i += blockDim.x) {
            const int px = i >> 2, part4 = i & 3;
            const int ly = px / 18, lx = px - ly * 18;
            const int gy = y0 - 1 + ly, gx = x0 - 1 + lx;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (gy >= 0 && gy < h && gx >= 0 && gx < w) v = ((const uint4*)(a + ((long)img * HW + (long)gy * w + gx) * 32))[part4];
            ((uint4*)ta)[i] = v;
        }
        __syncthreads();
        for (int px = 0; px < 256; ++px) {
            const int ly = px >> 4, lx = px & 15;
            const uint4 dv = *(const uint4*)&tdz[px * 32 + cb * 8];
            const uint2 av = *(const uint2*)&ta[((ly + dy) * 18 + lx + dx) * 32 + ib * 4];
            const uint32_t du[4] = {dv.x, dv.y, dv.z, dv.w};
            const uint32_t au[2] = {av.x, av.y};
            float dzf[8], af[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                dzf[2 * k] = __uint_as_float(du[k] << 16);
                dzf[2 * k + 1] = __uint_as_float(du[k] & 0xffff0000u);
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                af[2 * k] = __uint_as_float(au[k] << 16);
                af[2 * k + 1] = __uint_as_float(au[k] & 0xffff0000u);
            }
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int k = 0; k < 4; ++k) acc[i][k] += dzf[i] * af[k];
        }
        if (tid < 32) {
            for (int px = 0; px < 256; ++px) db += bf2f(tdz[px * 32 + tid]);
        }
    }
    /* partial row: [32][32][9] weights then [32] biases */
    float* pr = part + (long)blockIdx.x * (32 * 32 * 9 + 32);
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int k = 0; k < 4; ++k) pr[((cb * 8 + i) * 32 + ib * 4 + k) * 9 + t] = acc[i][k];
    if (tid < 32) pr[32 * 32 * 9 + tid] = db;
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
    c.launch("sur_conv1", sur_conv_first, dim3(ceil_div(P, kBlock)), dim3(kBlock), 0, x, ps, is, scale,
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
    IQM_CHECK(nimg <= 65535, IQM_ERR_INVALID, "too many images for one fused launch");
    if (std::getenv("IQM_SUR_FUSED_V1"))
        c.launch("sur_fused", sur_fused_v1_kernel, dim3(ceil_div(w, kFuSW), nimg), dim3(kFuThreads), kFuLdsV1, x,
                 ps, is, scale, (const float*)m.master, (const uint16_t*)m.wbf, out, h, w);
    else
        c.launch("sur_fused", sur_fused_kernel, dim3(ceil_div(w, kFuSW), nimg), dim3(kFuThreads), kFuLds, x, ps,
                 is, scale, (const float*)m.master, (const uint16_t*)m.wbf, out, h, w);
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
    c_.launch("sur_loss", sur_loss, dim3(lb), dim3(kBlock), 0, (const float*)acts_.pred, d_target,
              acts_.dpred, d_loss_, P);
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
        c_.launch("sur_dw_mid", sur_dw_mid, dim3(np), dim3(288), 0, (const uint16_t*)acts_.dz[cur],
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
