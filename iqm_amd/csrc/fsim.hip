// FSIM / FSIMc on MI355X.
//
// Pipeline per batch of same-size pairs (reference: FSIM::computeMetric,
// src/gpu/fsim/fsim.cpp:80-163), 7 launches:
//   fsim_prep          F2: YIQ of 255-scaled RGB, F x F box average
//   fsim_rowfft        forward row FFT of Y_in + i*Y_ref, split into the two
//                      row spectra (F4-F5, first half)
//   fsim_colpass       forward column FFT -> spectrum, times H_{s,o} (F9),
//                      inverse column FFT (F11, first half); 16 planes/image
//   fsim_rowpass       inverse row FFT (F11, second half) + response epilogue
//                      (F12) + |EO_{s=0}|^2 keys and their top-bit histogram
//   fsim_select        exact upper median by radix select (F13) -> noise
//                      power and threshold T (F15) in double
//   fsim_pcfinal       Scharr gradient (F3), phase congruency (F15), final
//                      multiply (F16), FP64 block partial sums (F17)
//   fsim_finalize      fixed-order sum of the partials, FSIM and FSIMc
// Per-size constants (filter planes, EM_n of F10, est of F14) are computed
// once per (wd, hd) by Ctx::fsim_plan and cached.
#include "fsim_body.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace iqm {

/* ---- filter bank, FP32 like the reference shaders --------------------- */

/* fsim_lowpassfilter.glsl:32-35 (same in log_gabor and angular): ifftshift
 * order, u = ((x + n/2) mod n - n/2) / n */
__device__ __forceinline__ float freq_coord(int x, int n) {
    unsigned shifted = ((unsigned)x + (unsigned)n / 2u) % (unsigned)n;
    return (float)((int)shifted - n / 2) / (float)n;
}

__global__ __launch_bounds__(kBlock) void fsim_tables_kernel(float* lg, float* ang, float* bank,
                                                             int wd, int hd) {
    const long N = (long)wd * hd;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const int y = (int)(idx / wd), x = (int)(idx - (long)y * wd);
    const float u = freq_coord(x, wd), v = freq_coord(y, hd);
    const float radius = sqrtf(u * u + v * v);
    /* fsim_lowpassfilter.glsl:37-38, cutoff 0.45, order 15 */
    const float lowpass = 1.0f / (1.0f + powf(radius / 0.45f, 2.0f * 15.0f));
    /* fsim_log_gabor.glsl:24-48 */
    const float ls = logf(0.55f);
    for (int s = 0; s < 4; ++s) {
        const float wavelength = 6.0f * (float)(1 << s);
        const float frequency = 1.0f / wavelength;
        const float l = logf(radius / frequency);
        float value = expf(-(l * l) / (2.0f * (ls * ls))) * lowpass;
        if (x == 0 && y == 0) value = 0.0f;
        lg[s * N + idx] = value;
    }
    /* fsim_angular.glsl:24-46 */
    const float pi = 3.141592653589f;
    const float thetaSigma = pi / 4.0f / 1.2f;
    const float theta = atan2f(-v, u);
    const float st = sinf(theta), ct = cosf(theta);
    for (int o = 0; o < 4; ++o) {
        const float angle = ((float)o * pi) / 4.0f;
        const float ds = st * cosf(angle) - ct * sinf(angle);
        const float dc = ct * cosf(angle) + st * sinf(angle);
        const float dtheta = fabsf(atan2f(ds, dc));
        ang[o * N + idx] = expf(-(dtheta * dtheta) / (2.0f * (thetaSigma * thetaSigma)));
    }
    /* fsim_filter_combinations.glsl:51: H_z = gabor * angular, z = 4 o + s */
    for (int z = 0; z < 16; ++z) bank[z * N + idx] = lg[(z & 3) * N + idx] * ang[(z >> 2) * N + idx];
}

/* Partial sums for EM_n[o] = sum_k (LG_0 * A_o)^2 over every bin (F10). */
__global__ __launch_bounds__(kBlock) void fsim_emn_kernel(const float* lg, const float* ang,
                                                          long N, double* partial) {
    __shared__ double sh[kBlock];
    const int o = blockIdx.y;
    double acc = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long)gridDim.x * blockDim.x) {
        float hv = lg[i] * ang[o * N + i];
        acc += (double)hv * (double)hv;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[o * gridDim.x + blockIdx.x] = sh[0];
}

/* Plane z = 4 o + s holds (LG_s * A_o, 0) (fsim_filter_combinations.glsl:51-52). */
__global__ __launch_bounds__(kBlock) void fsim_bank_planes_kernel(const float* lg,
                                                                  const float* ang, long N,
                                                                  float2* planes) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const int z = blockIdx.y, s = z & 3, o = z >> 2;
    planes[z * N + idx] = make_float2(lg[s * N + idx] * ang[o * N + idx], 0.0f);
}

/* est (F14): the reference multiplies by N and only covers complex elements
 * t < ceil(N/2) (fsim_mult_filters.glsl:30-33 bounds x = 2*gid by N); the
 * other half of its buffer is never written and is read as zero here. */
__global__ __launch_bounds__(kBlock) void fsim_est_kernel(const float2* h, long N, long half,
                                                          double* partial) {
    __shared__ double sh[2][kBlock];
    const int o = blockIdx.y;
    double a2 = 0, cr = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < half;
         t += (long)gridDim.x * blockDim.x) {
        double A[4];
        for (int s = 0; s < 4; ++s) A[s] = (double)h[(4 * o + s) * N + t].x;
        double s2 = 0, c = 0;
        for (int s = 0; s < 4; ++s) s2 += A[s] * A[s] * (double)N;
        for (int i = 0; i < 3; ++i)
            for (int j = i + 1; j < 4; ++j) c += A[i] * A[j] * (double)N;
        a2 += s2;
        cr += c;
    }
    sh[0][threadIdx.x] = a2;
    sh[1][threadIdx.x] = cr;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[(2 * o) * gridDim.x + blockIdx.x] = sh[0][0];
        partial[(2 * o + 1) * gridDim.x + blockIdx.x] = sh[1][0];
    }
}

const FsimSizePlan& Ctx::fsim_plan(int wd, int hd) {
    auto key = std::make_pair(wd, hd);
    auto it = fsim_plans_.find(key);
    if (it != fsim_plans_.end()) return it->second;
    FsimSizePlan P;
    P.wd = wd;
    P.hd = hd;
    P.N = (long)wd * hd;
    const long N = P.N;
    IQM_HIP(hipMalloc(&P.lg, sizeof(float) * 4 * N));
    IQM_HIP(hipMalloc(&P.ang, sizeof(float) * 4 * N));
    IQM_HIP(hipMalloc(&P.bank, sizeof(float) * 16 * N));
    launch("fsim_tables", fsim_tables_kernel, dim3(ceil_div(N, kBlock)), dim3(kBlock), 0, P.lg,
           P.ang, P.bank, wd, hd);

    const int nb = 64;
    double* partial = (double*)workspace(WS_PLAN_TMP, sizeof(double) * 8 * nb + sizeof(float2) * 16 * N + 256);
    float2* planes = (float2*)((char*)partial + ((sizeof(double) * 8 * nb + 255) / 256) * 256);
    launch("fsim_emn", fsim_emn_kernel, dim3(nb, 4), dim3(kBlock), 0, (const float*)P.lg,
           (const float*)P.ang, N, partial);
    std::vector<double> hp(8 * nb);
    IQM_HIP(hipMemcpyAsync(hp.data(), partial, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, stream_));
    IQM_HIP(hipStreamSynchronize(stream_));
    for (int o = 0; o < 4; ++o) {
        double s = 0;
        for (int b = 0; b < nb; ++b) s += hp[o * nb + b];
        P.em_n[o] = s;
    }
    launch("fsim_bank_planes", fsim_bank_planes_kernel, dim3(ceil_div(N, kBlock), 16),
           dim3(kBlock), 0, (const float*)P.lg, (const float*)P.ang, N, planes);
    fft2d_dev(planes, wd, hd, 16, +1, true);
    launch("fsim_est", fsim_est_kernel, dim3(nb, 4), dim3(kBlock), 0, (const float2*)planes, N,
           (N + 1) / 2, partial);
    IQM_HIP(hipMemcpyAsync(hp.data(), partial, sizeof(double) * 8 * nb, hipMemcpyDeviceToHost, stream_));
    IQM_HIP(hipStreamSynchronize(stream_));
    for (int j = 0; j < 8; ++j) {
        double s = 0;
        for (int b = 0; b < nb; ++b) s += hp[j * nb + b];
        P.est[j] = s;
    }
    return fsim_plans_.emplace(key, P).first->second;
}


/* ---- same-size stage kernels (bodies in fsim_body.hpp) ------------------ */

template <int FT>
__global__ __launch_bounds__(kBlock) void fsim_prep2_kernel(const uint8_t* __restrict__ img,
                                                            const uint8_t* __restrict__ ref, long pair_stride,
                                                            int w, int h, int ch, int wd, int hd, int cols,
                                                            float* __restrict__ yiq, unsigned* __restrict__ hist,
                                                            const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    fsim_prep2_body<FT>(img, ref, pair_stride, w, h, ch, wd, hd, cols, yiq, hist, blockIdx.x, blockIdx.y,
                        blockIdx.z);
}

template <int FT>
__global__ __launch_bounds__(kBlock) void fsim_prep_kernel(const uint8_t* __restrict__ img,
                                                           const uint8_t* __restrict__ ref,
                                                           long pair_stride, int w, int h,
                                                           int ch, int F, int wd, int hd, int cols, int prb,
                                                           float* __restrict__ yiq,
                                                           unsigned* __restrict__ hist,
                                                           const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    fsim_prep_body<FT>(img, ref, pair_stride, w, h, ch, F, wd, hd, cols, prb, yiq, hist, blockIdx.x, blockIdx.y,
                       blockIdx.z);
}

__global__ __launch_bounds__(1024) void fsim_select_kernel(const float* __restrict__ keys,
                                                           const unsigned* __restrict__ hist,
                                                           float* __restrict__ thr,
                                                           double* __restrict__ stats, long N,
                                                           NoiseArgs na, const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    fsim_select_body(keys, hist, thr, stats, N, na, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(kBlock) void fsim_pcfinal_kernel(
    const float* __restrict__ yiq, float* __restrict__ grad, const float* __restrict__ an_e,
    const float* __restrict__ thr, double* __restrict__ partial, int wd, int hd, float* simc_map,
    float* pc_map, float* pcm_map, float* sim_map, const float* __restrict__ pc_src, const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    fsim_pcfinal_body(yiq, grad, an_e, thr, partial, wd, hd, simc_map, pc_map, pcm_map, sim_map, pc_src,
                      blockIdx.x, blockIdx.y, gridDim.x, gridDim.y, blockIdx.z);
}

__global__ __launch_bounds__(kBlock) void fsim_finalize_kernel(const double* __restrict__ partial,
                                                               int nblk, double* __restrict__ res,
                                                               const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    fsim_finalize_body(partial, nblk, res, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void fsim_scores_kernel(const double* __restrict__ res, int n,
                                                             double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out[2 * i] = res[(long)i * 8 + 3];
        out[2 * i + 1] = res[(long)i * 8 + 4];
    }
}

/* ---- host orchestration ------------------------------------------------- */

Ctx::FsimLayout Ctx::fsim_layout(long N, int n, char* base) const {
    FsimRegions r = fsim_regions(N, n, last_chunk_, last_nblk_);
    FsimLayout L;
    L.N = N;
    L.yiq = (float*)(base + r.yiq);
    L.grad = (float*)(base + r.grad);
    L.an_e = (float*)(base + r.an_e);
    L.keys = (float*)(base + r.keys);
    L.stats = (double*)(base + r.stats);
    L.res = (double*)(base + r.res);
    return L;
}

void Ctx::launch_prep(const uint8_t* d_img, const uint8_t* d_ref, long pair_stride, int w, int h, int ch,
                      int F, int wd, int hd, int n, float* yiq, unsigned* hist) {
    IQM_CHECK(F <= kPrepMaxRows, IQM_ERR_UNSUPPORTED, "image too large for the FSIM prep tile");
    int prb = std::max(1, env_int("IQM_FSIM_PREP_RB", fsim_tune(wd, hd).prep_rb));
    while (prb > 1 && prb * F > kPrepMaxRows) --prb;
    const int cols = prep_cols(F, ch, prb);
    const size_t bytes = prep_lds_bytes(F, ch, cols, prb);
    IQM_CHECK(bytes <= (size_t)kMaxLdsBytes, IQM_ERR_UNSUPPORTED, "image too large for the FSIM prep tile");
    /* boxes of 4..8 over rows a multiple of 16 bytes apart: the register
     * column-sum kernel (each image's alignment is taken per workgroup) */
    if (F >= env_int("IQM_FSIM_PREP2_MINF", kPrep2MinF) && F <= 8 && ((long)w * ch) % 16 == 0 &&
        env_int("IQM_FSIM_PREP2", 1)) {
        const int c2 = std::min(kPrep2Cols, wd);
        const dim3 g2(ceil_div(wd, c2), hd, n);
        const size_t b2 = prep2_lds_bytes(F, ch, c2);
        switch (F) {
            case 2: launch("fsim_prep", fsim_prep2_kernel<2>, g2, dim3(kBlock), b2, d_img, d_ref, pair_stride, w, h, ch, wd, hd, c2, yiq, hist, gate_); break;
            case 3: launch("fsim_prep", fsim_prep2_kernel<3>, g2, dim3(kBlock), b2, d_img, d_ref, pair_stride, w, h, ch, wd, hd, c2, yiq, hist, gate_); break;
            case 4: launch("fsim_prep", fsim_prep2_kernel<4>, g2, dim3(kBlock), b2, d_img, d_ref, pair_stride, w, h, ch, wd, hd, c2, yiq, hist, gate_); break;
            case 5: launch("fsim_prep", fsim_prep2_kernel<5>, g2, dim3(kBlock), b2, d_img, d_ref, pair_stride, w, h, ch, wd, hd, c2, yiq, hist, gate_); break;
            case 6: launch("fsim_prep", fsim_prep2_kernel<6>, g2, dim3(kBlock), b2, d_img, d_ref, pair_stride, w, h, ch, wd, hd, c2, yiq, hist, gate_); break;
            case 7: launch("fsim_prep", fsim_prep2_kernel<7>, g2, dim3(kBlock), b2, d_img, d_ref, pair_stride, w, h, ch, wd, hd, c2, yiq, hist, gate_); break;
            default: launch("fsim_prep", fsim_prep2_kernel<8>, g2, dim3(kBlock), b2, d_img, d_ref, pair_stride, w, h, ch, wd, hd, c2, yiq, hist, gate_); break;
        }
        return;
    }
    const dim3 grid(ceil_div(wd, cols), ceil_div(hd, prb), n);
    switch (F) {
#define IQM_PREP_CASE(f)                                                                                      \
    case f:                                                                                                   \
        launch("fsim_prep", fsim_prep_kernel<f>, grid, dim3(kBlock), bytes, d_img, d_ref, pair_stride, w, h, ch, \
               F, wd, hd, cols, prb, yiq, hist, gate_);                                                       \
        break;
        IQM_PREP_CASE(1) IQM_PREP_CASE(2) IQM_PREP_CASE(3) IQM_PREP_CASE(4)
        IQM_PREP_CASE(5) IQM_PREP_CASE(6) IQM_PREP_CASE(7) IQM_PREP_CASE(8)
#undef IQM_PREP_CASE
        default:
            launch("fsim_prep", fsim_prep_kernel<0>, grid, dim3(kBlock), bytes, d_img, d_ref, pair_stride, w, h, ch,
                   F, wd, hd, cols, prb, yiq, hist, gate_);
    }
}

double* Ctx::fsim_batch(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch,
                        long pair_stride, int n, const FsimBatchOut& out) {
    int F = std::max(1, (int)std::round(std::min(w, h) / 256.0));
    const int wd = (int)std::round((float)w / (float)F);
    const int hd = (int)std::round((float)h / (float)F);
    IQM_CHECK(wd >= 2 && hd >= 2, IQM_ERR_TOO_SMALL, "image too small for FSIM");
    IQM_CHECK(fsim_planes_fit32((long)wd * hd), IQM_ERR_UNSUPPORTED, "downscaled image too large for FSIM");
    const FsimSizePlan& plan = fsim_plan(wd, hd);
    const LinePlan& Px = line_plan_for(wd, LINE_ROW);
    const LinePlan& Py = line_plan_for(hd, LINE_COL);
    mark("FFT library initialized"); /* fsim.cpp:95: the per-size plans are ready */
    const long N = (long)wd * hd;
    const int nblk = ceil_div(wd, kPcTx) * ceil_div(hd, kPcTy);
    const FsimTune tune = fsim_tune(wd, hd);
    const FsimTiles t = pick_tiles(*this, Px, Py, N, tune);
    const int chunk = std::min(t.chunk, n);
    last_chunk_ = chunk;
    last_nblk_ = nblk;
    FsimRegions r = fsim_regions(N, n, chunk, nblk);
    char* base = (char*)workspace(WS_FSIM, r.total);
    fsim_ws_ = base;
    float* yiq = (float*)(base + r.yiq);
    float* grad = (float*)(base + r.grad);
    float2* A = (float2*)(base + r.A);
    float2* T = (float2*)(base + r.T);
    float* an_e = (float*)(base + r.an_e);
    float* keys = (float*)(base + r.keys);
    unsigned* hist = (unsigned*)(base + r.hist);
    float* thr = (float*)(base + r.thr);
    double* stats = (double*)(base + r.stats);
    double* partial = (double*)(base + r.partial);
    double* res = (double*)(base + r.res);

    /* F2 */
    launch_prep(d_img, d_ref, pair_stride, w, h, ch, F, wd, hd, n, yiq, hist);
    /* row half of F5 (F3 is evaluated in fsim_pcfinal) */
    const bool tl = env_int("IQM_FSIM_GENERIC", tune.two_level ? 0 : 1) == 0;
    const size_t tl_budget = std::min<size_t>(lds_per_block_, (size_t)env_int("IQM_FSIM_TL_LDS", 160 * 1024));
    if (tl && tl_supported(wd)) {
        TlGradArgs a{yiq, A, Px.tw, hd, n, env_int("IQM_FSIM_TL_GRB", tune.fft_rb), tl_budget, gate_};
        tl_rowfft(*this, wd, a);
    } else {
        launch("fsim_rowfft", fsim_rowfft_kernel, dim3(ceil_div(hd, t.grad_rb), n), dim3(kBlock),
               line_lds_bytes(Px, t.grad_rb, 2), (const float*)yiq, A, wd, hd, t.grad_rb, Px, gate_);
    }
    /* per chunk: column half of F5, F9, F11, then the row half of F11 + F12 */
    const int nbx = ceil_div(wd, t.col_bx);
    const int nbx_pad = ceil_div(nbx, xcds_) * xcds_;
    const float invN = (float)(1.0 / (double)N);
    const size_t row_sh = rowpass_lds_bytes(Px, t.row_rb, wd);
    const bool tl_col = tl && tl_supported(hd);
    const bool tl_row = tl && tl_supported(wd);
    for (int p0 = 0; p0 < n; p0 += chunk) {
        const int np = std::min(chunk, n - p0);
        float2* Ach = A + (long)p0 * 2 * N;
        if (tl_col) {
            TlColArgs a{Ach, T, plan.bank, Py.tw, wd, np, env_int("IQM_FSIM_TL_BX", tune.col_bx), xcds_, tl_budget,
                        gate_};
            tl_colpass(*this, hd, a);
        } else {
            launch("fsim_colpass", fsim_colpass_kernel, dim3(nbx_pad * 2 * np), dim3(kBlockFft),
                   line_lds_bytes(Py, t.col_bx, colpass_bufs(Py)), Ach, T, (const float*)plan.bank, wd, hd,
                   t.col_bx, nbx_pad, Py, gate_);
        }
        if (tl_row) {
            TlRowArgs a{T, an_e + (long)p0 * 2 * kAnePlanes * N, keys + (long)p0 * 8 * N,
                        hist + (long)p0 * 8 * kHistBins, Px.tw, hd, np, env_int("IQM_FSIM_TL_RB", tune.row_rb),
                        invN, tl_budget, gate_};
            tl_rowpass(*this, wd, a);
        } else {
            launch("fsim_rowpass", fsim_rowpass_kernel, dim3(ceil_div(hd, t.row_rb), 2, np),
                   dim3(kBlockFft), row_sh, (const float2*)T, an_e + (long)p0 * 2 * kAnePlanes * N,
                   keys + (long)p0 * 8 * N, hist + (long)p0 * 8 * kHistBins, wd, hd, t.row_rb, Px,
                   invN, gate_);
        }
    }
    mark("pre median work recorded"); /* fsim.cpp:138 */
    /* F13 + thresholds */
    NoiseArgs na;
    for (int i = 0; i < 4; ++i) na.em_n[i] = plan.em_n[i];
    for (int i = 0; i < 8; ++i) na.est[i] = plan.est[i];
    launch("fsim_select", fsim_select_kernel, dim3(8, n), dim3(1024), sizeof(unsigned) * kSelectCap,
           (const float*)keys,
           (const unsigned*)hist, thr, stats, N, na, gate_);
    mark("noise powers computed"); /* fsim.cpp:141 */
    /* F15 + F16 + F17 */
    launch("fsim_pcfinal", fsim_pcfinal_kernel, dim3(ceil_div(wd, kPcTx), ceil_div(hd, kPcTy), n),
           dim3(kBlock), 0, (const float*)yiq,
           out.want_grad ? grad : (float*)nullptr, (const float*)an_e, (const float*)thr, partial, wd,
           hd, out.d_simc, out.d_pc, out.d_pcm, out.d_sim, (const float*)nullptr, gate_);
    launch("fsim_finalize", fsim_finalize_kernel, dim3(n), dim3(kBlock), 0, (const double*)partial,
           nblk, res, gate_);
    return res;
}

void Ctx::fsim_scores(const double* res, int n, double* d_scores) {
    launch("fsim_scores", fsim_scores_kernel, dim3(ceil_div(n, kBlock)), dim3(kBlock), 0, res, n, d_scores);
}

double* Ctx::fsim_batch_pcfn(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch,
                             long pair_stride, int n, const FsimBatchOut& out, const PcFn& pc_fn) {
    int F = std::max(1, (int)std::round(std::min(w, h) / 256.0));
    const int wd = (int)std::round((float)w / (float)F);
    const int hd = (int)std::round((float)h / (float)F);
    IQM_CHECK(wd >= 2 && hd >= 2, IQM_ERR_TOO_SMALL, "image too small for FSIM");
    IQM_CHECK(fsim_planes_fit32((long)wd * hd), IQM_ERR_UNSUPPORTED, "downscaled image too large for FSIM");
    const long N = (long)wd * hd;
    const int nblk = ceil_div(wd, kPcTx) * ceil_div(hd, kPcTy);
    size_t off = 0;
    auto take = [&](size_t b) { size_t o = off; off += align256(b); return o; };
    const size_t o_yiq = take(sizeof(float) * 6 * N * n), o_pc = take(sizeof(float) * 2 * N * n),
                 o_hist = take(sizeof(unsigned) * 8 * kHistBins * n), o_part = take(sizeof(double) * 4 * (size_t)nblk * n),
                 o_res = take(sizeof(double) * 8 * n);
    char* base = (char*)workspace(WS_SUR, off);
    float* yiq = (float*)(base + o_yiq);
    float* pc = (float*)(base + o_pc);
    double* partial = (double*)(base + o_part);
    double* res = (double*)(base + o_res);
    launch_prep(d_img, d_ref, pair_stride, w, h, ch, F, wd, hd, n, yiq, (unsigned*)(base + o_hist));
    pc_fn(yiq, N, n, pc);
    launch("fsim_pcfinal", fsim_pcfinal_kernel, dim3(ceil_div(wd, kPcTx), ceil_div(hd, kPcTy), n),
           dim3(kBlock), 0, (const float*)yiq, (float*)nullptr, (const float*)nullptr, (const float*)nullptr,
           partial, wd, hd, out.d_simc, out.d_pc, out.d_pcm, out.d_sim, (const float*)pc, gate_);
    launch("fsim_finalize", fsim_finalize_kernel, dim3(n), dim3(kBlock), 0, (const double*)partial, nblk, res,
           gate_);
    return res;
}

}  // namespace iqm
