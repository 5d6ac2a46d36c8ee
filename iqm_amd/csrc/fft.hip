// Generic batched 2-D FFT kernels (rows, then columns) and the host planner.
// They back iqm_fft2d and the per-size constants of FSIM; the FSIM hot path
// uses fused variants of the same LDS line transform (fsim.hip).
#include "fft.hpp"
#include "ctx.hpp"

#include <cmath>

namespace iqm {

bool plan_radices(int n, std::vector<int>& rad) {
    rad.clear();
    if (n < 1) return false;
    int r = n;
    while (r % 8 == 0) { rad.push_back(8); r /= 8; }
    while (r % 4 == 0) { rad.push_back(4); r /= 4; }
    while (r % 2 == 0) { rad.push_back(2); r /= 2; }
    for (int p : {3, 5, 7, 11, 13})
        while (r % p == 0) { rad.push_back(p); r /= p; }
    if (r != 1 || (int)rad.size() > kMaxStages) return false;
    return true;
}

/* In-place transform of rows (length w, contiguous) of `planes` planes;
 * each workgroup handles `lines` consecutive rows of one plane. */
template <int S>
__global__ __launch_bounds__(kBlock) void fft_rows_kernel(float2* data, int w, int h,
                                                          long plane_stride, int lines,
                                                          LinePlan P, float scale) {
    extern __shared__ float2 lds[];
    const int ld = lds_pitch(w);
    float2* b0 = lds;
    float2* b1 = lds + lines * ld;
    float2* base = data + (long)blockIdx.y * plane_stride;
    const int y0 = blockIdx.x * lines;
    const int nl = min(lines, h - y0);
    for (int idx = threadIdx.x; idx < nl * w; idx += blockDim.x) {
        int l = idx / w, x = idx - l * w;
        b0[l * ld + x] = base[(long)(y0 + l) * w + x];
    }
    __syncthreads();
    float2* r = lds_fft<S>(b0, b1, ld, nl, P);
    for (int idx = threadIdx.x; idx < nl * w; idx += blockDim.x) {
        int l = idx / w, x = idx - l * w;
        base[(long)(y0 + l) * w + x] = cscale(r[l * ld + x], scale);
    }
}

/* In-place transform of columns (length h, stride w) of `planes` planes;
 * each workgroup handles `cols` adjacent columns. */
template <int S>
__global__ __launch_bounds__(kBlock) void fft_cols_kernel(float2* data, int w, int h,
                                                          long plane_stride, int cols,
                                                          LinePlan P, float scale) {
    extern __shared__ float2 lds[];
    const int ld = lds_pitch(h);
    float2* b0 = lds;
    float2* b1 = lds + cols * ld;
    float2* base = data + (long)blockIdx.y * plane_stride;
    const int x0 = blockIdx.x * cols;
    const int nc = min(cols, w - x0);
    for (int idx = threadIdx.x; idx < nc * h; idx += blockDim.x) {
        int y = idx / nc, c = idx - y * nc;
        b0[c * ld + y] = base[(long)y * w + x0 + c];
    }
    __syncthreads();
    float2* r = lds_fft<S>(b0, b1, ld, nc, P);
    for (int idx = threadIdx.x; idx < nc * h; idx += blockDim.x) {
        int y = idx / nc, c = idx - y * nc;
        base[(long)y * w + x0 + c] = cscale(r[c * ld + y], scale);
    }
}

const LinePlanHost& Ctx::line_plan(int n) {
    auto it = line_plans_.find(n);
    if (it != line_plans_.end()) return it->second;
    LinePlanHost lp;
    std::vector<int> rad;
    IQM_CHECK(plan_radices(n, rad), IQM_ERR_UNSUPPORTED,
              "FFT length " + std::to_string(n) + " has a prime factor above 13");
    std::vector<float2> tw(n);
    const double two_pi = 6.283185307179586476925286766559;
    for (int j = 0; j < n; ++j) {
        double a = two_pi * (double)j / (double)n;
        tw[j] = make_float2((float)std::cos(a), (float)(-std::sin(a)));
    }
    float2* d = nullptr;
    IQM_HIP(hipMalloc(&d, sizeof(float2) * n));
    IQM_HIP(hipMemcpy(d, tw.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    lp.dev_tw = d;
    lp.plan.n = n;
    lp.plan.nst = (int)rad.size();
    for (int i = 0; i < kMaxStages; ++i) lp.plan.rad[i] = i < (int)rad.size() ? rad[i] : 1;
    lp.plan.tw = d;
    return line_plans_.emplace(n, lp).first->second;
}

int rows_per_block(int n) {
    int per = kMaxLdsBytes / (2 * lds_pitch(n) * (int)sizeof(float2));
    return per < 1 ? 0 : (per > 8 ? 8 : per);
}

void Ctx::fft2d_dev(float2* d, int w, int h, int planes, int dir, bool normalize) {
    const LinePlan& px = line_plan(w).plan;
    const LinePlan& py = line_plan(h).plan;
    const int lr = rows_per_block(w), lc = rows_per_block(h);
    IQM_CHECK(lr > 0 && lc > 0, IQM_ERR_UNSUPPORTED, "FFT line too long for LDS");
    const long ps = (long)w * h;
    const float s = normalize ? (float)(1.0 / ((double)w * (double)h)) : 1.0f;
    dim3 gr(ceil_div(h, lr), planes), gc(ceil_div(w, lc), planes);
    size_t shr = 2 * lr * lds_pitch(w) * sizeof(float2);
    size_t shc = 2 * lc * lds_pitch(h) * sizeof(float2);
    if (dir < 0) {
        launch("fft_rows_fwd", fft_rows_kernel<-1>, gr, dim3(kBlock), shr, d, w, h, ps, lr, px, 1.0f);
        launch("fft_cols_fwd", fft_cols_kernel<-1>, gc, dim3(kBlock), shc, d, w, h, ps, lc, py, s);
    } else {
        launch("fft_rows_inv", fft_rows_kernel<1>, gr, dim3(kBlock), shr, d, w, h, ps, lr, px, 1.0f);
        launch("fft_cols_inv", fft_cols_kernel<1>, gc, dim3(kBlock), shc, d, w, h, ps, lc, py, s);
    }
}

}  // namespace iqm
