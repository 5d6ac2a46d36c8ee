// Generic batched 2-D FFT kernels (rows, then columns) and the host planner.
// They back iqm_fft2d and the per-size constants of FSIM; the FSIM hot path
// uses fused variants of the same LDS line transform (fsim.hip).
#include "fft_line.hpp"
#include "ctx.hpp"

#include <algorithm>
#include <cmath>
#include <complex>

namespace iqm {

namespace {

/* recursive mixed-radix DFT in double (plan-time only), sign -1: splits by
 * the smallest prime factor p, direct O(n^2) sums for a prime length */
void host_dft(std::vector<std::complex<double>>& a) {
    const size_t n = a.size();
    if (n <= 1) return;
    const double two_pi = 6.283185307179586476925286766559;
    size_t p = 2;
    while (p * p <= n && n % p) ++p;
    if (n % p) p = n;
    if (p == n) {
        std::vector<std::complex<double>> o(n);
        for (size_t k = 0; k < n; ++k) {
            std::complex<double> acc = 0.0;
            for (size_t j = 0; j < n; ++j) {
                const double ang = -two_pi * (double)((j * k) % n) / (double)n;
                acc += a[j] * std::complex<double>(std::cos(ang), std::sin(ang));
            }
            o[k] = acc;
        }
        a.swap(o);
        return;
    }
    const size_t q = n / p;
    std::vector<std::vector<std::complex<double>>> sub(p, std::vector<std::complex<double>>(q));
    for (size_t r = 0; r < p; ++r)
        for (size_t k = 0; k < q; ++k) sub[r][k] = a[k * p + r];
    for (auto& v : sub) host_dft(v);
    for (size_t k = 0; k < q; ++k)
        for (size_t s = 0; s < p; ++s) {
            const size_t kk = k + q * s;
            std::complex<double> acc = 0.0;
            for (size_t r = 0; r < p; ++r) {
                const double ang = -two_pi * (double)((r * kk) % n) / (double)n;
                acc += sub[r][k] * std::complex<double>(std::cos(ang), std::sin(ang));
            }
            a[kk] = acc;
        }
}

/* smallest length >= n that the direct transforms plan: the smallest with a
 * two-level split when one exists up to 32 * 32, else the smallest the
 * Stockham schedule plans */
int next_plannable(int n) {
    int a, b;
    for (int m = n; m <= 32 * 32; ++m)
        if (tl2_split(m, a, b)) return m;
    std::vector<int> rad;
    for (int m = n;; ++m)
        if (plan_radices(m, rad)) return m;
}

template <class T>
T* upload_table(const std::vector<T>& v) {
    T* d = nullptr;
    IQM_HIP(hipMalloc(&d, sizeof(T) * v.size()));
    IQM_HIP(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return d;
}

std::vector<float2> twiddle_table(int n) {
    std::vector<float2> tw(n);
    const double two_pi = 6.283185307179586476925286766559;
    for (int j = 0; j < n; ++j) {
        const double a = two_pi * (double)j / (double)n;
        tw[j] = make_float2((float)std::cos(a), (float)(-std::sin(a)));
    }
    return tw;
}

}  // namespace

/* Inner length of a plan: the length its two-level split factors. */
static int inner_length(const LinePlan& P) { return P.rader ? P.rp - 1 : P.blue ? P.m : P.n; }

std::vector<int> line_split_candidates(const LinePlan& P) {
    std::vector<int> c;
    const int L = inner_length(P);
    const bool has = P.rader || (P.blue ? P.tm1 > 0 : P.tn1 > 0);
    if (!has) return c;
    for (int a = 2; a <= 32 && a < L; ++a)
        if (L % a == 0 && tl2_size(a) && tl2_size(L / a)) c.push_back(a);
    return c;
}

const LinePlan& Ctx::line_plan_for(int n, int role) {
    auto it = line_tunes_.find({n, role});
    return line_plan(n, it == line_tunes_.end() ? 0 : it->second.split1).plan;
}

const LinePlanHost& Ctx::line_plan(int n, int split1) {
    auto it = line_plans_.find({n, split1});
    if (it != line_plans_.end()) return it->second;
    IQM_CHECK(n >= 1 && n <= 8192, IQM_ERR_UNSUPPORTED,
              "FFT length " + std::to_string(n) + " outside [1, 8192]");
    if (split1 != 0) {
        /* the default plan with another split of its inner length; its
         * tables stay with the default plan (a Rader variant owns its input
         * slot table, which depends on the sub-line pitch) */
        const LinePlan& B = line_plan(n, 0).plan;
        const std::vector<int> cand = line_split_candidates(B);
        IQM_CHECK(std::find(cand.begin(), cand.end(), split1) != cand.end(), IQM_ERR_INVALID,
                  "FFT length " + std::to_string(n) + ": " + std::to_string(split1) +
                      " is not a factor of a two-level split of its plan");
        LinePlanHost lv;
        LinePlan& P = lv.plan;
        P = B;
        const int L = inner_length(B), s2 = L / split1;
        if (P.rader) {
            P.tm1 = split1;
            P.tm2 = s2;
            P.rpitch = rader_pitch(P.rp, split1, s2);
            P.lds_len = std::max(n, P.rq * P.rpitch);
            std::vector<int> rin, rout;
            rader_tables(P.rp, P.rpitch, rin, rout);
            int* ri = upload_table(rin);
            P.rin = ri;
            lv.owned = {ri};
        } else if (P.blue) {
            P.tm1 = split1;
            P.tm2 = s2;
            P.lds_len = std::max({P.m, split1 * (s2 + 1), s2 * (split1 + 1)});
        } else {
            P.tn1 = split1;
            P.tn2 = s2;
            P.lds_len = std::max(n, split1 * (s2 + 1));
        }
        return line_plans_.emplace(std::make_pair(n, split1), lv).first->second;
    }
    LinePlanHost lp;
    LinePlan& P = lp.plan;
    P = LinePlan{};
    P.n = n;
    std::vector<int> rad;
    const bool radices = plan_radices(n, rad);
    int s1, s2, rq, rp, m1, m2;
    if (radices || tl2_split(n, s1, s2)) {
        /* a direct plan: the two-level split where there is one, else the
         * Stockham radices */
        P.nst = radices ? (int)rad.size() : 0;
        for (int i = 0; i < kMaxStages; ++i) P.rad[i] = i < P.nst ? rad[i] : 1;
        float2* tw = upload_table(twiddle_table(n));
        P.tw = tw;
        P.lds_len = n;
        if (tl2_split(n, P.tn1, P.tn2)) P.lds_len = std::max(n, P.tn1 * (P.tn2 + 1));
        lp.owned = {tw};
    } else if (rader_split(n, rq, rp, m1, m2)) {
        P.rader = 1;
        P.rq = rq;
        P.rp = rp;
        P.m = rp - 1;
        P.tm1 = m1;
        P.tm2 = m2;
        P.rpitch = rader_pitch(rp, m1, m2);
        P.lds_len = std::max(n, rq * P.rpitch);
        std::vector<int> rin, rout;
        rader_tables(rp, P.rpitch, rin, rout);
        /* convolution kernels v_c = exp(S 2 pi i g^c / p) of the forward
         * (S = -1) and inverse (S = +1) transform, as V = FFT(v) / (p - 1) */
        const double two_pi = 6.283185307179586476925286766559;
        std::vector<float2> vf(rp - 1), vi(rp - 1);
        for (int dir = -1; dir <= 1; dir += 2) {
            std::vector<std::complex<double>> v(rp - 1);
            for (int c = 0; c < rp - 1; ++c) {
                const double a = dir * two_pi * (double)rout[c] / (double)rp;
                v[c] = std::complex<double>(std::cos(a), std::sin(a));
            }
            host_dft(v);
            std::vector<float2>& dst = dir < 0 ? vf : vi;
            for (int c = 0; c < rp - 1; ++c)
                dst[c] = make_float2((float)(v[c].real() / (rp - 1)), (float)(v[c].imag() / (rp - 1)));
        }
        float2* tw = upload_table(twiddle_table(n));
        float2* mtw = upload_table(twiddle_table(rp - 1));
        float2* g1 = upload_table(vf);
        float2* g2 = upload_table(vi);
        int* ri = upload_table(rin);
        int* ro = upload_table(rout);
        P.tw = tw;
        P.mtw = mtw;
        P.gf = g1;
        P.gi = g2;
        P.rin = ri;
        P.rout = ro;
        lp.owned = {tw, mtw, g1, g2, ri, ro};
    } else {
        /* any plannable m >= 2n - 1 holds the linear convolution */
        const int m = next_plannable(2 * n - 1);
        std::vector<int> mr;
        const bool mradices = plan_radices(m, mr);
        IQM_CHECK(mradices || tl2_split(m, s1, s2), IQM_ERR_UNSUPPORTED, "Bluestein length not plannable");
        P.blue = 1;
        P.m = m;
        P.mst = mradices ? (int)mr.size() : 0;
        for (int i = 0; i < kMaxStages; ++i) P.mrad[i] = i < P.mst ? mr[i] : 1;
        const double pi = 3.14159265358979323846;
        std::vector<float2> chirp(n);
        std::vector<std::complex<double>> cd(n);
        for (int j = 0; j < n; ++j) {
            const long long q = ((long long)j * j) % (2LL * n);
            const double a = pi * (double)q / (double)n;
            cd[j] = std::complex<double>(std::cos(a), -std::sin(a));
            chirp[j] = make_float2((float)cd[j].real(), (float)cd[j].imag());
        }
        /* forward kernel g_d = conj(c_d), inverse kernel g_d = c_d, d in (-n, n) */
        std::vector<std::complex<double>> gf(m, 0.0), gi(m, 0.0);
        for (int d = 0; d < n; ++d) {
            gf[d] = std::conj(cd[d]);
            gi[d] = cd[d];
            if (d) {
                gf[m - d] = gf[d];
                gi[m - d] = gi[d];
            }
        }
        host_dft(gf);
        host_dft(gi);
        std::vector<float2> gff(m), gif(m);
        for (int j = 0; j < m; ++j) {
            gff[j] = make_float2((float)gf[j].real(), (float)gf[j].imag());
            gif[j] = make_float2((float)gi[j].real(), (float)gi[j].imag());
        }
        float2* mtw = upload_table(twiddle_table(m));
        float2* ch = upload_table(chirp);
        float2* g1 = upload_table(gff);
        float2* g2 = upload_table(gif);
        P.mtw = mtw;
        P.chirp = ch;
        P.gf = g1;
        P.gi = g2;
        P.lds_len = m;
        /* the forward exchange area (tm1 rows of tm2 + 1) and the inverse
         * one (the split the other way round) */
        if (tl2_split(m, P.tm1, P.tm2)) P.lds_len = std::max({m, P.tm1 * (P.tm2 + 1), P.tm2 * (P.tm1 + 1)});
        lp.owned = {mtw, ch, g1, g2};
    }
    return line_plans_.emplace(std::make_pair(n, 0), lp).first->second;
}

/* LDS bytes of a kernel holding `bufs` buffers of `lines` lines of plan P,
 * plus the staged twiddles. */
size_t line_lds_bytes(const LinePlan& P, int lines, int bufs) {
    return (size_t)bufs * lines * lds_pitch(P.lds_len) * sizeof(float2) +
           (size_t)lds_twiddles(P) * sizeof(float2);
}

/* Largest line count in [1, cap] whose LDS footprint fits the budget; 0 if
 * even one line does not fit. */
int lines_that_fit(const LinePlan& P, int bufs, size_t budget, int cap) {
    int best = 0;
    for (int l = 1; l <= cap; ++l)
        if (line_lds_bytes(P, l, bufs) <= budget) best = l;
    return best;
}

/* In-place transform of rows (length w, contiguous) of every plane; each
 * workgroup handles `lines` consecutive rows of one plane. */
template <int S>
__global__ __launch_bounds__(kBlock) void fft_rows_kernel(float2* data, int w, int h,
                                                          long plane_stride, int lines,
                                                          LinePlan P, float scale) {
    extern __shared__ float2 lds[];
    const int ld = lds_pitch(P.lds_len);
    float2* b0 = lds;
    float2* b1 = lds + lines * ld;
    const float2* tw = stage_twiddles(P, lds + 2 * lines * ld);
    float2* base = data + (long)blockIdx.y * plane_stride;
    const int y0 = blockIdx.x * lines;
    const int nl = min(lines, h - y0);
    for (int idx = threadIdx.x; idx < nl * w; idx += blockDim.x) {
        int l = idx / w, x = idx - l * w;
        b0[l * ld + x] = base[(long)(y0 + l) * w + x];
    }
    __syncthreads();
    float2* r = line_fft<S>(b0, b1, ld, nl, P, tw);
    for (int idx = threadIdx.x; idx < nl * w; idx += blockDim.x) {
        int l = idx / w, x = idx - l * w;
        base[(long)(y0 + l) * w + x] = cscale(r[l * ld + x], scale);
    }
}

/* In-place transform of columns (length h, stride w) of every plane; each
 * workgroup handles `cols` adjacent columns. */
template <int S>
__global__ __launch_bounds__(kBlock) void fft_cols_kernel(float2* data, int w, int h,
                                                          long plane_stride, int cols,
                                                          LinePlan P, float scale) {
    extern __shared__ float2 lds[];
    const int ld = lds_pitch(P.lds_len);
    float2* b0 = lds;
    float2* b1 = lds + cols * ld;
    const float2* tw = stage_twiddles(P, lds + 2 * cols * ld);
    float2* base = data + (long)blockIdx.y * plane_stride;
    const int x0 = blockIdx.x * cols;
    const int nc = min(cols, w - x0);
    for (int idx = threadIdx.x; idx < nc * h; idx += blockDim.x) {
        int y = idx / nc, c = idx - y * nc;
        b0[c * ld + y] = base[(long)y * w + x0 + c];
    }
    __syncthreads();
    float2* r = line_fft<S>(b0, b1, ld, nc, P, tw);
    for (int idx = threadIdx.x; idx < nc * h; idx += blockDim.x) {
        int y = idx / nc, c = idx - y * nc;
        base[(long)y * w + x0 + c] = cscale(r[c * ld + y], scale);
    }
}

void Ctx::fft2d_dev(float2* d, int w, int h, int planes, int dir, bool normalize) {
    const LinePlan& px = line_plan(w).plan;
    const LinePlan& py = line_plan(h).plan;
    auto fit = [&](const LinePlan& P) {
        const int l = lines_that_fit(P, 2, std::min<size_t>(lds_per_block_, 65536), 8);
        return l > 0 ? l : lines_that_fit(P, 2, lds_per_block_, 8);
    };
    const int lr = fit(px), lc = fit(py);
    IQM_CHECK(lr > 0 && lc > 0, IQM_ERR_UNSUPPORTED, "FFT line too long for LDS");
    const long ps = (long)w * h;
    const float s = normalize ? (float)(1.0 / ((double)w * (double)h)) : 1.0f;
    dim3 gr(ceil_div(h, lr), planes), gc(ceil_div(w, lc), planes);
    const size_t shr = line_lds_bytes(px, lr, 2), shc = line_lds_bytes(py, lc, 2);
    if (dir < 0) {
        launch("fft_rows_fwd", fft_rows_kernel<-1>, gr, dim3(kBlock), shr, d, w, h, ps, lr, px, 1.0f);
        launch("fft_cols_fwd", fft_cols_kernel<-1>, gc, dim3(kBlock), shc, d, w, h, ps, lc, py, s);
    } else {
        launch("fft_rows_inv", fft_rows_kernel<1>, gr, dim3(kBlock), shr, d, w, h, ps, lr, px, 1.0f);
        launch("fft_cols_inv", fft_cols_kernel<1>, gc, dim3(kBlock), shc, d, w, h, ps, lc, py, s);
    }
}

}  // namespace iqm
