/*
 * CPU oracle for the SSIM and FSIM/FSIMc path -- TEST INFRASTRUCTURE ONLY.
 * See iqm_oracle.h.  Each function names the reference file:line it restates.
 *
 * Templated on the arithmetic type T: double is the accuracy oracle, float
 * mirrors the reference's FP32 shader arithmetic.
 */
#include "iqm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* Constants exactly as the shaders spell them (ssim.glsl:9-10,
 * fsim_angular.glsl:8, fsim_phase_congruency.glsl:8). */
constexpr double REF_E = 2.71828182846;
constexpr double REF_PI = 3.141592653589;

template <class T>
struct Cx {
    T re, im;
};

template <class T>
inline Cx<T> cmul(Cx<T> a, Cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

/* ------------------------------------------------------------------------
 * 1-D DFT of any length.
 *
 * Stands in for VkFFT (fsim.cpp:370-437,527-536,541-550), which is an
 * exact-length C2C DFT: forward sign -1 unnormalised, inverse normalised by
 * the element count (fsim.cpp:420,533,547).  Mixed radix by recursive
 * decimation in time for lengths whose prime factors are <= 31, Bluestein's
 * chirp-z for the rest.  Twiddles are computed in double and rounded to T.
 * ---------------------------------------------------------------------- */
template <class T>
struct Fft1 {
    int n = 0;
    std::vector<int> radix;
    std::vector<Cx<T>> w;  /* w[j] = exp(-2 pi i j / n) */
    bool blue = false;
    int m = 0;
    std::unique_ptr<Fft1<T>> inner;
    std::vector<Cx<T>> chirp;      /* exp(-pi i j^2 / n), j < n */
    std::vector<Cx<T>> gfwd, ginv; /* FFT_m of the conjugate / plain chirp kernel */

    explicit Fft1(int n_) : n(n_) {
        int r = n, p = 2;
        std::vector<int> f;
        while (r > 1) {
            if (p * p > r) { f.push_back(r); break; }
            if (r % p == 0) { f.push_back(p); r /= p; } else { p += (p == 2) ? 1 : 2; }
        }
        int maxp = f.empty() ? 1 : *std::max_element(f.begin(), f.end());
        if (maxp > 31) {
            blue = true;
            m = 1;
            while (m < 2 * n - 1) m <<= 1;
            inner.reset(new Fft1<T>(m));
            chirp.resize(n);
            for (int j = 0; j < n; ++j) {
                long long q = ((long long)j * j) % (2LL * n);
                double a = REF_PI_EXACT() * (double)q / (double)n;
                chirp[j] = {(T)std::cos(a), (T)(-std::sin(a))};
            }
            /* kernel g_d = conj(chirp_d) for forward, chirp_d for inverse */
            std::vector<Cx<T>> g(m, Cx<T>{0, 0});
            for (int d = 0; d < n; ++d) {
                g[d] = {chirp[d].re, -chirp[d].im};
                if (d) g[m - d] = g[d];
            }
            gfwd = g;
            inner->run(gfwd.data(), -1);
            for (auto& v : g) v.im = -v.im;
            ginv = g;
            inner->run(ginv.data(), -1);
        } else {
            /* radix 4 first where possible: fewer recursion levels */
            std::vector<int> rr;
            int twos = 0;
            for (int v : f) if (v == 2) ++twos;
            for (int k = 0; k < twos / 2; ++k) rr.push_back(4);
            if (twos % 2) rr.push_back(2);
            for (int v : f) if (v != 2) rr.push_back(v);
            radix = rr;
            w.resize(n);
            for (int j = 0; j < n; ++j) {
                double a = 2.0 * REF_PI_EXACT() * (double)j / (double)n;
                w[j] = {(T)std::cos(a), (T)(-std::sin(a))};
            }
        }
    }
    static double REF_PI_EXACT() { return 3.14159265358979323846; }

    /* twiddle exp(dir * 2 pi i e / n) for 0 <= e < n */
    inline Cx<T> tw(long e, int dir) const {
        Cx<T> v = w[e];
        if (dir > 0) v.im = -v.im;
        return v;
    }

    void rec(const Cx<T>* in, long istride, Cx<T>* out, int len, int level, int dir) const {
        if (len == 1) { out[0] = in[0]; return; }
        int p = radix[level];
        int mm = len / p;
        for (int q = 0; q < p; ++q) rec(in + q * istride, istride * p, out + (long)q * mm, mm, level + 1, dir);
        const long scale = n / len;
        const long pstep = n / p;
        Cx<T> x[32], y[32];
        for (int k = 0; k < mm; ++k) {
            for (int q = 0; q < p; ++q) {
                Cx<T> v = out[(long)q * mm + k];
                x[q] = (q && k) ? cmul(v, tw((long)q * k * scale, dir)) : v;
            }
            for (int t = 0; t < p; ++t) {
                T sr = 0, si = 0;
                for (int q = 0; q < p; ++q) {
                    Cx<T> c = tw(((long)q * t % p) * pstep, dir);
                    Cx<T> prod = cmul(x[q], c);
                    sr += prod.re;
                    si += prod.im;
                }
                y[t] = {sr, si};
            }
            for (int t = 0; t < p; ++t) out[(long)t * mm + k] = y[t];
        }
    }

    /* in-place transform of one contiguous line; dir = -1 forward, +1 inverse */
    void run(Cx<T>* data, int dir) const {
        if (n == 1) return;
        if (!blue) {
            std::vector<Cx<T>> tmp(data, data + n);
            rec(tmp.data(), 1, data, n, 0, dir);
            return;
        }
        /* X_k = c_k * sum_j (x_j c_j) conj(c_{k-j}), c_j = exp(dir pi i j^2 / n) */
        std::vector<Cx<T>> a(m, Cx<T>{0, 0});
        for (int j = 0; j < n; ++j) {
            Cx<T> c = chirp[j];
            if (dir > 0) c.im = -c.im;
            a[j] = cmul(data[j], c);
        }
        inner->run(a.data(), -1);
        const std::vector<Cx<T>>& g = (dir < 0) ? gfwd : ginv;
        for (int j = 0; j < m; ++j) a[j] = cmul(a[j], g[j]);
        inner->run(a.data(), +1);
        const T inv_m = (T)(1.0 / (double)m);
        for (int k = 0; k < n; ++k) {
            Cx<T> c = chirp[k];
            if (dir > 0) c.im = -c.im;
            Cx<T> v = {a[k].re * inv_m, a[k].im * inv_m};
            data[k] = cmul(v, c);
        }
    }
};

/* 2-D transform, x (width, contiguous) first, then y.  Same layout as the
 * reference's VkFFT configuration: size[0] = width, size[1] = height
 * (fsim.cpp:377-379). */
template <class T>
void fft2(Cx<T>* d, int w, int h, int dir, bool normalize) {
    Fft1<T> px(w), py(h);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) px.run(d + (long)y * w, dir);
#pragma omp parallel
    {
        std::vector<Cx<T>> col(h);
#pragma omp for schedule(static)
        for (int x = 0; x < w; ++x) {
            for (int y = 0; y < h; ++y) col[y] = d[(long)y * w + x];
            py.run(col.data(), dir);
            for (int y = 0; y < h; ++y) d[(long)y * w + x] = col[y];
        }
    }
    if (normalize) {
        const T s = (T)(1.0 / ((double)w * (double)h));
        const long N = (long)w * h;
        for (long i = 0; i < N; ++i) { d[i].re *= s; d[i].im *= s; }
    }
}

/* ------------------------------------------------------------------------
 * Pixel access at the boundary.  The reference forces every image to RGBA8
 * (main.cpp:37-40, stbi_load(..., 4)); gray becomes R=G=B.  Alpha is never
 * read by the SSIM or FSIM shaders.
 * ---------------------------------------------------------------------- */
inline void rgb_at(const uint8_t* p, int ch, long i, int& r, int& g, int& b) {
    const uint8_t* q = p + i * ch;
    if (ch == 1) { r = g = b = q[0]; }
    else { r = q[0]; g = q[1]; b = q[2]; }
}

/* ------------------------------------------------------------------------
 * SSIM
 * ---------------------------------------------------------------------- */

/* ssim_lumapack.glsl:18-20 -- Y = 0.299 R + 0.581 G + 0.114 B on UNORM
 * values (note 0.581, not Rec.601's 0.587). */
template <class T>
inline T ssim_luma(int r, int g, int b) {
    const T fr = (T)r / (T)255, fg = (T)g / (T)255, fb = (T)b / (T)255;
    return (T)0.299 * fr + (T)0.581 * fg + (T)0.114 * fb;
}

/* ssim_gaussinput.glsl:25-28 and ssim.glsl:25-28: unnormalised weight
 * pow(E, -(dx^2+dy^2) / (2 sigma^2)). */
template <class T>
inline T ssim_weight(int dx, int dy, T sigma) {
    T dist = (T)(dx * dx + dy * dy);
    return std::pow((T)REF_E, -(dist / ((T)2 * std::pow(sigma, (T)2))));
}

template <class T>
int ssim_impl(const uint8_t* a, const uint8_t* b, int w, int h, int ch, int ks, double k1d,
              double k2d, double sigmad, double* map_out, double* mssim_out) {
    const long N = (long)w * h;
    const T sigma = (T)sigmad, k1 = (T)k1d, k2 = (T)k2d;
    const int start = -(ks - 1) / 2, end = (ks - 1) / 2;
    const int side = end - start + 1;
    std::vector<T> wt((size_t)side * side);
    for (int xo = start; xo <= end; ++xo)
        for (int yo = start; yo <= end; ++yo)
            wt[(size_t)(xo - start) * side + (yo - start)] = ssim_weight<T>(xo, yo, sigma);

    /* ssim_lumapack.glsl:28-35 */
    std::vector<T> lx(N), ly(N);
    for (long i = 0; i < N; ++i) {
        int r, g, bb;
        rgb_at(a, ch, i, r, g, bb);
        lx[i] = ssim_luma<T>(r, g, bb);
        rgb_at(b, ch, i, r, g, bb);
        ly[i] = ssim_luma<T>(r, g, bb);
    }

    /* ssim_gaussinput.glsl:36-56: weighted mean over in-bounds taps,
     * renormalised by the in-bounds weight sum; xOffset outer, yOffset inner */
    std::vector<T> mx(N), my(N);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            T tx = 0, ty = 0, tw = 0;
            for (int xo = start; xo <= end; ++xo) {
                for (int yo = start; yo <= end; ++yo) {
                    int xx = x + xo, yy = y + yo;
                    if (xx >= w || yy >= h || xx < 0 || yy < 0) continue;
                    T wgt = wt[(size_t)(xo - start) * side + (yo - start)];
                    long j = (long)yy * w + xx;
                    tx += lx[j] * wgt;
                    ty += ly[j] * wgt;
                    tw += wgt;
                }
            }
            long i = (long)y * w + x;
            mx[i] = tx / tw;
            my[i] = ty / tw;
        }
    }

    /* ssim.glsl:43-79: moments about the centre pixel's means, same
     * renormalisation, then the SSIM formula with C1 = k1^2, C2 = k2^2 */
    const T c1 = std::pow(k1, (T)2), c2 = std::pow(k2, (T)2);
    std::vector<T> map(N);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            long i = (long)y * w + x;
            T mi = mx[i], mr = my[i];
            T vi = 0, vr = 0, cv = 0, tw = 0;
            for (int xo = start; xo <= end; ++xo) {
                for (int yo = start; yo <= end; ++yo) {
                    int xx = x + xo, yy = y + yo;
                    if (xx >= w || yy >= h || xx < 0 || yy < 0) continue;
                    T wgt = wt[(size_t)(xo - start) * side + (yo - start)];
                    long j = (long)yy * w + xx;
                    T di = lx[j] - mi, dr = ly[j] - mr;
                    vi += di * di * wgt; /* pow(x, 2.0) read as x*x */
                    vr += dr * dr * wgt;
                    cv += di * dr * wgt;
                    tw += wgt;
                }
            }
            vi /= tw;
            vr /= tw;
            cv /= tw;
            map[i] = (((T)2 * mi * mr + c1) * ((T)2 * cv + c2)) /
                     ((mi * mi + mr * mr + c1) * (vi + vr + c2));
        }
    }
    if (map_out)
        for (long i = 0; i < N; ++i) map_out[i] = (double)map[i];

    /* SSIM::computeMSSIM (ssim.cpp:386-401): x in [5, W-5], y in [5, H-5]
     * inclusive, double accumulator.  The reference wraps for W or H < 5;
     * those sizes are rejected here. */
    if (mssim_out) {
        const int s = (ks - 1) / 2;
        const int we = w - (ks - 1) / 2, he = h - (ks - 1) / 2;
        if (we < s || he < s) return -2;
        double sum = 0;
        for (int y = s; y <= he; ++y)
            for (int x = s; x <= we; ++x) sum += (double)map[(long)y * w + x];
        *mssim_out = sum / (double)((long)(we - s + 1) * (he - s + 1));
    }
    return 0;
}

/* ------------------------------------------------------------------------
 * FSIM
 * ---------------------------------------------------------------------- */

/* fsim.cpp:165-168 and :91-92 */
void fsim_dims(int w, int h, int& F, int& wd, int& hd) {
    int smaller = std::min(w, h);
    F = std::max(1, (int)std::round(smaller / 256.0));
    wd = (int)std::round((float)w / (float)F);
    hd = (int)std::round((float)h / (float)F);
}

/* frequency coordinate used by fsim_lowpassfilter.glsl:32-35,
 * fsim_log_gabor.glsl:34-37, fsim_angular.glsl:31-34: ifftshift order,
 * u = ((x + n/2) mod n - n/2) / n (the numpy fftfreq grid) */
template <class T>
inline T freq_coord(int x, int n) {
    unsigned shifted = ((unsigned)x + (unsigned)n / 2u) % (unsigned)n;
    return (T)((int)shifted - n / 2) / (T)n;
}

/* fsim_lowpassfilter.glsl:21-41 with cutoff 0.45, order 15
 * (fsim_lowpass_filter.cpp:44-45) */
template <class T>
inline T lowpass_at(T u, T v) {
    T radius = std::sqrt(u * u + v * v);
    return (T)1 / ((T)1 + std::pow(radius / (T)0.45, (T)2 * (T)15));
}

/* fsim_log_gabor.glsl:24-48: minWaveLength 6, mult 2, sigmaOnf 0.55; the DC
 * bin is forced to 0 */
template <class T>
inline T loggabor_at(int x, int y, T u, T v, int s, T lowpass) {
    if (x == 0 && y == 0) return 0;
    const T wavelength = (T)6 * std::pow((T)2, (T)s);
    T radius = std::sqrt(u * u + v * v);
    T frequency = (T)1 / wavelength;
    T lr = std::log(radius / frequency);
    T ls = std::log((T)0.55);
    T value = std::exp(-(lr * lr) / ((T)2 * (ls * ls)));
    return value * lowpass;
}

/* fsim_angular.glsl:24-46: dThetaOnSigma 1.2; theta = atan2(-v, u).  At DC the
 * GLSL atan(-0, 0) is undefined; the log-Gabor factor is 0 there, so every
 * filter product is 0 whatever this returns (C atan2 semantics here). */
template <class T>
inline T angular_at(T u, T v, int o) {
    const T pi = (T)REF_PI;
    const T thetaSigma = pi / (T)4 / (T)1.2;
    T theta = std::atan2(-v, u);
    T st = std::sin(theta), ct = std::cos(theta);
    T angle = ((T)o * pi) / (T)4;
    T ds = st * std::cos(angle) - ct * std::sin(angle);
    T dc = ct * std::cos(angle) + st * std::sin(angle);
    T dtheta = std::fabs(std::atan2(ds, dc));
    return std::exp(-(dtheta * dtheta) / ((T)2 * (thetaSigma * thetaSigma)));
}

template <class T>
void build_filters(int wd, int hd, std::vector<T>& lp, std::vector<T>& lg, std::vector<T>& ang) {
    const long N = (long)wd * hd;
    lp.assign(N, 0);
    lg.assign(4 * N, 0);
    ang.assign(4 * N, 0);
    for (int y = 0; y < hd; ++y) {
        for (int x = 0; x < wd; ++x) {
            long i = (long)y * wd + x;
            T u = freq_coord<T>(x, wd), v = freq_coord<T>(y, hd);
            lp[i] = lowpass_at<T>(u, v);
            for (int s = 0; s < 4; ++s) lg[s * N + i] = loggabor_at<T>(x, y, u, v, s, lp[i]);
            for (int o = 0; o < 4; ++o) ang[o * N + i] = angular_at<T>(u, v, o);
        }
    }
}

/* H_z with z = 4 o + s (fsim_filter_combinations.glsl:33-34: gabor index
 * z % SCALES, angular index z / ORIENTATIONS) */
template <class T>
void build_bank(int wd, int hd, std::vector<T>& H) {
    std::vector<T> lp, lg, ang;
    build_filters<T>(wd, hd, lp, lg, ang);
    const long N = (long)wd * hd;
    H.assign(16 * N, 0);
    for (int z = 0; z < 16; ++z) {
        int s = z % 4, o = z / 4;
        for (long i = 0; i < N; ++i) H[z * N + i] = lg[s * N + i] * ang[o * N + i];
    }
}

/* F10 (fsim_filter_combinations.cpp:83-134, fsim_filter_noise.glsl): EM_n[o]
 * = sum over every bin of the squared scale-0 filter (its zero imaginary
 * parts included).  The reference's in-place multi-pass tree is racy; the
 * intended full sum is computed here.
 *
 * F14 (fsim_estimate_energy.cpp:57-123, fsim_mult_filters.glsl:26-58): A_s =
 * Re(ifft2(H_{4o+s})); the shader indexes x = 2*gid and stops at x >= N, so
 * only complex elements t < ceil(N/2) (row-major) are written and summed; the
 * rest of the freshly allocated buffer is modelled as zero.  est[2o] =
 * N * sum_t sum_s A_s^2, est[2o+1] = N * sum_t sum_{i<j} A_i A_j. */
template <class T>
void noise_consts(int wd, int hd, const std::vector<T>& H, double em_n[4], double est[8]) {
    const long N = (long)wd * hd;
    for (int o = 0; o < 4; ++o) {
        double acc = 0;
        const T* h0 = &H[(long)(4 * o) * N];
        for (long i = 0; i < N; ++i) acc += (double)h0[i] * (double)h0[i];
        em_n[o] = acc;
    }
    const long half = (N + 1) / 2;
    std::vector<std::vector<T>> A(4, std::vector<T>(N));
    for (int o = 0; o < 4; ++o) {
        for (int s = 0; s < 4; ++s) {
            std::vector<Cx<T>> buf(N);
            const T* hz = &H[(long)(4 * o + s) * N];
            for (long i = 0; i < N; ++i) buf[i] = {hz[i], (T)0};
            fft2<T>(buf.data(), wd, hd, +1, true);
            for (long i = 0; i < N; ++i) A[s][i] = buf[i].re;
        }
        double an2 = 0, cross = 0;
        const T nN = (T)N;
        for (long t = 0; t < half; ++t) {
            T s2 = 0;
            for (int s = 0; s < 4; ++s) s2 += A[s][t] * A[s][t] * nN;
            T c = 0;
            for (int i = 0; i < 3; ++i)
                for (int j = i + 1; j < 4; ++j) c += A[i][t] * A[j][t] * nN;
            an2 += (double)s2;
            cross += (double)c;
        }
        est[2 * o] = an2;
        est[2 * o + 1] = cross;
    }
}

/* Per-size constants (the filter bank and the F10/F14 noise terms depend only
 * on (wd, hd)), kept across calls: computed once per size and precision,
 * exactly as a fresh computation would give them.  A few sizes are kept. */
template <class T>
struct SizeConsts {
    int wd = 0, hd = 0;
    std::vector<T> H;
    double em_n[4] = {0, 0, 0, 0};
    double est[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

template <class T>
std::shared_ptr<const SizeConsts<T>> size_consts(int wd, int hd) {
    static std::mutex mu;
    static std::vector<std::shared_ptr<const SizeConsts<T>>> cache;
    std::lock_guard<std::mutex> lk(mu);
    for (auto& c : cache)
        if (c->wd == wd && c->hd == hd) return c;
    auto c = std::make_shared<SizeConsts<T>>();
    c->wd = wd;
    c->hd = hd;
    build_bank<T>(wd, hd, c->H);
    noise_consts<T>(wd, hd, c->H, c->em_n, c->est);
    if (cache.size() >= 8) cache.erase(cache.begin());
    cache.push_back(c);
    return c;
}

template <class T>
int fsim_impl(const uint8_t* a, const uint8_t* b, int w, int h, int ch, orc_fsim_out* out) {
    int F, wd, hd;
    fsim_dims(w, h, F, wd, hd);
    const long N = (long)wd * hd;
    out->F = F;
    out->wd = wd;
    out->hd = hd;

    /* F2: fsim_downsample.glsl:21-54 -- YIQ of 255-scaled RGB, box sum over
     * the in-bounds part of [xF - F/2, xF - F/2 + F)^2, divided by F^2 */
    std::vector<T> yiq[2];
    const uint8_t* srcs[2] = {a, b};
    const T scaler = std::pow((T)F, (T)2);
    for (int im = 0; im < 2; ++im) {
        yiq[im].assign(3 * N, 0);
#pragma omp parallel for schedule(static)
        for (int y = 0; y < hd; ++y) {
            for (int x = 0; x < wd; ++x) {
                T si = 0, sq = 0, sy = 0;
                int xs = x * F - F / 2, ys = y * F - F / 2;
                for (int j = xs; j < xs + F; ++j) {
                    for (int k = ys; k < ys + F; ++k) {
                        if (j >= 0 && j < w && k >= 0 && k < h) {
                            int r, g, bb;
                            rgb_at(srcs[im], ch, (long)k * w + j, r, g, bb);
                            T R = (T)r, G = (T)g, B = (T)bb;
                            si += R * (T)0.596 - G * (T)0.274 - B * (T)0.322;
                            sq += R * (T)0.211 - G * (T)0.523 + B * (T)0.312;
                            sy += R * (T)0.299 + G * (T)0.587 + B * (T)0.114;
                        }
                    }
                }
                long i = (long)y * wd + x;
                yiq[im][i] = si / scaler;
                yiq[im][N + i] = sq / scaler;
                yiq[im][2 * N + i] = sy / scaler;
            }
        }
    }

    /* F3: fsim_gradientmap.glsl:16-57 -- Scharr pair as a correlation on Y,
     * zero padding, G = sqrt((Gx/16)^2 + (Gy/16)^2) */
    static const T vert[9] = {3, 0, -3, 10, 0, -10, 3, 0, -3};
    static const T hori[9] = {3, 10, 3, 0, 0, 0, -3, -10, -3};
    std::vector<T> grad[2];
    for (int im = 0; im < 2; ++im) {
        grad[im].assign(N, 0);
        const T* Y = &yiq[im][2 * N];
#pragma omp parallel for schedule(static)
        for (int y = 0; y < hd; ++y) {
            for (int x = 0; x < wd; ++x) {
                T vs = 0, hs = 0;
                for (int j = -1; j < 2; ++j) {
                    for (int k = -1; k < 2; ++k) {
                        int px = x + j, py = y + k;
                        if (px >= 0 && px < wd && py >= 0 && py < hd) {
                            T v = Y[(long)py * wd + px];
                            int idx = (j + 1) + (k + 1) * 3;
                            vs += v * vert[idx];
                            hs += v * hori[idx];
                        }
                    }
                }
                T gv = vs / (T)16, gh = hs / (T)16;
                grad[im][(long)y * wd + x] = std::sqrt(gv * gv + gh * gh);
            }
        }
    }

    /* F6-F9: filter bank H_z, z = 4o + s; F10 + F14 (per-size constants) */
    const auto consts = size_consts<T>(wd, hd);
    const std::vector<T>& H = consts->H;
    for (int i = 0; i < 4; ++i) out->em_n[i] = consts->em_n[i];
    for (int i = 0; i < 8; ++i) out->est[i] = consts->est[i];

    /* F4-F5: (Y, 0) per image, forward 2-D FFT, unnormalised */
    std::vector<Cx<T>> spec[2];
    for (int im = 0; im < 2; ++im) {
        spec[im].resize(N);
        for (long i = 0; i < N; ++i) spec[im][i] = {yiq[im][2 * N + i], (T)0};
        fft2<T>(spec[im].data(), wd, hd, -1, false);
    }

    /* F11 + F12 + the packing of F13: per image and orientation, the four
     * scales' responses EO = ifft2(H * F) (normalised), then
     * fsim_sum_filter_responses.glsl:33-83 */
    std::vector<T> An(8 * N), En(8 * N), keys(8 * N);
    /* the 32 (image, orientation, scale) responses are independent: one task
     * each (a transform inside runs on its task's thread) */
    std::vector<std::vector<Cx<T>>> resp(32);
#pragma omp parallel for schedule(dynamic, 1)
    for (int task = 0; task < 32; ++task) {
        const int im = task / 16, o = (task / 4) % 4, s = task % 4;
        const T* hz = &H[(long)(4 * o + s) * N];
        std::vector<Cx<T>>& eo = resp[task];
        eo.resize(N);
        for (long i = 0; i < N; ++i) eo[i] = {hz[i] * spec[im][i].re, hz[i] * spec[im][i].im};
        fft2<T>(eo.data(), wd, hd, +1, true);
    }
    for (int im = 0; im < 2; ++im) {
        for (int o = 0; o < 4; ++o) {
            const std::vector<Cx<T>>* EO = &resp[im * 16 + o * 4];
            const int io = im * 4 + o;
#pragma omp parallel for schedule(static)
            for (long i = 0; i < N; ++i) {
                T amp = 0, sr = 0, si = 0;
                for (int s = 0; s < 4; ++s) {
                    T re = EO[s][i].re, imv = EO[s][i].im;
                    amp += std::sqrt(re * re + imv * imv);
                    sr += re;
                    si += imv;
                }
                T xe = std::sqrt(sr * sr + si * si) + (T)0.0001;
                T me = sr / xe, mo = si / xe;
                T energy = 0;
                for (int s = 0; s < 4; ++s) {
                    T re = EO[s][i].re, imv = EO[s][i].im;
                    energy += re * me + imv * mo - std::fabs(re * mo - imv * me);
                }
                An[io * N + i] = amp;
                En[io * N + i] = energy;
                /* fsim_pack_for_median.glsl:33-44: |EO_{s=0}|^2 */
                T r0 = EO[0][i].re, i0 = EO[0][i].im;
                keys[io * N + i] = r0 * r0 + i0 * i0;
            }
        }
    }

    /* F13: fsim_noise_power.cpp:193-203 -- median = ascending element N/2,
     * mean = -median / ln(0.5), noisePower = mean / EM_n[o] */
    for (int io = 0; io < 8; ++io) {
        std::vector<T> tmp(keys.begin() + io * N, keys.begin() + (io + 1) * N);
        std::nth_element(tmp.begin(), tmp.begin() + N / 2, tmp.end());
        T median = tmp[N / 2];
        out->median[io] = (double)median;
        T mean = (T)(-(double)median / std::log(0.5));
        T emn = (T)out->em_n[io % 4];
        T np = mean / emn;
        out->noise_power[io] = (double)np;
        /* F15: fsim_phase_congruency.glsl:45-53 */
        const int o = io % 4;
        const T pi = (T)REF_PI;
        T e2 = (T)2 * np * (T)out->est[2 * o] + (T)4 * np * (T)out->est[2 * o + 1];
        T tau = std::sqrt(e2 / (T)2);
        T ene = tau * std::sqrt(pi / (T)2);
        T sig = std::sqrt(((T)2 - pi / (T)2) * (tau * tau));
        T thr = (ene + (T)2 * sig) / (T)1.7;
        out->T[io] = (double)thr;
    }

    /* F15: PC = sum_o max(E_o - T_o, 0) / sum_o An_o (no epsilon) */
    std::vector<T> pc[2];
    for (int im = 0; im < 2; ++im) {
        pc[im].assign(N, 0);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N; ++i) {
            T et = 0, at = 0;
            for (int o = 0; o < 4; ++o) {
                int io = im * 4 + o;
                T thr = (T)out->T[io];
                et += std::max(En[io * N + i] - thr, (T)0);
                at += An[io * N + i];
            }
            pc[im][i] = et / at;
        }
    }

    /* F16: fsim_final_multiply.glsl:26-57 */
    const T T1 = (T)0.85, T2 = (T)160, T3 = (T)200, T4 = (T)200, lambda = (T)0.03;
    std::vector<T> pcmv(N), simv(N), simcv(N);
    double spcm = 0, ssim = 0, ssimc = 0;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
        T p1 = pc[0][i], p2 = pc[1][i];
        T pcm = std::max(p1, p2);
        T pcsim = ((T)2 * p1 * p2 + T1) / (p1 * p1 + p2 * p2 + T1);
        T g1 = grad[0][i], g2 = grad[1][i];
        T gsim = ((T)2 * g1 * g2 + T2) / (g1 * g1 + g2 * g2 + T2);
        T sim = gsim * pcsim * pcm;
        T i1 = yiq[0][i], i2 = yiq[1][i];
        T q1 = yiq[0][N + i], q2 = yiq[1][N + i];
        T isim = ((T)2 * i1 * i2 + T3) / (i1 * i1 + i2 * i2 + T3);
        T qsim = ((T)2 * q1 * q2 + T4) / (q1 * q1 + q2 * q2 + T4);
        T simc = gsim * pcsim * std::pow(std::fabs(isim * qsim), lambda) * pcm;
        pcmv[i] = pcm;
        simv[i] = sim;
        simcv[i] = simc;
    }
    /* sums in index order (the same order at any thread count) */
    for (long i = 0; i < N; ++i) {
        spcm += (double)pcmv[i];
        ssim += (double)simv[i];
        ssimc += (double)simcv[i];
    }
    /* F17: fsim_final_multiply.cpp:294-300 -- sums over every pixel */
    out->sums[0] = spcm;
    out->sums[1] = ssim;
    out->sums[2] = ssimc;
    out->fsim = ssim / spcm;
    out->fsimc = ssimc / spcm;

    auto put = [N](double* dst, const std::vector<T>& src, long off, long cnt) {
        if (!dst) return;
        for (long i = 0; i < cnt; ++i) dst[i] = (double)src[off + i];
        (void)N;
    };
    put(out->yiq_in, yiq[0], 0, 3 * N);
    put(out->yiq_ref, yiq[1], 0, 3 * N);
    put(out->grad_in, grad[0], 0, N);
    put(out->grad_ref, grad[1], 0, N);
    if (out->an_e)
        for (int io = 0; io < 8; ++io)
            for (long i = 0; i < N; ++i) {
                out->an_e[(2 * io) * N + i] = (double)An[io * N + i];
                out->an_e[(2 * io + 1) * N + i] = (double)En[io * N + i];
            }
    put(out->keys, keys, 0, 8 * N);
    put(out->pc_in, pc[0], 0, N);
    put(out->pc_ref, pc[1], 0, N);
    put(out->pcm, pcmv, 0, N);
    put(out->sim, simv, 0, N);
    put(out->simc, simcv, 0, N);
    return 0;
}

bool valid_image_args(const void* a, const void* b, int w, int h, int ch) {
    return a && b && w > 0 && h > 0 && (ch == 1 || ch == 3 || ch == 4);
}

}  // namespace

extern "C" {

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_fft1d(int prec, double* data, int n, int dir, int normalize) {
    if (!data || n <= 0 || (dir != 1 && dir != -1)) return -1;
    if (prec == 64) {
        Fft1<double> p(n);
        p.run(reinterpret_cast<Cx<double>*>(data), dir);
        if (normalize)
            for (int i = 0; i < 2 * n; ++i) data[i] /= (double)n;
    } else {
        std::vector<Cx<float>> v(n);
        for (int i = 0; i < n; ++i) v[i] = {(float)data[2 * i], (float)data[2 * i + 1]};
        Fft1<float> p(n);
        p.run(v.data(), dir);
        for (int i = 0; i < n; ++i) {
            float s = normalize ? 1.0f / (float)n : 1.0f;
            data[2 * i] = v[i].re * s;
            data[2 * i + 1] = v[i].im * s;
        }
    }
    return 0;
}

int orc_fft2d(int prec, double* data, int w, int h, int dir, int normalize) {
    if (!data || w <= 0 || h <= 0 || (dir != 1 && dir != -1)) return -1;
    if (prec == 64) {
        fft2<double>(reinterpret_cast<Cx<double>*>(data), w, h, dir, normalize != 0);
    } else {
        long N = (long)w * h;
        std::vector<Cx<float>> v(N);
        for (long i = 0; i < N; ++i) v[i] = {(float)data[2 * i], (float)data[2 * i + 1]};
        fft2<float>(v.data(), w, h, dir, normalize != 0);
        for (long i = 0; i < N; ++i) { data[2 * i] = v[i].re; data[2 * i + 1] = v[i].im; }
    }
    return 0;
}

int orc_ssim(int prec, const uint8_t* img, const uint8_t* ref, int w, int h, int ch,
             int kernel_size, double k1, double k2, double sigma, double* map_out,
             double* mssim_out) {
    if (!valid_image_args(img, ref, w, h, ch) || kernel_size < 1) return -1;
    if (prec == 64) return ssim_impl<double>(img, ref, w, h, ch, kernel_size, k1, k2, sigma, map_out, mssim_out);
    return ssim_impl<float>(img, ref, w, h, ch, kernel_size, k1, k2, sigma, map_out, mssim_out);
}

int orc_fsim(int prec, const uint8_t* img, const uint8_t* ref, int w, int h, int ch,
             orc_fsim_out* out) {
    if (!valid_image_args(img, ref, w, h, ch) || !out) return -1;
    if (prec == 64) return fsim_impl<double>(img, ref, w, h, ch, out);
    return fsim_impl<float>(img, ref, w, h, ch, out);
}

int orc_fsim_batch(int prec, int n, const uint8_t* const* imgs, const uint8_t* const* refs, int w, int h, int ch,
                   orc_fsim_out* outs) {
    if (n <= 0 || !imgs || !refs || !outs) return -1;
    for (int i = 0; i < n; ++i)
        if (!valid_image_args(imgs[i], refs[i], w, h, ch)) return -1;
    int rc = 0;
    /* one pair per thread; each pair's own parallel loops run on that thread */
#pragma omp parallel for schedule(dynamic, 1) reduction(| : rc)
    for (int i = 0; i < n; ++i)
        rc |= prec == 64 ? fsim_impl<double>(imgs[i], refs[i], w, h, ch, &outs[i])
                         : fsim_impl<float>(imgs[i], refs[i], w, h, ch, &outs[i]);
    return rc;
}

int orc_fsim_filters(int prec, int wd, int hd, double* lowpass, double* loggabor, double* angular) {
    if (wd <= 0 || hd <= 0) return -1;
    const long N = (long)wd * hd;
    auto emit = [&](auto& lp, auto& lg, auto& an) {
        for (long i = 0; i < N; ++i) if (lowpass) lowpass[i] = (double)lp[i];
        for (long i = 0; i < 4 * N; ++i) {
            if (loggabor) loggabor[i] = (double)lg[i];
            if (angular) angular[i] = (double)an[i];
        }
    };
    if (prec == 64) {
        std::vector<double> lp, lg, an;
        build_filters<double>(wd, hd, lp, lg, an);
        emit(lp, lg, an);
    } else {
        std::vector<float> lp, lg, an;
        build_filters<float>(wd, hd, lp, lg, an);
        emit(lp, lg, an);
    }
    return 0;
}

int orc_fsim_noise_consts(int prec, int wd, int hd, double* em_n4, double* est8) {
    if (wd <= 0 || hd <= 0 || !em_n4 || !est8) return -1;
    if (prec == 64) {
        std::vector<double> H;
        build_bank<double>(wd, hd, H);
        noise_consts<double>(wd, hd, H, em_n4, est8);
    } else {
        std::vector<float> H;
        build_bank<float>(wd, hd, H);
        noise_consts<float>(wd, hd, H, em_n4, est8);
    }
    return 0;
}

void orc_fsim_dims(int w, int h, int* F, int* wd, int* hd) {
    int f, a, b;
    fsim_dims(w, h, f, a, b);
    if (F) *F = f;
    if (wd) *wd = a;
    if (hd) *hd = b;
}

}  // extern "C"
