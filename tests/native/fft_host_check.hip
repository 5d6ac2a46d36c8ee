// Host-side check of the FFT index math (no GPU needed): runs the same
// __host__ __device__ building blocks the kernels use -- Stockham stages,
// register DFTs, the compile-time two-level line split, the run-time
// two-level split with its size dispatch, Rader's identity with its plan
// tables and Bluestein's in three phases -- sequentially on the CPU and
// prints the largest error against a direct double-precision DFT.
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "fft_line.hpp"

using namespace iqm;
using cd = std::complex<double>;

static std::vector<cd> dft(const std::vector<cd>& x, int S) {
    const int n = (int)x.size();
    std::vector<cd> y(n);
    for (int k = 0; k < n; ++k) {
        cd acc = 0;
        for (int j = 0; j < n; ++j) {
            const double a = 2.0 * M_PI * (double)((long)j * k % n) / n;
            acc += x[j] * cd(std::cos(a), S * std::sin(a));
        }
        y[k] = acc;
    }
    return y;
}

static std::vector<cd> rnd(int n, int seed) {
    std::mt19937 g(seed);
    std::normal_distribution<double> d;
    std::vector<cd> x(n);
    for (auto& v : x) v = cd(d(g), d(g));
    return x;
}

static double relerr(const std::vector<float2>& got, const std::vector<cd>& ref) {
    double e = 0, m = 0;
    for (size_t i = 0; i < ref.size(); ++i) {
        e = std::max(e, std::abs(cd(got[i].x, got[i].y) - ref[i]));
        m = std::max(m, std::abs(ref[i]));
    }
    return e / m;
}

static std::vector<float2> twiddles(int n) {
    std::vector<float2> tw(n);
    for (int j = 0; j < n; ++j) {
        double a = 2.0 * M_PI * j / n;
        tw[j] = make_float2((float)std::cos(a), (float)-std::sin(a));
    }
    return tw;
}

template <int S>
static double check_stockham(int n) {
    std::vector<int> rad;
    if (!plan_radices(n, rad)) return -1;
    auto x = rnd(n, n);
    std::vector<float2> a(n), b(n);
    for (int i = 0; i < n; ++i) a[i] = make_float2((float)x[i].real(), (float)x[i].imag());
    auto tw = twiddles(n);
    float2 *src = a.data(), *dst = b.data();
    int p = 1;
    for (int r : rad) {
        for (int tid = 0; tid < 7; ++tid) stockham_dispatch<S>(r, src, dst, n, 1, n, p, tw.data(), tid, 7);
        std::swap(src, dst);
        p *= r;
    }
    return relerr(std::vector<float2>(src, src + n), dft(x, S));
}

template <int M, int S>
static double check_reg() {
    auto x = rnd(M, 1000 + M);
    float2 v[M];
    for (int i = 0; i < M; ++i) v[i] = make_float2((float)x[i].real(), (float)x[i].imag());
    RegDft<M, S>::run(v);
    return relerr(std::vector<float2>(v, v + M), dft(x, S));
}

template <int N1, int N2, int S>
static double check_two_level() {
    constexpr int n = N1 * N2;
    auto x = rnd(n, 7 * n);
    auto tw = twiddles(n);
    std::vector<float2> z(Exch<N1, N2>::line), out(n);
    for (int j2 = 0; j2 < N2; ++j2) {
        float2 v[N1];
        for (int j1 = 0; j1 < N1; ++j1) {
            const cd e = x[N2 * j1 + j2];
            v[j1] = make_float2((float)e.real(), (float)e.imag());
        }
        two_level_step1<N1, N2, S>(v, j2, tw.data(), z.data());
    }
    for (int k1 = 0; k1 < N1; ++k1) {
        float2 y[N2];
        two_level_step2<N1, N2, S>(z.data(), k1, y);
        for (int k2 = 0; k2 < N2; ++k2) out[k1 + N1 * k2] = y[k2];
    }
    return relerr(out, dft(x, S));
}

/* the run-time split as tl2_lines runs it: every step-1 item, then every
 * step-2 item, through an exchange area of n1 * (n2 + 1) points */
template <int S>
static double check_tl2(int n) {
    int n1, n2;
    if (!tl2_split(n, n1, n2)) return -1;
    auto x = rnd(n, 11 * n);
    auto tw = twiddles(n);
    std::vector<float2> in(n), z(n1 * (n2 + 1));
    for (int i = 0; i < n; ++i) in[i] = make_float2((float)x[i].real(), (float)x[i].imag());
    for (int j2 = 0; j2 < n2; ++j2) tl2_item1<S>(n1, in.data(), j2, n2, tw.data(), z.data());
    for (int k1 = 0; k1 < n1; ++k1) tl2_item2<S>(n2, z.data(), k1, n1, in.data());
    return relerr(in, dft(x, S));
}

/* Rader's identity as the kernels run it (fft_line.hpp line_rader and the
 * Rader branch of line_fft_io), sequentially over every work item, for
 * three lines of n points: the outputs stored through RaderOutStore to a
 * separate array and to the line buffer must both equal the DFT of each
 * line. */
template <int S>
static double check_rader(int n) {
    int q, p, m1, m2;
    if (!rader_split(n, q, p, m1, m2)) return -1;
    const int rp = rader_pitch(p, m1, m2);
    std::vector<int> rin, rout;
    rader_tables(p, rp, rin, rout);
    /* the plan's kernel V = FFT(v) / (p - 1), v_c = exp(S 2 pi i g^c / p) */
    std::vector<cd> v(p - 1);
    for (int c = 0; c < p - 1; ++c) v[c] = std::polar(1.0, S * 2.0 * M_PI * rout[c] / p);
    const auto Vd = dft(v, -1);
    std::vector<float2> V(p - 1);
    for (int c = 0; c < p - 1; ++c) V[c] = make_float2((float)(Vd[c].real() / (p - 1)), (float)(Vd[c].imag() / (p - 1)));
    const int nlines = 3, ld = lds_pitch(std::max(n, q * rp)), nl2 = nlines * q;
    auto tw = twiddles(n), mtw = twiddles(p - 1);
    std::vector<std::vector<cd>> xs;
    std::vector<float2> b0(nlines * ld), b1(nlines * ld), viaio(nlines * n);
    for (int l = 0; l < nlines; ++l) {
        xs.push_back(rnd(n, 13 * n + l));
        for (int j = 0; j < n; ++j) b0[l * ld + j] = make_float2((float)xs[l][j].real(), (float)xs[l][j].imag());
    }
    for (int l = 0; l < nlines; ++l)
        for (int j2 = 0; j2 < p; ++j2)
            rader_item1_f<S>(q, PtrLoad{&b0[l * ld]}, j2, p, tw.data(), &b1[l * rp + rin[j2]], nlines * rp);
    for (int l2 = 0; l2 < nl2; ++l2)
        for (int j2 = 0; j2 < m2; ++j2) tl2_item1<-1>(m1, &b1[l2 * rp], j2, m2, mtw.data(), &b0[l2 * rp]);
    /* forward step 2, V and inverse step 1 (split m2 * m1) in one phase */
    for (int l2 = 0; l2 < nl2; ++l2)
        for (int k1 = 0; k1 < m1; ++k1)
            conv_mid_item(m2, &b0[l2 * rp], k1, m1, V.data(), mtw.data(), &b1[l2 * rp], &b1[l2 * rp + rp - 2]);
    /* the store-functor path first (it leaves b0 and b1 as they are) */
    for (int l2 = 0; l2 < nl2; ++l2) {
        const int k1 = l2 / nlines, l = l2 - k1 * nlines;
        for (int kk = 0; kk < m2; ++kk)
            tl2_item2_f<1>(m1, &b1[l2 * rp], kk, m2,
                           RaderOutStore<PtrStore>{PtrStore{&viaio[l * n]}, &b1[l2 * rp], rout.data(), k1, q, rp});
    }
    /* line_rader: the same inverse step 2 storing into the line buffer b0 */
    for (int l2 = 0; l2 < nl2; ++l2) {
        const int k1 = l2 / nlines, l = l2 - k1 * nlines;
        for (int kk = 0; kk < m2; ++kk)
            tl2_item2_f<1>(m1, &b1[l2 * rp], kk, m2,
                           RaderOutStore<PtrStore>{PtrStore{&b0[l * ld]}, &b1[l2 * rp], rout.data(), k1, q, rp});
    }
    double e = 0;
    for (int l = 0; l < nlines; ++l) {
        std::vector<float2> out(b0.begin() + l * ld, b0.begin() + l * ld + n);
        const auto ref = dft(xs[l], S);
        e = std::max(e, relerr(out, ref));
        e = std::max(e, relerr(std::vector<float2>(viaio.begin() + l * n, viaio.begin() + (l + 1) * n), ref));
    }
    return e;
}

/* Bluestein in the three phases the kernels run (fft_line.hpp line_fft and
 * the Bluestein branch of line_fft_io), sequentially over every work item,
 * for two lines of n points, with the plan's tables made as Ctx::line_plan
 * makes them: chirp c_j = exp(-pi i (j^2 mod 2n) / n), m the smallest
 * length >= 2n - 1 with a two-level split, and the spectra of the
 * convolution kernels conj(c_d) (forward) and c_d (inverse), d in (-n, n),
 * from a double-precision DFT.  Both the in-place path (line_fft: input in
 * b0, output in b1) and the load/store path must equal the DFT of each
 * line. */
template <int S>
static double check_bluestein(int n) {
    int m = 2 * n - 1, m1 = 0, m2 = 0;
    while (m <= 32 * 32 && !tl2_split(m, m1, m2)) ++m;
    if (m1 == 0) return -1;
    std::vector<cd> cdd(n);
    std::vector<float2> chirp(n);
    for (int j = 0; j < n; ++j) {
        const double a = M_PI * (double)((long long)j * j % (2LL * n)) / n;
        cdd[j] = cd(std::cos(a), -std::sin(a));
        chirp[j] = make_float2((float)cdd[j].real(), (float)cdd[j].imag());
    }
    std::vector<cd> kf(m, 0.0), ki(m, 0.0);
    for (int d = 0; d < n; ++d) {
        kf[d] = std::conj(cdd[d]);
        ki[d] = cdd[d];
        if (d) {
            kf[m - d] = kf[d];
            ki[m - d] = ki[d];
        }
    }
    const auto Kf = dft(kf, -1), Ki = dft(ki, -1);
    std::vector<float2> g(m);
    for (int j = 0; j < m; ++j) {
        const cd v = S < 0 ? Kf[j] : Ki[j];
        g[j] = make_float2((float)v.real(), (float)v.imag());
    }
    const auto mtw = twiddles(m);
    const int nlines = 2, ld = lds_pitch(std::max({m, m1 * (m2 + 1), m2 * (m1 + 1)}));
    std::vector<std::vector<cd>> xs;
    std::vector<float2> b0(nlines * ld), b1(nlines * ld), in(nlines * n), viaio(nlines * n);
    for (int l = 0; l < nlines; ++l) {
        xs.push_back(rnd(n, 29 * n + l));
        for (int j = 0; j < n; ++j) in[l * n + j] = b0[l * ld + j] = make_float2((float)xs[l][j].real(), (float)xs[l][j].imag());
    }
    const float inv_m = 1.0f / (float)m;
    auto run = [&](auto load, auto store) {
        for (int l = 0; l < nlines; ++l)
            for (int j2 = 0; j2 < m2; ++j2)
                tl2_item1_f<-1>(m1, ChirpLoad<S, decltype(load(0))>{load(l), chirp.data(), n}, j2, m2, mtw.data(),
                                &b1[l * ld]);
        for (int l = 0; l < nlines; ++l)
            for (int k1 = 0; k1 < m1; ++k1) conv_mid_item(m2, &b1[l * ld], k1, m1, g.data(), mtw.data(), &b0[l * ld], nullptr);
        for (int l = 0; l < nlines; ++l)
            for (int kk = 0; kk < m2; ++kk)
                tl2_item2_f<1>(m1, &b0[l * ld], kk, m2,
                               ChirpStore<S, decltype(store(0))>{store(l), chirp.data(), n, inv_m});
    };
    /* load/store path: input from `in`, output to `viaio` */
    run([&](int l) { return PtrLoad{&in[l * n]}; }, [&](int l) { return PtrStore{&viaio[l * n]}; });
    /* in-place path: input in b0 (reloaded), output in b1 */
    for (int l = 0; l < nlines; ++l)
        for (int j = 0; j < n; ++j) b0[l * ld + j] = in[l * n + j];
    run([&](int l) { return PtrLoad{&b0[l * ld]}; }, [&](int l) { return PtrStore{&b1[l * ld]}; });
    double e = 0;
    for (int l = 0; l < nlines; ++l) {
        const auto ref = dft(xs[l], S);
        e = std::max(e, relerr(std::vector<float2>(viaio.begin() + l * n, viaio.begin() + (l + 1) * n), ref));
        e = std::max(e, relerr(std::vector<float2>(b1.begin() + l * ld, b1.begin() + l * ld + n), ref));
    }
    return e;
}

int main() {
    double worst = 0;
    auto rep = [&](const char* what, int n, double e) {
        std::printf("%s %d %.3e\n", what, n, e);
        if (e < 0 || e > worst) worst = e < 0 ? 1.0 : e;
    };
    for (int n : {2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 16, 27, 45, 64, 96, 270, 480, 1024, 4095}) {
        rep("stockham_fwd", n, check_stockham<-1>(n));
        rep("stockham_inv", n, check_stockham<1>(n));
    }
    rep("reg_fwd", 15, check_reg<15, -1>());
    rep("reg_fwd", 16, check_reg<16, -1>());
    rep("reg_fwd", 18, check_reg<18, -1>());
    rep("reg_fwd", 20, check_reg<20, -1>());
    rep("reg_inv", 24, check_reg<24, 1>());
    rep("reg_inv", 30, check_reg<30, 1>());
    rep("reg_fwd", 32, check_reg<32, -1>());
    rep("reg_fwd", 17, check_reg<17, -1>());
    rep("reg_inv", 19, check_reg<19, 1>());
    rep("reg_fwd", 23, check_reg<23, -1>());
    rep("reg_inv", 29, check_reg<29, 1>());
    rep("reg_fwd", 31, check_reg<31, -1>());
    rep("reg_inv", 31, check_reg<31, 1>());
    rep("two_level_fwd", 270, check_two_level<15, 18, -1>());
    rep("two_level_inv", 270, check_two_level<15, 18, 1>());
    rep("two_level_fwd", 480, check_two_level<20, 24, -1>());
    rep("two_level_inv", 480, check_two_level<20, 24, 1>());
    int nsplit = 0;
    for (int n = 4; n <= 1024; ++n) {
        int a, b;
        if (!tl2_split(n, a, b)) continue;
        if (n % 3 != 0 && n % 7 != 1 && n != 1024 && n != 338) continue; /* a spread of lengths and splits */
        ++nsplit;
        rep("tl2_fwd", n, check_tl2<-1>(n));
        rep("tl2_inv", n, check_tl2<1>(n));
    }
    for (int n : {221, 289, 323, 527, 589, 961, 46, 62}) { /* splits by the primes 17..31 */
        ++nsplit;
        rep("tl2_fwd", n, check_tl2<-1>(n));
        rep("tl2_inv", n, check_tl2<1>(n));
    }
    std::printf("run-time splits checked: %d\n", nsplit);
    /* Rader: q = 1 (17, 37, 61, 97, 929), q = 2..17 (74, 202, 611, 629, 1184) */
    for (int n : {17, 37, 61, 97, 929, 74, 202, 611, 629, 1184}) {
        rep("rader_fwd", n, check_rader<-1>(n));
        rep("rader_inv", n, check_rader<1>(n));
    }
    /* Bluestein: primes whose p - 1 has no two-level split (263, 269, 431, 83)
     * and lengths with such a prime factor (3 * 107, 223, 511 = 7 * 73); 509 takes m = 1024 */
    for (int n : {83, 223, 263, 269, 321, 431, 509, 511}) {
        rep("bluestein_fwd", n, check_bluestein<-1>(n));
        rep("bluestein_inv", n, check_bluestein<1>(n));
    }
    std::printf("worst %.3e\n", worst);
    return worst < 2e-6 ? 0 : 1;
}
