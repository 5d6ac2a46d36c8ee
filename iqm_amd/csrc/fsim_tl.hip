// Two-level (register) versions of the FSIM column and row passes for line
// lengths with a compile-time split n = N1 * N2 (fft_reg.hpp).  Each line
// transform is two in-register DFTs with one LDS exchange, instead of one
// LDS round trip per radix stage.  Same inputs, outputs and plane layout as
// fsim_colpass_kernel / fsim_rowpass_kernel in fsim.hip.
#include "ctx.hpp"
#include "fft_reg.hpp"
#include "fsim_tl.hpp"

namespace iqm {

constexpr int kHistBinsTl = 1024; /* as kHistBins in fsim.hip */

/* Column pass: forward column FFT of `bx` columns of one image, then for
 * each of the 16 filters z the product H_z * F inverse transformed along the
 * columns and stored to plane img*16 + z.  Grid as fsim_colpass_kernel
 * (column blocks padded to a multiple of 8 for the XCD mapping). */
template <int N1, int N2>
__global__ __launch_bounds__(512) void fsim_colpass_tl_kernel(const float2* __restrict__ A,
                                                               float2* __restrict__ Tout,
                                                               const float* __restrict__ bank,
                                                               const float2* __restrict__ twg,
                                                               int wd, int bx, int nbx_pad,
                                                               const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    constexpr int n = N1 * N2;
    constexpr int ZL = Exch<N1, N2>::line;
    constexpr int LS = n + 1; /* spectrum line pitch (odd) */
    extern __shared__ float2 lds[];
    /* read once: a read after an LDS barrier's memory clobber would be a
     * load whose wait (vmcnt) also drains this thread's pending stores */
    const int nthr = blockDim.x;
    const int cb = blockIdx.x % nbx_pad;
    const int rest = blockIdx.x / nbx_pad;
    const int img = rest & 1, pair = rest >> 1;
    const int x0 = cb * bx;
    if (x0 >= wd) return;
    const int nc = min(bx, wd - x0);
    float2* S = lds;
    float2* Z = S + bx * LS;
    float2* tw = Z + bx * ZL;
    for (int i = threadIdx.x; i < n; i += nthr) tw[i] = twg[i];
    const long N = (long)wd * n;
    const float2* src = A + ((long)pair * 2 + img) * N + x0;
    __syncthreads();
    /* forward: step 1, items (c, j2), c fastest */
    for (int it = threadIdx.x; it < bx * N2; it += nthr) {
        const int c = it % bx, j2 = it / bx;
        float2 x[N1];
        if (c < nc) {
            const unsigned o0 = (unsigned)j2 * wd + c, rs = (unsigned)wd * N2;
#pragma unroll
            for (int j1 = 0; j1 < N1; ++j1) x[j1] = *at32(src, o0 + j1 * rs);
        } else {
#pragma unroll
            for (int j1 = 0; j1 < N1; ++j1) x[j1] = make_float2(0.f, 0.f);
        }
        two_level_step1<N1, N2, -1>(x, j2, tw, Z + c * ZL);
    }
    lds_barrier();
    for (int it = threadIdx.x; it < bx * N1; it += nthr) {
        const int c = it % bx, k1 = it / bx;
        float2 y[N2];
        two_level_step2<N1, N2, -1>(Z + c * ZL, k1, y);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) S[c * LS + k1 + N1 * k2] = y[k2];
    }
    lds_barrier();
    /* plane offsets below N = wd * n are 32-bit unsigned, added to
     * workgroup-uniform base pointers */
    const unsigned rstride = (unsigned)wd * N2; /* N2 rows */
    for (int z = 0; z < 16; ++z) {
        const float* hz = bank + (long)z * N + x0;
        for (int it = threadIdx.x; it < bx * N2; it += nthr) {
            const int c = it % bx, j2 = it / bx;
            float2 x[N1];
            if (c < nc) {
                const unsigned o0 = (unsigned)j2 * wd + c;
                /* all N1 filter values are requested before the first is used */
                float h[N1];
#pragma unroll
                for (int j1 = 0; j1 < N1; ++j1) h[j1] = *at32(hz, o0 + j1 * rstride);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j1 = 0; j1 < N1; ++j1) x[j1] = cscale(S[c * LS + N2 * j1 + j2], h[j1]);
            } else {
#pragma unroll
                for (int j1 = 0; j1 < N1; ++j1) x[j1] = make_float2(0.f, 0.f);
            }
            two_level_step1<N1, N2, 1>(x, j2, tw, Z + c * ZL);
        }
        lds_barrier();
        float2* dst = Tout + ((long)pair * 32 + img * 16 + z) * N + x0;
        for (int it = threadIdx.x; it < bx * N1; it += nthr) {
            const int c = it % bx, k1 = it / bx;
            float2 y[N2];
            two_level_step2<N1, N2, 1>(Z + c * ZL, k1, y);
            if (c < nc) {
                /* streamed once to the row pass: keep it out of the L2 so the
                 * filter bank stays resident */
                const unsigned o0 = (unsigned)k1 * wd + c, kstride = (unsigned)wd * N1;
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) nt_store(at32(dst, o0 + k2 * kstride), y[k2]);
            }
        }
        lds_barrier(); /* the next filter overwrites Z */
    }
}

/* Row pass: for each orientation o = 0..3 in turn, the inverse row FFT of
 * the 4 scales of `rb` rows of one image, normalised by 1/N, then the
 * response epilogue, the median keys and their top-bit histogram (as
 * fsim_rowpass_kernel).  Each thread owns one step-1 item (the launch keeps
 * 4 * rb * max(N1, N2) <= blockDim), whose inputs for the next orientation are loaded
 * into registers before the current one's step 2 and epilogue, so the loads
 * overlap that work (the barriers in the loop order LDS only).  The
 * amplitudes are summed over o in registers, in the order o = 0..3, and
 * written once as the plane sum_o An. */
/* epilogue pixels per thread: the launch keeps blockDim >= 4 * rb * max(N1, N2),
 * so rb * N1 * N2 <= ceil(min(N1, N2) / 4) * blockDim */
template <int N1, int N2>
constexpr int row_slots() { return ((N1 < N2 ? N1 : N2) + 3) / 4; }

template <int N1, int N2>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(3))) void fsim_rowpass_tl_kernel(const float2* __restrict__ Tin,
                                                               float* __restrict__ an_e,
                                                               float* __restrict__ keys,
                                                               unsigned* __restrict__ hist,
                                                               const float2* __restrict__ twg,
                                                               int hd, int rb, float invN,
                                                               const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    constexpr int n = N1 * N2; /* = wd */
    constexpr int ZL = Exch<N1, N2>::line > n + 1 ? Exch<N1, N2>::line : n + 1;
    constexpr int LE = ZL; /* E shares Z's line pitch */
    extern __shared__ float2 lds[];
    const int nl = 4 * rb;
    float2* Z = lds;
    float2* E = lds; /* aliases Z, see step 2 */
    float2* tw = Z + nl * ZL;
    unsigned* hl = (unsigned*)(tw + n);
    /* pairs newest first: the last pairs the column pass wrote are read while
     * they may still be in the memory-side cache, and the PC kernel, which
     * walks pairs oldest first, then finds this pass's last outputs there */
    const int img = blockIdx.y, pair = gridDim.z - 1 - blockIdx.z;
    const long N = (long)n * hd;
    const int y0 = blockIdx.x * rb;
    const int nr = min(rb, hd - y0);
    /* this thread's step-1 item: line l1 = s1 * rb + r1, column j21 */
    const int it1 = threadIdx.x;
    const bool has1 = it1 < nl * N2;
    const int l1 = it1 / N2, j21 = it1 - l1 * N2;
    const int s1 = l1 / rb, r1 = l1 - s1 * rb;
    const bool ld1 = has1 && r1 < nr;
    float2 x[N1];
    auto fetch = [&](int o) {
        if (ld1) {
            const float2* src = Tin + ((long)pair * 32 + (img * 4 + o) * 4 + s1) * N + (long)(y0 + r1) * n;
#pragma unroll
            for (int j1 = 0; j1 < N1; ++j1) x[j1] = nt_load(&src[N2 * j1 + j21]);
        } else {
#pragma unroll
            for (int j1 = 0; j1 < N1; ++j1) x[j1] = make_float2(0.f, 0.f);
        }
    };
    fetch(0);
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
    for (int i = threadIdx.x; i < kHistBinsTl; i += blockDim.x) hl[i] = 0u;
    lds_barrier();
    float* ap = an_e + ((long)pair * 2 + img) * kAnePlanes * N;
    const float epsN = 0.0001f / invN;
    constexpr int kSlots = row_slots<N1, N2>();
    float sa[kSlots];
    for (int o = 0; o < 4; ++o) {
        const int io = img * 4 + o;
        if (has1) two_level_step1<N1, N2, 1>(x, j21, tw, Z + l1 * ZL);
        if (o < 3) fetch(o + 1);
        lds_barrier();
        /* step 2 in rounds of blockDim items: every item of a round reads its
         * exchange row into registers before any item writes outputs over the
         * same LDS (E aliases Z) */
        const int items2 = nl * N1;
        for (int base = 0; base < items2; base += blockDim.x) {
            const int it = base + threadIdx.x;
            float2 y[N2];
            const bool act = it < items2;
            const int l = act ? it / N1 : 0, k1 = act ? it - l * N1 : 0;
            if (act) {
#pragma unroll
                for (int j2 = 0; j2 < N2; ++j2) y[j2] = Z[l * ZL + k1 * Exch<N1, N2>::pitch + j2];
            }
            lds_barrier();
            if (act) {
                RegDft<N2, 1>::run(y);
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) E[l * LE + k1 + N1 * k2] = y[k2];
            }
            lds_barrier();
        }
        float* enp = ap + (long)(1 + o) * N;
        float* kp = keys + ((long)pair * 8 + io) * N;
        const int sstride = rb * LE; /* E lines of one scale */
#pragma unroll
        for (int q = 0; q < kSlots; ++q) {
            const int idx = threadIdx.x + q * blockDim.x;
            if (idx < nr * n) {
                const int r = idx / n, xx = idx - r * n;
                const float2* er = E + r * LE + xx;
                float2 eo[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) eo[s] = er[s * sstride];
                float amp, energy, key;
                fsim_response_epilogue(eo, invN, epsN, amp, energy, key);
                const unsigned g = (unsigned)(y0 * n + idx); /* < N */
                /* sum over o, in the order o = 0..3 */
                if (o == 0) sa[q] = amp;
                else if (o < 3) sa[q] += amp;
                else *at32(ap, g) = sa[q] + amp;
                *at32(enp, g) = energy;
                *at32(kp, g) = key;
                atomicAdd(&hl[(__float_as_uint(key) >> 21) & 1023u], 1u);
            }
        }
        lds_barrier(); /* histogram complete; E read (the next step 1 overwrites it) */
        unsigned* hg = hist + ((long)pair * 8 + io) * kHistBinsTl;
        for (int i = threadIdx.x; i < kHistBinsTl; i += blockDim.x) {
            const unsigned c = hl[i];
            if (c) atomicAdd(&hg[i], c);
            hl[i] = 0u; /* the next orientation's atomics follow several barriers */
        }
    }
}

/* Two-level version of fsim_rowfft_kernel (fsim.hip): the forward row FFT
 * of the packed rows Y_in + i*Y_ref and the split into the two row spectra. */
template <int N1, int N2>
__global__ __launch_bounds__(512) void fsim_rowfft_tl_kernel(const float* __restrict__ yiq,
                                                             float2* __restrict__ A,
                                                                  const float2* __restrict__ twg,
                                                                  int hd, int rb,
                                                                  const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    constexpr int n = N1 * N2; /* = wd */
    constexpr int ZL = Exch<N1, N2>::line > n + 1 ? Exch<N1, N2>::line : n + 1;
    extern __shared__ float2 lds[];
    float2* Z = lds;
    float2* E = lds; /* aliases Z */
    float2* tw = Z + rb * ZL;
    const int pair = blockIdx.y;
    const long N = (long)n * hd;
    const int y0 = blockIdx.x * rb;
    const int nr = min(rb, hd - y0);
    const float* Yin = yiq + (long)pair * 6 * N;
    const float* Yref = Yin + N;
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = twg[i];
    __syncthreads();
    for (int it = threadIdx.x; it < rb * N2; it += blockDim.x) {
        const int r = it / N2, j2 = it - r * N2;
        float2 x[N1];
        if (r < nr) {
            const long o = (long)(y0 + r) * n + j2;
#pragma unroll
            for (int j1 = 0; j1 < N1; ++j1) x[j1] = make_float2(Yin[o + N2 * j1], Yref[o + N2 * j1]);
        } else {
#pragma unroll
            for (int j1 = 0; j1 < N1; ++j1) x[j1] = make_float2(0.f, 0.f);
        }
        two_level_step1<N1, N2, -1>(x, j2, tw, Z + r * ZL);
    }
    lds_barrier();
    const int items2 = rb * N1;
    for (int base = 0; base < items2; base += blockDim.x) {
        const int it = base + threadIdx.x;
        float2 y[N2];
        const bool act = it < items2;
        const int r = act ? it / N1 : 0, k1 = act ? it - r * N1 : 0;
        if (act) {
#pragma unroll
            for (int j2 = 0; j2 < N2; ++j2) y[j2] = Z[r * ZL + k1 * Exch<N1, N2>::pitch + j2];
        }
        lds_barrier();
        if (act) {
            RegDft<N2, -1>::run(y);
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) E[r * ZL + k1 + N1 * k2] = y[k2];
        }
        lds_barrier();
    }
    float2* Ain = A + (long)pair * 2 * N;
    float2* Aref = Ain + N;
    for (int idx = threadIdx.x; idx < nr * n; idx += blockDim.x) {
        const int r = idx / n, k = idx - r * n;
        const int m = (k == 0) ? 0 : n - k;
        const float2 zk = E[r * ZL + k], zm = E[r * ZL + m];
        const long g = (long)(y0 + r) * n + k;
        Ain[g] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        Aref[g] = make_float2(0.5f * (zk.y + zm.y), 0.5f * (zm.x - zk.x));
    }
}

namespace {

template <int N1, int N2>
size_t col_lds(int bx) {
    return sizeof(float2) * ((size_t)bx * (N1 * N2 + 1) + (size_t)bx * Exch<N1, N2>::line + N1 * N2);
}
template <int N1, int N2>
size_t row_lds(int rb) {
    const size_t line = Exch<N1, N2>::line > N1 * N2 + 1 ? Exch<N1, N2>::line : N1 * N2 + 1;
    return sizeof(float2) * ((size_t)4 * rb * line + N1 * N2) + sizeof(unsigned) * kHistBinsTl;
}

int threads_for(int items) {
    int t = ((items + 63) / 64) * 64;
    return t < 64 ? 64 : (t > 512 ? 512 : t);
}

template <int N1, int N2>
void col_launch(Ctx& c, const TlColArgs& a) {
    const int n = N1 * N2;
    int bx = a.bx;
    while (bx > 1 && col_lds<N1, N2>(bx) > a.lds_budget) --bx;
    const int nbx = ceil_div(a.wd, bx);
    const int nbx_pad = ceil_div(nbx, a.xcds) * a.xcds;
    const int thr = threads_for(bx * (N1 > N2 ? N1 : N2));
    c.launch("fsim_colpass", fsim_colpass_tl_kernel<N1, N2>, dim3(nbx_pad * 2 * a.npairs), dim3(thr),
             col_lds<N1, N2>(bx), a.A, a.T, a.bank, a.tw, a.wd, bx, nbx_pad, a.gate);
    (void)n;
}

template <int N1, int N2>
void row_launch(Ctx& c, const TlRowArgs& a) {
    int rb = a.rb;
    /* one step-1 item per thread and at most row_slots() epilogue pixels per thread */
    constexpr int mx = N1 > N2 ? N1 : N2;
    while (rb > 1 && (row_lds<N1, N2>(rb) > a.lds_budget || 4 * rb * mx > 512)) --rb;
    const int thr = threads_for(4 * rb * mx);
    constexpr int slots = row_slots<N1, N2>();
    IQM_CHECK(4 * rb * N2 <= thr && rb * N1 * N2 <= slots * thr, IQM_ERR_UNSUPPORTED,
              "row pass tile does not fit its thread mapping");
    c.launch("fsim_rowpass", fsim_rowpass_tl_kernel<N1, N2>, dim3(ceil_div(a.hd, rb), 2, a.npairs),
             dim3(thr), row_lds<N1, N2>(rb), a.T, a.an_e, a.keys, a.hist, a.tw, a.hd, rb, a.invN, a.gate);
}

template <int N1, int N2>
void grad_launch(Ctx& c, const TlGradArgs& a) {
    constexpr int n = N1 * N2;
    constexpr size_t line = Exch<N1, N2>::line > n + 1 ? Exch<N1, N2>::line : n + 1;
    int rb = a.rb;
    auto lds = [&](int r) { return sizeof(float2) * ((size_t)r * line + n); };
    while (rb > 1 && lds(rb) > a.lds_budget) --rb;
    const int thr = threads_for(rb * (N1 > N2 ? N1 : N2));
    c.launch("fsim_rowfft", fsim_rowfft_tl_kernel<N1, N2>, dim3(ceil_div(a.hd, rb), a.npairs),
             dim3(thr), lds(rb), a.yiq, a.A, a.tw, a.hd, rb, a.gate);
}

}  // namespace

/* Compile-time splits available to the two-level kernels. */
bool tl_supported(int n) { return n == 270 || n == 480 || n == 256; }

void tl_colpass(Ctx& c, int hd, const TlColArgs& a) {
    switch (hd) {
        case 270: col_launch<15, 18>(c, a); break;
        case 480: col_launch<20, 24>(c, a); break;
        case 256: col_launch<16, 16>(c, a); break;
        default: throw Error(IQM_ERR_UNSUPPORTED, "no two-level column split");
    }
}

void tl_rowfft(Ctx& c, int wd, const TlGradArgs& a) {
    switch (wd) {
        case 270: grad_launch<15, 18>(c, a); break;
        case 480: grad_launch<20, 24>(c, a); break;
        case 256: grad_launch<16, 16>(c, a); break;
        default: throw Error(IQM_ERR_UNSUPPORTED, "no two-level row split");
    }
}

void tl_rowpass(Ctx& c, int wd, const TlRowArgs& a) {
    switch (wd) {
        case 270: row_launch<15, 18>(c, a); break;
        case 480: row_launch<20, 24>(c, a); break;
        case 256: row_launch<16, 16>(c, a); break;
        default: throw Error(IQM_ERR_UNSUPPORTED, "no two-level row split");
    }
}

}  // namespace iqm
