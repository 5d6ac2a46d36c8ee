// The generic line-transform passes of FSIM for one size per launch: any
// line length, through the run-time two-level, Stockham, Rader and Bluestein
// transforms of fft_line.hpp (bodies in fsim_body.hpp).  Lengths with a
// compile-time split use fsim_tl.hip instead.
#include "fsim_body.hpp"

namespace iqm {

__global__ __launch_bounds__(kBlock) void fsim_rowfft_kernel(const float* __restrict__ yiq,
                                                             float2* __restrict__ A, int wd, int hd,
                                                             int rb, LinePlan P, const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    fsim_rowfft_body(yiq, A, wd, hd, rb, P, blockIdx.x, blockIdx.y);
}

__global__ __launch_bounds__(kBlockFft) void fsim_colpass_kernel(float2* __restrict__ A,
                                                              float2* __restrict__ Tout,
                                                              const float* __restrict__ bank,
                                                              int wd, int hd, int bx, int nbx_pad,
                                                              LinePlan P, const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    fsim_colpass_body(A, Tout, bank, wd, hd, bx, nbx_pad, P, blockIdx.x);
}

__global__ __launch_bounds__(kBlockFft) void fsim_rowpass_kernel(const float2* __restrict__ Tin,
                                                              float* __restrict__ an_e,
                                                              float* __restrict__ keys,
                                                              unsigned* __restrict__ hist,
                                                              int wd, int hd, int rb, LinePlan P,
                                                              float invN, const int* __restrict__ gate) {
    if (gate_closed(gate)) return;
    fsim_rowpass_body(Tin, an_e, keys, hist, wd, hd, rb, P, invN, blockIdx.x, blockIdx.y, blockIdx.z);
}

/* ---- timing of one pass under a plan (line-plan autotuner) --------------- */

/* pseudo-random floats in [-1, 1) from a multiplicative hash of the index:
 * varied response values, so that the keys of the row pass spread over the
 * histogram bins as real data does */
__global__ __launch_bounds__(kBlock) void line_noise_kernel(float* p, long n, unsigned seed) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        unsigned v = (unsigned)i * 2654435761u ^ seed;
        v ^= v >> 15;
        v *= 2246822519u;
        v ^= v >> 13;
        p[i] = (float)(v >> 8) * (1.0f / 8388608.0f) - 1.0f;
    }
}

/* A workload that fills the device like a batch does: one pair, up to 512
 * workgroups of the pass (two per CU) with the tiles pick_tiles gives
 * without a tune -- for the columns up to 256 column blocks per image of P.n
 * rows, for the rows up to 256 row blocks per image of P.n columns.  Inputs are noise
 * (the filter planes too).  One warm-up launch, then the fastest of three,
 * each between the two events. */
double Ctx::time_line_pass(const LinePlan& P, int role, hipEvent_t e0, hipEvent_t e1) {
    const FsimTiles t0 = pick_tiles(*this, P, P, (long)P.n * P.n, FsimTune{});
    /* 256 blocks per image, fewer for long lines: at most 4 Mi pixels
     * (1 GiB of response planes) */
    const long tile_px = (long)P.n * (role == LINE_COL ? t0.col_bx : t0.row_rb);
    const int blocks = (int)std::max(16L, std::min(256L, (4L << 20) / tile_px));
    const int wd = role == LINE_COL ? blocks * t0.col_bx : P.n;
    const int hd = role == LINE_COL ? P.n : blocks * t0.row_rb;
    const long N = (long)wd * hd;
    const FsimTiles t = pick_tiles(*this, P, P, N, FsimTune{});
    const size_t tbytes = sizeof(float2) * 32 * N;
    const size_t abytes = role == LINE_COL ? sizeof(float2) * 2 * N + sizeof(float) * 16 * N
                                           : sizeof(float) * (2 * kAnePlanes + 8) * N + sizeof(unsigned) * 8 * kHistBins;
    char* ws = (char*)workspace(WS_TUNE, tbytes + abytes);
    float2* T = (float2*)ws;
    char* rest = ws + tbytes;
    launch("line_noise", line_noise_kernel, dim3(4096), dim3(kBlock), 0, (float*)ws,
           (long)((tbytes + abytes) / sizeof(float)), 977u);
    auto one = [&] {
        if (role == LINE_COL) {
            const int nbx = ceil_div(wd, t.col_bx), nbx_pad = ceil_div(nbx, xcds_) * xcds_;
            float2* A = (float2*)rest;
            const float* bank = (const float*)(rest + sizeof(float2) * 2 * N);
            launch("line_tune", fsim_colpass_kernel, dim3(nbx_pad * 2), dim3(kBlockFft),
                   line_lds_bytes(P, t.col_bx, colpass_bufs(P)), A, T, bank, wd, hd, t.col_bx, nbx_pad, P,
                   (const int*)nullptr);
        } else {
            float* an_e = (float*)rest;
            float* keys = an_e + 2 * kAnePlanes * N;
            unsigned* hist = (unsigned*)(keys + 8 * N);
            launch("line_tune", fsim_rowpass_kernel, dim3(ceil_div(hd, t.row_rb), 2, 1), dim3(kBlockFft),
                   rowpass_lds_bytes(P, t.row_rb, wd), (const float2*)T, an_e, keys, hist, wd, hd, t.row_rb, P,
                   (float)(1.0 / (double)N), (const int*)nullptr);
        }
    };
    one();
    float best = 1e30f;
    for (int r = 0; r < 3; ++r) {
        IQM_HIP(hipEventRecord(e0, stream_));
        one();
        IQM_HIP(hipEventRecord(e1, stream_));
        IQM_HIP(hipEventSynchronize(e1));
        float ms = 0;
        IQM_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
    }
    return 1e3 * best;
}

}  // namespace iqm
