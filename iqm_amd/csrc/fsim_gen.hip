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

__global__ __launch_bounds__(kBlockFft) void fsim_colpass_kernel(const float2* __restrict__ A,
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

}  // namespace iqm
