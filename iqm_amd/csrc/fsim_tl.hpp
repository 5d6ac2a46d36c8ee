// Host entry points of the two-level FSIM column and row passes (fsim_tl.hip).
#pragma once

#include "ctx.hpp"

namespace iqm {

struct TlColArgs {
    const float2* A;      /* row spectra of the chunk, [pair][2][N] */
    float2* T;            /* response planes of the chunk, [pair][32][N] */
    const float* bank;    /* 16 filter planes */
    const float2* tw;     /* exp(-2 pi i j / hd) */
    int wd, npairs, bx, xcds;
    size_t lds_budget;
};

struct TlRowArgs {
    const float2* T;
    float* an_e;
    float* keys;
    unsigned* hist;
    const float2* tw;     /* exp(-2 pi i j / wd) */
    int hd, npairs, rb;
    float invN;
    size_t lds_budget;
};

struct TlGradArgs {
    const float* yiq;
    float2* A;
    const float2* tw;     /* exp(-2 pi i j / wd) */
    int hd, npairs, rb;
    size_t lds_budget;
};

bool tl_supported(int n);
void tl_rowfft(Ctx& c, int wd, const TlGradArgs& a);
void tl_colpass(Ctx& c, int hd, const TlColArgs& a);
void tl_rowpass(Ctx& c, int wd, const TlRowArgs& a);

}  // namespace iqm
