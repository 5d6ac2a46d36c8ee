// FSIM over mixed batches: one launch per stage over pairs of any sizes,
// each workgroup finding its pair in a device job table (bodies in
// fsim_body.hpp, the batch around them in batch.hip).
#include "fsim_body.hpp"

namespace iqm {

/* ---- mixed batches: one launch per stage over pairs of any sizes ---------- */

/* pair j owns workgroups [st[j], st[j+1]) of a stage; st[n] is the total */
__device__ __forceinline__ int mix_find(const int* __restrict__ st, int n, int b) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (st[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void mix_prep_kernel(const MixJob* __restrict__ jobs, const int* __restrict__ st,
                                                          int n) {
    const int j = mix_find(st, n, blockIdx.x);
    const MixJob& J = jobs[j];
    const int local = blockIdx.x - st[j];
    const int gx = (J.wd + J.cols - 1) / J.cols;
    switch (J.F) {
#define IQM_MIXPREP_CASE(f)                                                                                  \
    case f:                                                                                                  \
        fsim_prep_body<f>(J.img, J.ref, 0, J.w, J.h, J.ch, J.F, J.wd, J.hd, J.cols, J.prep_rb, J.yiq, J.hist, \
                          local % gx, local / gx, 0);                                                        \
        break;
        IQM_MIXPREP_CASE(1) IQM_MIXPREP_CASE(2) IQM_MIXPREP_CASE(3) IQM_MIXPREP_CASE(4)
        IQM_MIXPREP_CASE(5) IQM_MIXPREP_CASE(6) IQM_MIXPREP_CASE(7) IQM_MIXPREP_CASE(8)
#undef IQM_MIXPREP_CASE
        default:
            fsim_prep_body<0>(J.img, J.ref, 0, J.w, J.h, J.ch, J.F, J.wd, J.hd, J.cols, J.prep_rb, J.yiq, J.hist,
                              local % gx, local / gx, 0);
    }
}

__global__ __launch_bounds__(kBlock) void mix_rowfft_kernel(const MixJob* __restrict__ jobs,
                                                            const int* __restrict__ st, int n) {
    const int j = mix_find(st, n, blockIdx.x);
    const MixJob& J = jobs[j];
    fsim_rowfft_body(J.yiq, J.A, J.wd, J.hd, J.grad_rb, J.Px, blockIdx.x - st[j], 0);
}

__global__ __launch_bounds__(kBlockFft) void mix_colpass_kernel(const MixJob* __restrict__ jobs,
                                                                const int* __restrict__ st, int n) {
    const int j = mix_find(st, n, blockIdx.x);
    const MixJob& J = jobs[j];
    fsim_colpass_body(J.A, J.T, J.bank, J.wd, J.hd, J.col_bx, J.nbx, J.Py, blockIdx.x - st[j]);
}

__global__ __launch_bounds__(kBlockFft) void mix_rowpass_kernel(const MixJob* __restrict__ jobs,
                                                                const int* __restrict__ st, int n) {
    const int j = mix_find(st, n, blockIdx.x);
    const MixJob& J = jobs[j];
    const int local = blockIdx.x - st[j];
    const int gx = (J.hd + J.row_rb - 1) / J.row_rb;
    fsim_rowpass_body(J.T, J.an_e, J.keys, J.hist, J.wd, J.hd, J.row_rb, J.Px, J.invN, local % gx, local / gx, 0);
}

__global__ __launch_bounds__(1024) void mix_select_kernel(const MixJob* __restrict__ jobs, const int* __restrict__ st,
                                                          int n) {
    const int j = mix_find(st, n, blockIdx.x);
    const MixJob& J = jobs[j];
    fsim_select_body(J.keys, J.hist, J.thr, J.stats, J.N, J.na, blockIdx.x - st[j], 0);
}

__global__ __launch_bounds__(kBlock) void mix_pcfinal_kernel(const MixJob* __restrict__ jobs,
                                                             const int* __restrict__ st, int n) {
    const int j = mix_find(st, n, blockIdx.x);
    const MixJob& J = jobs[j];
    const int local = blockIdx.x - st[j];
    fsim_pcfinal_body(J.yiq, nullptr, J.an_e, J.thr, J.partial, J.wd, J.hd, J.simc_map, nullptr, nullptr, nullptr,
                      nullptr, local % J.pc_gx, local / J.pc_gx, J.pc_gx, J.pc_gy, 0);
}

/* finalize, then the pair's scalars into its row of the batch result */
__global__ __launch_bounds__(kBlock) void mix_finalize_kernel(const MixJob* __restrict__ jobs,
                                                              const int* __restrict__ st, int n) {
    const int j = mix_find(st, n, blockIdx.x);
    const MixJob& J = jobs[j];
    fsim_finalize_body(J.partial, J.pc_gx * J.pc_gy, J.res, 0);
    if (threadIdx.x == 0) {
        J.out[1] = J.res[3];
        J.out[2] = J.res[4];
    }
}

void Ctx::mix_fsim_fill(const PairJob& pj, MixJob& mj, char* base, size_t& off, MixCounts& cnt, int* blk,
                        int* cls) {
    const int F = std::max(1, (int)std::round(std::min(pj.w, pj.h) / 256.0));
    const int wd = (int)std::round((float)pj.w / (float)F);
    const int hd = (int)std::round((float)pj.h / (float)F);
    IQM_CHECK(F <= kPrepMaxRows, IQM_ERR_UNSUPPORTED, "image too large for the FSIM prep tile");
    IQM_CHECK(fsim_planes_fit32((long)wd * hd), IQM_ERR_UNSUPPORTED, "downscaled image too large for FSIM");
    const FsimSizePlan& plan = fsim_plan(wd, hd);
    const LinePlan& Px = line_plan_for(wd, LINE_ROW);
    const LinePlan& Py = line_plan_for(hd, LINE_COL);
    const long N = (long)wd * hd;
    const FsimTiles t = pick_tiles(*this, Px, Py, N, fsim_tune(wd, hd));
    mj.F = F;
    mj.wd = wd;
    mj.hd = hd;
    mj.N = N;
    mj.invN = (float)(1.0 / (double)N);
    /* prep: 4 downscaled rows per workgroup where kPrepMaxRows full rows
     * allow it (256-pair batch: 1.62 ms with 1 row, 1.20 with 2, 1.06 with
     * 4: fewer, larger workgroups, each finding its pair once) */
    int prb = std::max(1, env_int("IQM_MIX_PREP_RB", 4));
    while (prb > 1 && prb * F > kPrepMaxRows) --prb;
    mj.prep_rb = prb;
    mj.cols = prep_cols(F, pj.ch, prb);
    mj.grad_rb = t.grad_rb;
    mj.col_bx = t.col_bx;
    mj.nbx = ceil_div(wd, t.col_bx);
    mj.row_rb = t.row_rb;
    mj.pc_gx = ceil_div(wd, kPcTx);
    mj.pc_gy = ceil_div(hd, kPcTy);
    mj.Px = Px;
    mj.Py = Py;
    mj.bank = plan.bank;
    for (int i = 0; i < 4; ++i) mj.na.em_n[i] = plan.em_n[i];
    for (int i = 0; i < 8; ++i) mj.na.est[i] = plan.est[i];
    const FsimRegions r = fsim_regions(N, 1, 1, mj.pc_gx * mj.pc_gy);
    char* b = base ? base + off : nullptr;
    mj.yiq = (float*)(b + r.yiq);
    mj.A = (float2*)(b + r.A);
    mj.T = (float2*)(b + r.T);
    mj.an_e = (float*)(b + r.an_e);
    mj.keys = (float*)(b + r.keys);
    mj.hist = (unsigned*)(b + r.hist);
    mj.thr = (float*)(b + r.thr);
    mj.stats = (double*)(b + r.stats);
    mj.partial = (double*)(b + r.partial);
    mj.res = (double*)(b + r.res);
    mj.simc_map = pj.simc_map;
    off += r.total;
    /* workgroups and LDS per stage */
    const size_t kSmall = 65536;
    auto add = [&](int stage, int nb, size_t lds) {
        const int c = lds > kSmall ? 1 : 0;
        blk[stage] = nb;
        cls[stage] = c;
        cnt.blocks[stage][c] += nb;
        cnt.lds[stage][c] = std::max(cnt.lds[stage][c], lds);
    };
    add(MIX_PREP, ceil_div(wd, mj.cols) * ceil_div(hd, prb), prep_lds_bytes(F, pj.ch, mj.cols, prb));
    add(MIX_ROWFFT, ceil_div(hd, t.grad_rb), line_lds_bytes(Px, t.grad_rb, 2));
    add(MIX_COL, mj.nbx * 2, line_lds_bytes(Py, t.col_bx, colpass_bufs(Py)));
    add(MIX_ROW, ceil_div(hd, t.row_rb) * 2, rowpass_lds_bytes(Px, t.row_rb, wd));
    add(MIX_SELECT, 8, sizeof(unsigned) * kSelectCap);
    add(MIX_PCF, mj.pc_gx * mj.pc_gy, 0);
    add(MIX_FIN, 1, 0);
}

void Ctx::mix_fsim_launch(const MixJob* d_jobs, const int* d_starts, int n, const MixCounts& cnt) {
    struct K {
        int stage;
        const char* tag;
        void (*fn)(const MixJob*, const int*, int);
        int threads;
    };
    const K ks[] = {{MIX_PREP, "fsim_prep", mix_prep_kernel, kBlock},
                    {MIX_ROWFFT, "fsim_rowfft", mix_rowfft_kernel, kBlock},
                    {MIX_COL, "fsim_colpass", mix_colpass_kernel, kBlockFft},
                    {MIX_ROW, "fsim_rowpass", mix_rowpass_kernel, kBlockFft},
                    {MIX_SELECT, "fsim_select", mix_select_kernel, 1024},
                    {MIX_PCF, "fsim_pcfinal", mix_pcfinal_kernel, kBlock},
                    {MIX_FIN, "fsim_finalize", mix_finalize_kernel, kBlock}};
    for (const K& k : ks)
        for (int c = 0; c < 2; ++c)
            if (cnt.blocks[k.stage][c] > 0)
                launch(k.tag, k.fn, dim3(cnt.blocks[k.stage][c]), dim3(k.threads), cnt.lds[k.stage][c], d_jobs,
                       d_starts + ((size_t)k.stage * 2 + c) * (n + 1), n);
}


}  // namespace iqm
