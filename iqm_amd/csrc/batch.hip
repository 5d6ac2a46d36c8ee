// Mixed batches: pairs of any sizes, SSIM and/or FSIM per pair.  Every stage
// is one launch over all pairs of the batch: the host writes a device job
// table (per pair: sizes, tiles, FFT line plans, noise constants, buffers)
// and, per stage, the prefix sums of the pairs' workgroup counts; a
// workgroup finds its pair by binary search and runs the same stage body as
// a single-pair launch, so the scores are bit-identical to the single-pair
// calls -- except for pairs whose downscaled lines have a compile-time
// two-level kernel (480, 270 and 256 long: 1080p, 4K and 512 x 512 inputs),
// which the single-pair call runs and the mixed batch does not (it uses the
// run-time transforms; the scores then differ by FP32 rounding, ~1e-8).
// Pairs whose line transforms need more than 64 KiB of LDS run in a second launch of the
// stage, so they do not cut the occupancy of the others.  The sequence is
// captured once as a hipGraph with a job table of its own; a repeated batch
// costs one graph launch.
#include "ctx.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace iqm {

__global__ void batch_fill_kernel(double* out, int n, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v;
}


static void check_jobs(const PairJob* jobs, int n, const SsimArgs& sp) {
    IQM_CHECK(jobs && n > 0, IQM_ERR_INVALID, "empty batch");
    const int R = (sp.ks - 1) / 2;
    for (int j = 0; j < n; ++j) {
        const PairJob& b = jobs[j];
        const std::string at = " (pair " + std::to_string(j) + ")";
        IQM_CHECK(b.img && b.ref && b.w > 0 && b.h > 0 && (b.ch == 1 || b.ch == 3 || b.ch == 4) &&
                      (b.methods & 3) != 0 && (b.methods & ~3) == 0,
                  IQM_ERR_INVALID, "invalid batch entry" + at);
        if (b.methods & 1)
            IQM_CHECK(b.w - R >= R && b.h - R >= R, IQM_ERR_TOO_SMALL, "image smaller than the SSIM window" + at);
        if (b.methods & 2) {
            const int F = std::max(1, (int)std::round(std::min(b.w, b.h) / 256.0));
            IQM_CHECK(std::round((float)b.w / F) >= 2 && std::round((float)b.h / F) >= 2, IQM_ERR_TOO_SMALL,
                      "image too small for FSIM" + at);
        }
    }
}

double* Ctx::mixed_batch(const PairJob* jobs, int n, const SsimArgs& sp, void* table, bool upload) {
    check_jobs(jobs, n, sp);
    /* per-pair setup: sizes, tiles, plans (built here, outside any capture:
     * a plan build allocates and synchronises), workgroup counts; then the
     * same again with the real workspace base */
    std::vector<MixJob> mj((size_t)n);
    std::vector<int> blk((size_t)n * MIX_STAGES), cls((size_t)n * MIX_STAGES);
    MixCounts cnt;
    auto fill = [&](char* base) {
        cnt = MixCounts{};
        std::fill(blk.begin(), blk.end(), 0);
        std::fill(cls.begin(), cls.end(), 0);
        size_t off = 0;
        for (int j = 0; j < n; ++j) {
            std::memset(&mj[j], 0, sizeof(MixJob));
            mj[j].img = jobs[j].img;
            mj[j].ref = jobs[j].ref;
            mj[j].w = jobs[j].w;
            mj[j].h = jobs[j].h;
            mj[j].ch = jobs[j].ch;
            if (jobs[j].methods & 1) mix_ssim_fill(jobs[j], mj[j], base, off, cnt, &blk[(size_t)j * MIX_STAGES],
                                                   &cls[(size_t)j * MIX_STAGES], sp);
            if (jobs[j].methods & 2) mix_fsim_fill(jobs[j], mj[j], base, off, cnt, &blk[(size_t)j * MIX_STAGES],
                                                   &cls[(size_t)j * MIX_STAGES]);
        }
        return off;
    };
    const size_t ws = fill(nullptr);
    char* base = (char*)workspace(WS_MIX, ws);
    fill(base);
    double* out = (double*)workspace(WS_BATCH, sizeof(double) * 4 * n);
    for (int j = 0; j < n; ++j) mj[j].out = out + 4 * j;
    /* block prefix per (stage, LDS class): pairs outside a class own none */
    const size_t nst = (size_t)MIX_STAGES * 2 * (n + 1);
    std::vector<int> starts(nst, 0);
    for (int sg = 0; sg < MIX_STAGES; ++sg)
        for (int c = 0; c < 2; ++c) {
            int* st = &starts[((size_t)sg * 2 + c) * (n + 1)];
            for (int j = 0; j < n; ++j) {
                const int b = (cls[(size_t)j * MIX_STAGES + sg] == c) ? blk[(size_t)j * MIX_STAGES + sg] : 0;
                st[j + 1] = st[j] + b;
            }
        }
    const size_t tab_bytes = sizeof(MixJob) * n + sizeof(int) * nst;
    char* tab = table ? (char*)table : (char*)workspace(WS_MIXTAB, tab_bytes);
    if (upload) {
        IQM_CHECK(!capturing_, IQM_ERR_DEVICE, "job table upload during graph capture");
        IQM_HIP(hipMemcpyAsync(tab, mj.data(), sizeof(MixJob) * n, hipMemcpyHostToDevice, stream_));
        IQM_HIP(hipMemcpyAsync(tab + sizeof(MixJob) * n, starts.data(), sizeof(int) * nst, hipMemcpyHostToDevice,
                               stream_));
        IQM_HIP(hipStreamSynchronize(stream_)); /* the host vectors die with this call */
    }
    const MixJob* d_jobs = (const MixJob*)tab;
    const int* d_starts = (const int*)(tab + sizeof(MixJob) * n);
    launch("batch_fill", batch_fill_kernel, dim3(ceil_div(4 * n, kBlock)), dim3(kBlock), 0, out, 4 * n,
           (double)NAN);
    mix_ssim_launch(d_jobs, d_starts, n, cnt, sp);
    mix_fsim_launch(d_jobs, d_starts, n, cnt);
    return out;
}

size_t mixed_table_bytes(int n) { return sizeof(MixJob) * n + sizeof(int) * (size_t)MIX_STAGES * 2 * (n + 1); }

double* Ctx::mixed_batch_graph(const PairJob* jobs, int n, const SsimArgs& sp) {
    if (!graphs_ || timing_) return mixed_batch(jobs, n, sp);
    check_jobs(jobs, n, sp);
    std::string key((const char*)jobs, sizeof(PairJob) * (size_t)n);
    key.append((const char*)&sp, sizeof(sp));
    auto it = mixed_graphs_.find(key);
    if (it != mixed_graphs_.end()) {
        it->second.last_use = ++graph_tick_;
        IQM_HIP(hipGraphLaunch(it->second.exec, stream_));
        return it->second.res;
    }
    /* the graph reads a job table of its own; an eager pass fills it, sizes
     * every workspace and builds every per-size plan */
    void* table = nullptr;
    IQM_HIP(hipMalloc(&table, mixed_table_bytes(n)));
    double* res = nullptr;
    hipGraph_t g = nullptr;
    try {
        res = mixed_batch(jobs, n, sp, table, true);
        IQM_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        capturing_ = true;
        double* res2 = mixed_batch(jobs, n, sp, table, false);
        capturing_ = false;
        IQM_HIP(hipStreamEndCapture(stream_, &g));
        IQM_CHECK(res2 == res, IQM_ERR_DEVICE, "batch workspace moved during capture");
    } catch (...) {
        if (capturing_) {
            capturing_ = false;
            (void)hipStreamEndCapture(stream_, &g);
        }
        if (g) (void)hipGraphDestroy(g);
        (void)hipStreamSynchronize(stream_);
        (void)hipFree(table);
        throw;
    }
    GraphEntry e;
    e.res = res;
    e.table = table;
    e.last_use = ++graph_tick_;
    const hipError_t ie = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        (void)hipFree(table);
        IQM_HIP(ie);
    }
    evict_lru(mixed_graphs_);
    mixed_graphs_.emplace(std::move(key), e);
    return res;
}

}  // namespace iqm
