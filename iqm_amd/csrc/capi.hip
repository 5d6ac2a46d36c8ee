// Context implementation and the extern "C" entry points of include/iqm/iqm.h.
#include "ctx.hpp"
#include "surrogate.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

namespace iqm {

Ctx::Ctx(int device) : device_(device) {
    int count = 0;
    IQM_HIP(hipGetDeviceCount(&count));
    IQM_CHECK(device >= 0 && device < count, IQM_ERR_INVALID,
              "device ordinal " + std::to_string(device) + " out of range");
    IQM_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    IQM_HIP(hipGetDeviceProperties(&prop, device));
    name_ = std::string(prop.name) + " (" + prop.gcnArchName + ", " +
            std::to_string(prop.multiProcessorCount) + " CUs, " +
            std::to_string(prop.sharedMemPerBlock / 1024) + " KiB LDS/block)";
    lds_per_block_ = prop.sharedMemPerBlock;
    cus_ = std::max(1, prop.multiProcessorCount);
    if (const char* g = std::getenv("IQM_GRAPHS")) graphs_ = std::atoi(g) != 0;
    ssim_noalloc_setup(device);
    IQM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
}

Ctx::~Ctx() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graphs();
    for (auto& kv : ws_) (void)hipFree(kv.second.first);
    for (auto& kv : line_plans_)
        for (void* p : kv.second.owned) (void)hipFree(p);
    for (auto& kv : fsim_plans_) {
        (void)hipFree(kv.second.lg);
        (void)hipFree(kv.second.ang);
        (void)hipFree(kv.second.bank);
    }
    for (auto& p : pending_) {
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    for (auto e : free_events_) (void)hipEventDestroy(e);
    if (pinned_) (void)hipHostFree(pinned_);
    if (copy_stream_) (void)hipStreamSynchronize(copy_stream_);
    for (int i = 0; i < 2; ++i) {
        if (h2d_done_[i]) (void)hipEventDestroy(h2d_done_[i]);
        if (comp_done_[i]) (void)hipEventDestroy(comp_done_[i]);
    }
    if (stage_) (void)hipHostFree(stage_);
    if (copy_stream_) (void)hipStreamDestroy(copy_stream_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void* Ctx::host_pinned(size_t bytes) {
    if (pinned_bytes_ < bytes) {
        if (pinned_) {
            IQM_HIP(hipStreamSynchronize(stream_));
            IQM_HIP(hipHostFree(pinned_));
            pinned_ = nullptr;
            pinned_bytes_ = 0;
        }
        const size_t want = std::max(bytes, (size_t)4096);
        if (hipHostMalloc(&pinned_, want, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            pinned_ = nullptr;
            throw Error(IQM_ERR_NOMEM, "pinned host allocation of " + std::to_string(want) + " bytes failed");
        }
        pinned_bytes_ = want;
    }
    return pinned_;
}

void* Ctx::workspace(int slot, size_t bytes) {
    auto& e = ws_[slot];
    if (e.second < bytes) {
        IQM_CHECK(!capturing_, IQM_ERR_DEVICE, "workspace growth during graph capture");
        drop_graphs(); /* captured graphs hold the old addresses */
        if (e.first) {
            IQM_HIP(hipStreamSynchronize(stream_));
            IQM_HIP(hipFree(e.first));
            e.first = nullptr;
            e.second = 0;
        }
        size_t want = std::max(bytes, (size_t)256);
        if (hipMalloc(&e.first, want) != hipSuccess) {
            (void)hipGetLastError();
            e.first = nullptr;
            throw Error(IQM_ERR_NOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
        }
        e.second = want;
    }
    return e.first;
}

void Ctx::release_graph(GraphEntry& e) {
    /* a launch of this graph may still be running on the stream */
    (void)hipStreamSynchronize(stream_);
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (e.table) (void)hipFree(e.table);
    e.exec = nullptr;
    e.table = nullptr;
}

void Ctx::drop_graphs() {
    ++graph_gen_;
    for (auto& kv : graphs_cache_) release_graph(kv.second);
    graphs_cache_.clear();
    for (auto& kv : mixed_graphs_) release_graph(kv.second);
    mixed_graphs_.clear();
}

double* Ctx::fsim_batch_graph(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch,
                              long pair_stride, int n, const FsimBatchOut& out) {
    if (!graphs_ || timing_ || gate_ || capturing_) return fsim_batch(d_img, d_ref, w, h, ch, pair_stride, n, out);
    GraphKey key{d_img, d_ref, out.d_simc, out.d_pc, out.d_pcm, out.d_sim, pair_stride, w, h, ch, n};
    auto it = graphs_cache_.find(key);
    if (it == graphs_cache_.end()) {
        /* one eager run sizes the workspaces and builds the per-size plans
         * (both allocate and synchronise, which a capture may not) */
        fsim_batch(d_img, d_ref, w, h, ch, pair_stride, n, out);
        hipGraph_t g = nullptr;
        IQM_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        capturing_ = true;
        double* res = nullptr;
        try {
            res = fsim_batch(d_img, d_ref, w, h, ch, pair_stride, n, out);
        } catch (...) {
            capturing_ = false;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        capturing_ = false;
        IQM_HIP(hipStreamEndCapture(stream_, &g));
        GraphEntry e;
        e.res = res;
        e.last_use = ++graph_tick_;
        const hipError_t ie = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        IQM_HIP(ie);
        evict_lru(graphs_cache_);
        it = graphs_cache_.emplace(key, e).first;
        return res; /* the capture ran nothing: the eager run above produced it */
    }
    it->second.last_use = ++graph_tick_;
    IQM_HIP(hipGraphLaunch(it->second.exec, stream_));
    return it->second.res;
}

void Ctx::sync() {
    IQM_HIP(hipStreamSynchronize(stream_));
    if (timing_ || !pending_.empty()) collect_timing();
}

hipEvent_t Ctx::take_event() {
    if (!free_events_.empty()) {
        hipEvent_t e = free_events_.back();
        free_events_.pop_back();
        return e;
    }
    hipEvent_t e;
    IQM_HIP(hipEventCreate(&e));
    return e;
}

void Ctx::collect_timing() {
    for (auto& p : pending_) {
        IQM_HIP(hipEventSynchronize(p.second.second));
        float ms = 0.f;
        IQM_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        auto& t = totals_[p.first];
        t.first += ms;
        t.second += 1;
        free_events_.push_back(p.second.first);
        free_events_.push_back(p.second.second);
    }
    pending_.clear();
}

void Ctx::set_timing(bool on) {
    if (!on) sync();
    timing_ = on;
}

int Ctx::stage_times(int max, const char** names, double* ms, int64_t* launches) {
    sync();
    names_cache_.clear();
    for (auto& kv : totals_) names_cache_.push_back(kv.first);
    int i = 0;
    for (auto& kv : totals_) {
        if (i >= max) break;
        if (names) names[i] = names_cache_[i].c_str();
        if (ms) ms[i] = kv.second.first;
        if (launches) launches[i] = kv.second.second;
        ++i;
    }
    return (int)totals_.size();
}

void Ctx::reset_times() {
    sync();
    totals_.clear();
}

}  // namespace iqm

/* ------------------------------------------------------------------------ */

struct iqm_ctx {
    std::unique_ptr<iqm::Ctx> c;
};

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return IQM_OK;
    } catch (const iqm::Error& e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return fail(IQM_ERR_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(IQM_ERR_DEVICE, e.what());
    }
}

/* Validates an image view and returns its tight row size. */
void check_image(const iqm_image* im, const char* what) {
    IQM_CHECK(im && im->data, IQM_ERR_INVALID, std::string("null ") + what + " image");
    IQM_CHECK(im->width > 0 && im->height > 0, IQM_ERR_INVALID, std::string("empty ") + what + " image");
    IQM_CHECK(im->channels == 1 || im->channels == 3 || im->channels == 4, IQM_ERR_INVALID,
              std::string(what) + " image must have 1, 3 or 4 channels");
    const int tight = im->width * im->channels;
    IQM_CHECK(im->row_stride == 0 || im->row_stride >= tight, IQM_ERR_INVALID,
              std::string(what) + " row_stride smaller than width * channels");
}

void check_pair(const iqm_image* a, const iqm_image* b) {
    check_image(a, "input");
    check_image(b, "reference");
    IQM_CHECK(a->width == b->width && a->height == b->height, IQM_ERR_SIZE_MISMATCH,
              "Compared images must have the same size");
    IQM_CHECK(a->channels == b->channels, IQM_ERR_INVALID,
              "input and reference must have the same channel count");
}

/* Copies an image view into a tight device buffer. */
void upload(iqm::Ctx& c, int slot, const iqm_image* im, uint8_t** dev) {
    const size_t row = (size_t)im->width * im->channels;
    const size_t stride = im->row_stride ? (size_t)im->row_stride : row;
    uint8_t* d = (uint8_t*)c.workspace(slot, row * im->height);
    if (stride == row) {
        IQM_HIP(hipMemcpyAsync(d, im->data, row * im->height, hipMemcpyHostToDevice, c.stream()));
    } else {
        IQM_HIP(hipMemcpy2DAsync(d, row, im->data, stride, row, im->height, hipMemcpyHostToDevice,
                                 c.stream()));
    }
    *dev = d;
}

}  // namespace

extern "C" {

const char* iqm_last_error(void) { return g_last_error.c_str(); }

int iqm_create(int device, iqm_ctx** out) {
    if (!out) return fail(IQM_ERR_INVALID, "null output pointer");
    *out = nullptr;
    return guarded([&] {
        auto* h = new iqm_ctx;
        try {
            h->c.reset(new iqm::Ctx(device));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void iqm_destroy(iqm_ctx* ctx) { delete ctx; }

const char* iqm_device_name(iqm_ctx* ctx) { return ctx ? ctx->c->name().c_str() : ""; }

void iqm_ssim_default_params(iqm_ssim_params* p) {
    if (!p) return;
    p->kernel_size = 11; /* ssim.h:31-34 */
    p->k1 = 0.01f;
    p->k2 = 0.03f;
    p->sigma = 1.5f;
}

int iqm_ssim(iqm_ctx* ctx, const iqm_image* img, const iqm_image* ref,
             const iqm_ssim_params* params, float* map_out, double* mssim_out) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        check_pair(img, ref);
        iqm_ssim_params p;
        iqm_ssim_default_params(&p);
        if (params) p = *params;
        iqm::Ctx& c = *ctx->c;
        roctxRangePushA("iqm_ssim");
        c.begin_marks();
        uint8_t *da, *db;
        upload(c, iqm::WS_IN_IMG, img, &da);
        upload(c, iqm::WS_IN_REF, ref, &db);
        c.mark("start GPU pipeline"); /* ssim.cpp:78 (after the upload here) */
        const long N = (long)img->width * img->height;
        float* dmap = map_out ? (float*)c.workspace(iqm::WS_MAPS, sizeof(float) * N) : nullptr;
        double* d_ms = c.ssim_batch(da, db, img->width, img->height, img->channels, 0, 1, p.kernel_size, p.k1,
                                    p.k2, p.sigma, dmap, nullptr);
        c.mark("end GPU pipeline"); /* ssim.cpp:227 */
        if (map_out) {
            IQM_HIP(hipMemcpyAsync(map_out, dmap, sizeof(float) * N, hipMemcpyDeviceToHost, c.stream()));
        }
        c.mark("end copy from GPU"); /* ssim.cpp:232 */
        double ms = 0.0;
        IQM_HIP(hipMemcpyAsync(&ms, d_ms, sizeof(double), hipMemcpyDeviceToHost, c.stream()));
        c.sync();
        c.mark("end MSSIM compute"); /* ssim.cpp:235: the cropped mean, computed on the device */
        roctxRangePop();
        if (mssim_out) *mssim_out = ms;
    });
}

int iqm_ssim_batch_dev(iqm_ctx* ctx, const uint8_t* d_img, const uint8_t* d_ref, int32_t w,
                       int32_t h, int32_t ch, int64_t pair_stride, int32_t n,
                       const iqm_ssim_params* params, float* d_maps, double* mssim_out) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_CHECK(d_img && d_ref && w > 0 && h > 0 && n > 0, IQM_ERR_INVALID, "invalid batch arguments");
        IQM_CHECK(ch == 1 || ch == 3 || ch == 4, IQM_ERR_INVALID, "channels must be 1, 3 or 4");
        IQM_CHECK(n == 1 || pair_stride >= (int64_t)w * h * ch, IQM_ERR_INVALID, "pair_stride too small");
        iqm_ssim_params p;
        iqm_ssim_default_params(&p);
        if (params) p = *params;
        std::vector<double> ms(n);
        ctx->c->ssim_batch(d_img, d_ref, w, h, ch, pair_stride, n, p.kernel_size, p.k1, p.k2,
                           p.sigma, d_maps, ms.data());
        if (mssim_out) std::memcpy(mssim_out, ms.data(), sizeof(double) * n);
    });
}

int iqm_ssim_batch_dev_noalloc(iqm_ctx* ctx, const uint8_t* d_img, const uint8_t* d_ref, int32_t w, int32_t h,
                               int32_t ch, int64_t pair_stride, int32_t n, const iqm_ssim_params* params,
                               float* d_maps, double* d_mssim) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_CHECK(d_img && d_ref && d_mssim && w > 0 && h > 0 && n > 0, IQM_ERR_INVALID, "invalid batch arguments");
        IQM_CHECK(ch == 1 || ch == 3 || ch == 4, IQM_ERR_INVALID, "channels must be 1, 3 or 4");
        IQM_CHECK(n == 1 || pair_stride >= (int64_t)w * h * ch, IQM_ERR_INVALID, "pair_stride too small");
        iqm_ssim_params p;
        iqm_ssim_default_params(&p);
        if (params) p = *params;
        ctx->c->ssim_batch_noalloc(d_img, d_ref, w, h, ch, pair_stride, n, p.kernel_size, p.k1, p.k2, p.sigma,
                                   d_maps, d_mssim);
    });
}

void iqm_fsim_dims(int32_t w, int32_t h, int32_t* F, int32_t* wd, int32_t* hd) {
    int f = std::max(1, (int)std::round(std::min(w, h) / 256.0));
    if (F) *F = f;
    if (wd) *wd = (int)std::round((float)w / (float)f);
    if (hd) *hd = (int)std::round((float)h / (float)f);
}

int iqm_fsim(iqm_ctx* ctx, const iqm_image* img, const iqm_image* ref, double* fsim_out,
             double* fsimc_out, iqm_fsim_maps* maps, iqm_fsim_detail* detail) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        check_pair(img, ref);
        iqm::Ctx& c = *ctx->c;
        roctxRangePushA("iqm_fsim");
        c.begin_marks();
        int F, wd, hd;
        iqm_fsim_dims(img->width, img->height, &F, &wd, &hd);
        c.mark("downscale factor computed"); /* fsim.cpp:85 */
        uint8_t *da, *db;
        upload(c, iqm::WS_IN_IMG, img, &da);
        upload(c, iqm::WS_IN_REF, ref, &db);
        c.mark("images sent to gpu"); /* fsim.cpp:89 */
        const long N = (long)wd * hd;
        iqm::FsimBatchOut out;
        out.want_grad = detail != nullptr;
        float* dm = nullptr;
        if (maps) {
            dm = (float*)c.workspace(iqm::WS_MAPS, sizeof(float) * 5 * N);
            out.d_simc = dm;
            out.d_pcm = dm + N;
            out.d_sim = dm + 2 * N;
            out.d_pc = dm + 3 * N;
        }
        double* dres = c.fsim_batch(da, db, img->width, img->height, img->channels, 0, 1, out);
        double res[8];
        IQM_HIP(hipMemcpyAsync(res, dres, sizeof(res), hipMemcpyDeviceToHost, c.stream()));
        auto d2h = [&](float* dst, const float* src, long cnt) {
            if (dst) IQM_HIP(hipMemcpyAsync(dst, src, sizeof(float) * cnt, hipMemcpyDeviceToHost, c.stream()));
        };
        if (maps) {
            d2h(maps->simc, dm, N);
            d2h(maps->pcm, dm + N, N);
            d2h(maps->sim, dm + 2 * N, N);
            d2h(maps->pc_in, dm + 3 * N, N);
            d2h(maps->pc_ref, dm + 4 * N, N);
        }
        double stats[32];
        if (detail) {
            iqm::Ctx::FsimLayout L = c.fsim_layout(N, 1, c.fsim_ws_base());
            if (detail->yiq_in) {
                d2h(detail->yiq_in, L.yiq + 2 * N, N);      /* I */
                d2h(detail->yiq_in + N, L.yiq + 3 * N, N);  /* Q */
                d2h(detail->yiq_in + 2 * N, L.yiq, N);      /* Y */
            }
            if (detail->yiq_ref) {
                d2h(detail->yiq_ref, L.yiq + 4 * N, N);
                d2h(detail->yiq_ref + N, L.yiq + 5 * N, N);
                d2h(detail->yiq_ref + 2 * N, L.yiq + N, N);
            }
            d2h(detail->grad_in, L.grad, N);
            d2h(detail->grad_ref, L.grad + N, N);
            d2h(detail->an_e, L.an_e, 2 * iqm::kAnePlanes * N);
            d2h(detail->keys, L.keys, 8 * N);
            IQM_HIP(hipMemcpyAsync(stats, L.stats, sizeof(stats), hipMemcpyDeviceToHost, c.stream()));
        }
        c.sync();
        c.mark("FSIM, FSIMc computed"); /* fsim.cpp:155 */
        roctxRangePop();
        if (fsim_out) *fsim_out = res[3];
        if (fsimc_out) *fsimc_out = res[4];
        if (detail) {
            const iqm::FsimSizePlan& P = c.fsim_plan(wd, hd);
            for (int i = 0; i < 4; ++i) detail->em_n[i] = P.em_n[i];
            for (int i = 0; i < 8; ++i) {
                detail->est[i] = P.est[i];
                detail->median[i] = stats[i * 4 + 0];
                detail->noise_power[i] = stats[i * 4 + 1];
                detail->T[i] = stats[i * 4 + 2];
            }
            for (int i = 0; i < 3; ++i) detail->sums[i] = res[i];
        }
    });
}

int iqm_fsim_batch_dev(iqm_ctx* ctx, const uint8_t* d_img, const uint8_t* d_ref, int32_t w,
                       int32_t h, int32_t ch, int64_t pair_stride, int32_t n,
                       float* d_simc_maps, double* fsim_out, double* fsimc_out) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_CHECK(d_img && d_ref && w > 0 && h > 0 && n > 0, IQM_ERR_INVALID, "invalid batch arguments");
        IQM_CHECK(ch == 1 || ch == 3 || ch == 4, IQM_ERR_INVALID, "channels must be 1, 3 or 4");
        IQM_CHECK(n == 1 || pair_stride >= (int64_t)w * h * ch, IQM_ERR_INVALID, "pair_stride too small");
        iqm::Ctx& c = *ctx->c;
        iqm::FsimBatchOut out;
        out.d_simc = d_simc_maps;
        double* dres = c.fsim_batch_graph(d_img, d_ref, w, h, ch, pair_stride, n, out);
        if (fsim_out || fsimc_out) {
            /* pinned staging: the copy is a direct transfer, not a staged one */
            double* res = (double*)c.host_pinned(sizeof(double) * 8 * n);
            IQM_HIP(hipMemcpyAsync(res, dres, sizeof(double) * 8 * n, hipMemcpyDeviceToHost, c.stream()));
            c.sync();
            for (int i = 0; i < n; ++i) {
                if (fsim_out) fsim_out[i] = res[(size_t)8 * i + 3];
                if (fsimc_out) fsimc_out[i] = res[(size_t)8 * i + 4];
            }
        }
    });
}

int iqm_fsim_batch_dev_async(iqm_ctx* ctx, const uint8_t* d_img, const uint8_t* d_ref, int32_t w,
                             int32_t h, int32_t ch, int64_t pair_stride, int32_t n,
                             float* d_simc_maps, double* d_scores) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_CHECK(d_img && d_ref && d_scores && w > 0 && h > 0 && n > 0, IQM_ERR_INVALID,
                  "invalid batch arguments");
        IQM_CHECK(ch == 1 || ch == 3 || ch == 4, IQM_ERR_INVALID, "channels must be 1, 3 or 4");
        IQM_CHECK(n == 1 || pair_stride >= (int64_t)w * h * ch, IQM_ERR_INVALID, "pair_stride too small");
        iqm::Ctx& c = *ctx->c;
        iqm::FsimBatchOut out;
        out.d_simc = d_simc_maps;
        const double* dres = c.fsim_batch_graph(d_img, d_ref, w, h, ch, pair_stride, n, out);
        c.fsim_scores(dres, n, d_scores);
    });
}

int iqm_fsim_batch_host(iqm_ctx* ctx, const uint8_t* h_img, const uint8_t* h_ref, int32_t w, int32_t h,
                        int32_t ch, int64_t pair_stride, int32_t n, int32_t chunk, double* fsim_out,
                        double* fsimc_out) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_CHECK(h_img && h_ref && w > 0 && h > 0 && n > 0 && chunk > 0, IQM_ERR_INVALID,
                  "invalid batch arguments");
        IQM_CHECK(ch == 1 || ch == 3 || ch == 4, IQM_ERR_INVALID, "channels must be 1, 3 or 4");
        IQM_CHECK(n == 1 || pair_stride >= (int64_t)w * h * ch, IQM_ERR_INVALID, "pair_stride too small");
        ctx->c->fsim_host_stream(h_img, h_ref, w, h, ch, pair_stride, n, chunk, fsim_out, fsimc_out);
    });
}

int iqm_host_alloc(iqm_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(IQM_ERR_INVALID, "null argument");
    return guarded([&] {
        IQM_HIP(hipSetDevice(ctx->c->device()));
        if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            *out = nullptr;
            throw iqm::Error(IQM_ERR_NOMEM, "pinned host allocation failed");
        }
    });
}

int iqm_host_free(iqm_ctx* ctx, void* p) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        ctx->c->sync();
        IQM_HIP(hipHostFree(p));
    });
}

int iqm_batch_dev(iqm_ctx* ctx, const iqm_pair_job* jobs, int32_t n, const iqm_ssim_params* ssim,
                  iqm_pair_result* results) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_CHECK(jobs && results && n > 0, IQM_ERR_INVALID, "invalid batch arguments");
        iqm_ssim_params p;
        iqm_ssim_default_params(&p);
        if (ssim) p = *ssim;
        iqm::SsimArgs sa;
        sa.ks = p.kernel_size;
        sa.k1 = p.k1;
        sa.k2 = p.k2;
        sa.sigma = p.sigma;
        IQM_CHECK(sa.ks >= 3 && sa.ks <= 31 && (sa.ks % 2) == 1, IQM_ERR_INVALID,
                  "kernel_size must be odd and in [3, 31]");
        std::vector<iqm::PairJob> v((size_t)n);
        for (int i = 0; i < n; ++i) {
            v[i].img = jobs[i].d_img;
            v[i].ref = jobs[i].d_ref;
            v[i].w = jobs[i].w;
            v[i].h = jobs[i].h;
            v[i].ch = jobs[i].ch;
            v[i].methods = jobs[i].methods;
            v[i].ssim_map = jobs[i].d_ssim_map;
            v[i].simc_map = jobs[i].d_simc_map;
        }
        iqm::Ctx& c = *ctx->c;
        double* d = c.mixed_batch_graph(v.data(), n, sa);
        std::vector<double> r((size_t)4 * n);
        IQM_HIP(hipMemcpyAsync(r.data(), d, sizeof(double) * r.size(), hipMemcpyDeviceToHost, c.stream()));
        c.sync();
        for (int i = 0; i < n; ++i) {
            results[i].mssim = r[(size_t)4 * i];
            results[i].fsim = r[(size_t)4 * i + 1];
            results[i].fsimc = r[(size_t)4 * i + 2];
        }
    });
}

struct iqm_sur {
    iqm_ctx* ctx;
    iqm::Surrogate* s;
};

int32_t iqm_sur_param_count(void) { return iqm::SurLayout::total; }

int iqm_sur_create(iqm_ctx* ctx, uint64_t seed, iqm_sur** out) {
    if (!ctx || !out) return fail(IQM_ERR_INVALID, "null argument");
    *out = nullptr;
    return guarded([&] {
        IQM_HIP(hipSetDevice(ctx->c->device()));
        auto* h = new iqm_sur{ctx, nullptr};
        try {
            h->s = new iqm::Surrogate(*ctx->c, seed);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void iqm_sur_destroy(iqm_sur* s) {
    if (!s) return;
    guarded([&] { delete s->s; });
    delete s;
}

#define SUR_GUARD(s) \
    if (!(s)) return fail(IQM_ERR_INVALID, "null surrogate")

int iqm_sur_get_params(iqm_sur* s, float* host_params) {
    SUR_GUARD(s);
    return guarded([&] {
        IQM_CHECK(host_params, IQM_ERR_INVALID, "null buffer");
        s->s->get_params(host_params);
    });
}

int iqm_sur_set_params(iqm_sur* s, const float* host_params) {
    SUR_GUARD(s);
    return guarded([&] {
        IQM_CHECK(host_params, IQM_ERR_INVALID, "null buffer");
        s->ctx->c->sync();
        s->s->set_params(host_params);
        s->ctx->c->sync();
    });
}

int iqm_sur_get_grad(iqm_sur* s, float* host_grad) {
    SUR_GUARD(s);
    return guarded([&] {
        IQM_CHECK(host_grad, IQM_ERR_INVALID, "null buffer");
        s->s->get_grad(host_grad);
    });
}

int iqm_sur_state(iqm_sur* s, int64_t* step, uint64_t* data_cursor) {
    SUR_GUARD(s);
    if (step) *step = s->s->m.step;
    if (data_cursor) *data_cursor = s->s->m.data_cursor;
    return IQM_OK;
}

int iqm_sur_make_patches(iqm_sur* s, int32_t n, int32_t size, uint64_t seed, float* d_x, float* d_pc) {
    SUR_GUARD(s);
    return guarded([&] {
        IQM_CHECK(d_x && d_pc, IQM_ERR_INVALID, "null buffer");
        IQM_CHECK(size <= 4096, IQM_ERR_INVALID, "patch size above 4096");
        s->s->make_patches(n, size, seed, d_x, d_pc);
    });
}

int iqm_sur_make_patches_scaled(iqm_sur* s, int32_t n, int32_t size, int32_t scale, uint64_t seed, float* d_x,
                                float* d_pc) {
    SUR_GUARD(s);
    return guarded([&] {
        IQM_CHECK(d_x && d_pc, IQM_ERR_INVALID, "null buffer");
        IQM_CHECK(scale >= 1 && (long)size * scale <= 8192, IQM_ERR_INVALID, "scaled patch size above 8192");
        s->s->make_patches(n, size, seed, d_x, d_pc, scale);
    });
}

static void sur_check_shape(const float* a, const float* b, int n, int h, int w) {
    IQM_CHECK(a && b, IQM_ERR_INVALID, "null buffer");
    IQM_CHECK(n > 0 && h > 0 && w > 0, IQM_ERR_INVALID, "invalid image count or size");
}

int iqm_sur_loss_grad(iqm_sur* s, const float* d_x, const float* d_target, int32_t n, int32_t h,
                      int32_t w, double* loss) {
    SUR_GUARD(s);
    return guarded([&] {
        sur_check_shape(d_x, d_target, n, h, w);
        const double l = s->s->loss_and_grad(d_x, d_target, n, h, w);
        if (loss) *loss = l;
    });
}

int iqm_sur_train_step(iqm_sur* s, const float* d_x, const float* d_target, int32_t n, int32_t h,
                       int32_t w, float lr, double* loss) {
    SUR_GUARD(s);
    return guarded([&] {
        sur_check_shape(d_x, d_target, n, h, w);
        IQM_CHECK(lr > 0.f && std::isfinite(lr), IQM_ERR_INVALID, "learning rate must be positive");
        const double l = s->s->train_step(d_x, d_target, n, h, w, lr);
        if (loss) *loss = l;
    });
}

int iqm_sur_infer(iqm_sur* s, const float* d_x, float* d_out, int32_t n, int32_t h, int32_t w, int32_t fused) {
    SUR_GUARD(s);
    return guarded([&] {
        sur_check_shape(d_x, d_out, n, h, w);
        s->s->infer(d_x, d_out, n, h, w, fused != 0);
        s->ctx->c->sync();
    });
}

int64_t iqm_sur_save(iqm_sur* s, uint8_t* buf, int64_t cap) {
    if (!s) return -1;
    int64_t len = -1;
    if (guarded([&] {
            const std::vector<uint8_t> b = s->s->save();
            len = (int64_t)b.size();
            if (buf && cap >= len) std::memcpy(buf, b.data(), b.size());
        }) != IQM_OK)
        return -1;
    return len;
}

int iqm_sur_load(iqm_sur* s, const uint8_t* buf, int64_t len) {
    SUR_GUARD(s);
    return guarded([&] {
        IQM_CHECK(buf && len > 0, IQM_ERR_INVALID, "null checkpoint");
        s->s->load(std::vector<uint8_t>(buf, buf + len));
    });
}

int iqm_fsim_surrogate_batch_dev(iqm_ctx* ctx, iqm_sur* s, const uint8_t* d_img, const uint8_t* d_ref,
                                 int32_t w, int32_t h, int32_t ch, int64_t pair_stride, int32_t n,
                                 int32_t spot_every, float spot_tol, double* fsim_out,
                                 double* fsimc_out, float* spot_err, int32_t* fell_back) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    SUR_GUARD(s);
    return guarded([&] {
        IQM_CHECK(s->ctx == ctx, IQM_ERR_INVALID, "surrogate belongs to another context");
        IQM_CHECK(d_img && d_ref && w > 0 && h > 0 && n > 0 && spot_every >= 0, IQM_ERR_INVALID,
                  "invalid batch arguments");
        IQM_CHECK(ch == 1 || ch == 3 || ch == 4, IQM_ERR_INVALID, "channels must be 1, 3 or 4");
        IQM_CHECK(n == 1 || pair_stride >= (int64_t)w * h * ch, IQM_ERR_INVALID, "pair_stride too small");
        s->s->fsim_batch(d_img, d_ref, w, h, ch, pair_stride, n, spot_every, spot_tol, fsim_out, fsimc_out,
                         spot_err, fell_back);
    });
}

int iqm_sur_set_failure_memory(iqm_sur* s, int32_t on) {
    SUR_GUARD(s);
    return guarded([&] { s->s->set_remember(on != 0); });
}

int iqm_sur_set_loss_weight(iqm_sur* s, float eps) {
    SUR_GUARD(s);
    return guarded([&] { s->s->set_loss_weight(eps); });
}

int iqm_fft2d(iqm_ctx* ctx, float* data, int32_t w, int32_t h, int32_t batch, int32_t dir,
              int32_t normalize) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_CHECK(data && w > 0 && h > 0 && batch > 0 && (dir == 1 || dir == -1), IQM_ERR_INVALID,
                  "invalid FFT arguments");
        iqm::Ctx& c = *ctx->c;
        const size_t bytes = sizeof(float2) * (size_t)w * h * batch;
        float2* d = (float2*)c.workspace(iqm::WS_FFT, bytes);
        IQM_HIP(hipMemcpyAsync(d, data, bytes, hipMemcpyHostToDevice, c.stream()));
        c.fft2d_dev(d, w, h, batch, dir, normalize != 0);
        IQM_HIP(hipMemcpyAsync(data, d, bytes, hipMemcpyDeviceToHost, c.stream()));
        c.sync();
    });
}

int iqm_dev_alloc(iqm_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(IQM_ERR_INVALID, "null argument");
    return guarded([&] {
        IQM_HIP(hipSetDevice(ctx->c->device()));
        if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) {
            (void)hipGetLastError();
            throw iqm::Error(IQM_ERR_NOMEM, "device allocation failed");
        }
    });
}

int iqm_dev_free(iqm_ctx* ctx, void* p) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        ctx->c->sync();
        IQM_HIP(hipFree(p));
    });
}

int iqm_memcpy_h2d(iqm_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->c->stream()));
        ctx->c->sync();
    });
}

int iqm_memcpy_d2h(iqm_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->c->stream()));
        ctx->c->sync();
    });
}

int iqm_memcpy2d_dev(iqm_ctx* ctx, void* dst, int64_t dst_pitch, const void* src, int64_t src_pitch,
                     int64_t width_bytes, int64_t rows) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_CHECK(dst && src && width_bytes >= 0 && rows >= 0 && dst_pitch >= width_bytes &&
                      src_pitch >= width_bytes, IQM_ERR_INVALID, "invalid 2-D copy");
        IQM_HIP(hipMemcpy2DAsync(dst, (size_t)dst_pitch, src, (size_t)src_pitch, (size_t)width_bytes,
                                 (size_t)rows, hipMemcpyDeviceToDevice, ctx->c->stream()));
    });
}

int iqm_synth_pair_dev(iqm_ctx* ctx, int32_t w, int32_t h, uint64_t seed, uint8_t* d_img, uint8_t* d_ref) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_CHECK(d_img && d_ref, IQM_ERR_INVALID, "null buffer");
        ctx->c->synth_reference(w, h, seed, d_ref);
        ctx->c->synth_distort(d_ref, 3L * w, w, h, seed + 1000003ull, d_img);
    });
}

int iqm_synth_reference_dev(iqm_ctx* ctx, int32_t w, int32_t h, uint64_t seed, uint8_t* d_out) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_CHECK(d_out, IQM_ERR_INVALID, "null buffer");
        ctx->c->synth_reference(w, h, seed, d_out);
    });
}

int iqm_synth_distort_dev(iqm_ctx* ctx, const uint8_t* d_src, int64_t src_row_stride, int32_t w,
                          int32_t h, uint64_t seed, uint8_t* d_out) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_CHECK(d_src && d_out, IQM_ERR_INVALID, "null buffer");
        ctx->c->synth_distort(d_src, src_row_stride, w, h, seed, d_out);
    });
}

void* iqm_stream(iqm_ctx* ctx) { return ctx ? (void*)ctx->c->stream() : nullptr; }

int iqm_sync(iqm_ctx* ctx) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] { ctx->c->sync(); });
}

int iqm_autotune_fsim(iqm_ctx* ctx, int32_t w, int32_t h, int32_t ch, int32_t n,
                      double* us_per_pair) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        iqm::FsimTune t = ctx->c->autotune_fsim(w, h, ch, n);
        if (us_per_pair) *us_per_pair = t.us_per_pair;
    });
}

int iqm_autotune_lines(iqm_ctx* ctx, const iqm_pair_job* jobs, int32_t n, int32_t* n_tuned) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        IQM_CHECK(jobs && n > 0, IQM_ERR_INVALID, "invalid batch arguments");
        std::vector<iqm::PairJob> v((size_t)n);
        for (int i = 0; i < n; ++i) {
            v[i].img = jobs[i].d_img;
            v[i].ref = jobs[i].d_ref;
            v[i].w = jobs[i].w;
            v[i].h = jobs[i].h;
            v[i].ch = jobs[i].ch;
            v[i].methods = jobs[i].methods;
        }
        const int k = ctx->c->autotune_lines_batch(v.data(), n);
        if (n_tuned) *n_tuned = k;
    });
}

int64_t iqm_tune_export(iqm_ctx* ctx, char* buf, int64_t cap) {
    if (!ctx) return -1;
    int64_t len = -1;
    if (guarded([&] {
            const std::string s = ctx->c->export_tunes();
            len = (int64_t)s.size();
            if (buf && cap > 0) {
                const int64_t n = std::min<int64_t>(len, cap - 1);
                std::memcpy(buf, s.data(), (size_t)n);
                buf[n] = '\0';
            }
        }) != IQM_OK)
        return -1;
    return len;
}

int iqm_tune_import(iqm_ctx* ctx, const char* json) {
    if (!ctx || !json) return fail(IQM_ERR_INVALID, "null argument");
    return guarded([&] { ctx->c->import_tunes(json); });
}

int iqm_graph_cache_sizes(iqm_ctx* ctx, int32_t* same_size, int32_t* mixed) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] {
        int a = 0, b = 0;
        ctx->c->graph_cache_sizes(&a, &b);
        if (same_size) *same_size = a;
        if (mixed) *mixed = b;
    });
}

int iqm_set_graphs(iqm_ctx* ctx, int32_t on) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] { ctx->c->set_graphs(on != 0); });
}

int iqm_set_marks(iqm_ctx* ctx, int32_t on) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] { ctx->c->set_marks(on != 0); });
}

int32_t iqm_marks(iqm_ctx* ctx, int32_t max, const char** names, int64_t* ns) {
    if (!ctx) return -1;
    const auto& m = ctx->c->marks();
    for (int32_t i = 0; i < max && i < (int32_t)m.size(); ++i) {
        if (names) names[i] = m[i].first.c_str();
        if (ns) ns[i] = m[i].second;
    }
    return (int32_t)m.size();
}

int iqm_set_stage_timing(iqm_ctx* ctx, int32_t on) {
    if (!ctx) return fail(IQM_ERR_INVALID, "null context");
    return guarded([&] { ctx->c->set_timing(on != 0); });
}

int32_t iqm_stage_times(iqm_ctx* ctx, int32_t max, const char** names, double* total_ms,
                        int64_t* launches) {
    if (!ctx) return -1;
    int r = -1;
    if (guarded([&] { r = ctx->c->stage_times(max, names, total_ms, launches); }) != IQM_OK) return -1;
    return r;
}

void iqm_reset_stage_times(iqm_ctx* ctx) {
    if (ctx) guarded([&] { ctx->c->reset_times(); });
}

}  // extern "C"
