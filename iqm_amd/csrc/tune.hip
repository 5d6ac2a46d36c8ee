// Autotuner of the FSIM launch sequence: times candidate tile shapes for a
// downscaled size on synthetic pairs and keeps the fastest per (wd, hd).  The
// table round-trips through a small JSON plan file (export/import).
#include "ctx.hpp"
#include "fsim_tl.hpp"

#include <cstdio>
#include <cstring>
#include <sstream>
#include <vector>

namespace iqm {

FsimTune Ctx::fsim_tune(int wd, int hd) const {
    auto it = tunes_.find({wd, hd});
    return it == tunes_.end() ? FsimTune{} : it->second;
}

/* Seeded bytes (a multiplicative hash) for tuning inputs. */
__global__ __launch_bounds__(kBlock) void tune_fill_kernel(uint8_t* p, long n, unsigned seed) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        unsigned v = (unsigned)i * 2654435761u ^ seed;
        v ^= v >> 15;
        v *= 2246822519u;
        v ^= v >> 13;
        p[i] = (uint8_t)(v >> 24);
    }
}

FsimTune Ctx::autotune_fsim(int w, int h, int ch, int n) {
    IQM_CHECK(w > 0 && h > 0 && n > 0 && (ch == 1 || ch == 3 || ch == 4), IQM_ERR_INVALID,
              "invalid autotune arguments");
    const int F = std::max(1, (int)std::round(std::min(w, h) / 256.0));
    const int wd = (int)std::round((float)w / (float)F);
    const int hd = (int)std::round((float)h / (float)F);
    const long pair_bytes = (long)w * h * ch;
    uint8_t* buf = nullptr;
    IQM_HIP(hipMalloc(&buf, 2 * pair_bytes * n));
    struct Free {
        uint8_t* p;
        ~Free() { (void)hipFree(p); }
    } guard{buf};
    launch("tune_fill", tune_fill_kernel, dim3(1024), dim3(kBlock), 0, buf, 2 * pair_bytes * n, 12345u);
    const uint8_t* a = buf;
    const uint8_t* b = buf + pair_bytes * n;

    std::vector<FsimTune> cand;
    const bool split = tl_supported(wd) || tl_supported(hd);
    /* two-level kernels: small tiles; generic transforms: tiles that fill
     * the workgroup's threads */
    const std::vector<int> bx_tl{8, 16}, rb_tl{1, 2, 4}, fr_tl{2, 4, 8};
    const std::vector<int> bx_g{16, 24, 32}, rb_g{2, 3, 4}, fr_g{8, 16};
    for (int tl = split ? 1 : 0; tl >= 0; --tl)
        for (int bx : tl ? bx_tl : bx_g)
            for (int rb : tl ? rb_tl : rb_g)
                for (int fr : tl ? fr_tl : fr_g)
                    for (int pr : {1, 2}) {
                        FsimTune t;
                        t.two_level = tl;
                        t.col_bx = bx;
                        t.row_rb = rb;
                        t.fft_rb = fr;
                        t.prep_rb = pr;
                        t.us_per_pair = -1; /* explicit tiles (not the built-in defaults) */
                        cand.push_back(t);
                    }
    /* state restored however the loop ends: the size's previous tiles (or
     * none), the stage timer, the two events */
    const auto key = std::make_pair(wd, hd);
    auto prev_it = tunes_.find(key);
    const bool had_prev = prev_it != tunes_.end();
    const FsimTune prev = had_prev ? prev_it->second : FsimTune{};
    struct Restore {
        Ctx& c;
        bool was_timing;
        hipEvent_t e0, e1;
        ~Restore() {
            c.timing_ = was_timing;
            c.free_events_.push_back(e0);
            c.free_events_.push_back(e1);
        }
    } restore{*this, timing_, take_event(), take_event()};
    timing_ = false;
    FsimBatchOut out;
    FsimTune best;
    double best_ms = 1e30;
    int failed = 0;
    for (FsimTune t : cand) {
        tunes_[key] = t;
        try {
            fsim_batch(a, b, w, h, ch, pair_bytes, n, out); /* warm-up: plans and workspaces */
            IQM_HIP(hipEventRecord(restore.e0, stream_));
            const int reps = 3;
            for (int r = 0; r < reps; ++r) fsim_batch(a, b, w, h, ch, pair_bytes, n, out);
            IQM_HIP(hipEventRecord(restore.e1, stream_));
            IQM_HIP(hipEventSynchronize(restore.e1));
            float ms = 0;
            IQM_HIP(hipEventElapsedTime(&ms, restore.e0, restore.e1));
            ms /= reps;
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        } catch (const Error&) {
            /* a candidate the device refuses (launch resources, memory) is
             * skipped; the stream is drained before the next one */
            ++failed;
            (void)hipGetLastError();
            (void)hipStreamSynchronize(stream_);
        } catch (...) {
            if (had_prev) tunes_[key] = prev;
            else tunes_.erase(key);
            throw;
        }
    }
    if (best_ms >= 1e30) {
        if (had_prev) tunes_[key] = prev;
        else tunes_.erase(key);
        throw Error(IQM_ERR_DEVICE, "autotune: every one of " + std::to_string(failed) + " candidates failed");
    }
    best.us_per_pair = best_ms * 1e3 / n;
    set_fsim_tune(wd, hd, best);
    return best;
}

std::string Ctx::export_tunes() const {
    std::ostringstream o;
    o << "{\"fsim\": [";
    bool first = true;
    for (auto& kv : tunes_) {
        const FsimTune& t = kv.second;
        o << (first ? "" : ", ") << "{\"wd\": " << kv.first.first << ", \"hd\": " << kv.first.second
          << ", \"two_level\": " << t.two_level << ", \"col_bx\": " << t.col_bx
          << ", \"row_rb\": " << t.row_rb << ", \"fft_rb\": " << t.fft_rb << ", \"prep_rb\": " << t.prep_rb
          << ", \"us_per_pair\": " << t.us_per_pair << "}";
        first = false;
    }
    o << "]}";
    return o.str();
}

/* Reads the entries export_tunes writes (one object per size, fields in any
 * order). */
void Ctx::import_tunes(const std::string& json) {
    size_t pos = 0;
    int count = 0;
    auto field = [](const std::string& obj, const char* key, double& v) {
        const std::string k = std::string("\"") + key + "\"";
        size_t p = obj.find(k);
        if (p == std::string::npos) return false;
        p = obj.find(':', p);
        if (p == std::string::npos) return false;
        return std::sscanf(obj.c_str() + p + 1, "%lf", &v) == 1;
    };
    while ((pos = json.find('{', pos + 1)) != std::string::npos) {
        const size_t end = json.find('}', pos);
        IQM_CHECK(end != std::string::npos, IQM_ERR_INVALID, "tune plan: unterminated entry");
        const std::string obj = json.substr(pos, end - pos + 1);
        double wd, hd, tl, bx, rb, fr, us = 0, pr = 1;
        if (!field(obj, "wd", wd)) continue; /* the outer object */
        IQM_CHECK(field(obj, "hd", hd) && field(obj, "two_level", tl) && field(obj, "col_bx", bx) &&
                      field(obj, "row_rb", rb) && field(obj, "fft_rb", fr),
                  IQM_ERR_INVALID, "tune plan: entry misses a field");
        field(obj, "us_per_pair", us);
        field(obj, "prep_rb", pr); /* absent in older plan files */
        IQM_CHECK(bx >= 1 && bx <= 64 && rb >= 1 && rb <= 16 && fr >= 1 && fr <= 64 && pr >= 1 && pr <= 16,
                  IQM_ERR_INVALID, "tune plan: tile out of range");
        FsimTune t;
        t.two_level = (int)tl;
        t.col_bx = (int)bx;
        t.row_rb = (int)rb;
        t.fft_rb = (int)fr;
        t.prep_rb = (int)pr;
        t.us_per_pair = us;
        set_fsim_tune((int)wd, (int)hd, t);
        ++count;
        pos = end;
    }
    (void)count;
}

}  // namespace iqm
