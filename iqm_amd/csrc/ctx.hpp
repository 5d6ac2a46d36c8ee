// Device context: one HIP device, one stream, cached device workspaces, FFT
// line plans and per-size FSIM constants.  Replaces the reference's
// VulkanRuntime (src/gpu/base/vulkan_runtime.h:15-82) for this path.
#pragma once

#include "common.hpp"
#include "fft.hpp"

#include <functional>
#include <map>
#include <tuple>
#include <string>
#include <utility>
#include <vector>

namespace iqm {

struct LinePlanHost {
    LinePlan plan{};
    std::vector<float2*> owned; /* device tables of the plan */
};

/* Per-(wd, hd) constants of FSIM: the filter planes and the image-independent
 * noise terms (EM_n of F10, est of F14). */
struct FsimSizePlan {
    int wd = 0, hd = 0;
    long N = 0;
    float* lg = nullptr;   /* 4 planes: log-Gabor x lowpass per scale (F6-F7) */
    float* ang = nullptr;  /* 4 planes: angular spread per orientation (F8)  */
    float* bank = nullptr; /* 16 planes H_z = LG_s * A_o, z = 4 o + s (F9)  */
    double em_n[4] = {0, 0, 0, 0};
    double est[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

/* Noise constants handed to the median/threshold kernel by value. */
struct FsimNoiseArgs {
    double em_n[4];
    double est[8];
};

/* Tile choices of the FSIM launch sequence for one (wd, hd). */
struct FsimTune {
    int two_level = 1; /* use the register two-level passes when the size has a split */
    int col_bx = 16;   /* columns per column-pass workgroup */
    int row_rb = 2;    /* rows per row-pass workgroup */
    int fft_rb = 4;    /* rows per row-FFT workgroup */
    double us_per_pair = 0; /* measured by the autotuner (0 = default, not measured) */
};

struct FsimBatchOut {
    float* d_simc = nullptr;  /* nullable device maps, n * N */
    float* d_pc = nullptr;    /* nullable device, n * 2 * N (pc_in, pc_ref) */
    float* d_pcm = nullptr;   /* nullable device, n * N */
    float* d_sim = nullptr;   /* nullable device, n * N */
    bool want_grad = false;   /* keep the F3 gradient planes in the workspace */
};

class Ctx {
   public:
    explicit Ctx(int device);
    ~Ctx();
    Ctx(const Ctx&) = delete;
    Ctx& operator=(const Ctx&) = delete;

    int device() const { return device_; }
    /* LDS bytes one workgroup may allocate on this device */
    size_t lds_per_block() const { return lds_per_block_; }
    int xcds() const { return xcds_; }
    hipStream_t stream() const { return stream_; }
    const std::string& name() const { return name_; }

    /* device workspace slot of at least `bytes` (contents not preserved) */
    void* workspace(int slot, size_t bytes);
    void sync();

    const LinePlanHost& line_plan(int n);
    const FsimSizePlan& fsim_plan(int wd, int hd);

    /* batched in-place 2-D FFT of device planes */
    void fft2d_dev(float2* d, int w, int h, int planes, int dir, bool normalize);

    /* FSIM over n same-size device-resident pairs; results land in the
     * returned device array of n x 8 doubles:
     * [sum pcm, sum sim, sum simc, fsim, fsimc, -, -, -]. */
    double* fsim_batch(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch,
                       long pair_stride, int n, const FsimBatchOut& out);
    /* F2 of n pairs into the yiq planes; clears the median histograms */
    void launch_prep(const uint8_t* d_img, const uint8_t* d_ref, long pair_stride, int w, int h, int ch,
                     int F, int wd, int hd, int n, float* yiq, unsigned* hist);
    /* FSIM with phase congruency supplied by pc_fn instead of the FFT stages:
     * pc_fn(yiq, N, n, pc) fills pc[pair][2][N] from the Y planes (plane 0 and
     * 1 of every 6-plane pair block of yiq).  Results as fsim_batch. */
    using PcFn = std::function<void(const float* yiq, long N, int n, float* pc)>;
    double* fsim_batch_pcfn(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch,
                            long pair_stride, int n, const FsimBatchOut& out, const PcFn& pc_fn);
    /* pointer to the device workspace of the last fsim_batch (stage buffers) */
    struct FsimLayout {
        long N;
        float* yiq;    /* [pair][6][N]: Y_in, Y_ref, I_in, Q_in, I_ref, Q_ref */
        float* grad;   /* [pair][2][N] */
        float* an_e;   /* [pair][8][2][N] */
        float* keys;   /* [pair][8][N] */
        double* stats; /* [pair][8][4]: median, noise power, T, - */
        double* res;   /* [pair][8] */
    };
    FsimLayout fsim_layout(long N, int n, char* base) const;
    char* fsim_ws_base() { return fsim_ws_; }

    /* SSIM over n same-size device-resident pairs; returns n mssim values
     * on the host. */
    void ssim_batch(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch,
                    long pair_stride, int n, int ks, float k1, float k2, float sigma,
                    float* d_maps, double* mssim_host);

    template <class Kernel, class... Args>
    void launch(const char* tag, Kernel k, dim3 grid, dim3 block, size_t shmem, Args... args) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timing_) {
            e0 = take_event();
            e1 = take_event();
            IQM_HIP(hipEventRecord(e0, stream_));
        }
        hipLaunchKernelGGL(k, grid, block, shmem, stream_, args...);
        IQM_HIP(hipGetLastError());
        if (timing_) {
            IQM_HIP(hipEventRecord(e1, stream_));
            pending_.push_back({tag, {e0, e1}});
        }
    }

    /* Replays the launch sequence of a repeated fsim_batch call (same
     * arguments) as a captured hipGraph.  Returns the result array. */
    double* fsim_batch_graph(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch,
                             long pair_stride, int n, const FsimBatchOut& out);
    void set_graphs(bool on) { graphs_ = on; }

    /* Tile table: tuned entries per (wd, hd); defaults otherwise. */
    FsimTune fsim_tune(int wd, int hd) const;
    void set_fsim_tune(int wd, int hd, const FsimTune& t) {
        tunes_[{wd, hd}] = t;
        drop_graphs();
    }
    /* Times candidate tiles on n synthetic pairs of w x h x ch and keeps the
     * fastest for the downscaled size; returns it. */
    FsimTune autotune_fsim(int w, int h, int ch, int n);
    std::string export_tunes() const;
    void import_tunes(const std::string& json);
    bool graphs() const { return graphs_; }

    void set_timing(bool on);
    int stage_times(int max, const char** names, double* ms, int64_t* launches);
    void reset_times();

   private:
    hipEvent_t take_event();
    void collect_timing();

    int device_ = 0;
    size_t lds_per_block_ = 65536;
    int xcds_ = 8;
    hipStream_t stream_ = nullptr;
    std::string name_;
    std::map<int, std::pair<void*, size_t>> ws_;
    std::map<int, LinePlanHost> line_plans_;
    std::map<std::pair<int, int>, FsimSizePlan> fsim_plans_;
    std::map<std::pair<int, int>, FsimTune> tunes_;
    char* fsim_ws_ = nullptr;
    int last_chunk_ = 1;
    int last_nblk_ = 1;

    struct GraphKey {
        const void *a, *b, *o1, *o2, *o3, *o4;
        long stride;
        int w, h, ch, n;
        bool operator<(const GraphKey& k) const {
            return std::tie(a, b, o1, o2, o3, o4, stride, w, h, ch, n) <
                   std::tie(k.a, k.b, k.o1, k.o2, k.o3, k.o4, k.stride, k.w, k.h, k.ch, k.n);
        }
    };
    struct GraphEntry {
        hipGraphExec_t exec = nullptr;
        double* res = nullptr;
    };
    std::map<GraphKey, GraphEntry> graphs_cache_;
    void drop_graphs();
    bool graphs_ = true;
    bool capturing_ = false;

    bool timing_ = false;
    std::vector<hipEvent_t> free_events_;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending_;
    std::map<std::string, std::pair<double, int64_t>> totals_;
    std::vector<std::string> names_cache_;
};

/* workspace slots */
enum { WS_IN_IMG = 0, WS_IN_REF, WS_FSIM, WS_SSIM, WS_MAPS, WS_FFT, WS_PLAN_TMP, WS_SUR };

size_t line_lds_bytes(const LinePlan& P, int lines, int bufs);
int lines_that_fit(const LinePlan& P, int bufs, size_t budget, int cap);

}  // namespace iqm
