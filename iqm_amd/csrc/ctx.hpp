// Device context: one HIP device, one stream, cached device workspaces, FFT
// line plans and per-size FSIM constants.  Replaces the reference's
// VulkanRuntime (src/gpu/base/vulkan_runtime.h:15-82) for this path.
#pragma once

#include "common.hpp"
#include "fft.hpp"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <chrono>

#include <functional>
#include <map>
#include <tuple>
#include <string>
#include <utility>
#include <vector>

namespace iqm {

struct LinePlanHost {
    LinePlan plan{};
    std::vector<void*> owned; /* device tables of the plan */
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

/* Planes per image of the row-pass output read by the PC kernel:
 * sum over orientations of An, then E_o for o = 0..3. */
constexpr int kAnePlanes = 5;

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
    int fft_rb = 8;    /* rows per row-FFT workgroup */
    int prep_rb = 1;   /* downscaled rows per prep workgroup */
    double us_per_pair = 0; /* measured by the autotuner; 0: built-in defaults, < 0: explicit tiles */
};

/* One pair of a mixed batch (device pointers, packed rows). */
struct PairJob {
    const uint8_t* img = nullptr;
    const uint8_t* ref = nullptr;
    int w = 0, h = 0, ch = 3;
    int methods = 3;            /* bit 0 SSIM, bit 1 FSIM */
    float* ssim_map = nullptr;  /* nullable device, w * h */
    float* simc_map = nullptr;  /* nullable device, wd * hd */
};

struct SsimArgs {
    int ks = 11;
    float k1 = 0.01f, k2 = 0.03f, sigma = 1.5f;
};

/* One pair of a mixed batch as the batched kernels read it from the device
 * job table: sizes, tiles, line plans and per-pair buffers. */
struct MixJob {
    const uint8_t* img;
    const uint8_t* ref;
    int w, h, ch;
    /* SSIM */
    int ssim_gx, ssim_gy;
    double ssim_count;
    float* ssim_map;
    double* ssim_partial;
    double* ssim_out;
    /* FSIM */
    int F, wd, hd, cols, prep_rb, grad_rb, col_bx, nbx, row_rb, pc_gx, pc_gy;
    long N;
    float invN;
    LinePlan Px, Py;
    const float* bank;
    FsimNoiseArgs na;
    float* yiq;
    float2* A;
    float2* T;
    float* an_e;
    float* keys;
    unsigned* hist;
    float* thr;
    double* stats;
    double* partial;
    double* res;
    float* simc_map;
    double* out; /* [mssim, fsim, fsimc, -] */
};

/* Stages of a mixed batch; each is one launch per LDS class (0: the
 * workgroup fits 64 KiB, 1: it needs more). */
enum { MIX_SSIM = 0, MIX_SSIMFIN, MIX_PREP, MIX_ROWFFT, MIX_COL, MIX_ROW, MIX_SELECT, MIX_PCF, MIX_FIN,
       MIX_STAGES };
struct MixCounts {
    int blocks[MIX_STAGES][2] = {};
    size_t lds[MIX_STAGES][2] = {};
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
    int cus() const { return cus_; } /* compute units of the device */
    /* LDS bytes one workgroup may allocate on this device */
    size_t lds_per_block() const { return lds_per_block_; }
    int xcds() const { return xcds_; }
    hipStream_t stream() const { return stream_; }
    const std::string& name() const { return name_; }

    /* device workspace slot of at least `bytes` (contents not preserved) */
    void* workspace(int slot, size_t bytes);
    /* pinned host buffer of at least `bytes` for result copies (contents not
     * preserved; growing it waits for the stream) */
    void* host_pinned(size_t bytes);
    void sync();

    /* Line plan of length n; split1 > 0 picks the two-level split
     * (split1, L / split1) of the plan's inner length L (n for a direct
     * plan, p - 1 for Rader's, m for Bluestein's) instead of the most
     * balanced one (split1 = 0). */
    const LinePlanHost& line_plan(int n, int split1 = 0);
    /* The plan of length n for the column (LINE_COL) or row (LINE_ROW)
     * passes of FSIM: the tuned split where autotune_line has run, else the
     * default. */
    const LinePlan& line_plan_for(int n, int role);
    const FsimSizePlan& fsim_plan(int wd, int hd);

    /* batched in-place 2-D FFT of device planes */
    void fft2d_dev(float2* d, int w, int h, int planes, int dir, bool normalize);

    /* FSIM over n same-size device-resident pairs; results land in the
     * returned device array of n x 8 doubles:
     * [sum pcm, sum sim, sum simc, fsim, fsimc, -, -, -]. */
    double* fsim_batch(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch,
                       long pair_stride, int n, const FsimBatchOut& out);
    /* [fsim, fsimc] of each pair from a result array of fsim_batch into
     * d_scores (n x 2 doubles, device), on the stream */
    void fsim_scores(const double* res, int n, double* d_scores);
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
        float* an_e;   /* [pair][2][kAnePlanes][N]: sum_o An, E_0..E_3 */
        float* keys;   /* [pair][8][N] */
        double* stats; /* [pair][8][4]: median, noise power, T, - */
        double* res;   /* [pair][8] */
    };
    FsimLayout fsim_layout(long N, int n, char* base) const;
    char* fsim_ws_base() { return fsim_ws_; }

    /* FSIM over n host-resident pairs (pair_stride bytes apart): chunks of
     * `chunk` pairs are copied to the device on a copy stream while the
     * previous chunk is scored on the context's stream (two device slots;
     * pageable inputs pass through two pinned staging slots).  Scores to the
     * host arrays (nullable). */
    void fsim_host_stream(const uint8_t* h_img, const uint8_t* h_ref, int w, int h, int ch, long pair_stride,
                          int n, int chunk, double* fsim_out, double* fsimc_out);

    /* Seeded synthetic RGB images on the device (recipe of iqm_amd/synth.py):
     * the colour-correlated 1/f^2 reference of w x h, and the distortion of a
     * w x h RGB view (row_stride bytes) into packed rows. */
    void synth_reference(int w, int h, uint64_t seed, uint8_t* d_out);
    void synth_distort(const uint8_t* d_src, long row_stride, int w, int h, uint64_t seed, uint8_t* d_out);

    /* SSIM over n same-size device-resident pairs; returns n mssim values
     * on the host. */
    double* ssim_batch(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch,
                       long pair_stride, int n, int ks, float k1, float k2, float sigma,
                       float* d_maps, double* mssim_host);
    /* The same batch with its workgroup partials in the library's static
     * device scratch (no allocation; see ssim_noalloc_setup): mssim of each
     * pair to d_mssim (device, n doubles), enqueued without waiting. */
    void ssim_batch_noalloc(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch, long pair_stride,
                            int n, int ks, float k1, float k2, float sigma, float* d_maps, double* d_mssim);
    /* the launches shared by both: partials to `partial`, results to `out` */
    void ssim_enqueue(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch, long pair_stride, int n,
                      int ks, float k1, float k2, float sigma, float* d_maps, double* partial, double* out,
                      bool timed);

    /* Pairs of any sizes and methods in one launch sequence on the stream:
     * returns the device array of n x 4 doubles [mssim, fsim, fsimc, -] (NaN
     * for a method not asked for).  The _graph variant replays a captured
     * hipGraph of the whole sequence when called again with the same jobs. */
    double* mixed_batch(const PairJob* jobs, int n, const SsimArgs& sp, void* table = nullptr,
                        bool upload = true);
    double* mixed_batch_graph(const PairJob* jobs, int n, const SsimArgs& sp);
    /* per-pair setup of a mixed batch: fills the method's fields of mj with
     * buffers carved from base + off (base may be null to size the
     * workspace), adds its workgroups to cnt (blk: per-stage counts of this
     * pair, cls: its LDS class per stage) */
    void mix_fsim_fill(const PairJob& pj, MixJob& mj, char* base, size_t& off, MixCounts& cnt, int* blk, int* cls);
    void mix_ssim_fill(const PairJob& pj, MixJob& mj, char* base, size_t& off, MixCounts& cnt, int* blk, int* cls,
                       const SsimArgs& sp);
    /* launches of the batched stages over the device job table */
    void mix_fsim_launch(const MixJob* d_jobs, const int* d_starts, int n, const MixCounts& cnt);
    void mix_ssim_launch(const MixJob* d_jobs, const int* d_starts, int n, const MixCounts& cnt, const SsimArgs& sp);

    template <class Kernel, class... Args>
    void launch(const char* tag, Kernel k, dim3 grid, dim3 block, size_t shmem, Args... args) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timing_) {
            e0 = take_event();
            e1 = take_event();
            IQM_HIP(hipEventRecord(e0, stream_));
        }
        roctxRangePushA(tag); /* names the launch for rocprofv3 --marker-trace */
        hipLaunchKernelGGL(k, grid, block, shmem, stream_, args...);
        roctxRangePop();
        IQM_HIP(hipGetLastError());
        if (timing_) {
            IQM_HIP(hipEventRecord(e1, stream_));
            pending_.push_back({tag, {e0, e1}});
        }
    }

    /* Launch gate of the exact FSIM kernels (see gate_closed): while set,
     * every kernel fsim_batch / fsim_batch_pcfn enqueues checks the flag. */
    void set_gate(const int* g) { gate_ = g; }
    const int* gate() const { return gate_; }
    /* bumped whenever the captured graphs are dropped (workspaces moved) */
    uint64_t graph_generation() const { return graph_gen_; }

    /* Replays the launch sequence of a repeated fsim_batch call (same
     * arguments) as a captured hipGraph.  Returns the result array. */
    double* fsim_batch_graph(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch,
                             long pair_stride, int n, const FsimBatchOut& out);
    void set_graphs(bool on) { graphs_ = on; }
    /* marks a stream capture another component runs on this context's
     * stream (workspace growth is refused while it is on) */
    void set_capturing(bool on) { capturing_ = on; }
    bool timing() const { return timing_; }

    /* Tile table: tuned entries per (wd, hd); defaults otherwise. */
    FsimTune fsim_tune(int wd, int hd) const;
    void set_fsim_tune(int wd, int hd, const FsimTune& t) {
        tunes_[{wd, hd}] = t;
        drop_graphs();
    }
    /* Times candidate tiles on n synthetic pairs of w x h x ch and keeps the
     * fastest for the downscaled size; returns it. */
    FsimTune autotune_fsim(int w, int h, int ch, int n);
    /* Times the generic pass of the role on synthetic lines of length n
     * under every two-level split of the plan's inner length and keeps the
     * fastest; returns its first factor (0: the default split). */
    int autotune_line(int n, int role);
    /* Line plans of a mixed batch chosen in the batch: every (length, role)
     * of its FSIM pairs timed alone (autotune_line), then whole policies --
     * the default split of every length, the default with its factors
     * swapped, the per-length winners -- timed over the batch's own column
     * and row passes; each role keeps its fastest policy.  Returns the
     * number of (length, role) entries. */
    int autotune_lines_batch(const PairJob* jobs, int n);
    /* Mean microseconds of one launch of the generic column or row pass of
     * plan P over a synthetic workload (fsim_gen.hip). */
    double time_line_pass(const LinePlan& P, int role, hipEvent_t e0, hipEvent_t e1);
    /* a line plan choice: the split's first factor (0: the default split)
     * and, when timed here, the pass's microseconds under it and under the
     * default */
    struct LineTune {
        int split1 = 0;
        double us = 0, us_default = 0;
    };
    void set_line_tune(int n, int role, const LineTune& t) {
        line_tunes_[{n, role}] = t;
        drop_graphs();
    }
    std::string export_tunes() const;
    void import_tunes(const std::string& json);
    bool graphs() const { return graphs_; }

    /* Host time marks at stage boundaries (the reference's Timestamps,
     * src/timestamps.h:18-51): when on, mark() waits for the stream and
     * records (name, high_resolution_clock nanoseconds); begin_marks() starts
     * the list of a new top-level call.  Off: mark() costs nothing. */
    void set_marks(bool on) { marks_ = on; }
    void begin_marks() { mark_list_.clear(); }
    void mark(const char* name) {
        if (!marks_ || capturing_) return;
        IQM_HIP(hipStreamSynchronize(stream_));
        const auto t = std::chrono::high_resolution_clock::now().time_since_epoch();
        mark_list_.emplace_back(name, (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t).count());
    }
    const std::vector<std::pair<std::string, int64_t>>& marks() const { return mark_list_; }

    void set_timing(bool on);
    int stage_times(int max, const char** names, double* ms, int64_t* launches);
    void reset_times();

   private:
    hipEvent_t take_event();
    void collect_timing();
    void ensure_copy_stream();
    void* host_staging(size_t bytes);

    int device_ = 0;
    int cus_ = 1;
    size_t lds_per_block_ = 65536;
    int xcds_ = 8;
    hipStream_t stream_ = nullptr;
    std::string name_;
    std::map<int, std::pair<void*, size_t>> ws_;
    void* pinned_ = nullptr; /* hipHostMalloc'd result staging */
    size_t pinned_bytes_ = 0;
    /* input pipeline of fsim_host_stream */
    hipStream_t copy_stream_ = nullptr;
    hipEvent_t h2d_done_[2] = {nullptr, nullptr};
    hipEvent_t comp_done_[2] = {nullptr, nullptr};
    void* stage_ = nullptr; /* pinned staging of pageable inputs */
    size_t stage_bytes_ = 0;
    std::map<std::pair<int, int>, LinePlanHost> line_plans_; /* (n, split1) */
    std::map<std::pair<int, int>, LineTune> line_tunes_;      /* (n, role) */
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
        void* table = nullptr; /* device job table a mixed-batch graph reads */
        uint64_t last_use = 0; /* graph_tick_ of its last launch */
    };
    std::map<GraphKey, GraphEntry> graphs_cache_;
    std::map<std::string, GraphEntry> mixed_graphs_;
    uint64_t graph_tick_ = 0;
    uint64_t graph_gen_ = 0;
    const int* gate_ = nullptr;
    void drop_graphs();
    void release_graph(GraphEntry& e);
    /* makes room for one more entry: evicts the least recently launched
     * graph (and its job table) when the cache holds kGraphCacheCap */
    template <class Map>
    void evict_lru(Map& m) {
        if (m.size() < (size_t)kGraphCacheCap) return;
        auto victim = m.begin();
        for (auto it = m.begin(); it != m.end(); ++it)
            if (it->second.last_use < victim->second.last_use) victim = it;
        release_graph(victim->second);
        m.erase(victim);
    }

   public:
    /* captured graphs held per cache (same-size FSIM batches, mixed batches);
     * each holds at most kGraphCacheCap */
    static constexpr int kGraphCacheCap = 16;
    void graph_cache_sizes(int* same_size, int* mixed) const {
        if (same_size) *same_size = (int)graphs_cache_.size();
        if (mixed) *mixed = (int)mixed_graphs_.size();
    }

   private:
    bool graphs_ = true;
    bool capturing_ = false;

    bool marks_ = false;
    std::vector<std::pair<std::string, int64_t>> mark_list_;
    bool timing_ = false;
    std::vector<hipEvent_t> free_events_;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending_;
    std::map<std::string, std::pair<double, int64_t>> totals_;
    std::vector<std::string> names_cache_;
};

/* Static device scratch of the allocation-free SSIM entry point: FP64
 * workgroup partials (1 MiB per GPU, held from library load to process end),
 * and the device ordinals it is set up for. */
constexpr int kNoAllocPartials = 1 << 17;
constexpr int kMaxNoAllocDevices = 64;
/* Takes the device address of the scratch on `device` (which must be
 * current) once; called by every context's constructor. */
void ssim_noalloc_setup(int device);

/* workspace slots */
enum { WS_IN_IMG = 0, WS_IN_REF, WS_FSIM, WS_SSIM, WS_MAPS, WS_FFT, WS_PLAN_TMP, WS_SUR, WS_GEN, WS_BATCH,
       WS_MIX, WS_MIXTAB, WS_STREAM, WS_STREAM_SCORES, WS_TUNE };

/* roles of a line plan in FSIM: the column passes (line length hd) and the
 * row passes (line length wd) */
enum { LINE_COL = 0, LINE_ROW = 1 };

size_t line_lds_bytes(const LinePlan& P, int lines, int bufs);
int lines_that_fit(const LinePlan& P, int bufs, size_t budget, int cap);
/* First factors a of the two-level splits (a, L / a) of the plan's inner
 * length L, both orders, ascending (empty: the plan has none). */
std::vector<int> line_split_candidates(const LinePlan& P);

}  // namespace iqm
