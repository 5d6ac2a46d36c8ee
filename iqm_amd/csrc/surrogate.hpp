// Learned phase-congruency surrogate: a 6-layer 3x3 convolutional network,
// 32 channels, ReLU, from the downscaled luminance (Y / 255) to the PC map.
// BF16 weights and activations with FP32 accumulation; FP32 master weights
// and Adam moments for training.  Host side of surrogate.hip.
#pragma once

#include "ctx.hpp"

#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

namespace iqm {

constexpr int kSurCh = 32;
constexpr int kSurLayers = 6;

/* Parameter offsets in the flat FP32 parameter vector (PyTorch OIHW order). */
struct SurLayout {
    static constexpr int w1 = 0;                       /* [32][1][3][3]  */
    static constexpr int b1 = w1 + 32 * 9;             /* [32]           */
    static constexpr int wmid = b1 + 32;               /* 4 x ([32][32][3][3] + [32]) */
    static constexpr int mid_stride = 32 * 32 * 9 + 32;
    static constexpr int w6 = wmid + 4 * mid_stride;   /* [1][32][3][3]  */
    static constexpr int b6 = w6 + 32 * 9;             /* [1]            */
    static constexpr int total = b6 + 1;               /* 37,601 */
    static constexpr int wl(int l) { return wmid + (l - 2) * mid_stride; }           /* l = 2..5 */
    static constexpr int bl(int l) { return wmid + (l - 2) * mid_stride + 32 * 32 * 9; }
};

/* Device verdict of a surrogate batch's spot check. */
struct SpotFlags {
    int use_exact; /* 1: the batch fell back to exact PC */
    float err;     /* largest spot-check tile error */
};

struct SurModel {
    float* master = nullptr;   /* FP32 parameters (SurLayout)                */
    float* grad = nullptr;     /* FP32 gradient of the last step             */
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    uint16_t* wbf = nullptr;   /* BF16 mid-layer weights, [layer][tap][co][ci] */
    uint16_t* wbf_t = nullptr; /* flipped transpose for the input gradient, [layer][tap][ci][co] */
    long step = 0;
    uint64_t seed = 0;
    uint64_t data_cursor = 0;  /* patches consumed (for resume)              */
};

/* Device workspace of one forward/backward over nimg images of h x w. */
struct SurActs {
    int nimg = 0, h = 0, w = 0;
    float* x = nullptr;        /* input Y/255, [img][h][w]                   */
    uint16_t* a[5] = {};       /* post-ReLU activations of layers 1..5, NHWC bf16 */
    float* pred = nullptr;     /* output, [img][h][w]                        */
    float* dpred = nullptr;
    uint16_t* dz[2] = {};      /* ping-pong gradient buffers, NHWC bf16      */
    float* partial = nullptr;  /* weight-gradient partials                   */
    int npart = 0;
};

class Surrogate {
   public:
    Surrogate(Ctx& c, uint64_t seed);
    ~Surrogate();
    Surrogate(const Surrogate&) = delete;
    Surrogate& operator=(const Surrogate&) = delete;

    /* forward of nimg images of h x w (device FP32 Y/255 in, FP32 PC out):
     * fused = one kernel for all layers (inference), else the layer-by-layer
     * kernels that training uses */
    void infer(const float* d_x, float* d_out, int nimg, int h, int w, bool fused = true);
    /* fused forward reading image j at d_x + (j/2)*pstride + (j%2)*istride, scaled */
    void infer_strided(const float* d_x, long pstride, long istride, float scale, float* d_out,
                       int nimg, int h, int w);
    /* one Adam step on a minibatch; returns the MSE loss (synchronises) */
    double train_step(const float* d_x, const float* d_target, int nimg, int h, int w, float lr);
    /* loss and gradient into grad (no update), for gradient checks */
    double loss_and_grad(const float* d_x, const float* d_target, int nimg, int h, int w);
    /* generation of seeded training patches and their exact PC targets */
    void make_patches(int n, int size, uint64_t seed, float* d_x, float* d_pc, int scale = 1);

    /* FSIM/FSIMc over n device-resident pairs with phase congruency from the
     * network.  Pairs i with i % spot_every == 0 (spot_every > 0) are also
     * evaluated exactly; the error is the largest |PC_net - PC_exact| over 4
     * seeded 32 x 32 tiles of each of their two images.  When it is not
     * within tol the whole batch is recomputed exactly.  The verdict is taken
     * on the device and the exact kernels are gated by it, so the sequence
     * has no host round trip (a repeated batch replays one captured graph).
     * With the failure memory on (the default), a downscaled size whose
     * check failed once goes straight to exact PC afterwards (fell_back = 1,
     * spot_err = the failing error).  Scores and verdict to the host
     * (all nullable). */
    void fsim_batch(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch, long pair_stride, int n,
                    int spot_every, float tol, double* fsim_out, double* fsimc_out, float* spot_err,
                    int* fell_back);
    /* the launch sequence of one such batch on the stream (no host sync);
     * returns the device result array (n x 8 doubles, as Ctx::fsim_batch) */
    double* enqueue(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch, long pair_stride, int n,
                    int spot_every, float tol);
    double* enqueue_graph(const uint8_t* d_img, const uint8_t* d_ref, int w, int h, int ch, long pair_stride,
                          int n, int spot_every, float tol);
    /* training loss: eps < 0 mean squared error; eps >= 0 each pixel's
     * squared error weighted by max(PC of the pair's two images) + eps */
    void set_loss_weight(float eps) { loss_eps_ = eps; }
    /* failure memory on/off (clears what was remembered) */
    void set_remember(bool on) {
        remember_ = on;
        failed_.clear();
    }

    void set_params(const float* host);   /* FP32 master, SurLayout order */
    void get_params(float* host) const;
    void get_grad(float* host) const;
    std::vector<uint8_t> save() const;
    void load(const std::vector<uint8_t>& blob);

    SurModel m;

   private:
    void refresh_bf16();
    void ensure_acts(int nimg, int h, int w);
    void forward(int nimg, int h, int w);
    double backward(const float* d_target);
    Ctx& c_;
    SurActs acts_;
    double* d_loss_ = nullptr;
    float* spot_pc_ = nullptr;   /* exact PC maps of the spot-checked pairs */
    float* spot_err_ = nullptr;  /* per (pair, image, tile) max-abs errors */
    size_t spot_cap_ = 0;        /* pairs spot_pc_/spot_err_ hold */
    const float* pc_net_ = nullptr; /* network PC maps of the last enqueue */
    double* d_out_ = nullptr;    /* selected result array, n x 8 */
    size_t out_cap_ = 0;
    SpotFlags* d_flags_ = nullptr;
    bool remember_ = true;
    float loss_eps_ = -1.0f;
    std::map<std::pair<int, int>, float> failed_; /* (wd, hd) -> failing spot error */
    void ensure_batch_buffers(int n, int m, long N);
    void drop_graphs();
    struct GraphKey {
        const void *a, *b;
        long stride;
        int w, h, ch, n, every;
        uint32_t tol;
        bool operator<(const GraphKey& k) const {
            return std::tie(a, b, stride, w, h, ch, n, every, tol) <
                   std::tie(k.a, k.b, k.stride, k.w, k.h, k.ch, k.n, k.every, k.tol);
        }
    };
    struct GraphEntry {
        hipGraphExec_t exec = nullptr;
        double* res = nullptr;
        uint64_t last_use = 0;
    };
    std::map<GraphKey, GraphEntry> graphs_;
    uint64_t graph_tick_ = 0, graph_gen_ = 0;
};

}  // namespace iqm
