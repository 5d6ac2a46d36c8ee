/*
 * iqm.h -- C ABI of the MI355X SSIM and FSIM/FSIMc library (libiqm_hip.so).
 *
 * Drop-in boundary for the reference's per-method interface (cadik/IQM):
 *   src/gpu/ssim/ssim.h:25-35   class SSIM { SSIMResult computeMetric(rt, image, ref); }
 *   src/gpu/fsim/fsim.h:33-36   class FSIM { FSIMResult computeMetric(rt, image, ref); }
 *   src/input_image.h:11-15     struct InputImage { int width, height; vector<u8> data; }
 * Plain pointers and sizes only; no C++ or torch types cross this boundary.
 * The C++ classes in iqm/iqm.hpp wrap these calls with the reference's
 * exceptions and result structs.
 *
 * Conventions:
 *   - every function returns IQM_OK (0) or a negative-free status code below;
 *     iqm_last_error() gives the message of the calling thread's last failure;
 *   - one context = one device + one HIP stream; a context is not thread safe,
 *     separate contexts may be used from separate threads;
 *   - host pointers are owned by the caller; the context owns device memory,
 *     FFT plans and per-size constants;
 *   - images are interleaved 8-bit with 1 (gray), 3 (RGB) or 4 (RGBA)
 *     channels; gray reads as R=G=B and alpha is ignored, as in the
 *     reference's forced-RGBA loading (src/main.cpp:37-40).
 */
#ifndef IQM_IQM_H
#define IQM_IQM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    IQM_OK = 0,
    IQM_ERR_INVALID = 1,        /* null pointer, bad channel count, bad params     */
    IQM_ERR_SIZE_MISMATCH = 2,  /* main.cpp:74-76 "Compared images must have the same size" */
    IQM_ERR_TOO_SMALL = 3,      /* SSIM crop needs W,H >= kernel_size (ssim.cpp:390-400) */
    IQM_ERR_UNSUPPORTED = 4,    /* size outside the FFT planner's range             */
    IQM_ERR_DEVICE = 5,         /* HIP runtime failure                              */
    IQM_ERR_NOMEM = 6
};

typedef struct iqm_ctx iqm_ctx;

/* Host image view (src/input_image.h:11-15, generalised to 1/3/4 channels and
 * a row stride in bytes; row_stride 0 means width * channels). */
typedef struct {
    const uint8_t* data;
    int32_t width, height, channels, row_stride;
} iqm_image;

/* SSIM parameters, defaults of src/gpu/ssim/ssim.h:31-34. */
typedef struct {
    int32_t kernel_size; /* odd, 3..31 */
    float k1, k2, sigma;
} iqm_ssim_params;

/* Per-pixel FSIM outputs on the downscaled wd x hd grid (all nullable host
 * pointers, wd*hd floats each).  The reference computes these in
 * fsim_phase_congruency.glsl and fsim_final_multiply.glsl:38,48,57 but
 * returns none of them. */
typedef struct {
    float* pc_in;
    float* pc_ref;
    float* pcm;
    float* sim;
    float* simc;
} iqm_fsim_maps;

/* Stage-level detail of one FSIM evaluation (all nullable host pointers). */
typedef struct {
    float* yiq_in;      /* 3 planes I, Q, Y of the downscaled input (F2)       */
    float* yiq_ref;
    float* grad_in;     /* Scharr magnitude (F3)                               */
    float* grad_ref;
    float* an_e;        /* 2 x 5 planes per image: sum_o An, E_0..E_3 (F12)    */
    float* keys;        /* 8 planes |EO_{img, s=0, o}|^2 (F13 median input)    */
    double em_n[4];     /* F10 */
    double est[8];      /* F14 */
    double median[8];   /* F13, [img*4 + o] */
    double noise_power[8];
    double T[8];        /* F15 noise thresholds */
    double sums[3];     /* F17: sum pcm, sum sim, sum simc */
} iqm_fsim_detail;

/* ---- context -------------------------------------------------------- */

/* Creates a context on HIP device `device` (replaces VulkanRuntime,
 * src/gpu/base/vulkan_runtime.cpp:22-67). */
int iqm_create(int device, iqm_ctx** out);
void iqm_destroy(iqm_ctx* ctx);
const char* iqm_last_error(void);
const char* iqm_device_name(iqm_ctx* ctx);
void iqm_ssim_default_params(iqm_ssim_params* p);

/* ---- SSIM ------------------------------------------------------------ */

/* SSIM::computeMetric (src/gpu/ssim/ssim.cpp:71-243) + computeMSSIM
 * (:386-401).  map_out (nullable): width*height floats, row-major. */
int iqm_ssim(iqm_ctx* ctx, const iqm_image* img, const iqm_image* ref,
             const iqm_ssim_params* params, float* map_out, double* mssim_out);

/* Batched SSIM on device-resident inputs: n pairs of w x h x ch images at
 * d_img + i*pair_stride and d_ref + i*pair_stride (bytes, tightly packed
 * rows).  d_maps (nullable, device): n*w*h floats.  mssim_out: n host doubles. */
int iqm_ssim_batch_dev(iqm_ctx* ctx, const uint8_t* d_img, const uint8_t* d_ref, int32_t w,
                       int32_t h, int32_t ch, int64_t pair_stride, int32_t n,
                       const iqm_ssim_params* params, float* d_maps, double* mssim_out);

/* Allocation-free variant of iqm_ssim_batch_dev.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It allocates nothing: the FP64 workgroup partials go to a static device
 * scratch of IQM_SSIM_NOALLOC_MAX_PARTIALS doubles (1 MiB per GPU, held until
 * the process ends), and the caller passes every input and output.  A batch
 * needs n * ceil(w / 256) * ceil(h / 64) partials for kernel_size <= 13
 * (ceil(w / 64) * ceil(h / 16) for 15, ceil(w / 32) * ceil(h / 8) above);
 * more is rejected with IQM_ERR_INVALID.  d_mssim: n device doubles; the call
 * is enqueued on the context's stream and returns without waiting (iqm_sync
 * waits).  All calls on one GPU share its scratch: calls on one context are
 * ordered by its stream, calls that overlap on one GPU from different
 * contexts or streams race. */
#define IQM_SSIM_NOALLOC_MAX_PARTIALS 131072
int iqm_ssim_batch_dev_noalloc(iqm_ctx* ctx, const uint8_t* d_img, const uint8_t* d_ref, int32_t w,
                               int32_t h, int32_t ch, int64_t pair_stride, int32_t n,
                               const iqm_ssim_params* params, float* d_maps, double* d_mssim);

/* ---- FSIM / FSIMc ------------------------------------------------------ */

/* Downscale factor and grid (src/gpu/fsim/fsim.cpp:83,91-92,165-168). */
void iqm_fsim_dims(int32_t w, int32_t h, int32_t* F, int32_t* wd, int32_t* hd);

/* FSIM::computeMetric (src/gpu/fsim/fsim.cpp:80-163): returns FSIM and FSIMc
 * (double; the reference's FSIMResult holds floats).  maps and detail are
 * nullable. */
int iqm_fsim(iqm_ctx* ctx, const iqm_image* img, const iqm_image* ref, double* fsim_out,
             double* fsimc_out, iqm_fsim_maps* maps, iqm_fsim_detail* detail);

/* Batched FSIM on device-resident inputs (same size per call), laid out as
 * for iqm_ssim_batch_dev.  d_simc_maps (nullable, device): n*wd*hd floats.
 * fsim_out / fsimc_out: n host doubles each (nullable). */
int iqm_fsim_batch_dev(iqm_ctx* ctx, const uint8_t* d_img, const uint8_t* d_ref, int32_t w,
                       int32_t h, int32_t ch, int64_t pair_stride, int32_t n,
                       float* d_simc_maps, double* fsim_out, double* fsimc_out);
/* As iqm_fsim_batch_dev, but enqueued on the context's stream without waiting:
 * d_scores (device) receives n x 2 doubles [fsim, fsimc] when the stream
 * reaches them (iqm_sync waits).  Successive calls pipeline; the inputs and
 * d_scores must stay valid until then.  No counterpart in the reference, whose
 * computeMetric returns after each pair; it is for callers that keep batches
 * in flight. */
int iqm_fsim_batch_dev_async(iqm_ctx* ctx, const uint8_t* d_img, const uint8_t* d_ref, int32_t w,
                             int32_t h, int32_t ch, int64_t pair_stride, int32_t n,
                             float* d_simc_maps, double* d_scores);

/* Host-fed FSIM: n pairs in host memory (pair_stride bytes apart, packed
 * rows) scored in chunks of `chunk` pairs.  The input pipeline the reference
 * runs per call (src/gpu/fsim/fsim.cpp:170-252: staging buffer, upload, wait)
 * becomes a two-slot ring: chunk k is copied to the device on a copy stream
 * while chunk k-1 is scored on the context's stream.  Inputs in pinned memory
 * (iqm_host_alloc) are copied by DMA directly; pageable inputs pass through
 * pinned staging filled by the calling thread.  Returns when all scores are
 * in fsim_out / fsimc_out (n host doubles each, nullable); scores equal
 * iqm_fsim_batch_dev's on the same pairs. */
int iqm_fsim_batch_host(iqm_ctx* ctx, const uint8_t* h_img, const uint8_t* h_ref, int32_t w, int32_t h,
                        int32_t ch, int64_t pair_stride, int32_t n, int32_t chunk, double* fsim_out,
                        double* fsimc_out);
/* Pinned host memory for iqm_fsim_batch_host's direct path (hipHostMalloc). */
int iqm_host_alloc(iqm_ctx* ctx, size_t bytes, void** out);
int iqm_host_free(iqm_ctx* ctx, void* p);

/* ---- mixed batches ------------------------------------------------------ */
/* Pairs of uneven sizes, each scored with SSIM and/or FSIM/FSIMc, issued as
 * one launch sequence and replayed as one captured hipGraph when the same
 * batch (same pointers, sizes, methods and SSIM parameters) is submitted
 * again.  The reference scores one pair per call (src/main.cpp:69-113,
 * 180-214); this is the batched form of those calls. */
#define IQM_JOB_SSIM 1
#define IQM_JOB_FSIM 2
typedef struct {
    const uint8_t* d_img;   /* device, packed rows of w * ch bytes */
    const uint8_t* d_ref;
    int32_t w, h, ch;       /* ch 1, 3 or 4 */
    int32_t methods;        /* IQM_JOB_SSIM | IQM_JOB_FSIM */
    float* d_ssim_map;      /* nullable device, w * h floats */
    float* d_simc_map;      /* nullable device, wd * hd floats (iqm_fsim_dims) */
} iqm_pair_job;
typedef struct {
    double mssim, fsim, fsimc; /* NaN for a method not asked for */
} iqm_pair_result;
/* ssim: nullable (defaults). results: n host structs. */
int iqm_batch_dev(iqm_ctx* ctx, const iqm_pair_job* jobs, int32_t n, const iqm_ssim_params* ssim,
                  iqm_pair_result* results);

/* ---- learned phase-congruency surrogate -------------------------------- */
/* No counterpart in the reference: a 6-layer 3x3 convolutional network
 * (32 channels, ReLU, BF16 weights and activations, FP32 accumulation on the
 * matrix cores) from the downscaled luminance Y/255 to the PC map of F15
 * (fsim_phase_congruency.glsl:26-64), trained on this device against exact
 * PC.  A surrogate belongs to one context and must be destroyed before it. */
typedef struct iqm_sur iqm_sur;

/* Number of FP32 parameters (OIHW weights and biases of the 6 layers). */
int32_t iqm_sur_param_count(void);
/* New network with He-uniform weights from `seed` (SplitMix64) and zero
 * biases and Adam moments. */
int iqm_sur_create(iqm_ctx* ctx, uint64_t seed, iqm_sur** out);
void iqm_sur_destroy(iqm_sur* s);
int iqm_sur_get_params(iqm_sur* s, float* host_params);
int iqm_sur_set_params(iqm_sur* s, const float* host_params);
/* Gradient of the last iqm_sur_loss_grad / iqm_sur_train_step. */
int iqm_sur_get_grad(iqm_sur* s, float* host_grad);
/* Adam steps taken and training images consumed (the resume cursor). */
int iqm_sur_state(iqm_sur* s, int64_t* step, uint64_t* data_cursor);
/* n seeded size x size training images (even n: images 2i and 2i+1 form an
 * FSIM pair; 1/f^2 fields, half of them distorted): d_x gets Y/255 and
 * d_pc the exact PC, n*size*size device floats each. */
int iqm_sur_make_patches(iqm_sur* s, int32_t n, int32_t size, uint64_t seed, float* d_x, float* d_pc);
/* The same with the fields made at scale*size and brought to size x size by
 * FSIM's own box downscale (the distortions are applied at full
 * resolution), as a large input reaches the network; scale*size must
 * downscale to size (e.g. 256 x 4). */
int iqm_sur_make_patches_scaled(iqm_sur* s, int32_t n, int32_t size, int32_t scale, uint64_t seed, float* d_x,
                                float* d_pc);
/* Mean squared error over n images of h x w (device) and its gradient
 * (weighted as iqm_sur_set_loss_weight sets). */
int iqm_sur_loss_grad(iqm_sur* s, const float* d_x, const float* d_target, int32_t n, int32_t h,
                      int32_t w, double* loss);
/* One Adam step (beta 0.9/0.999, eps 1e-8) at learning rate lr. */
int iqm_sur_train_step(iqm_sur* s, const float* d_x, const float* d_target, int32_t n, int32_t h,
                       int32_t w, float lr, double* loss);
/* Network output for n images of h x w (device in, device out).  fused != 0:
 * one kernel for all six layers (activations stay in LDS; what the FSIM path
 * uses); 0: the layer-by-layer kernels of training. */
int iqm_sur_infer(iqm_sur* s, const float* d_x, float* d_out, int32_t n, int32_t h, int32_t w,
                  int32_t fused);
/* Checkpoint (weights, Adam moments, step, seed, cursor) as a byte blob;
 * returns its length and writes it when cap allows.  -1 on error. */
int64_t iqm_sur_save(iqm_sur* s, uint8_t* buf, int64_t cap);
int iqm_sur_load(iqm_sur* s, const uint8_t* buf, int64_t len);
/* FSIM/FSIMc as iqm_fsim_batch_dev with PC from the network.  Pairs with
 * i % spot_every == 0 (spot_every > 0) are also run exactly; spot_err gets the
 * largest |PC_net - PC_exact| over 4 seeded 32x32 tiles of their images, and
 * when it is not within spot_tol the batch is recomputed exactly
 * (*fell_back = 1).  The verdict is taken on the device: the exact kernels
 * are enqueued gated by it, so a batch has no host round trip before its
 * results, and a repeated batch (same pointers, sizes and spot settings)
 * replays one captured graph.  A downscaled size whose spot check failed is
 * remembered by the surrogate: later spot-checked batches of that size skip
 * the network and run exactly (fell_back = 1, spot_err = the failing error);
 * iqm_sur_set_failure_memory(s, 0) turns that off.  spot_err and fell_back
 * are nullable.
 * Cost with the committed 6-layer network at 1920x1080: about the exact
 * path's pairs/s without a spot check; a failing check costs network + spot
 * pairs + the exact batch once, then exact speed. */
int iqm_fsim_surrogate_batch_dev(iqm_ctx* ctx, iqm_sur* s, const uint8_t* d_img, const uint8_t* d_ref,
                                 int32_t w, int32_t h, int32_t ch, int64_t pair_stride, int32_t n,
                                 int32_t spot_every, float spot_tol, double* fsim_out,
                                 double* fsimc_out, float* spot_err, int32_t* fell_back);
/* Failure memory of iqm_fsim_surrogate_batch_dev (default on); setting it
 * clears what was remembered. */
int iqm_sur_set_failure_memory(iqm_sur* s, int32_t on);
/* Training loss of iqm_sur_loss_grad / iqm_sur_train_step.  eps < 0 (the
 * default): mean squared error.  eps >= 0: each pixel's squared error
 * weighted by max(PC_a, PC_b) + eps of the FSIM pair it belongs to (images
 * 2i and 2i + 1; n must be even), the weight FSIM gives that pixel's
 * similarity (the loss is the weighted sum over the pixel count).  No
 * counterpart in the reference (the network is this library's). */
int iqm_sur_set_loss_weight(iqm_sur* s, float eps);

/* ---- seeded synthetic workloads on the device -------------------------- */
/* No counterpart in the reference: the benchmark inputs of iqm_amd/synth.py
 * generated on the device.  Reference = three channels of colour-correlated
 * (0.8) 1/f^2 noise, joint min-max to u8; test = the reference blurred
 * (sigma U[0.5,3]), contrast-scaled about the mean (U[0.7,1.3]), plus
 * N(0, sigma_n^2) noise (sigma_n U[2,20]), quantised to 16 levels, clipped.
 * Packed RGB rows.  iqm_synth_pair_dev writes the pair of `seed` (the test
 * image uses seed + 1000003); iqm_synth_distort_dev distorts a w x h RGB view
 * with row stride src_row_stride bytes.  Both run on the context's stream. */
int iqm_synth_pair_dev(iqm_ctx* ctx, int32_t w, int32_t h, uint64_t seed, uint8_t* d_img, uint8_t* d_ref);
/* The reference image of `seed` alone (sides up to 8192). */
int iqm_synth_reference_dev(iqm_ctx* ctx, int32_t w, int32_t h, uint64_t seed, uint8_t* d_out);
int iqm_synth_distort_dev(iqm_ctx* ctx, const uint8_t* d_src, int64_t src_row_stride, int32_t w,
                          int32_t h, uint64_t seed, uint8_t* d_out);

/* ---- FFT (replaces VkFFT: fsim.cpp:370-437 plan, :527-550 append) ------- */

/* In-place batched 2-D C2C FFT of `batch` interleaved complex float planes
 * of w x h (w contiguous), host memory (line lengths up to 8192 that fit the LDS).  dir = -1 forward (unnormalised),
 * +1 inverse; normalize != 0 divides by w*h. */
int iqm_fft2d(iqm_ctx* ctx, float* data, int32_t w, int32_t h, int32_t batch, int32_t dir,
              int32_t normalize);

/* ---- device memory and stream helpers for callers without their own ------ */

int iqm_dev_alloc(iqm_ctx* ctx, size_t bytes, void** out);
int iqm_dev_free(iqm_ctx* ctx, void* p);
int iqm_memcpy_h2d(iqm_ctx* ctx, void* dst, const void* src, size_t bytes);
int iqm_memcpy_d2h(iqm_ctx* ctx, void* dst, const void* src, size_t bytes);
/* 2-D copy of `rows` rows of width_bytes on the context's stream (device to
 * device, asynchronous: ordered before the context's later work). */
int iqm_memcpy2d_dev(iqm_ctx* ctx, void* dst, int64_t dst_pitch, const void* src, int64_t src_pitch,
                     int64_t width_bytes, int64_t rows);
/* The context's hipStream_t (every kernel of the context runs on it). */
void* iqm_stream(iqm_ctx* ctx);
int iqm_sync(iqm_ctx* ctx);

/* Batched device entry points replay a captured hipGraph when called again
 * with the same arguments (default on; IQM_GRAPHS=0 in the environment turns
 * it off at context creation). */
int iqm_set_graphs(iqm_ctx* ctx, int32_t on);
/* Captured graphs the context holds for same-size FSIM batches and for mixed
 * batches (no counterpart in the reference; for monitoring and tests).  Each
 * cache holds at most 16: a call with new buffers beyond that evicts the
 * least recently launched graph and frees its job table. */
int iqm_graph_cache_sizes(iqm_ctx* ctx, int32_t* same_size, int32_t* mixed);

/* ---- autotuner (tile shapes and FFT plans of the FSIM launch sequence) -- */

/* Times candidate tile shapes for images of w x h x ch on n synthetic pairs
 * and keeps the fastest for that downscaled size in the context (used by every
 * later FSIM call of that size).  us_per_pair (nullable) gets its time. */
int iqm_autotune_fsim(iqm_ctx* ctx, int32_t w, int32_t h, int32_t ch, int32_t n,
                      double* us_per_pair);
/* FFT plans of a mixed batch (replaces VkFFT's plan choice, which the
 * reference rebuilds on every call: src/gpu/fsim/fsim.cpp:94,160 and the
 * configuration at :370-437).  For every downscaled line length of the
 * batch's FSIM pairs (IQM_JOB_FSIM), as columns (length hd) and as rows
 * (length wd), times the generic pass alone under each two-level split of
 * the length's plan (the length itself, Rader's p - 1 or Bluestein's m, both
 * factor orders); then times the batch's own column and row passes under
 * three policies -- the default split of every length, the default with its
 * factors swapped, the per-length winners -- and keeps per pass the fastest
 * (the default unless another is 1% faster), for every later FSIM call of
 * those lengths.  n_tuned (nullable) gets the number of (length, role)
 * entries.  Scores change only by FP32 rounding. */
int iqm_autotune_lines(iqm_ctx* ctx, const iqm_pair_job* jobs, int32_t n, int32_t* n_tuned);
/* The context's tile table as JSON (a plan file); returns its length without
 * the terminating NUL and writes at most cap bytes (NUL-terminated when it
 * fits).  Returns -1 on error. */
int64_t iqm_tune_export(iqm_ctx* ctx, char* buf, int64_t cap);
/* Loads a plan file written by iqm_tune_export into the context. */
int iqm_tune_import(iqm_ctx* ctx, const char* json);

/* ---- stage marks (src/timestamps.h:18-51) ------------------------------- */

/* When on, iqm_ssim and iqm_fsim record host time marks at the stage
 * boundaries the reference's computeMetric marks (src/gpu/ssim/ssim.cpp:78,
 * 227,232,235; src/gpu/fsim/fsim.cpp:85,89,95,138,141,155), waiting for the
 * stream at each so that a mark times its stage (as the reference's
 * synchronous stages do).  Off (the default) costs nothing.  Every kernel
 * launch and every iqm_ssim / iqm_fsim call is also a roctx range named after
 * it, for rocprofv3 --marker-trace. */
int iqm_set_marks(iqm_ctx* ctx, int32_t on);
/* Marks of the context's last iqm_ssim / iqm_fsim call: returns their count
 * and copies up to max (name, nanoseconds of the C++ high_resolution_clock
 * since its epoch) pairs; names stay valid until the next call. */
int32_t iqm_marks(iqm_ctx* ctx, int32_t max, const char** names, int64_t* ns);

/* ---- stage timing ------------------------------------------------------- */

/* When on, every kernel launch is bracketed by hipEvents on the context's
 * stream and its duration is accumulated per kernel name. */
int iqm_set_stage_timing(iqm_ctx* ctx, int32_t on);
/* Copies up to `max` (name, total ms, launches) triples; returns the count. */
int32_t iqm_stage_times(iqm_ctx* ctx, int32_t max, const char** names, double* total_ms,
                        int64_t* launches);
void iqm_reset_stage_times(iqm_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
