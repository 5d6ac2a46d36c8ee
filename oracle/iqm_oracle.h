/*
 * CPU oracle for the SSIM and FSIM/FSIMc path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is a from-source CPU restatement of the reference's GPU
 * algorithms (cadik/IQM: shaders/ssim/ and shaders/fsim/ GLSL files and the host
 * code in src/gpu/ssim, src/gpu/fsim).  The reference has no CPU SSIM/FSIM of
 * its own, so this is the checker every HIP result is compared with.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it; the
 * product library never links or calls it.
 *
 * Every entry point exists in two precisions: prec == 64 is the accuracy
 * oracle (all arithmetic in double), prec == 32 mirrors the reference's FP32
 * per-pixel arithmetic (reductions still accumulate in double).
 *
 * Parity status of each part is stated next to its function in
 * iqm_oracle.cpp; the noise-energy estimate (F14) carries "parity unpinned"
 * for its dependence on the reference reading never-written device memory.
 */
#ifndef IQM_ORACLE_H
#define IQM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of OpenMP threads used by the oracle (0 = library default). */
void orc_set_threads(int n);
int orc_get_threads(void);

/* 1-D and 2-D complex DFT, interleaved (re, im) doubles.  dir = -1 forward
 * (unnormalised), +1 inverse; normalize != 0 divides by the length. */
int orc_fft1d(int prec, double* data, int n, int dir, int normalize);
int orc_fft2d(int prec, double* data, int w, int h, int dir, int normalize);

/* SSIM (the shaders/ssim/ GLSL files + SSIM::computeMSSIM).  Images are interleaved
 * 8-bit with ch in {1,3,4} channels, tightly packed.  map_out (nullable) gets
 * w*h values, row-major. */
int orc_ssim(int prec, const uint8_t* img, const uint8_t* ref, int w, int h, int ch,
             int kernel_size, double k1, double k2, double sigma,
             double* map_out, double* mssim_out);

typedef struct {
    int F, wd, hd;
    double fsim, fsimc;
    double em_n[4];          /* sum_k H_{s=0,o}(k)^2 over all bins             */
    double est[8];           /* [2o] = sum An^2 * N, [2o+1] = sum AiAj * N      */
    double median[8];        /* [img*4+o] upper median of |EO_{img,s=0,o}|^2   */
    double noise_power[8];   /* [img*4+o]                                      */
    double T[8];             /* noise threshold per [img*4+o]                  */
    double sums[3];          /* sum pcm, sum sim, sum simc                     */
    /* optional per-pixel outputs (nullable), row-major wd*hd each */
    double* yiq_in;          /* 3 planes: I, Q, Y                              */
    double* yiq_ref;
    double* grad_in;
    double* grad_ref;
    double* an_e;            /* 8 x 2 planes: (An, E) for [img*4+o]           */
    double* keys;            /* 8 planes: |EO_{img,s=0,o}|^2                   */
    double* pc_in;
    double* pc_ref;
    double* pcm;
    double* sim;
    double* simc;
} orc_fsim_out;

int orc_fsim(int prec, const uint8_t* img, const uint8_t* ref, int w, int h, int ch,
             orc_fsim_out* out);
/* orc_fsim over n pairs of one size (arrays of n image pointers and n
 * outputs), one pair per thread: the CPU baseline's throughput form. */
int orc_fsim_batch(int prec, int n, const uint8_t* const* imgs, const uint8_t* const* refs, int w, int h, int ch,
                   orc_fsim_out* outs);

/* Frequency-domain filters on a wd x hd grid (fsim_lowpassfilter.glsl,
 * fsim_log_gabor.glsl, fsim_angular.glsl).  loggabor/angular hold 4 planes. */
int orc_fsim_filters(int prec, int wd, int hd, double* lowpass, double* loggabor,
                     double* angular);

/* Per-size constants of the noise model: EM_n (F10) and est (F14). */
int orc_fsim_noise_consts(int prec, int wd, int hd, double* em_n4, double* est8);

/* Downscale factor and downscaled size (fsim.cpp:91-92,165-168). */
void orc_fsim_dims(int w, int h, int* F, int* wd, int* hd);

#ifdef __cplusplus
}
#endif
#endif
