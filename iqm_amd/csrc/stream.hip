// Host-fed FSIM batches: the input pipeline in front of the device path.
//
// The reference stages every image of every call through a fresh host-visible
// staging buffer and waits for the device between steps
// (src/gpu/fsim/fsim.cpp:170-252, src/gpu/ssim/ssim.cpp:245-346).  Here a
// run of n host pairs is cut into chunks; chunk k is copied host -> device on
// a copy stream while chunk k-1 is scored on the context's stream:
//
//   copy stream:   [H2D k]--ev h2d[k%2]--> [H2D k+1] ...
//                      ^ waits ev done[k%2] (chunk k-2 has finished reading the slot)
//   ctx stream:              [FSIM k (graph)][scores k -> slot k]--ev done[k%2]--> ...
//
// Two device slots alternate.  Pinned host inputs (iqm_host_alloc) are copied
// by DMA straight from the caller's memory; pageable inputs first go through
// two pinned staging slots, filled by the calling thread while the device
// works on the previous chunks.  Scores of each chunk go to their own part of
// one device array and come back in one copy at the end.
#include "ctx.hpp"

#include <algorithm>
#include <cstring>

namespace iqm {

/* true when [p, p + bytes) starts and ends in pinned host memory the runtime
 * knows (hipHostMalloc); pageable memory answers with an error, cleared here */
static bool pinned_range(const void* p, size_t bytes) {
    for (const char* q : {(const char*)p, (const char*)p + (bytes ? bytes - 1 : 0)}) {
        hipPointerAttribute_t a;
        if (hipPointerGetAttributes(&a, q) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        if (a.type != hipMemoryTypeHost) return false;
    }
    return true;
}

void Ctx::ensure_copy_stream() {
    if (copy_stream_) return;
    IQM_HIP(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
        IQM_HIP(hipEventCreateWithFlags(&h2d_done_[i], hipEventDisableTiming));
        IQM_HIP(hipEventCreateWithFlags(&comp_done_[i], hipEventDisableTiming));
    }
}

void* Ctx::host_staging(size_t bytes) {
    if (stage_bytes_ < bytes) {
        if (stage_) {
            IQM_HIP(hipStreamSynchronize(copy_stream_ ? copy_stream_ : stream_));
            IQM_HIP(hipHostFree(stage_));
            stage_ = nullptr;
            stage_bytes_ = 0;
        }
        if (hipHostMalloc(&stage_, bytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            stage_ = nullptr;
            throw Error(IQM_ERR_NOMEM, "pinned staging allocation of " + std::to_string(bytes) + " bytes failed");
        }
        stage_bytes_ = bytes;
    }
    return stage_;
}

void Ctx::fsim_host_stream(const uint8_t* h_img, const uint8_t* h_ref, int w, int h, int ch, long pair_stride,
                           int n, int chunk, double* fsim_out, double* fsimc_out) {
    const size_t pair_bytes = (size_t)w * h * ch;
    if (n == 1) pair_stride = (long)pair_bytes;
    chunk = std::max(1, std::min(chunk, n));
    const size_t span = (size_t)(n - 1) * (size_t)pair_stride + pair_bytes;
    const bool direct = pinned_range(h_img, span) && pinned_range(h_ref, span);
    ensure_copy_stream();
    const size_t slot_bytes = pair_bytes * (size_t)chunk; /* one image side of one chunk, packed */
    uint8_t* dev = (uint8_t*)workspace(WS_STREAM, 4 * slot_bytes);
    double* d_scores = (double*)workspace(WS_STREAM_SCORES, sizeof(double) * 2 * (size_t)n);
    uint8_t* stage = direct ? nullptr : (uint8_t*)host_staging(4 * slot_bytes);
    bool staged_in_flight[2] = {false, false};
    /* the device slots were last used by an earlier call on the context's stream */
    IQM_HIP(hipEventRecord(comp_done_[0], stream_));
    IQM_HIP(hipEventRecord(comp_done_[1], stream_));
    const int nchunks = (n + chunk - 1) / chunk;
    for (int k = 0; k < nchunks; ++k) {
        const int s = k & 1;
        const int p0 = k * chunk, nb = std::min(chunk, n - p0);
        uint8_t* d_img = dev + (size_t)s * 2 * slot_bytes;
        uint8_t* d_ref = d_img + slot_bytes;
        IQM_HIP(hipStreamWaitEvent(copy_stream_, comp_done_[s], 0));
        for (int side = 0; side < 2; ++side) {
            const uint8_t* src = (side ? h_ref : h_img) + (size_t)p0 * (size_t)pair_stride;
            uint8_t* dst = side ? d_ref : d_img;
            if (direct) {
                IQM_HIP(hipMemcpy2DAsync(dst, pair_bytes, src, (size_t)pair_stride, pair_bytes, (size_t)nb,
                                         hipMemcpyHostToDevice, copy_stream_));
            } else {
                uint8_t* st = stage + ((size_t)s * 2 + side) * slot_bytes;
                if (side == 0 && staged_in_flight[s]) IQM_HIP(hipEventSynchronize(h2d_done_[s]));
                for (int i = 0; i < nb; ++i) std::memcpy(st + (size_t)i * pair_bytes, src + (size_t)i * pair_stride,
                                                         pair_bytes);
                IQM_HIP(hipMemcpyAsync(dst, st, pair_bytes * nb, hipMemcpyHostToDevice, copy_stream_));
            }
        }
        IQM_HIP(hipEventRecord(h2d_done_[s], copy_stream_));
        staged_in_flight[s] = true;
        IQM_HIP(hipStreamWaitEvent(stream_, h2d_done_[s], 0));
        FsimBatchOut out;
        const double* res = fsim_batch_graph(d_img, d_ref, w, h, ch, (long)pair_bytes, nb, out);
        fsim_scores(res, nb, d_scores + 2 * (size_t)p0);
        IQM_HIP(hipEventRecord(comp_done_[s], stream_));
    }
    double* hs = (double*)host_pinned(sizeof(double) * 2 * (size_t)n);
    IQM_HIP(hipMemcpyAsync(hs, d_scores, sizeof(double) * 2 * (size_t)n, hipMemcpyDeviceToHost, stream_));
    sync();
    for (int i = 0; i < n; ++i) {
        if (fsim_out) fsim_out[i] = hs[2 * (size_t)i];
        if (fsimc_out) fsimc_out[i] = hs[2 * (size_t)i + 1];
    }
}

}  // namespace iqm
