/*
 * iqm.hpp -- C++ classes over the C ABI, mirroring the reference's
 * per-method interface so that its drivers switch by changing a namespace:
 *
 *   reference (cadik/IQM)                         this header
 *   IQM::GPU::VulkanRuntime                       IQM::HIP::Runtime
 *     (src/gpu/base/vulkan_runtime.h:15-82)
 *   IQM::GPU::SSIM, SSIMResult                    IQM::HIP::SSIM, SSIMResult
 *     (src/gpu/ssim/ssim.h:17-35)
 *   IQM::GPU::FSIM, FSIMResult                    IQM::HIP::FSIM, FSIMResult
 *     (src/gpu/fsim/fsim.h:27-36)
 *   InputImage (src/input_image.h:11-15)          the same struct (RGBA8)
 *
 * Failures throw std::runtime_error carrying the library's message, which for
 * the reference's own checks is the reference's text ("Compared images must
 * have the same size", src/main.cpp:74-76).
 */
#ifndef IQM_IQM_HPP
#define IQM_IQM_HPP

#include <algorithm>
#include <chrono>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "iqm/iqm.h"

#ifndef INPUT_IMAGE_H
#define INPUT_IMAGE_H
/* src/input_image.h:11-15: interleaved RGBA8, row-major, tightly packed */
struct InputImage {
    int width;
    int height;
    std::vector<unsigned char> data;
};
#endif

namespace IQM::HIP {

inline void check(int rc) {
    if (rc != IQM_OK) throw std::runtime_error(iqm_last_error());
}

inline iqm_image view(const InputImage& im) {
    return iqm_image{im.data.data(), im.width, im.height, 4, im.width * 4};
}

/* Host chrono marks (src/timestamps.h:15-51): computeMetric fills `inner`
 * with the library's stage marks when the runtime has marks on
 * (Runtime::setMarks), and print() writes the reference's table: per mark the
 * time since `start` and since the previous mark, in microseconds, then the
 * total. */
class Timestamps {
   public:
    using clock = std::chrono::high_resolution_clock;
    std::vector<std::pair<std::string, clock::time_point>> inner;
    void mark(const std::string& name) { inner.emplace_back(name, clock::now()); }
    /* the marks the library recorded during the last call on ctx */
    void take(iqm_ctx* ctx) {
        const int32_t n = iqm_marks(ctx, 0, nullptr, nullptr);
        if (n <= 0) return;
        std::vector<const char*> names((size_t)n);
        std::vector<int64_t> ns((size_t)n);
        iqm_marks(ctx, n, names.data(), ns.data());
        for (int32_t i = 0; i < n; ++i)
            inner.emplace_back(names[i], clock::time_point(std::chrono::duration_cast<clock::duration>(
                                             std::chrono::nanoseconds(ns[i]))));
    }
    void print(clock::time_point start, clock::time_point end, std::ostream& os = std::cout) const {
        using us = std::chrono::microseconds;
        const auto exec = std::chrono::duration_cast<us>(end - start);
        int longest = 5; /* "TOTAL" */
        for (const auto& m : inner) longest = std::max(longest, (int)m.first.size());
        const int pad = (int)std::ceil(std::log10((double)std::max<long long>(exec.count(), 1))) + 2;
        for (size_t i = 0; i < inner.size(); ++i) {
            const auto since = std::chrono::duration_cast<us>(inner[i].second - start);
            const auto diff = std::chrono::duration_cast<us>(inner[i].second - (i ? inner[i - 1].second : start));
            os << std::setw(longest) << inner[i].first << ": " << std::setw(pad) << (std::to_string(since.count()) + "us")
               << " | " << std::setw(pad) << (std::to_string(diff.count()) + "us") << std::endl;
        }
        os << std::setw(longest) << "TOTAL" << ": " << std::setw(pad) << (std::to_string(exec.count()) + "us")
           << std::endl;
    }
};

class Runtime {
   public:
    explicit Runtime(int device = 0) { check(iqm_create(device, &ctx_)); }
    ~Runtime() { iqm_destroy(ctx_); }
    Runtime(const Runtime&) = delete;
    Runtime& operator=(const Runtime&) = delete;
    iqm_ctx* ctx() const { return ctx_; }
    std::string selectedDevice() const { return iqm_device_name(ctx_); }
    /* stage marks of computeMetric (waits for the device at every stage) */
    void setMarks(bool on) const { check(iqm_set_marks(ctx_, on ? 1 : 0)); }

   private:
    iqm_ctx* ctx_ = nullptr;
};

struct SSIMResult {
    std::vector<float> imageData;
    unsigned int width;
    unsigned int height;
    float mssim;
    Timestamps timestamps;
};

class SSIM {
   public:
    explicit SSIM(const Runtime&) {}
    SSIMResult computeMetric(const Runtime& rt, const InputImage& image, const InputImage& ref) {
        SSIMResult r;
        r.width = (unsigned)image.width;
        r.height = (unsigned)image.height;
        r.imageData.resize((size_t)image.width * image.height);
        iqm_image a = view(image), b = view(ref);
        iqm_ssim_params p{kernelSize, k_1, k_2, sigma};
        double mssim = 0;
        check(iqm_ssim(rt.ctx(), &a, &b, &p, r.imageData.data(), &mssim));
        r.mssim = (float)mssim;
        r.timestamps.take(rt.ctx());
        return r;
    }
    [[nodiscard]] double computeMSSIM(const float* buffer, unsigned width, unsigned height) const {
        /* ssim.cpp:386-401, for callers holding a map */
        double sum = 0;
        const unsigned start = (kernelSize - 1) / 2;
        const unsigned we = width - (kernelSize - 1) / 2, he = height - (kernelSize - 1) / 2;
        for (unsigned y = start; y <= he; y++)
            for (unsigned x = start; x <= we; x++) sum += buffer[y * width + x];
        return sum / static_cast<double>((we - start + 1) * (he - start + 1));
    }

    int kernelSize = 11;
    float k_1 = 0.01f;
    float k_2 = 0.03f;
    float sigma = 1.5f;
};

struct FSIMResult {
    float fsim;
    float fsimc;
    Timestamps timestamps;
    /* per-pixel outputs on the downscaled grid (filled when requested) */
    int wd = 0, hd = 0;
    std::vector<float> simc, pcm;
};

class FSIM {
   public:
    explicit FSIM(const Runtime&) {}
    FSIMResult computeMetric(const Runtime& rt, const InputImage& image, const InputImage& ref) {
        FSIMResult r;
        iqm_image a = view(image), b = view(ref);
        double fs = 0, fc = 0;
        iqm_fsim_maps maps{};
        if (wantMaps) {
            int32_t F;
            iqm_fsim_dims(image.width, image.height, &F, &r.wd, &r.hd);
            r.simc.resize((size_t)r.wd * r.hd);
            r.pcm.resize((size_t)r.wd * r.hd);
            maps.simc = r.simc.data();
            maps.pcm = r.pcm.data();
        }
        check(iqm_fsim(rt.ctx(), &a, &b, &fs, &fc, wantMaps ? &maps : nullptr, nullptr));
        r.fsim = (float)fs;
        r.fsimc = (float)fc;
        r.timestamps.take(rt.ctx());
        return r;
    }

    bool wantMaps = false;
};

}  // namespace IQM::HIP

#endif
