// iqm -- command line front end with the reference's surface
// (src/main.cpp:69-113, 180-214, 264-293; src/args.cpp:8-61;
// src/methods.{h,cpp}): the same options, method names, stdout lines
// ("MSSIM: ", "FSIM: "/"FSIMc: ") and error messages, with SSIM and FSIM
// computed by libiqm_hip.so on the GPU.
//
// Extra options (passed through the reference's generic "--key value" map):
//   --device N    HIP device ordinal (default 0)
//   --repeat N    evaluate the metric N times and print the mean time, the
//                 counterpart of the reference's profile loop (src/profile.cpp:
//                 218-290); each evaluation is a roctx range "iteration i"
//                 (every stage launch inside it is a range of its own), so
//                 rocprofv3 --marker-trace attributes the kernels per run
// With -v the per-stage Timestamps table of the reference is printed
// (src/timestamps.h:18-51, src/main.cpp:98-100,208-210).
#include <chrono>
#include <cstring>
#include <iostream>
#include <map>
#include <optional>
#include <stdexcept>
#include <string>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "iqm/iqm.hpp"
#include "image_io.hpp"

namespace IQM {

/* src/methods.h:13-19 */
enum class Method {
    SSIM = 1,
    CW_SSIM_CPU = 2,
    SVD = 3,
    FSIM = 4,
    FLIP = 5,
};

/* src/methods.cpp:9-23 */
std::string method_name(const Method& method) {
    switch (method) {
        case Method::SSIM: return "SSIM";
        case Method::CW_SSIM_CPU: return "CW-SSIM";
        case Method::SVD: return "SVD";
        case Method::FSIM: return "FSIM";
        case Method::FLIP: return "FLIP";
        default: throw std::runtime_error("unknown method");
    }
}

/* src/args.h / src/args.cpp:8-61 */
class Args {
   public:
    Args(unsigned argc, const char* argv[]) {
        bool parsedMethod = false, parsedInput = false, parsedReference = false;
        for (unsigned i = 0; i < argc; i++) {
            if (i + 1 < argc) {
                if (std::strcmp(argv[i], "--method") == 0) {
                    static const std::map<std::string, Method> names = {
                        {"SSIM", Method::SSIM}, {"CW_SSIM_CPU", Method::CW_SSIM_CPU},
                        {"SVD", Method::SVD},   {"FSIM", Method::FSIM},
                        {"FLIP", Method::FLIP}};
                    auto it = names.find(argv[i + 1]);
                    if (it == names.end()) throw std::runtime_error("Unknown method");
                    method = it->second;
                    parsedMethod = true;
                } else if (std::strcmp(argv[i], "--input") == 0) {
                    inputPath = argv[i + 1];
                    parsedInput = true;
                } else if (std::strcmp(argv[i], "--ref") == 0) {
                    refPath = argv[i + 1];
                    parsedReference = true;
                } else if (std::strcmp(argv[i], "--output") == 0) {
                    outputPath = std::string(argv[i + 1]);
                } else {
                    options.emplace(std::string(argv[i]), std::string(argv[i + 1]));
                }
            }
            if (std::strcmp(argv[i], "-v") == 0) verbose = true;
        }
        if (!parsedMethod) throw std::runtime_error("missing method");
        if (!parsedInput) throw std::runtime_error("missing input");
        if (!parsedReference) throw std::runtime_error("missing reference");
    }

    Method method = Method::SSIM;
    std::string inputPath, refPath;
    std::optional<std::string> outputPath;
    std::map<std::string, std::string> options;
    bool verbose = false;
};

}  // namespace IQM

namespace {

int opt_int(const IQM::Args& a, const char* key, int dflt) {
    auto it = a.options.find(key);
    return it == a.options.end() ? dflt : std::stoi(it->second);
}

/* convertFloatToChar (src/main.cpp:59-67): v * 255 truncated; values outside
 * [0, 1] are clamped here (the reference's cast is undefined for them). */
std::vector<unsigned char> convertFloatToChar(const std::vector<float>& data) {
    std::vector<unsigned char> out(data.size());
    for (size_t i = 0; i < data.size(); ++i) {
        float v = data[i] * 255.0f;
        v = v < 0.f ? 0.f : (v > 255.f ? 255.f : v);
        out[i] = static_cast<unsigned char>(v);
    }
    return out;
}

void print_timing(const char* what, double ms, int repeat) {
    std::cout << what << " mean time over " << repeat << " runs: " << ms << " ms" << std::endl;
}

void load_pair(const IQM::Args& args, InputImage& input, InputImage& reference) {
    input = iqmcli::load_image(args.inputPath);
    reference = iqmcli::load_image(args.refPath);
    if (input.width != reference.width || input.height != reference.height)
        throw std::runtime_error("Compared images must have the same size");
}

/* ssim() of src/main.cpp:69-113 */
void ssim(const IQM::Args& args) {
    InputImage input, reference;
    load_pair(args, input, reference);
    const IQM::HIP::Runtime rt(opt_int(args, "--device", 0));
    if (args.verbose) std::cout << "Selected device: " << rt.selectedDevice() << std::endl;
    IQM::HIP::SSIM ssim(rt);
    rt.setMarks(args.verbose);
    const int repeat = std::max(1, opt_int(args, "--repeat", 1));
    auto start = std::chrono::high_resolution_clock::now(), begin = start;
    IQM::HIP::SSIMResult result;
    for (int i = 0; i < repeat; ++i) {
        start = std::chrono::high_resolution_clock::now();
        roctxRangePushA(("iteration " + std::to_string(i)).c_str());
        result = ssim.computeMetric(rt, input, reference);
        roctxRangePop();
    }
    auto end = std::chrono::high_resolution_clock::now();
    std::cout << "MSSIM: " << result.mssim << std::endl;
    if (args.verbose) result.timestamps.print(start, end);
    if (repeat > 1)
        print_timing("SSIM", std::chrono::duration<double, std::milli>(end - begin).count() / repeat, repeat);
    if (args.outputPath.has_value()) {
        auto converted = convertFloatToChar(result.imageData);
        if (!iqmcli::write_png_gray(args.outputPath.value(), result.width, result.height, converted.data()))
            throw std::runtime_error("Failed to save output image");
    }
}

/* fsim() of src/main.cpp:180-214 */
void fsim(const IQM::Args& args) {
    InputImage input, reference;
    load_pair(args, input, reference);
    const IQM::HIP::Runtime rt(opt_int(args, "--device", 0));
    IQM::HIP::FSIM fsim(rt);
    if (args.verbose) std::cout << "Selected device: " << rt.selectedDevice() << std::endl;
    fsim.wantMaps = args.outputPath.has_value();
    rt.setMarks(args.verbose);
    const int repeat = std::max(1, opt_int(args, "--repeat", 1));
    auto start = std::chrono::high_resolution_clock::now(), begin = start;
    IQM::HIP::FSIMResult result;
    for (int i = 0; i < repeat; ++i) {
        start = std::chrono::high_resolution_clock::now();
        roctxRangePushA(("iteration " + std::to_string(i)).c_str());
        result = fsim.computeMetric(rt, input, reference);
        roctxRangePop();
    }
    auto end = std::chrono::high_resolution_clock::now();
    std::cout << "FSIM: " << result.fsim << std::endl << "FSIMc: " << result.fsimc << std::endl;
    if (args.verbose) result.timestamps.print(start, end);
    if (repeat > 1)
        print_timing("FSIM", std::chrono::duration<double, std::milli>(end - begin).count() / repeat, repeat);
    if (args.outputPath.has_value()) {
        auto converted = convertFloatToChar(result.simc);
        if (!iqmcli::write_png_gray(args.outputPath.value(), result.wd, result.hd, converted.data()))
            throw std::runtime_error("Failed to save output image");
    }
}

}  // namespace

/* main of src/main.cpp:264-293 */
int main(int argc, const char** argv) {
    try {
        auto args = IQM::Args(argc, argv);
        if (args.verbose) std::cout << "Selected method: " << IQM::method_name(args.method) << std::endl;
        switch (args.method) {
            case IQM::Method::SSIM: ssim(args); break;
            case IQM::Method::CW_SSIM_CPU: break;
            case IQM::Method::SVD: throw std::runtime_error("SVD support was not compiled");
            case IQM::Method::FSIM: fsim(args); break;
            case IQM::Method::FLIP: throw std::runtime_error("FLIP support was not compiled");
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        exit(-1);
    }
    return 0;
}
