// fdtd_gpu_shim.cpp — the reference's objective entry points with their exact
// C++ signatures, so that a search driver written against FDTD_GPU.cuh:8 and
// FDTD_CPU.h:8 (e.g. test.cxx:94-121, which passes them to TAO as function
// pointers) links against libtomo.so unchanged.
//
//   double FDTD_GPU(const std::vector<double>&)  : REF_GPU preset, 27 parameters,
//       fitness -sum (D - m)^2 / 40 (FDTD_GPU.cu:1004-1008);
//   double FDTD_CPU(const std::vector<double>&)  : REF_CPU preset, 162 parameters,
//       fitness -sum (D - m)^2 / (m^2 40) (FDTD_CPU.cxx:841-845), evaluated by
//       the same gfx950 kernels with the CPU-path semantics.
//
// The "measurement" m is read from a file (the reference hard-codes it at
// FDTD_GPU.cu:996 and FDTD_CPU.cxx:789): $TOMO_MEASUREMENT_GPU /
// $TOMO_MEASUREMENT_CPU, else data/measurement_{gpu,cpu}.f32 next to the
// package (raw little-endian float32, D[exc][freq][angle] then, for the GPU
// file, the phase P in the same layout).
//
// Differences from the reference, by design: a HIP failure throws
// std::runtime_error instead of exit(-1) (FDTD_GPU.cu:35-48); a short vector
// throws std::out_of_range exactly like its vector::at; the device context
// persists between calls (the reference mallocs and frees per call).
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tomo.h"

namespace {

std::mutex g_mu;
tomo_ctx* g_ctx = nullptr;

std::string data_dir() {
    Dl_info info;
    if (dladdr((void*)&data_dir, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        const size_t k = p.find_last_of('/');
        return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/data";
    }
    return "data";
}

std::vector<float> load_measurement(const char* env, const char* name, size_t n) {
    const char* e = std::getenv(env);
    const std::string path = e ? std::string(e) : data_dir() + "/" + name;
    std::vector<float> v;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open measurement file " + path);
    v.resize(n);
    const size_t got = std::fread(v.data(), sizeof(float), n, f);
    std::fclose(f);
    if (got < n) throw std::runtime_error("measurement file " + path + " is shorter than expected");
    return v;
}

tomo_ctx* ctx() {
    if (!g_ctx) {
        const char* d = std::getenv("TOMO_DEVICE");
        if (tomo_ctx_create(d ? std::atoi(d) : 0, 0, &g_ctx) != TOMO_OK)
            throw std::runtime_error(std::string("tomo_ctx_create: ") + tomo_last_error());
    }
    return g_ctx;
}

double run(int preset, const std::vector<double>& arguments, const char* env, const char* file, bool with_phase,
           const char* label) {
    std::lock_guard<std::mutex> g(g_mu);
    tomo_sim_config cfg;
    if (tomo_preset(preset, 0, &cfg) != TOMO_OK) throw std::runtime_error(tomo_last_error());
    if (arguments.size() < (size_t)cfg.n_params) throw std::out_of_range("vector::_M_range_check");
    const size_t L = (size_t)tomo_farfield_len(&cfg);
    static std::vector<float> meas_gpu, meas_cpu;
    std::vector<float>& m = with_phase ? meas_gpu : meas_cpu;
    if (m.empty()) m = load_measurement(env, file, with_phase ? 2 * L : L);
    double fit = 0;
    const int rc = tomo_eval(ctx(), &cfg, arguments.data(), arguments.size(), m.data(), m.size(), &fit, nullptr,
                             nullptr);
    if (rc == TOMO_E_SHORT) throw std::out_of_range("vector::_M_range_check");
    if (rc != TOMO_OK) throw std::runtime_error(std::string("tomo_eval: ") + tomo_last_error());
    std::cout << label << fit << std::endl;
    return fit;
}

}  // namespace

double FDTD_GPU(const std::vector<double>& arguments) {
    return run(TOMO_PRESET_REF_GPU, arguments, "TOMO_MEASUREMENT_GPU", "measurement_gpu.f32", true, "fitness is: ");
}

double FDTD_CPU(const std::vector<double>& arguments) {
    return run(TOMO_PRESET_REF_CPU, arguments, "TOMO_MEASUREMENT_CPU", "measurement_cpu.f32", false, " fitness = ");
}
