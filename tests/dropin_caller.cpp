// dropin_caller.cpp — a caller shaped like the reference's search driver
// (test.cxx:55-121): it declares the objective functions exactly as
// FDTD_GPU.cuh:8 and FDTD_CPU.h:8 do, hands them to a driver as plain
// function pointers (the role of TAO's ps.go(FDTD_CPU, FDTD_GPU, ...)), and
// links against libtomo.so with no other change.
//
// It checks, and prints one line per check:
//   * FDTD_GPU on the 27-parameter bounds of test.cxx:55-77 returns the same
//     double as tomo_fitness(REF_GPU) for the same vector on a context under
//     the shims' schedule (TOMO_DROPIN_SCHEDULE), and is <= 0;
//   * FDTD_CPU on a 162-parameter vector returns tomo_fitness(REF_CPU);
//   * a vector shorter than the path reads throws std::out_of_range, as the
//     reference's vector::at does (FDTD_GPU.cu:655-693, FDTD_CPU.cxx:681-698);
//   * a population evaluated through the function pointer gives the values
//     of direct calls.
// Exit status 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "../include/tomo.h"

using std::vector;

// the reference's declarations (FDTD_GPU.cuh:8, FDTD_CPU.h:8)
double FDTD_GPU(const vector<double>& arguments);
double FDTD_CPU(const vector<double>& arguments);

typedef double (*objective)(const vector<double>&);

// What a TAO worker does with the pointer it is given: evaluate each member.
static vector<double> evaluate_population(objective f, const vector<vector<double>>& pop) {
    vector<double> out;
    for (const vector<double>& x : pop) out.push_back(f(x));
    return out;
}

static int fails = 0;
static void check(bool ok, const char* what, double a, double b) {
    std::printf("%s %s: %.17g %.17g\n", ok ? "PASS" : "FAIL", what, a, b);
    if (!ok) ++fails;
}

int main() {
    std::setvbuf(stdout, nullptr, _IOLBF, 0);  // each check line reaches a pipe as it is printed
    // test.cxx:55-77: 9 eps_inf in [1, 7], 9 delta_eps in [1.5, 41], 9 sigma in [0, 0]
    vector<vector<double>> pop;
    for (int k = 0; k < 3; ++k) {
        vector<double> x;
        for (int i = 0; i < 9; ++i) x.push_back(1.0 + 6.0 * ((i * 7 + k * 3) % 10) / 9.0);
        for (int i = 0; i < 9; ++i) x.push_back(1.5 + 39.5 * ((i * 5 + k) % 10) / 9.0);
        for (int i = 0; i < 9; ++i) x.push_back(0.0);
        pop.push_back(x);
    }
    const vector<double> via_ptr = evaluate_population(FDTD_GPU, pop);
    tomo_ctx* ctx = nullptr;
    if (tomo_ctx_create(0, 0, &ctx) != TOMO_OK) {
        std::printf("FAIL tomo_ctx_create: %s\n", tomo_last_error());
        return 1;
    }
    // the shims' schedule: the DFT chunk sets the running DFT's partial sums
    const tomo_schedule sched = TOMO_DROPIN_SCHEDULE;
    if (tomo_set_schedule(ctx, &sched) != TOMO_OK) {
        std::printf("FAIL tomo_set_schedule: %s\n", tomo_last_error());
        return 1;
    }
    for (size_t k = 0; k < pop.size(); ++k) {
        const double direct = FDTD_GPU(pop[k]);
        double abi = 0;
        const int rc = tomo_fitness(ctx, TOMO_PRESET_REF_GPU, 0, pop[k].data(), pop[k].size(), nullptr, 0, &abi);
        check(rc == TOMO_OK && direct == abi, "FDTD_GPU == tomo_fitness(REF_GPU)", direct, abi);
        check(via_ptr[k] == direct, "function pointer == direct call", via_ptr[k], direct);
        check(direct <= 0 && std::isfinite(direct), "FDTD_GPU fitness finite and <= 0", direct, 0);
    }
    vector<double> x162;
    for (int i = 0; i < 81; ++i) x162.push_back(1.0 + 39.0 * ((i * 13) % 17) / 16.0);
    for (int i = 0; i < 81; ++i) x162.push_back(1.1 * ((i * 7) % 11) / 10.0);
    const double fc = FDTD_CPU(x162);
    double abi = 0;
    const int rc = tomo_fitness(ctx, TOMO_PRESET_REF_CPU, 0, x162.data(), x162.size(), nullptr, 0, &abi);
    check(rc == TOMO_OK && fc == abi, "FDTD_CPU == tomo_fitness(REF_CPU)", fc, abi);
    for (int which = 0; which < 2; ++which) {
        const vector<double> shortv(which ? 100 : 20, 2.0);
        bool threw = false;
        try {
            (void)(which ? FDTD_CPU(shortv) : FDTD_GPU(shortv));
        } catch (const std::out_of_range&) {
            threw = true;
        }
        check(threw, which ? "FDTD_CPU(short vector) throws std::out_of_range"
                           : "FDTD_GPU(short vector) throws std::out_of_range", threw, 1);
    }
    tomo_stats st{};
    if (tomo_get_stats(ctx, &st) == TOMO_OK)
        std::printf("INFO resident_refused %lld\n", (long long)st.resident_refused);
    tomo_ctx_destroy(ctx);
    std::printf("%s\n", fails ? "FAILED" : "ALL PASS");
    return fails ? 1 : 0;
}
