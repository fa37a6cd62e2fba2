// sanitize_main.cpp — runs the oracle restatement (fdtd_oracle.cpp) and the
// library's host-only set-up code (tomography_amd/csrc/setup.cpp: presets,
// parameter unpacking, rasteriser inputs, coefficient and CPML tables) on
// small cases of every mode, for a build under AddressSanitizer and
// UndefinedBehaviorSanitizer (oracle/Makefile target `sanitize`; host code
// only, g++).  Exit status 0 when every case ran; the sanitizers abort on
// the first finding.
//
// TEST INFRASTRUCTURE ONLY.
#include <cstdio>
#include <vector>

#include "../include/tomo.h"
#include "../tomography_amd/csrc/setup.h"

extern "C" int orc_eval(const tomo_sim_config* c, const double* p, size_t np, const float* target, size_t nt,
                        int fp64, double* fitness, double* D, double* P);

int main() {
    int failures = 0;
    // every preset through the host set-up, every excitation
    for (int id = TOMO_PRESET_REF_GPU; id <= TOMO_PRESET_CFG5; ++id) {
        tomo_sim_config c;
        if (tomo::preset(id, 0, &c) != TOMO_OK) { ++failures; continue; }
        std::vector<double> p;
        if (c.mode == TOMO_MODE_PHYSICAL) {
            p = {3.0, -2.0, 4.0, 20.0, 34.0, 13.0, 0.2, 1.0, -5.0, 6.0, 3.0, 25.0, 40.0, 10.0, 0.5, 0.0};
        } else {
            const int n2 = c.numpatches * c.numpatches;
            for (int k = 0; k < c.n_params; ++k) p.push_back(k < n2 ? 2.0 + k % 5 : (k < 2 * n2 ? 10.0 : 0.1));
        }
        for (int e = 0; e < c.n_exc; e += (c.n_exc > 4 ? 5 : 1)) {
            tomo::SimSetup s;
            if (tomo::build_setup(c, p.data(), p.size(), e, s) != TOMO_OK) ++failures;
        }
        // short vectors and bad values take the error paths
        tomo::SimSetup s;
        if (tomo::build_setup(c, p.data(), 3, 0, s) == TOMO_OK) ++failures;
    }
    // small evaluations of the oracle in both modes and precisions
    tomo_sim_config r;
    tomo::preset(TOMO_PRESET_REF_GPU, 0, &r);
    r.nx = r.ny = 61;
    r.ntff_offset = 12;
    r.steps = 60;
    r.n_exc = 2;
    std::vector<double> pr(27);
    for (int k = 0; k < 27; ++k) pr[k] = k < 9 ? 3.0 : (k < 18 ? 12.0 : 0.1);
    std::vector<float> tgt((size_t)r.n_exc * r.n_freq * r.n_angles * 2, 0.01f);
    std::vector<double> D(tgt.size()), P(tgt.size());
    for (int fp64 = 0; fp64 < 2; ++fp64) {
        double f = 0;
        if (orc_eval(&r, pr.data(), pr.size(), tgt.data(), tgt.size(), fp64, &f, D.data(), P.data())) ++failures;
        r.flags = TOMO_FLAG_PHASE_FITNESS;
        if (orc_eval(&r, pr.data(), pr.size(), tgt.data(), tgt.size(), fp64, &f, D.data(), P.data())) ++failures;
        r.flags = 0;
    }
    tomo_sim_config q;
    tomo::preset(TOMO_PRESET_CFG1, 0, &q);
    q.nx = q.ny = 72;
    q.steps = 80;
    q.n_freq = 3;
    q.n_angles = 12;
    q.breast_radius = 0.01;
    q.skin_thickness = 0.001;
    q.src_radius = 0.014;
    std::vector<double> pq = {2.0, -1.0, 3.0, 20.0, 34.0, 13.0, 0.2, 0.0};
    std::vector<double> Dq((size_t)q.n_exc * q.n_freq * q.n_angles), Pq(Dq.size());
    for (int fp64 = 0; fp64 < 2; ++fp64) {
        double f = 0;
        if (orc_eval(&q, pq.data(), pq.size(), nullptr, 0, fp64, &f, Dq.data(), Pq.data())) ++failures;
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "sanitized run complete", failures);
    return failures ? 1 : 0;
}
