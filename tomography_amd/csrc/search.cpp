// search.cpp — asynchronous population searches over the request loop: the
// replacement of the reference's TAO master/worker drivers (test.cxx:79-121,
// `ParticleSwarmMPI::go` / `DifferentialEvolutionMPI::go` handing FDTD_GPU to
// MPI workers).  TAO itself is an un-vendored submodule (.gitmodules:1-3, no
// pinned version); the algorithms here are the published ones:
//   - differential evolution DE/rand/1/bin (Storn & Price 1997), run
//     steady-state and asynchronously as TAO's MPI master does: a member's
//     trial is generated as soon as a result frees a slot, and replaces the
//     member when its fitness is at least as good (fitness: higher is better);
//   - particle swarm with the constriction coefficients of Clerc & Kennedy
//     (2002), asynchronous: a particle moves and is re-evaluated as soon as
//     its previous evaluation returns, against the best known so far.
// Every evaluation is a tomo_submit request, so the device keeps up to
// `inflight` evaluations resident and advances them together.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/tomo.h"

namespace {

struct Rng {  // splitmix64
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u01() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    int below(int n) { return (int)(u01() * n) % n; }
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Pending {
    int member;           // population index the evaluation belongs to
    std::vector<double> x;
};

}  // namespace

extern "C" int tomo_search_run(tomo_ctx* ctx, const tomo_sim_config* cfg, const double* lo, const double* hi,
                               const tomo_search_config* sc, double* best_x, tomo_search_result* res) {
    if (!ctx || !cfg || !lo || !hi || !sc || !best_x) return TOMO_E_INVALID;
    const int np = sc->pop, dim = sc->dim;
    if (np < 4 || dim < 8 || sc->max_evals < np || (sc->method != TOMO_SEARCH_DE && sc->method != TOMO_SEARCH_PSO))
        return TOMO_E_INVALID;
    for (int d = 0; d < dim; ++d)
        if (!(lo[d] <= hi[d]) || !std::isfinite(lo[d]) || !std::isfinite(hi[d])) return TOMO_E_INVALID;
    const int inflight = sc->inflight > 0 ? std::min(sc->inflight, np) : np;
    Rng rng(sc->seed);
    auto clamp = [&](std::vector<double>& v) {
        for (int d = 0; d < dim; ++d) v[d] = std::min(hi[d], std::max(lo[d], v[d]));
    };
    // population state
    std::vector<std::vector<double>> x(np, std::vector<double>(dim)), vel(np, std::vector<double>(dim, 0.0));
    std::vector<std::vector<double>> pbest(np);
    std::vector<double> fit(np, -INFINITY), pfit(np, -INFINITY);
    std::vector<char> known(np, 0);
    std::vector<double> gbest;
    double gfit = -INFINITY;
    for (int i = 0; i < np; ++i)
        for (int d = 0; d < dim; ++d) {
            x[i][d] = lo[d] + (hi[d] - lo[d]) * rng.u01();
            vel[i][d] = (sc->method == TOMO_SEARCH_PSO) ? (hi[d] - lo[d]) * (rng.u01() - 0.5) * 0.2 : 0.0;
        }
    const double F = sc->F > 0 ? sc->F : 0.5, CR = sc->CR > 0 ? sc->CR : 0.9;
    const double chi = 0.7298, c1 = 1.49618, c2 = 1.49618;  // constriction (Clerc & Kennedy)

    std::map<uint64_t, Pending> out;
    std::vector<int> next_init;  // members whose initial evaluation is not yet issued
    for (int i = np - 1; i >= 0; --i) next_init.push_back(i);
    std::vector<double> lat;
    int64_t issued = 0, done = 0, failed = 0;
    const double t0 = now_s();

    // a trial vector for member i (needs at least 4 known members for DE)
    auto make_trial = [&](int i, std::vector<double>& v) -> bool {
        v.assign(dim, 0.0);
        if (sc->method == TOMO_SEARCH_DE) {
            std::vector<int> kn;
            for (int j = 0; j < np; ++j)
                if (known[j] && j != i) kn.push_back(j);
            if ((int)kn.size() < 3) return false;
            int a = kn[rng.below((int)kn.size())], b, c;
            do b = kn[rng.below((int)kn.size())]; while (b == a);
            do c = kn[rng.below((int)kn.size())]; while (c == a || c == b);
            const int jr = rng.below(dim);
            for (int d = 0; d < dim; ++d) {
                const bool take = (rng.u01() < CR) || d == jr;
                v[d] = take ? x[a][d] + F * (x[b][d] - x[c][d]) : x[i][d];
            }
        } else {
            if (gbest.empty()) return false;
            for (int d = 0; d < dim; ++d) {
                const double r1 = rng.u01(), r2 = rng.u01();
                const double pb = pbest[i].empty() ? x[i][d] : pbest[i][d];
                vel[i][d] = chi * (vel[i][d] + c1 * r1 * (pb - x[i][d]) + c2 * r2 * (gbest[d] - x[i][d]));
                v[d] = x[i][d] + vel[i][d];
            }
        }
        clamp(v);
        return true;
    };
    std::vector<int> ready;  // members whose last evaluation returned and that wait for a new one
    int rc = TOMO_OK;
    while (done < sc->max_evals && rc == TOMO_OK) {
        // issue while there is room
        while ((int)out.size() < inflight && issued < sc->max_evals) {
            Pending p;
            if (!next_init.empty()) {
                p.member = next_init.back();
                next_init.pop_back();
                p.x = x[p.member];
            } else {
                bool ok = false;
                for (size_t k = 0; k < ready.size() && !ok; ++k) {
                    if (make_trial(ready[k], p.x)) {
                        p.member = ready[k];
                        ready.erase(ready.begin() + k);
                        ok = true;
                    }
                }
                if (!ok) break;
            }
            uint64_t tk = 0;
            rc = tomo_submit(ctx, cfg, p.x.data(), (size_t)dim, &tk);
            if (rc) break;
            out.emplace(tk, std::move(p));
            ++issued;
        }
        if (rc || out.empty()) break;
        int active = 0;
        rc = tomo_progress(ctx, &active);
        if (rc) break;
        for (auto it = out.begin(); it != out.end();) {
            double f = 0, l = 0;
            const int pr = tomo_poll(ctx, it->first, &f, &l);
            if (pr == TOMO_E_PENDING) { ++it; continue; }
            ++done;
            const int i = it->second.member;
            if (pr != TOMO_OK || !std::isfinite(f)) {
                ++failed;
                f = -INFINITY;
            } else {
                lat.push_back(l);
            }
            const std::vector<double>& xv = it->second.x;
            if (sc->method == TOMO_SEARCH_DE) {
                if (!known[i] || f >= fit[i]) { x[i] = xv; fit[i] = f; }
            } else {
                x[i] = xv;
                fit[i] = f;
                if (f >= pfit[i]) { pfit[i] = f; pbest[i] = xv; }
            }
            known[i] = 1;
            if (f > gfit || gbest.empty()) { gfit = f; gbest = xv; }
            ready.push_back(i);
            it = out.erase(it);
        }
    }
    // drain what is still resident (only after an error): cancel, then poll
    // once so that the context drops the request records
    for (auto& kv : out) {
        if (tomo_cancel(ctx, kv.first) == TOMO_OK) (void)tomo_poll(ctx, kv.first, nullptr, nullptr);
    }
    if (rc) return rc;
    if (gbest.empty()) return TOMO_E_NONFINITE;
    std::copy(gbest.begin(), gbest.end(), best_x);
    if (res) {
        const double wall = now_s() - t0;
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double q) { return lat.empty() ? 0.0 : lat[std::min(lat.size() - 1, (size_t)(q * lat.size()))]; };
        res->best_fitness = gfit;
        res->evals = done;
        res->failed = failed;
        res->wall_s = wall;
        res->evals_per_s = done / std::max(wall, 1e-9);
        res->p50_latency_s = pct(0.50);
        res->p99_latency_s = pct(0.99);
    }
    return TOMO_OK;
}
