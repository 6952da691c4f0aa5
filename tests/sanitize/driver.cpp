// Sanitizer driver (test infrastructure): exercises the host library (the
// instance parsers, the knapsack / independent-set / TSP diagram builders on
// one and on eight host threads, with and without state-dominance partners,
// the frontier sort) and, unless built with -DNO_ORACLE, the oracle's
// enumerators, so that builds with -fsanitize=address,undefined or
// -fsanitize=thread can check them on the CPU.  Built and run by
// tests/test_sanitizers.py through tests/sanitize/Makefile.
//
//   driver FILE:TYPE ...   TYPE is the problem type (1 knapsack, 2 set
//                          packing, 3 TSP) of the instance file
// Prints "ok" and exits 0 when every run agreed.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pmdd_engine.h"
#include "../../include/pmdd_host.h"

#ifndef NO_ORACLE
extern "C" {
struct oracle_stats {
    int64_t v[6];
    int32_t layer_coupling, reserved;
    int64_t w[2];
};
int oracle_enumerate(const pmdd_dd_view *v, int method, int problem_type, int state_dominance,
                     int threads, int score_rule, int32_t **out_points, int64_t *out_count,
                     oracle_stats *st, double *std_cand, double *std_surv);
void oracle_free(void *p);
}
#endif

namespace {

// A checksum over every array of the view, so each element is read.
uint64_t view_sum(const pmdd_dd_view *v) {
    uint64_t h = 1469598103934665603ULL;
    auto mix = [&](int64_t x) { h = (h ^ static_cast<uint64_t>(x)) * 1099511628211ULL; };
    const int64_t nodes = v->layer_off[v->num_layers];
    const int64_t arcs = v->in_off[nodes];
    for (int l = 0; l <= v->num_layers; ++l) mix(v->layer_off[l]);
    for (int64_t g = 0; g <= nodes; ++g) mix(v->in_off[g]);
    for (int64_t a = 0; a < arcs; ++a) {
        mix(v->arc_src[a]);
        mix(v->arc_val[a]);
        for (int o = 0; o < v->d; ++o) mix(v->arc_w[a * v->d + o]);
    }
    if (v->min_weight != nullptr) {
        for (int64_t g = 0; g < nodes; ++g) mix(v->min_weight[g]);
    }
    if (v->sd_partner != nullptr) {
        for (int64_t g = 0; g < nodes; ++g) mix(v->sd_partner[g]);
    }
    return h;
}

int fails = 0;

void check(bool ok, const std::string &what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s\n", what.c_str());
        ++fails;
    }
}

}  // namespace

int main(int argc, char **argv) {
    for (int i = 1; i < argc; ++i) {
        std::string arg(argv[i]);
        const size_t colon = arg.rfind(':');
        if (colon == std::string::npos) {
            std::fprintf(stderr, "usage: driver FILE:TYPE ...\n");
            return 2;
        }
        const std::string path = arg.substr(0, colon);
        const int type = std::atoi(arg.c_str() + colon + 1);
        for (int flags = 0; flags <= 1; ++flags) {
            uint64_t sums[2] = {0, 0};
            for (int t = 0; t < 2; ++t) {
                setenv("PMDD_HOST_THREADS", t == 0 ? "1" : "8", 1);
                pmdd_dd *dd = nullptr;
                char err[256] = {0};
                const int rc = pmdd_dd_build_from_file(
                    path.c_str(), type, 0, flags ? PMDD_BUILD_STATE_DOMINANCE : 0, &dd, err,
                    sizeof(err));
                check(rc == PMDD_OK, path + ": build failed: " + err);
                if (rc != PMDD_OK) {
                    continue;
                }
                const pmdd_dd_view *v = pmdd_dd_get_view(dd);
                sums[t] = view_sum(v);
                pmdd_dd_structure s{};
                pmdd_dd_get_structure(dd, &s);
                int64_t nodes = 0;
                for (int l = 0; l < v->num_layers; ++l) nodes += pmdd_dd_layer_size(dd, l);
                check(nodes == v->layer_off[v->num_layers], path + ": layer sizes");
#ifndef NO_ORACLE
                if (t == 0) {
                    // methods 1, 2, 3 (and state dominance where the build has
                    // its data) give one frontier; the sort orders it
                    std::vector<int32_t> ref;
                    int64_t ref_n = -1;
                    for (int method = 1; method <= 3; ++method) {
                        for (int sd = 0; sd <= flags; ++sd) {
                            int32_t *pts = nullptr;
                            int64_t n = 0;
                            oracle_stats st{};
                            std::vector<double> sc(v->num_layers), ss(v->num_layers);
                            const int orc = oracle_enumerate(v, method, type, sd, 2, 1, &pts, &n,
                                                             &st, sc.data(), ss.data());
                            check(orc == PMDD_OK, path + ": oracle failed");
                            pmdd_frontier fr{v->d, 0, n, pts};
                            check(pmdd_frontier_sort(&fr) == PMDD_OK, path + ": sort failed");
                            std::vector<int32_t> got(pts, pts + n * v->d);
                            if (ref_n < 0) {
                                ref = got;
                                ref_n = n;
                            } else {
                                check(n == ref_n && got == ref,
                                      path + ": method " + std::to_string(method) + " sd " +
                                          std::to_string(sd) + " differs");
                            }
                            oracle_free(pts);
                        }
                    }
                }
#endif
                pmdd_dd_free(dd);
            }
            check(sums[0] == sums[1], path + ": 1 and 8 host threads built different arrays");
        }
    }
    if (fails != 0) {
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
