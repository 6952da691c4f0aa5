// pmdd -- command-line entry point with the reference's argv surface
// (src/pmdd/main.cpp:67-467, src/pmdd/util/cli_parser.cpp:181-488).
//
//   pmdd <input> <problem_type 1|2|3> <method 1|2|3> <state_dominance 0|1>
//        [--backend gpu | gpu] [--max-cand N] [--max-prod N]
//        [--save-frontier] [--frontier-out PATH] [--save-stats] [--stats-out PATH]
//
// Invoked through a link named multiobj_nobjs<N>, the run also requires the
// instance to have N objectives (the reference builds one binary per N,
// src/pmdd/makefile:54-58).  The diagram is built by libpmdd_host and
// enumerated by the HIP engine behind include/pmdd_engine.h; the CPU backend
// of the reference is not part of this build.
#include <sys/resource.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iostream>
#include <numeric>
#include <string>
#include <vector>

#include "../../../include/pmdd_engine.h"
#include "../../../include/pmdd_host.h"
#include "output.hpp"

using namespace pmdd_cli;

namespace {

int objectives_from_program_name(const char *argv0) {
    std::string name(argv0 != nullptr ? argv0 : "");
    const size_t slash = name.find_last_of('/');
    if (slash != std::string::npos) {
        name = name.substr(slash + 1);
    }
    const std::string prefix = "multiobj_nobjs";
    if (name.compare(0, prefix.size(), prefix) != 0) {
        return 0;
    }
    const int n = atoi(name.c_str() + prefix.size());
    return n > 0 ? n : 0;
}

long long peak_rss_bytes() {
    struct rusage ru;
    if (getrusage(RUSAGE_SELF, &ru) != 0) {
        return 0;
    }
    return static_cast<long long>(ru.ru_maxrss) * 1024LL;
}

double seconds_since(const std::chrono::steady_clock::time_point &t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace

int main(int argc, char *argv[]) {
    Options opts;
    std::string parse_error;
    if (!parse_cli_args(argc, argv, &opts, &parse_error)) {
        if (!parse_error.empty()) {
            std::cout << parse_error << std::endl;
        }
        print_usage(argv[0]);
        return 1;
    }
    if (opts.backend != BACKEND_GPU) {
        std::cout << "Error - the CPU backend is not part of this build; run with --backend gpu"
                  << std::endl;
        return 1;
    }
    if (opts.problem_type == 3 && opts.method != 1 && opts.method != 3) {
        std::cout << "Error - method " << opts.method << " not valid for TSP" << std::endl;
        return 1;
    }

    // --- decision diagram (main.cpp:113-199) ---
    const auto wall_c0 = std::chrono::steady_clock::now();
    const clock_t cpu_c0 = clock();
    pmdd_dd *dd = nullptr;
    char err[1024] = {0};
    int rc = pmdd_dd_build_from_file(opts.input_path.c_str(), opts.problem_type,
                                     objectives_from_program_name(argv[0]),
                                     opts.state_dominance > 0 ? PMDD_BUILD_STATE_DOMINANCE : 0,
                                     &dd, err, sizeof(err));
    if (rc != PMDD_OK) {
        std::cout << err << std::endl;
        return 1;
    }
    RunStats st;
    st.cpu_compile_s = static_cast<double>(clock() - cpu_c0) / CLOCKS_PER_SEC;
    st.wall_compile_s = seconds_since(wall_c0);

    pmdd_dd_structure ds;
    pmdd_dd_get_structure(dd, &ds);
    Structure rec;
    rec.original_width = ds.original_width;
    rec.reduced_width = ds.reduced_width;
    rec.original_num_nodes = ds.original_num_nodes;
    rec.reduced_num_nodes = ds.reduced_num_nodes;
    rec.max_num_in_arcs = ds.max_num_in_arcs;
    rec.total_num_in_arcs = ds.total_num_in_arcs;
    for (int l = 0; l < ds.num_layers; ++l) {
        rec.max_num_nodes_per_layer.push_back(static_cast<long>(pmdd_dd_layer_size(dd, l)));
    }

    // --- enumeration on the GPU (main.cpp:328-401) ---
    const auto wall_e0 = std::chrono::steady_clock::now();
    const clock_t cpu_e0 = clock();
    const char *what = opts.method == 1 ? "top-down" : opts.method == 2 ? "bottom-up" : "coupled";
    pmdd_ctx *ctx = nullptr;
    rc = pmdd_ctx_create(0, &ctx, err, sizeof(err));
    pmdd_frontier fr{};
    std::vector<double> std_c(ds.num_layers, 0.0), std_s(ds.num_layers, 0.0);
    pmdd_stats es{};
    es.std_capacity = ds.num_layers;
    es.std_candidates_per_layer = std_c.data();
    es.std_frontier_survivors_per_layer = std_s.data();
    if (rc == PMDD_OK) {
        pmdd_opts eo{};
        eo.method = opts.method;
        eo.problem_type = opts.problem_type;
        eo.state_dominance = opts.state_dominance;
        eo.max_cand = opts.max_cand;
        eo.max_prod = opts.max_prod;
        // the stats line carries the kernel times, as the reference's does
        // (it times its kernels whenever it is given stats, topdown.cu:428-507)
        eo.phase_timing = 1;
        rc = pmdd_enumerate(ctx, pmdd_dd_get_view(dd), &eo, &fr, &es, err, sizeof(err));
    }
    if (rc != PMDD_OK) {
        std::cout << "Error - GPU backend requested but " << what << " enumeration failed";
        if (err[0] != '\0') {
            std::cout << ": " << err;
        }
        std::cout << std::endl;
        return 1;
    }
    st.cpu_enumeration_s = static_cast<double>(clock() - cpu_e0) / CLOCKS_PER_SEC;
    st.wall_enumeration_s = seconds_since(wall_e0);
    st.cpu_total_s = st.cpu_compile_s + st.cpu_enumeration_s;

    // lexicographically ascending, as written by the reference (main.cpp:403)
    const int d = fr.d;
    if (pmdd_frontier_sort(&fr) != PMDD_OK) {
        std::cout << "Error - out of host memory sorting the frontier" << std::endl;
        pmdd_frontier_free(&fr);
        return 1;
    }
    std::vector<int32_t> sorted(fr.points, fr.points + static_cast<size_t>(fr.count) * d);
    if (opts.save_frontier) {
        std::string save_error;
        if (!write_frontier_gzip_csv(sorted, d, opts.frontier_out_path, &save_error)) {
            std::cout << "Error - failed to save frontier to '" << opts.frontier_out_path << "'";
            if (!save_error.empty()) {
                std::cout << ": " << save_error;
            }
            std::cout << std::endl;
            return 1;
        }
    }

    st.num_solutions = fr.count;
    st.layer_coupling = es.layer_coupling;
    st.dominance_filtered_total = es.dominance_filtered_total;
    st.wall_pack_transfer_s = es.wall_pack_transfer_s;
    st.wall_join_s = es.wall_join_s;
    st.kernel_expand_td_s = es.kernel_expand_td_s;
    st.kernel_dominance_s = es.kernel_dominance_s;
    st.kernel_total_s = es.kernel_total_s;
    st.gpu_mem_peak_used_bytes = es.gpu_mem_peak_used_bytes;
    st.gpu_mem_peak_reserved_bytes = es.gpu_mem_peak_reserved_bytes;
    st.work_candidates_total = es.work_candidates_total;
    st.work_candidates_peak = es.work_candidates_peak;
    st.work_frontier_survivors_total = es.work_frontier_survivors_total;
    st.work_frontier_peak_points = es.work_frontier_peak_points;
    st.work_join_products_total = es.work_join_products_total;
    if (opts.method == 1 || opts.method == 2) {
        st.std_candidates_per_layer = std_c;
        st.std_frontier_survivors_per_layer = std_s;
    }
    st.cpu_mem_peak_bytes = peak_rss_bytes();
    print_and_save_run_summary(opts, st, rec);

    pmdd_frontier_free(&fr);
    pmdd_ctx_destroy(ctx);
    pmdd_dd_free(dd);
    return 0;
}
