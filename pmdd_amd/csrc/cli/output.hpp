// Run outputs of the CLI: 3-line stdout summary, gzip CSV frontier and the
// JSONL statistics record (src/pmdd/util/output_utils.cpp:87-374).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pmdd_cli {

enum Backend { BACKEND_CPU = 0, BACKEND_GPU = 1 };

struct Options {
    std::string input_path;
    int problem_type = 0;
    int method = 0;
    int state_dominance = 0;
    Backend backend = BACKEND_CPU;
    int cpu_threads = 1;
    long long max_cand = 20000000LL;
    long long max_prod = 625000LL;
    bool save_frontier = false;
    std::string frontier_out_path;
    bool save_stats = false;
    std::string stats_out_path;
};

// Mirror of EnumerationStats (src/pmdd/util/stats.hpp:18-85), the fields the
// outputs print.
struct RunStats {
    double cpu_state_dominance_s = 0.0;
    long long dominance_filtered_total = 0;
    int layer_coupling = 0;
    long long num_solutions = 0;
    double cpu_compile_s = 0.0, cpu_enumeration_s = 0.0, cpu_total_s = 0.0;
    double wall_compile_s = 0.0, wall_enumeration_s = 0.0;
    double wall_state_dominance_s = 0.0, wall_cutset_sort_s = 0.0;
    double wall_cutset_convolution_s = 0.0, wall_cutset_partial_merge_s = 0.0;
    double wall_pack_transfer_s = 0.0, wall_join_s = 0.0;
    double kernel_expand_td_s = 0.0, kernel_dominance_s = 0.0, kernel_total_s = 0.0;
    long long cpu_mem_peak_bytes = 0;
    long long gpu_mem_peak_used_bytes = 0, gpu_mem_peak_reserved_bytes = 0;
    long long work_candidates_total = 0, work_candidates_peak = 0;
    long long work_frontier_survivors_total = 0, work_frontier_peak_points = 0;
    long long work_join_products_total = 0;
    std::vector<double> std_candidates_per_layer, std_frontier_survivors_per_layer;
    int cpu_layers_td = 0, cpu_layers_bu = 0;
    long long cpu_nodes_expanded = 0;
    int cpu_cutset_size = 0;
};

// DDStats (stats.hpp:94-109)
struct Structure {
    long original_width = -1, reduced_width = -1;
    long original_num_nodes = -1, reduced_num_nodes = -1;
    std::vector<long> max_num_nodes_per_layer;
    long max_num_in_arcs = -1, total_num_in_arcs = -1;
    std::string status_state = "ok";
    std::string status_error_message;
};

const char *backend_to_string(Backend b);
void print_usage(const char *prog);
bool parse_cli_args(int argc, char **argv, Options *out, std::string *error);

bool write_frontier_gzip_csv(const std::vector<int32_t> &points, int d, const std::string &path,
                             std::string *error);
bool write_stats_jsonl(const std::string &path, const Options &opts, const RunStats &st,
                       const Structure &rec, std::string *error);
void print_and_save_run_summary(const Options &opts, const RunStats &st, const Structure &rec);

} // namespace pmdd_cli
