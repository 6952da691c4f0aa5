// CLI argument rules and run outputs.
//   parse_cli_args            src/pmdd/util/cli_parser.cpp:181-488
//   write_frontier_gzip_csv   src/pmdd/util/output_utils.cpp:87-152 (zlib in process,
//                             blocks of rows as gzip members on host threads, instead
//                             of a gzip child; decompressed bytes are the same)
//   write_stats_jsonl         src/pmdd/util/output_utils.cpp:154-300
//   print_and_save_run_summary src/pmdd/util/output_utils.cpp:302-374
#include "output.hpp"

#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <charconv>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <thread>

namespace pmdd_cli {

namespace {

std::string stem_of(const std::string &input_path, const char *fallback) {
    std::string base = input_path;
    const size_t slash = base.find_last_of("/\\");
    if (slash != std::string::npos) {
        base = base.substr(slash + 1);
    }
    const size_t dot = base.find_last_of('.');
    if (dot != std::string::npos) {
        base = base.substr(0, dot);
    }
    return base.empty() ? std::string(fallback) : base;
}

bool parse_positive_int(const std::string &value, int *out) {
    if (value.empty()) {
        return false;
    }
    errno = 0;
    char *end = nullptr;
    const long parsed = strtol(value.c_str(), &end, 10);
    if (errno != 0 || end == value.c_str() || *end != '\0' || parsed <= 0 || parsed > INT_MAX) {
        return false;
    }
    if (out != nullptr) {
        *out = static_cast<int>(parsed);
    }
    return true;
}

// positive integer with an optional K / M / B suffix
bool parse_budget(const std::string &value, long long *out) {
    if (value.empty()) {
        return false;
    }
    long long mult = 1;
    std::string digits = value;
    const char last = value.back();
    if (isalpha(static_cast<unsigned char>(last))) {
        digits = value.substr(0, value.size() - 1);
        const char s = static_cast<char>(toupper(static_cast<unsigned char>(last)));
        if (s == 'K') {
            mult = 1000LL;
        } else if (s == 'M') {
            mult = 1000000LL;
        } else if (s == 'B') {
            mult = 1000000000LL;
        } else {
            return false;
        }
    }
    if (digits.empty()) {
        return false;
    }
    for (char c : digits) {
        if (!isdigit(static_cast<unsigned char>(c))) {
            return false;
        }
    }
    errno = 0;
    char *end = nullptr;
    const long long parsed = strtoll(digits.c_str(), &end, 10);
    if (errno != 0 || end == digits.c_str() || *end != '\0' || parsed <= 0 ||
        parsed > LLONG_MAX / mult) {
        return false;
    }
    *out = parsed * mult;
    return true;
}

std::string json_escape(const std::string &v) {
    std::string e;
    for (unsigned char c : v) {
        switch (c) {
        case '"': e += "\\\""; break;
        case '\\': e += "\\\\"; break;
        case '\b': e += "\\b"; break;
        case '\f': e += "\\f"; break;
        case '\n': e += "\\n"; break;
        case '\r': e += "\\r"; break;
        case '\t': e += "\\t"; break;
        default:
            if (c < 0x20) {
                static const char hex[] = "0123456789abcdef";
                e += "\\u00";
                e += hex[(c >> 4) & 0xF];
                e += hex[c & 0xF];
            } else {
                e += static_cast<char>(c);
            }
        }
    }
    return e;
}

template <class T> void json_array(std::ostream &out, const std::vector<T> &v) {
    out << "[";
    for (size_t i = 0; i < v.size(); ++i) {
        if (i > 0) {
            out << ",";
        }
        out << v[i];
    }
    out << "]";
}

bool set_err(std::string *error, const std::string &msg) {
    if (error != nullptr) {
        *error = msg;
    }
    return false;
}

} // namespace

const char *backend_to_string(Backend b) { return b == BACKEND_GPU ? "gpu" : "cpu"; }

void print_usage(const char *prog) {
    std::cout << '\n';
    std::cout << "Usage: " << prog
              << " [input file] [problem type] [method] [state_dominance] [options]\n";
    std::cout << "\n\twhere:\n";
    std::cout << "\t\tproblem_type = 1: knapsack\n";
    std::cout << "\t\tproblem_type = 2: set packing\n";
    std::cout << "\t\tproblem_type = 3: TSP\n\n";
    std::cout << "\t\tmethod = 1: top-down BFS\n";
    std::cout << "\t\tmethod = 2: bottom-up BFS (knapsack and set packing)\n";
    std::cout << "\t\tmethod = 3: dynamic layer cutset\n\n";
    std::cout << "\t\tstate_dominance = 0:  disable state dominance\n";
    std::cout << "\t\tstate_dominance = 1:  state dominance strategy 1\n\n";
    std::cout << "\t\tNamed backend options:\n";
    std::cout << "\t\t\t--backend cpu|gpu\n";
    std::cout << "\t\t\t--cpu-threads <N>   (cpu only)\n";
    std::cout << "\t\tShorthand backend options:\n";
    std::cout << "\t\t\tcpu [N]\n";
    std::cout << "\t\t\tgpu\n";
    std::cout << "\t\tbackend omitted defaults to cpu; this build runs the gpu backend only\n\n";
    std::cout << "\t\t--save-frontier: save Pareto frontier to <input_stem>.frontier.csv.gz\n";
    std::cout << "\t\t--frontier-out <path>: save Pareto frontier to explicit gzip CSV path\n";
    std::cout << "\t\t--save-stats: save one JSONL record with run statistics\n";
    std::cout << "\t\t--stats-out <path>: save JSONL statistics to explicit path\n";
    std::cout << "\t\t--max-cand <N>: GPU candidate batch cap; supports K, M, B suffixes "
                 "(default 20M)\n";
    std::cout << "\t\t--max-prod <N>: GPU coupled batch product cap; supports K, M, B suffixes "
                 "(default 625K); auto: batches of 32K products doubling to 32M\n";
    std::cout << "\t\toptional arguments can be provided in any order\n";
    std::cout << std::endl;
}

bool parse_cli_args(int argc, char **argv, Options *out, std::string *error) {
    if (argc < 5) {
        return set_err(error, "");
    }
    Options o;
    o.input_path = argv[1];
    o.problem_type = atoi(argv[2]);
    o.method = atoi(argv[3]);
    o.state_dominance = atoi(argv[4]);
    if (o.problem_type < 1 || o.problem_type > 3) {
        return set_err(error, "Error - invalid problem type '" + std::string(argv[2]) +
                                  "'. Valid problem types are 1 (knapsack), 2 (set packing), 3 "
                                  "(TSP).");
    }
    bool backend_set = false, from_named = false, from_short = false;
    bool threads_set = false, cand_set = false, prod_set = false;
    for (int i = 5; i < argc; ++i) {
        const std::string tok(argv[i]);
        if (tok == "--backend") {
            if (from_short) {
                return set_err(error, "Error - cannot mix --backend with shorthand backend token.");
            }
            if (backend_set) {
                return set_err(error, "Error - backend provided multiple times.");
            }
            if (i + 1 >= argc) {
                return set_err(error, "Error - --backend requires a value (cpu or gpu).");
            }
            const std::string b(argv[++i]);
            if (b == "cpu") {
                o.backend = BACKEND_CPU;
            } else if (b == "gpu") {
                o.backend = BACKEND_GPU;
            } else if (b == "cuda") {
                return set_err(error, "Error - backend token 'cuda' is unsupported; use 'gpu'.");
            } else {
                return set_err(error, "Error - invalid backend '" + b + "'. Use cpu or gpu.");
            }
            backend_set = from_named = true;
        } else if (tok == "--cpu-threads") {
            if (threads_set) {
                return set_err(error, "Error - cpu thread count provided multiple times.");
            }
            if (i + 1 >= argc) {
                return set_err(error, "Error - --cpu-threads requires a positive integer.");
            }
            const std::string v(argv[++i]);
            if (!parse_positive_int(v, &o.cpu_threads)) {
                return set_err(error, "Error - invalid --cpu-threads value '" + v +
                                          "' (expected positive integer).");
            }
            threads_set = true;
        } else if (tok == "--max-cand") {
            if (cand_set) {
                return set_err(error, "Error - GPU candidate batch cap provided multiple times.");
            }
            if (i + 1 >= argc) {
                return set_err(error, "Error - --max-cand requires a positive integer, "
                                      "optionally using K, M, or B suffix.");
            }
            const std::string v(argv[++i]);
            if (!parse_budget(v, &o.max_cand)) {
                return set_err(error, "Error - invalid --max-cand value '" + v +
                                          "' (expected positive integer with optional K, M, or "
                                          "B suffix).");
            }
            cand_set = true;
        } else if (tok == "--max-prod") {
            if (prod_set) {
                return set_err(error,
                               "Error - GPU coupled batch product cap provided multiple times.");
            }
            if (i + 1 >= argc) {
                return set_err(error, "Error - --max-prod requires a positive integer, "
                                      "optionally using K, M, or B suffix.");
            }
            const std::string v(argv[++i]);
            // "auto" (an extension): adaptive coupling batches, 32K products
            // doubling to 32M; a number caps every batch
            if (v == "auto") {
                o.max_prod = 0;
            } else if (!parse_budget(v, &o.max_prod)) {
                return set_err(error, "Error - invalid --max-prod value '" + v +
                                          "' (expected positive integer with optional K, M, or "
                                          "B suffix).");
            }
            prod_set = true;
        } else if (tok == "cpu" || tok == "gpu") {
            if (from_named) {
                return set_err(error, "Error - cannot mix shorthand backend token with --backend.");
            }
            if (backend_set) {
                return set_err(error, "Error - backend provided multiple times.");
            }
            o.backend = tok == "cpu" ? BACKEND_CPU : BACKEND_GPU;
            backend_set = from_short = true;
            if (i + 1 < argc) {
                const std::string nxt(argv[i + 1]);
                errno = 0;
                char *end = nullptr;
                const long raw = strtol(nxt.c_str(), &end, 10);
                const bool is_int = errno == 0 && end != nxt.c_str() && *end == '\0';
                if (is_int && tok == "gpu") {
                    return set_err(error,
                                   "Error - gpu shorthand does not accept a kernel value; use 'gpu'.");
                }
                if (is_int && tok == "cpu") {
                    if (threads_set) {
                        return set_err(error, "Error - cpu threads provided multiple times.");
                    }
                    if (raw <= 0 || raw > INT_MAX) {
                        return set_err(error, "Error - invalid cpu shorthand thread count '" + nxt +
                                                  "' (expected positive integer).");
                    }
                    o.cpu_threads = static_cast<int>(raw);
                    threads_set = true;
                    ++i;
                }
            }
        } else if (tok == "cuda") {
            return set_err(error, "Error - backend token 'cuda' is unsupported; use 'gpu'.");
        } else if (tok == "--save-frontier") {
            o.save_frontier = true;
        } else if (tok == "--frontier-out") {
            if (i + 1 >= argc) {
                return set_err(error, "Error - --frontier-out requires a file path.");
            }
            o.frontier_out_path = argv[++i];
            if (o.frontier_out_path.empty()) {
                return set_err(error, "Error - --frontier-out path cannot be empty.");
            }
            o.save_frontier = true;
        } else if (tok == "--save-stats") {
            o.save_stats = true;
        } else if (tok == "--stats-out") {
            if (i + 1 >= argc) {
                return set_err(error, "Error - --stats-out requires a file path.");
            }
            o.stats_out_path = argv[++i];
            if (o.stats_out_path.empty()) {
                return set_err(error, "Error - --stats-out path cannot be empty.");
            }
            o.save_stats = true;
        } else {
            if (parse_positive_int(tok, nullptr)) {
                return set_err(error, "Error - positional numeric argument '" + tok +
                                          "' must immediately follow shorthand backend token "
                                          "cpu|gpu.");
            }
            return set_err(error, "Error - unrecognized optional argument '" + tok + "'.");
        }
    }
    if (o.save_frontier && o.frontier_out_path.empty()) {
        o.frontier_out_path = stem_of(o.input_path, "frontier") + ".frontier.csv.gz";
    }
    if (o.save_stats && o.stats_out_path.empty()) {
        o.stats_out_path = stem_of(o.input_path, "stats") + ".stats.jsonl";
    }
    if (o.backend == BACKEND_GPU && threads_set) {
        return set_err(error, "Error - cpu thread options are not valid with backend=gpu.");
    }
    if (o.backend == BACKEND_CPU && cand_set) {
        return set_err(error, "Error - --max-cand is only valid with backend=gpu.");
    }
    if (o.backend == BACKEND_CPU && prod_set) {
        return set_err(error, "Error - --max-prod is only valid with backend=gpu.");
    }
    if (o.backend == BACKEND_CPU && !threads_set) {
        const char *env = getenv("OMP_NUM_THREADS");
        int t = 0;
        o.cpu_threads = (env != nullptr && parse_positive_int(env, &t)) ? t : 1;
    }
    if (o.backend == BACKEND_GPU && o.method != 1 && o.method != 2 && o.method != 3) {
        return set_err(error, "Error - GPU backend is unsupported for method " +
                                  std::to_string(o.method) + ".");
    }
    *out = o;
    return true;
}

// Rows are cut into blocks; each block is formatted and deflated (level 6,
// gzip's default) as a gzip member of its own on a host thread, and the
// members are written in order.  A gzip file of several members decompresses
// to the concatenation of their data (RFC 1952, 2.2), so the decompressed
// bytes are those of one stream over all rows.
bool write_frontier_gzip_csv(const std::vector<int32_t> &points, int d, const std::string &path,
                             std::string *error) {
    if (path.empty()) {
        return set_err(error, "output path is empty");
    }
    if (d <= 0 || points.size() % static_cast<size_t>(d) != 0) {
        return set_err(error, "frontier has invalid dimension");
    }
    constexpr size_t kBlockRows = 32768;
    const size_t rows = points.size() / static_cast<size_t>(d);
    const size_t blocks = std::max<size_t>(1, (rows + kBlockRows - 1) / kBlockRows);
    std::vector<std::string> member(blocks);
    std::vector<char> failed(blocks, 0);
    auto compress_block = [&](size_t b) {
        std::string text;
        text.reserve(kBlockRows * static_cast<size_t>(d) * 8);
        char num[16];
        const size_t r1 = std::min(rows, (b + 1) * kBlockRows);
        for (size_t r = b * kBlockRows; r < r1; ++r) {
            for (int o = 0; o < d; ++o) {
                if (o > 0) {
                    text += ',';
                }
                const auto res = std::to_chars(num, num + sizeof(num),
                                               points[r * static_cast<size_t>(d) + o]);
                text.append(num, static_cast<size_t>(res.ptr - num));
            }
            text += '\n';
        }
        z_stream zs{};
        if (deflateInit2(&zs, 6, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK) {
            failed[b] = 1;
            return;
        }
        std::string &out = member[b];
        out.resize(deflateBound(&zs, static_cast<uLong>(text.size())));
        zs.next_in = reinterpret_cast<Bytef *>(text.data());
        zs.avail_in = static_cast<uInt>(text.size());
        zs.next_out = reinterpret_cast<Bytef *>(out.data());
        zs.avail_out = static_cast<uInt>(out.size());
        const int rc = deflate(&zs, Z_FINISH);
        out.resize(zs.total_out);
        deflateEnd(&zs);
        if (rc != Z_STREAM_END) {
            failed[b] = 1;
        }
    };
    const size_t threads = std::min<size_t>(
        {blocks, 8, std::max<size_t>(1, std::thread::hardware_concurrency())});
    std::atomic<size_t> next{0};
    auto worker = [&]() {
        for (size_t b = next++; b < blocks; b = next++) {
            compress_block(b);
        }
    };
    if (threads <= 1) {
        worker();
    } else {
        std::vector<std::thread> pool;
        for (size_t t = 0; t < threads; ++t) {
            pool.emplace_back(worker);
        }
        for (std::thread &th : pool) {
            th.join();
        }
    }
    for (char f : failed) {
        if (f != 0) {
            return set_err(error, "failed while compressing the frontier");
        }
    }
    std::FILE *f = std::fopen(path.c_str(), "wb");
    if (f == nullptr) {
        return set_err(error, "could not open gzip output");
    }
    bool ok = true;
    for (const std::string &m : member) {
        ok = ok && std::fwrite(m.data(), 1, m.size(), f) == m.size();
    }
    if (std::fclose(f) != 0 || !ok) {
        return set_err(error, "failed while writing compressed frontier");
    }
    return true;
}

bool write_stats_jsonl(const std::string &path, const Options &opts, const RunStats &st,
                       const Structure &rec, std::string *error) {
    std::ofstream out(path.c_str(), std::ios::out | std::ios::app);
    if (!out.is_open()) {
        return set_err(error, "could not open output path");
    }
    out << std::fixed << std::setprecision(6);
    out << "{";
    out << "\"schema_version\":1,";
    out << "\"identity\":{";
    out << "\"input_path\":\"" << json_escape(opts.input_path) << "\",";
    out << "\"problem_type\":" << opts.problem_type << ",";
    out << "\"method\":" << opts.method << ",";
    out << "\"state_dominance\":" << opts.state_dominance << ",";
    out << "\"backend\":\"" << backend_to_string(opts.backend) << "\",";
    out << "\"cpu_threads\":" << opts.cpu_threads;
    out << "},";
    out << "\"outputs\":{";
    out << "\"num_solutions\":" << st.num_solutions << ",";
    out << "\"save_frontier\":" << (opts.save_frontier ? "true" : "false") << ",";
    out << "\"frontier_out_path\":\"" << json_escape(opts.frontier_out_path) << "\"";
    out << "},";
    out << "\"timing\":{";
    out << "\"cpu\":{";
    out << "\"cpu_compile_s\":" << st.cpu_compile_s << ",";
    out << "\"cpu_enumeration_s\":" << st.cpu_enumeration_s << ",";
    out << "\"cpu_total_s\":" << st.cpu_total_s << ",";
    out << "\"cpu_state_dominance_s\":" << st.cpu_state_dominance_s;
    out << "},";
    out << "\"wall\":{";
    out << "\"wall_compile_s\":" << st.wall_compile_s << ",";
    out << "\"wall_enumeration_s\":" << st.wall_enumeration_s;
    out << "},";
    out << "\"kernel\":{";
    out << "\"kernel_expand_td_s\":" << st.kernel_expand_td_s << ",";
    out << "\"kernel_dominance_s\":" << st.kernel_dominance_s << ",";
    out << "\"kernel_total_s\":" << st.kernel_total_s;
    out << "}";
    out << "},";
    out << "\"memory\":{";
    out << "\"cpu\":{\"cpu_mem_peak_bytes\":" << st.cpu_mem_peak_bytes << "},";
    out << "\"gpu\":{";
    out << "\"gpu_mem_peak_used_bytes\":" << st.gpu_mem_peak_used_bytes << ",";
    out << "\"gpu_mem_peak_reserved_bytes\":" << st.gpu_mem_peak_reserved_bytes;
    out << "}";
    out << "},";
    out << "\"work\":{";
    out << "\"work_candidates_total\":" << st.work_candidates_total << ",";
    out << "\"work_candidates_peak\":" << st.work_candidates_peak << ",";
    out << "\"work_frontier_survivors_total\":" << st.work_frontier_survivors_total << ",";
    out << "\"work_frontier_peak_points\":" << st.work_frontier_peak_points << ",";
    out << "\"work_join_products_total\":" << st.work_join_products_total;
    out << "},";
    out << "\"dominance\":{\"dominance_filtered_total\":" << st.dominance_filtered_total << "},";
    out << "\"structure\":{";
    out << "\"is_tsp_branch\":" << (opts.problem_type == 3 ? "true" : "false") << ",";
    out << "\"postprocess_sort_applied\":true,";
    out << "\"original_width\":" << rec.original_width << ",";
    out << "\"reduced_width\":" << rec.reduced_width << ",";
    out << "\"original_num_nodes\":" << rec.original_num_nodes << ",";
    out << "\"reduced_num_nodes\":" << rec.reduced_num_nodes << ",";
    out << "\"max_num_nodes_per_layer\":";
    json_array(out, rec.max_num_nodes_per_layer);
    out << ",";
    out << "\"max_num_in_arcs\":" << rec.max_num_in_arcs << ",";
    out << "\"total_num_in_arcs\":" << rec.total_num_in_arcs << ",";
    out << "\"layer_coupling\":" << st.layer_coupling;
    out << "},";
    out << "\"metrics\":{";
    out << "\"wall_state_dominance_s\":" << st.wall_state_dominance_s << ",";
    out << "\"wall_cutset_sort_s\":" << st.wall_cutset_sort_s << ",";
    out << "\"wall_cutset_convolution_s\":" << st.wall_cutset_convolution_s << ",";
    out << "\"wall_cutset_partial_merge_s\":" << st.wall_cutset_partial_merge_s << ",";
    out << "\"wall_pack_transfer_s\":" << st.wall_pack_transfer_s << ",";
    out << "\"wall_join_s\":" << st.wall_join_s << ",";
    out << "\"cpu_layers_td\":" << st.cpu_layers_td << ",";
    out << "\"cpu_layers_bu\":" << st.cpu_layers_bu << ",";
    out << "\"cpu_nodes_expanded\":" << st.cpu_nodes_expanded << ",";
    out << "\"cpu_cutset_size\":" << st.cpu_cutset_size << ",";
    out << "\"std_candidates_per_layer\":";
    json_array(out, st.std_candidates_per_layer);
    out << ",";
    out << "\"std_frontier_survivors_per_layer\":";
    json_array(out, st.std_frontier_survivors_per_layer);
    out << "},";
    out << "\"status\":{";
    out << "\"status_state\":\"" << json_escape(rec.status_state) << "\",";
    out << "\"status_error_message\":\"" << json_escape(rec.status_error_message) << "\"";
    out << "}";
    out << "}\n";
    if (!out.good()) {
        return set_err(error, "failed while writing JSONL output");
    }
    return true;
}

void print_and_save_run_summary(const Options &opts, const RunStats &st, const Structure &rec) {
    std::cout << st.num_solutions << std::endl;
    std::cout << st.cpu_total_s << std::endl;
    if (opts.problem_type == 3) {
        std::cout << st.cpu_compile_s << "\t" << st.cpu_enumeration_s << "\t" << st.wall_compile_s
                  << "\t" << st.wall_enumeration_s << std::endl;
    } else {
        std::cout << opts.method << "\t" << opts.state_dominance << "\t" << rec.original_width
                  << "\t" << rec.reduced_width << "\t" << rec.original_num_nodes << "\t"
                  << rec.reduced_num_nodes << "\t" << st.cpu_compile_s << "\t"
                  << st.cpu_enumeration_s << "\t" << st.layer_coupling << "\t"
                  << st.dominance_filtered_total << "\t" << st.cpu_state_dominance_s << "\t"
                  << st.wall_compile_s << "\t" << st.wall_enumeration_s << std::endl;
    }
    if (opts.save_stats) {
        std::string err;
        if (!write_stats_jsonl(opts.stats_out_path, opts, st, rec, &err)) {
            std::cerr << "Error - failed to save stats to '" << opts.stats_out_path << "'";
            if (!err.empty()) {
                std::cerr << ": " << err;
            }
            std::cerr << '\n';
            exit(1);
        }
    }
}

} // namespace pmdd_cli
