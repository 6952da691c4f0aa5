// Instance parsers for the three reference .dat formats.
//
//   knapsack     src/pmdd/instances/knapsack_instance.cpp:17-47
//                "n_vars n_cons n_objs", then n_objs rows of n_vars profits,
//                then n_cons rows of n_vars coefficients followed by the rhs.
//   set packing  src/pmdd/instances/setpacking_instance.hpp:49-89
//                "n_vars n_cons", "n_objs", n_objs rows of n_vars coefficients,
//                then per constraint "k v_1 .. v_k" with 1-based variables.
//   TSP          src/pmdd/instances/tsp_instance.cpp:11-41
//                "n_objs n_cities", then n_objs n_cities x n_cities matrices.
//
// Unlike the reference (which trusts the stream), a short or non-numeric file
// is an error here.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pmdd {

struct KnapsackInstance {
    int n_vars = 0;
    int n_cons = 0;
    int n_objs = 0;
    std::vector<std::vector<int32_t>> obj_coeffs; // [var][obj]
    std::vector<std::vector<int32_t>> coeffs;     // [cons][var]
    std::vector<int32_t> rhs;                     // [cons]
};

struct SetPackingInstance {
    int n_vars = 0;
    int n_cons = 0;
    int n_objs = 0;
    std::vector<std::vector<int32_t>> objs;      // [obj][var]
    std::vector<std::vector<int32_t>> vars_cons; // [cons] 0-based variable ids
};

struct TspInstance {
    int n_objs = 0;
    int n_cities = 0;
    std::vector<int32_t> dist; // [obj][from][to], row-major
    int32_t at(int o, int i, int j) const {
        return dist[(static_cast<size_t>(o) * n_cities + i) * n_cities + j];
    }
};

// Each reader returns true on success; on failure *error holds the message.
bool read_knapsack(const std::string &path, KnapsackInstance *out, std::string *error);
bool read_setpacking(const std::string &path, SetPackingInstance *out, std::string *error);
bool read_tsp(const std::string &path, TspInstance *out, std::string *error);

} // namespace pmdd
