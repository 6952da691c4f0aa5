// Instance parsers (formats documented in instances.hpp).
#include "instances.hpp"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace pmdd {
namespace {

// Whitespace-separated integer reader over a whole file.
class IntReader {
  public:
    bool open(const std::string &path) {
        std::ifstream in(path.c_str(), std::ios::in | std::ios::binary);
        if (!in.is_open()) {
            return false;
        }
        std::ostringstream ss;
        ss << in.rdbuf();
        buf_ = ss.str();
        pos_ = 0;
        return true;
    }

    bool next(long long *v) {
        while (pos_ < buf_.size() && isspace(static_cast<unsigned char>(buf_[pos_]))) {
            ++pos_;
        }
        if (pos_ >= buf_.size()) {
            return false;
        }
        const char *begin = buf_.c_str() + pos_;
        char *end = nullptr;
        errno = 0;
        long long x = strtoll(begin, &end, 10);
        if (end == begin || errno != 0) {
            return false;
        }
        pos_ += static_cast<size_t>(end - begin);
        *v = x;
        return true;
    }

    bool next_i32(int32_t *v) {
        long long x = 0;
        if (!next(&x) || x < INT32_MIN || x > INT32_MAX) {
            return false;
        }
        *v = static_cast<int32_t>(x);
        return true;
    }

  private:
    std::string buf_;
    size_t pos_ = 0;
};

bool fail(std::string *error, const std::string &msg) {
    if (error != nullptr) {
        *error = msg;
    }
    return false;
}

} // namespace

bool read_knapsack(const std::string &path, KnapsackInstance *out, std::string *error) {
    IntReader r;
    if (!r.open(path)) {
        return fail(error, "Error - file " + path + " could not be found");
    }
    KnapsackInstance inst;
    int32_t nv = 0, nc = 0, no = 0;
    if (!r.next_i32(&nv) || !r.next_i32(&nc) || !r.next_i32(&no)) {
        return fail(error, "Error - malformed knapsack header in " + path);
    }
    if (nv <= 0 || nc <= 0 || no <= 0) {
        return fail(error, "Error - knapsack header must hold positive sizes in " + path);
    }
    inst.n_vars = nv;
    inst.n_cons = nc;
    inst.n_objs = no;
    inst.obj_coeffs.assign(nv, std::vector<int32_t>(no, 0));
    inst.coeffs.assign(nc, std::vector<int32_t>(nv, 0));
    inst.rhs.assign(nc, 0);
    for (int o = 0; o < no; ++o) {
        for (int i = 0; i < nv; ++i) {
            if (!r.next_i32(&inst.obj_coeffs[i][o])) {
                return fail(error, "Error - truncated knapsack objectives in " + path);
            }
        }
    }
    for (int c = 0; c < nc; ++c) {
        for (int i = 0; i < nv; ++i) {
            if (!r.next_i32(&inst.coeffs[c][i])) {
                return fail(error, "Error - truncated knapsack constraints in " + path);
            }
        }
        if (!r.next_i32(&inst.rhs[c])) {
            return fail(error, "Error - truncated knapsack right-hand side in " + path);
        }
    }
    *out = std::move(inst);
    return true;
}

bool read_setpacking(const std::string &path, SetPackingInstance *out, std::string *error) {
    IntReader r;
    if (!r.open(path)) {
        return fail(error, "Error: could not open file " + path);
    }
    SetPackingInstance inst;
    int32_t nv = 0, nc = 0, no = 0;
    if (!r.next_i32(&nv) || !r.next_i32(&nc) || !r.next_i32(&no)) {
        return fail(error, "Error - malformed set packing header in " + path);
    }
    if (nv <= 0 || nc < 0 || no <= 0) {
        return fail(error, "Error - set packing header must hold positive sizes in " + path);
    }
    inst.n_vars = nv;
    inst.n_cons = nc;
    inst.n_objs = no;
    inst.objs.assign(no, std::vector<int32_t>(nv, 0));
    for (int o = 0; o < no; ++o) {
        for (int i = 0; i < nv; ++i) {
            if (!r.next_i32(&inst.objs[o][i])) {
                return fail(error, "Error - truncated set packing objectives in " + path);
            }
        }
    }
    inst.vars_cons.resize(nc);
    for (int c = 0; c < nc; ++c) {
        int32_t k = 0;
        if (!r.next_i32(&k) || k < 0) {
            return fail(error, "Error - truncated set packing constraint in " + path);
        }
        inst.vars_cons[c].resize(k);
        for (int i = 0; i < k; ++i) {
            int32_t v = 0;
            if (!r.next_i32(&v) || v < 1 || v > nv) {
                return fail(error, "Error - set packing variable out of range in " + path);
            }
            inst.vars_cons[c][i] = v - 1;
        }
    }
    *out = std::move(inst);
    return true;
}

bool read_tsp(const std::string &path, TspInstance *out, std::string *error) {
    IntReader r;
    if (!r.open(path)) {
        return fail(error, "Error: could not open file " + path);
    }
    TspInstance inst;
    int32_t no = 0, n = 0;
    if (!r.next_i32(&no) || !r.next_i32(&n)) {
        return fail(error, "Error - malformed TSP header in " + path);
    }
    if (no <= 0 || n <= 0 || n > 62) {
        return fail(error, "Error - TSP header out of range (need 1 <= n_cities <= 62) in " + path);
    }
    inst.n_objs = no;
    inst.n_cities = n;
    inst.dist.assign(static_cast<size_t>(no) * n * n, 0);
    for (size_t k = 0; k < inst.dist.size(); ++k) {
        if (!r.next_i32(&inst.dist[k])) {
            return fail(error, "Error - truncated TSP distance matrices in " + path);
        }
    }
    *out = std::move(inst);
    return true;
}

} // namespace pmdd
