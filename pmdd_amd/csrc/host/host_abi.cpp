// C ABI of the host surface (declared in include/pmdd_host.h).
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../../include/pmdd_host.h"
#include "dd.hpp"

struct pmdd_dd {
    pmdd::FlatDD flat;
    pmdd_dd_view view;
};

static void write_err(char *err, size_t errlen, const std::string &msg) {
    if (err != nullptr && errlen > 0) {
        std::snprintf(err, errlen, "%s", msg.c_str());
    }
}

extern "C" int pmdd_dd_build_from_file(const char *path, int problem_type, int expect_d,
                                       int flags, pmdd_dd **out, char *err, size_t errlen) {
    if (out == nullptr || path == nullptr) {
        write_err(err, errlen, "pmdd_dd_build_from_file: NULL argument");
        return PMDD_E_INVALID;
    }
    *out = nullptr;
    try {
        pmdd_dd *dd = new pmdd_dd();
        std::string error;
        if (!pmdd::build_from_file(path, problem_type, expect_d,
                                   (flags & PMDD_BUILD_STATE_DOMINANCE) != 0, &dd->flat, &error)) {
            delete dd;
            write_err(err, errlen, error);
            return PMDD_E_INVALID;
        }
        dd->view = dd->flat.view();
        *out = dd;
        write_err(err, errlen, "");
        return PMDD_OK;
    } catch (const std::bad_alloc &) {
        write_err(err, errlen, "host allocation failed while building the decision diagram");
        return PMDD_E_NOMEM;
    } catch (const std::exception &e) {
        write_err(err, errlen, e.what());
        return PMDD_E_INVALID;
    }
}

extern "C" const pmdd_dd_view *pmdd_dd_get_view(const pmdd_dd *dd) {
    return dd != nullptr ? &dd->view : nullptr;
}

extern "C" void pmdd_dd_get_structure(const pmdd_dd *dd, pmdd_dd_structure *out) {
    if (dd != nullptr && out != nullptr) {
        *out = dd->flat.structure;
    }
}

extern "C" int64_t pmdd_dd_layer_size(const pmdd_dd *dd, int layer) {
    if (dd == nullptr || layer < 0 || layer >= dd->flat.num_layers) {
        return -1;
    }
    return dd->flat.layer_size(layer);
}

extern "C" void pmdd_dd_free(pmdd_dd *dd) { delete dd; }

namespace {

// Rows of D ints sorted as fixed-size arrays (lexicographic operator<).
template <int D> void sort_rows(int32_t *p, int64_t n) {
    std::vector<std::array<int32_t, D>> rows(static_cast<size_t>(n));
    std::memcpy(rows.data(), p, sizeof(int32_t) * D * static_cast<size_t>(n));
    std::sort(rows.begin(), rows.end());
    std::memcpy(p, rows.data(), sizeof(int32_t) * D * static_cast<size_t>(n));
}

} // namespace

// ParetoFrontier::sort_lexicographic_ascending (src/pmdd/enum/pareto_frontier.hpp:418-433),
// the order main() writes (src/pmdd/main.cpp:403).  The points change only
// once the sorted copy is complete.
static void sort_frontier(pmdd_frontier *fr) {
    if (fr == nullptr || fr->points == nullptr || fr->count < 2) {
        return;
    }
    const int64_t n = fr->count;
    switch (fr->d) {
    case 1: sort_rows<1>(fr->points, n); return;
    case 2: sort_rows<2>(fr->points, n); return;
    case 3: sort_rows<3>(fr->points, n); return;
    case 4: sort_rows<4>(fr->points, n); return;
    case 5: sort_rows<5>(fr->points, n); return;
    case 6: sort_rows<6>(fr->points, n); return;
    case 7: sort_rows<7>(fr->points, n); return;
    case 8: sort_rows<8>(fr->points, n); return;
    default: break;
    }
    const int d = fr->d;
    if (d <= 0) {
        return;
    }
    std::vector<int64_t> order(static_cast<size_t>(n));
    std::iota(order.begin(), order.end(), 0);
    const int32_t *pts = fr->points;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return std::lexicographical_compare(pts + a * d, pts + a * d + d, pts + b * d,
                                            pts + b * d + d);
    });
    std::vector<int32_t> sorted(static_cast<size_t>(n) * d);
    for (int64_t i = 0; i < n; ++i) {
        std::copy(pts + order[i] * d, pts + order[i] * d + d, sorted.begin() + i * d);
    }
    std::copy(sorted.begin(), sorted.end(), fr->points);
}

extern "C" int pmdd_frontier_sort(pmdd_frontier *fr) {
    try {
        sort_frontier(fr);
        return PMDD_OK;
    } catch (const std::bad_alloc &) {
        return PMDD_E_NOMEM;
    }
}
