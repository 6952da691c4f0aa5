// C ABI of the host surface (declared in include/pmdd_host.h).
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

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
