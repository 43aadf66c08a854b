// Internals shared by the files that implement the extern "C" boundary: the opaque operator handle,
// the last-error message and the exception-to-status-code guard.
#pragma once

#include "../../include/tandem_dg.h"

#include "host/operator.h"

#include <memory>
#include <stdexcept>
#include <string>

struct tdg_op {
    std::unique_ptr<tdg::Operator> impl;
};

namespace tdg_abi {
inline thread_local std::string g_error;

inline int fail(int code, std::string const& msg) {
    g_error = msg;
    return code;
}

template <class F> int guarded(F&& f) {
    try {
        f();
        return TDG_OK;
    } catch (std::invalid_argument const& e) {
        return fail(TDG_ERR_ARG, e.what());
    } catch (std::logic_error const& e) {
        return fail(TDG_ERR_STATE, e.what());
    } catch (std::exception const& e) {
        std::string m = e.what();
        return fail(m.rfind("HIP error", 0) == 0 ? TDG_ERR_HIP : TDG_ERR_INTERNAL, m);
    } catch (...) {
        return fail(TDG_ERR_INTERNAL, "unknown exception");
    }
}

inline void require(bool cond, char const* msg) {
    if (!cond) throw std::invalid_argument(msg);
}
// Looks up the device address of the allocation-free apply's scratch array on the current device.
void prepare_noalloc_scratch();

inline tdg::Operator& ready(tdg_op* op) {
    require(op && op->impl, "null operator");
    if (!op->impl->ready()) throw std::logic_error("tdg_set_material() has not been called");
    return *op->impl;
}
} // namespace tdg_abi
