// capi_util.hpp -- exception firewall for the C ABI: no C++ exception
// crosses lsdtt3_gpu.h; each entry point maps them to a status and records
// the message for lsdtt_gpu_last_error().
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "../../include/lsdtt3_gpu.h"
#include "common.hpp"

namespace lsdtt_gpu {

inline std::string& last_error_ref()
{
    static thread_local std::string e;
    return e;
}
inline const char* last_error_cstr() { return last_error_ref().c_str(); }
inline void set_error(const std::string& m) { last_error_ref() = m; }
inline int fail_arg(const std::string& m)
{
    set_error(m);
    return LSDTT_GPU_ERR_ARG;
}

template <class F>
int guarded(F&& fn)
{
    try {
        last_error_ref().clear();
        return fn();
    } catch (const std::invalid_argument& e) {
        set_error(e.what());
        return LSDTT_GPU_ERR_ARG;
    } catch (const std::bad_alloc&) {
        set_error("out of memory");
        return LSDTT_GPU_ERR_RUNTIME;
    } catch (const std::exception& e) {
        set_error(e.what());
        return LSDTT_GPU_ERR_RUNTIME;
    }
}

template <class T, class F>
T* guarded_ptr(F&& fn)
{
    try {
        last_error_ref().clear();
        return fn();
    } catch (const std::exception& e) {
        set_error(e.what());
        return nullptr;
    }
}

}  // namespace lsdtt_gpu
