// capi_lem.hpp -- conversions between the C ABI's LEM structs and the
// engine's (shared by the single-model and ensemble entry points).
#pragma once

#include <memory>
#include <vector>

#include "../../include/lsdtt3_gpu.h"
#include "lem.hpp"

struct lsdtt_gpu_lem {
    std::unique_ptr<lsdtt_gpu::LemEngine> eng;
    std::vector<int64_t> bf_cache;
    bool bf_set = false;
    bool z_set = false;
};

namespace lsdtt_gpu {

bool bc_ok(const char* s);
// throws std::invalid_argument on options the reference rejects
void validate_options(const lsdtt_gpu_lem_options& o);
LemParams to_params(const lsdtt_gpu_lem_options& o, double cell_size);
void to_c_stats(const LemStats& st, lsdtt_gpu_lem_stats* out);

}  // namespace lsdtt_gpu
