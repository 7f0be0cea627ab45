// internal.hpp -- error plumbing shared by the host translation units of libmoris_gpu.
// No exception crosses the C ABI: entry points catch, store the message per thread and return a
// status code (replacing MORIS_ERROR's throw, projects/MRS/ASR/src/fn_log_assert.hpp:41,92).
#pragma once
#include <exception>
#include <memory>
#include <new>
#include <string>

#include "../../../include/moris_gpu.h"

namespace mgpu {
mgpu_status set_error(mgpu_status code, const std::string& msg);
}

#define MGPU_GUARD_BEGIN try {
#define MGPU_GUARD_END                                                           \
    }                                                                            \
    catch (const std::bad_alloc&) {                                              \
        return mgpu::set_error(MGPU_ERR_INVALID_ARG, "host allocation failed");  \
    }                                                                            \
    catch (const std::exception& ex) {                                           \
        return mgpu::set_error(MGPU_ERR_INVALID_ARG, ex.what());                 \
    }                                                                            \
    catch (...) {                                                                \
        return mgpu::set_error(MGPU_ERR_INVALID_ARG, "unknown host exception");  \
    }
