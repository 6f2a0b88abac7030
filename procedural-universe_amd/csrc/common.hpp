// Internal helpers shared by the library's translation units: error
// reporting across the C ABI and the HIP error check.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/nbody.h"

namespace nbody {

// Thread-local message of the last failed call (nbody_last_error).
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
const char* last_error();
void clear_error();

// Evaluate a HIP call; on failure record where and why and return
// NBODY_ERR_DEVICE (or NBODY_ERR_OUT_OF_MEMORY) from the enclosing function.
#define NB_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t nb_e_ = (expr);                                                     \
        if (nb_e_ != hipSuccess) {                                                     \
            ::nbody::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(nb_e_), \
                               __FILE__, __LINE__);                                    \
            return nb_e_ == hipErrorOutOfMemory ? NBODY_ERR_OUT_OF_MEMORY              \
                                                : NBODY_ERR_DEVICE;                    \
        }                                                                              \
    } while (0)

#define NB_CHECK(cond, status, ...)            \
    do {                                       \
        if (!(cond)) {                         \
            ::nbody::set_error(__VA_ARGS__);   \
            return (status);                   \
        }                                      \
    } while (0)

#define NB_TRY(expr)                           \
    do {                                       \
        nbody_status nb_s_ = (expr);           \
        if (nb_s_ != NBODY_OK) return nb_s_;   \
    } while (0)

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

}  // namespace nbody
