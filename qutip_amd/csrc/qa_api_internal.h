// Internal host-side helpers shared by the translation units of the C ABI:
// status codes, the per-thread error message and the HIP error check macro.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>

#include "../../include/qutip_amd.h"

namespace qa {
void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);
}  // namespace qa

#define QA_HIP(expr)                                                          \
    do {                                                                      \
        hipError_t _e = (expr);                                               \
        if (_e != hipSuccess) return qa::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

#define QA_CHECK_LAUNCH() QA_HIP(hipGetLastError())

// propagate a non-zero status code of a library call
#define QA_CHECK(expr)                                                        \
    do {                                                                      \
        const int _rc = (expr);                                               \
        if (_rc != QA_OK) return _rc;                                         \
    } while (0)

#define QA_REQUIRE(cond, ...)                                                 \
    do {                                                                      \
        if (!(cond)) {                                                        \
            qa::set_error(__VA_ARGS__);                                       \
            return QA_ERR_ARG;                                                \
        }                                                                     \
    } while (0)
