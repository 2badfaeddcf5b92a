// util.cpp -- messages, fatal errors and process-wide configuration.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>

#include "common.h"

namespace ssa {

Config& config() {
    static Config c;
    return c;
}

// src/util/util.c:36-45: message to stderr, then exit(1)
void fatal(const char* fmt, ...) {
    if (fmt) {
        va_list ap;
        va_start(ap, fmt);
        vfprintf(stderr, fmt, ap);
        va_end(ap);
        fprintf(stderr, "\n");
    }
    exit(1);
}

static bool shown(int level) { return level <= config().output_mode; }

// src/util/util.c:54-64 (no newline added, as in the reference)
void print_info(const char* fmt, ...) {
    if (!shown(OUTPUT_INFO)) return;
    printf("libssa INFO: ");
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
}

// src/util/util.c:66-77
void print_warning(const char* fmt, ...) {
    if (!shown(OUTPUT_WARNING)) return;
    printf("libssa WARNING: ");
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    printf("\n");
}

// src/util/util.c:79-90
void print_error(const char* fmt, ...) {
    if (!shown(OUTPUT_ERROR)) return;
    printf("libssa ERROR: ");
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    printf("\n");
}

// The reference's 8-bit kernels exist for AVX2 (32 lanes) and SSE4.1 (16
// lanes), its 16-bit kernels for AVX2 (16) and SSE2 (8); the selection follows
// set_max_compute_capability (src/cpu_config.c:29-33, src/algo/8/search_8.c:36-60).
int lanes8() { return config().simd_mode >= COMPUTE_ON_AVX2 ? 32 : 16; }
int lanes16() { return config().simd_mode >= COMPUTE_ON_AVX2 ? 16 : 8; }

}  // namespace ssa
