// lekf_ktime.hpp -- per-kernel timing for measurement.  Off unless switched
// on by lekf_ktime_enable; then a KTime scope around a launch records a HIP
// event pair on the launch's stream, tagged with the kernel's name, and
// lekf_ktime_read sums the elapsed times of one name.
#pragma once
#include <hip/hip_runtime.h>

void lekf_ktime_mark(const char* name, hipStream_t st, bool begin);

struct KTime {
    const char* name;
    hipStream_t st;
    KTime(const char* n, hipStream_t s) : name(n), st(s) { lekf_ktime_mark(name, st, true); }
    ~KTime() { lekf_ktime_mark(name, st, false); }
    KTime(const KTime&) = delete;
    KTime& operator=(const KTime&) = delete;
};
