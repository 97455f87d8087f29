// Shared host-side helpers of libvortex_hip: error recording, the stream that
// reference-ABI launches go to, and launch bookkeeping.
//
// Error model (mirrors the reference wrappers, which return void and surface
// failures through the runtime's last-error query, ffi.rs:6-21): every entry
// point records the first failure in a thread-local slot read by
// vx_get_last_error()/vx_get_last_error_message(); nothing aborts the process.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/vortex_hip.h"

namespace vx {

void set_error(int code, const char* what);
int peek_error();  // current thread's recorded error, not cleared
// Moves the calling thread's recorded error out (0 when none) and clears it.
int take_error(std::string& msg);
// Records code and message as given, unless an error is already recorded.
void put_error(int code, const std::string& msg);
hipStream_t current_stream();

// Records launch/runtime failures after a kernel launch or runtime call.
inline bool check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        set_error((int)e, what);
        return false;
    }
    return true;
}
inline bool check_launch(const char* what) { return check(hipGetLastError(), what); }

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
inline uint32_t ilog2(uint64_t n) {
    uint32_t l = 0;
    while ((1ull << l) < n) l++;
    return l;
}
inline bool is_pow2(uint64_t n) { return n && !(n & (n - 1)); }

// Optional launch timeline of a context object: HIP events recorded on its
// stream after labelled launches; entry i > 0 spans marks i-1 .. i.
struct Timeline {
    bool on = false;
    std::vector<hipEvent_t> ev;
    std::vector<std::string> labels;
    size_t n = 0;
    void reset(bool enable) {
        on = enable;
        n = 0;
    }
    void mark(const std::string& label, hipStream_t s) {
        if (!on) return;
        if (n == ev.size()) {
            hipEvent_t e;
            if (!check(hipEventCreate(&e), "hipEventCreate")) return;
            ev.push_back(e);
            labels.emplace_back();
        }
        labels[n] = label;
        check(hipEventRecord(ev[n], s), "hipEventRecord");
        n++;
    }
    const char* label(uint32_t i) const { return i < n ? labels[i].c_str() : ""; }
    float ms(uint32_t i) {
        if (i == 0 || i >= n) return 0.f;
        float t = 0.f;
        if (!check(hipEventSynchronize(ev[i]), "hipEventSynchronize")) return -1.f;
        if (!check(hipEventElapsedTime(&t, ev[i - 1], ev[i]), "hipEventElapsedTime")) return -1.f;
        return t;
    }
    ~Timeline() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

}  // namespace vx
