// Runtime part of the C ABI: errors, streams, memory (replaces the CUDA runtime
// calls bound at src/cuda/ffi.rs:6-76 and used by src/device/buffer.rs) and the
// elementwise field kernels of cuda/fields.cu.
#include <cstring>
#include <mutex>
#include <string>

#include "common.hpp"
#include "field.hpp"

namespace vx {

namespace {
thread_local int t_err = 0;
thread_local std::string t_msg;
thread_local hipStream_t t_stream = nullptr;
}  // namespace

void set_error(int code, const char* what) {
    if (t_err == 0) {
        t_err = code ? code : -1;
        t_msg = what ? what : "";
        if (code > 0) {
            t_msg += ": ";
            t_msg += hipGetErrorString((hipError_t)code);
        }
    }
}
hipStream_t current_stream() { return t_stream; }
int take_error(std::string& msg) {
    const int e = t_err;
    msg = e ? t_msg : std::string();
    t_err = 0;
    return e;
}
void put_error(int code, const std::string& msg) {
    if (t_err == 0 && code) {
        t_err = code;
        t_msg = msg;
    }
}
int peek_error() { return t_err; }

__global__ void m31_add_kernel(const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = madd(a[i], b[i]);
}
__global__ void m31_mul_kernel(const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = mmul(a[i], b[i]);
}

}  // namespace vx

using namespace vx;

extern "C" {

int32_t vx_get_last_error(void) {
    int e = t_err;
    t_err = 0;
    return e;
}
const char* vx_get_last_error_message(void) { return t_msg.c_str(); }

int32_t vx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
int32_t vx_set_device(int32_t d) { return check(hipSetDevice(d), "hipSetDevice") ? 0 : -1; }
int32_t vx_malloc(void** p, size_t bytes) { return check(hipMalloc(p, bytes), "hipMalloc") ? 0 : -1; }
int32_t vx_free(void* p) { return check(hipFree(p), "hipFree") ? 0 : -1; }
int32_t vx_malloc_host(void** p, size_t bytes) { return check(hipHostMalloc(p, bytes, 0), "hipHostMalloc") ? 0 : -1; }
int32_t vx_free_host(void* p) { return check(hipHostFree(p), "hipHostFree") ? 0 : -1; }
int32_t vx_memcpy_h2d(void* d, const void* s, size_t b) {
    return check(hipMemcpy(d, s, b, hipMemcpyHostToDevice), "hipMemcpy H2D") ? 0 : -1;
}
int32_t vx_memcpy_d2h(void* d, const void* s, size_t b) {
    return check(hipMemcpy(d, s, b, hipMemcpyDeviceToHost), "hipMemcpy D2H") ? 0 : -1;
}
int32_t vx_memcpy_d2d(void* d, const void* s, size_t b) {
    return check(hipMemcpy(d, s, b, hipMemcpyDeviceToDevice), "hipMemcpy D2D") ? 0 : -1;
}
int32_t vx_memset(void* p, int32_t v, size_t b) { return check(hipMemset(p, v, b), "hipMemset") ? 0 : -1; }
int32_t vx_device_synchronize(void) { return check(hipDeviceSynchronize(), "hipDeviceSynchronize") ? 0 : -1; }
int32_t vx_mem_get_info(size_t* f, size_t* t) { return check(hipMemGetInfo(f, t), "hipMemGetInfo") ? 0 : -1; }
int32_t vx_stream_create(void** s) {
    hipStream_t st = nullptr;
    if (!check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate")) return -1;
    *s = (void*)st;
    return 0;
}
int32_t vx_stream_synchronize(void* s) {
    return check(hipStreamSynchronize((hipStream_t)s), "hipStreamSynchronize") ? 0 : -1;
}
int32_t vx_stream_destroy(void* s) { return check(hipStreamDestroy((hipStream_t)s), "hipStreamDestroy") ? 0 : -1; }
int32_t vx_event_create(void** e) {
    hipEvent_t ev = nullptr;
    if (!e || !check(hipEventCreate(&ev), "hipEventCreate")) return -1;
    *e = (void*)ev;
    return 0;
}
int32_t vx_event_record(void* e, void* s) {
    return check(hipEventRecord((hipEvent_t)e, (hipStream_t)s), "hipEventRecord") ? 0 : -1;
}
int32_t vx_event_elapsed_ms(void* start, void* end, float* ms) {
    if (!ms || !check(hipEventSynchronize((hipEvent_t)end), "hipEventSynchronize")) return -1;
    return check(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end), "hipEventElapsedTime") ? 0 : -1;
}
int32_t vx_event_destroy(void* e) { return check(hipEventDestroy((hipEvent_t)e), "hipEventDestroy") ? 0 : -1; }
void vx_set_stream(void* s) { t_stream = (hipStream_t)s; }
void* vx_get_stream(void) { return (void*)t_stream; }

void cuda_m31_add(const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n) {
    if (!n) return;
    m31_add_kernel<<<ceil_div(n, 256), 256, 0, current_stream()>>>(a, b, out, n);
    check_launch("cuda_m31_add");
}
void cuda_m31_mul(const uint32_t* a, const uint32_t* b, uint32_t* out, uint32_t n) {
    if (!n) return;
    m31_mul_kernel<<<ceil_div(n, 256), 256, 0, current_stream()>>>(a, b, out, n);
    check_launch("cuda_m31_mul");
}
void cuda_device_sync(void) { check(hipDeviceSynchronize(), "cuda_device_sync"); }

}  // extern "C"
