// Runtime part of the C-ABI: error state, device selection, memory moves and
// a graph-capture-safe zero fill.
#include <cstring>
#include <string>

#include "rsg_common.h"

namespace rsg {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int check_hip(hipError_t e, const char *what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return RSG_EHIP;
}

Tuning &tuning() {
    static Tuning t;
    return t;
}

Coefs scaled_coefs(float d) {
    static const float table[36] = {RSG_DER_TABLE};
    Coefs c;
    const float inv = 1.0f / d;
    for (int i = 0; i < 36; i++) c.c[i] = table[i] * inv;
    return c;
}

__global__ void fill_zero_kernel(float *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.f;
}

__global__ void fill_zero_bytes_kernel(unsigned char *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

int launch_zero(void *ptr, size_t bytes, hipStream_t st) {
    if (!bytes) return RSG_OK;
    if (!ptr) return fail(RSG_EINVAL, "rsg_zero: null pointer");
    const int threads = 256;
    if (bytes % 4 == 0 && ((size_t)ptr % 4) == 0) {
        size_t n = bytes / 4;
        int blocks = (int)((n + threads - 1) / threads);
        if (blocks > 4096) blocks = 4096;
        fill_zero_kernel<<<blocks, threads, 0, st>>>((float *)ptr, n);
    } else {
        int blocks = (int)((bytes + threads - 1) / threads);
        if (blocks > 4096) blocks = 4096;
        fill_zero_bytes_kernel<<<blocks, threads, 0, st>>>((unsigned char *)ptr, bytes);
    }
    RSG_LAUNCH_CHECK("fill_zero_kernel");
    return RSG_OK;
}

}  // namespace rsg

using namespace rsg;

extern "C" {

const char *rsg_last_error(void) { return g_last_error.c_str(); }

int rsg_device_count(int *n) {
    if (!n) return fail(RSG_EINVAL, "null n");
    RSG_HIP(hipGetDeviceCount(n));
    return RSG_OK;
}

int rsg_set_device(int dev) {
    RSG_HIP(hipSetDevice(dev));
    return RSG_OK;
}

int rsg_stream_sync(void *stream) {
    if (stream)
        RSG_HIP(hipStreamSynchronize((hipStream_t)stream));
    else
        RSG_HIP(hipDeviceSynchronize());
    return RSG_OK;
}

int rsg_malloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(RSG_EINVAL, "null ptr");
    *ptr = nullptr;
    if (!bytes) return RSG_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(RSG_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return RSG_OK;
}

int rsg_free(void *ptr) {
    if (ptr) RSG_HIP(hipFree(ptr));
    return RSG_OK;
}

int rsg_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return RSG_OK;
    RSG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    RSG_HIP(hipStreamSynchronize((hipStream_t)stream));
    return RSG_OK;
}

int rsg_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return RSG_OK;
    RSG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    RSG_HIP(hipStreamSynchronize((hipStream_t)stream));
    return RSG_OK;
}

int rsg_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return RSG_OK;
    RSG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return RSG_OK;
}

int rsg_set_tuning(const char *key, int value) {
    if (!key) return fail(RSG_EINVAL, "null key");
    std::string k(key);
    if (k == "rz2d") {
        if (value != 1 && value != 2 && value != 4 && value != 8)
            return fail(RSG_EINVAL, "rz2d must be 1, 2, 4 or 8");
        tuning().rz2d = value;
    } else if (k == "rz2d_ragged") {
        if (value != 1 && value != 2 && value != 4 && value != 8)
            return fail(RSG_EINVAL, "rz2d_ragged must be 1, 2, 4 or 8");
        tuning().rz2d_ragged = value;
    } else if (k == "zc3d") {
        if (value < 0 || value > 4096) return fail(RSG_EINVAL, "zc3d must be 0..4096");
        tuning().zc3d = value;
    } else if (k == "fuse_io") {
        if (value != 0 && value != 1) return fail(RSG_EINVAL, "fuse_io must be 0 or 1");
        tuning().fuse_io = value;
    } else if (k == "pml_cost") {
        if (value < 1 || value > 10000) return fail(RSG_EINVAL, "pml_cost must be 1..10000");
        tuning().pml_cost = value;
    } else if (k == "xcd3d") {
        if (value != 0 && value != 1) return fail(RSG_EINVAL, "xcd3d must be 0 or 1");
        tuning().xcd3d = value;
    } else {
        return fail(RSG_EINVAL, "unknown tuning key " + k);
    }
    return RSG_OK;
}

int rsg_get_tuning(const char *key, int *value) {
    if (!key || !value) return fail(RSG_EINVAL, "null argument");
    std::string k(key);
    if (k == "rz2d")
        *value = tuning().rz2d;
    else if (k == "rz2d_ragged")
        *value = tuning().rz2d_ragged;
    else if (k == "zc3d")
        *value = tuning().zc3d;
    else if (k == "fuse_io")
        *value = tuning().fuse_io;
    else if (k == "pml_cost")
        *value = tuning().pml_cost;
    else if (k == "xcd3d")
        *value = tuning().xcd3d;
    else
        return fail(RSG_EINVAL, "unknown tuning key " + k);
    return RSG_OK;
}

int rsg_zero(void *ptr, size_t bytes, void *stream) {
    return launch_zero(ptr, bytes, (hipStream_t)stream);
}

}  // extern "C"
