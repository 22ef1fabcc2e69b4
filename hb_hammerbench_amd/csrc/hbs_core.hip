// hbs_core.hip -- status codes, device check and input generators.
#include <string.h>

#include <mutex>

#include "hbs_internal.h"

namespace hbs {

int check_device() {
    static std::mutex mu;
    static int state[64];  // 0 unknown, 1 gfx950, 2 other
    int dev = 0;
    HBS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return HBS_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(mu);
    if (state[dev] == 0) {
        hipDeviceProp_t prop;
        HBS_HIP_TRY(hipGetDeviceProperties(&prop, dev));
        state[dev] = (strncmp(prop.gcnArchName, "gfx950", 6) == 0) ? 1 : 2;
    }
    return state[dev] == 1 ? HBS_SUCCESS : HBS_ERR_DEVICE;
}

int cu_count() {
    static std::mutex mu;
    static int cus[64];  // 0 unknown
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    std::lock_guard<std::mutex> lk(mu);
    if (cus[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n < 1)
            n = 256;
        cus[dev] = n;
    }
    return cus[dev];
}

}  // namespace hbs

extern "C" const char *hbs_strerror(int code) {
    switch (code) {
        case HBS_SUCCESS: return "success";
        case HBS_ERR_ARG: return "invalid size or argument";
        case HBS_ERR_NULL: return "null pointer argument";
        case HBS_ERR_ALIGN: return "pointer or leading dimension misaligned";
        case HBS_ERR_WORKSPACE: return "workspace missing or too small";
        case HBS_ERR_DEVICE: return "current device is not a gfx950 GPU";
        case HBS_ERR_UNSUPPORTED: return "request not supported by this build";
        case HBS_ERR_SCRATCH:
            return "problem exceeds the fixed scratch of an allocation-free entry point";
        case HBS_ERR_IN_USE: return "grouped plan still has executable graphs";
        default: break;
    }
    if (code > 0) return hipGetErrorString((hipError_t)code);
    return "unknown status code";
}

extern "C" int hbs_version(void) { return 500; }

// ---------------------------------------------------------------------------
// apps/sgemm/main.cpp:46-52 -- mat1[i] = i % 7, mat2[i] = i % 3.
// ---------------------------------------------------------------------------
__global__ void fill_pattern_kernel(float *__restrict__ a, float *__restrict__ b, int64_t n,
                                    int64_t first) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = first + k;
        a[k] = (float)(i % 7);
        b[k] = (float)(i % 3);
    }
}

extern "C" int hbs_fill_pattern(float *mat1, float *mat2, int64_t n, int64_t first,
                                void *stream) {
    if (n < 0 || first < 0) return HBS_ERR_ARG;
    if (n == 0) return HBS_SUCCESS;
    if (!mat1 || !mat2) return HBS_ERR_NULL;
    int st = hbs::check_device();
    if (st) return st;
    const int threads = 256;
    int64_t blocks = (n + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    fill_pattern_kernel<<<(unsigned)blocks, threads, 0, (hipStream_t)stream>>>(mat1, mat2, n, first);
    return hbs::launch_status();
}

// ---------------------------------------------------------------------------
// Counter-based generators.  Element i of seed's stream is the (i+1)-th output
// of SplitMix64 seeded with `seed`, so any slice is generated independently.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t rng64(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float gen_one(int kind, uint64_t seed, uint64_t i) {
    if (kind == HBS_GEN_UNIFORM) {
        const uint64_t r = rng64(seed, i);
        const float u = (float)(r >> 40) * 0x1p-24f;
        return 2.0f * u - 1.0f;
    }
    if (kind == HBS_GEN_LOGUNIFORM) {
        const uint64_t r = rng64(seed, i);
        const double u = (double)(r >> 11) * 0x1p-53;
        const float m = (float)exp(2.302585092994045684 * (-3.0 + 6.0 * u));
        return (r & 1ULL) ? -m : m;
    }
    // normal(0,1) by Box-Muller with a Bernoulli(0.01) x1e3 outlier mask
    const uint64_t r1 = rng64(seed, 3 * i), r2 = rng64(seed, 3 * i + 1), r3 = rng64(seed, 3 * i + 2);
    const double u1 = (double)((r1 >> 11) + 1) * 0x1p-53;
    const double u2 = (double)(r2 >> 11) * 0x1p-53;
    double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586477 * u2);
    const double u3 = (double)(r3 >> 11) * 0x1p-53;
    if (u3 < 0.01) z *= 1000.0;
    return (float)z;
}

__global__ void fill_random_kernel(float *__restrict__ dst, int64_t n, int kind, uint64_t seed,
                                   uint64_t first) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        dst[k] = gen_one(kind, seed, first + (uint64_t)k);
    }
}

extern "C" int hbs_fill_random(float *dst, int64_t n, int kind, uint64_t seed, uint64_t first,
                               void *stream) {
    if (n < 0) return HBS_ERR_ARG;
    if (kind < HBS_GEN_UNIFORM || kind > HBS_GEN_NORMAL_OUTLIER) return HBS_ERR_ARG;
    if (n == 0) return HBS_SUCCESS;
    if (!dst) return HBS_ERR_NULL;
    int st = hbs::check_device();
    if (st) return st;
    const int threads = 256;
    int64_t blocks = (n + threads - 1) / threads;
    if (blocks > 262144) blocks = 262144;
    fill_random_kernel<<<(unsigned)blocks, threads, 0, (hipStream_t)stream>>>(dst, n, kind, seed,
                                                                             first);
    return hbs::launch_status();
}
