// Context, memory, error reporting, launch timing, reductions and small
// element-wise kernels of libexactclosure.
#include "ec_internal.h"

#include <cmath>

int ec_fail(ec_ctx *ctx, int code, const char *fmt, ...) {
    if (ctx) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

extern "C" int ec_version(void) { return 1; }

extern "C" int ec_create(ec_ctx **out, int device, void *stream) {
    if (!out) return EC_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return EC_E_HIP;
    if (device < 0 || device >= ndev) return EC_E_INVALID;
    if (hipSetDevice(device) != hipSuccess) return EC_E_HIP;
    ec_ctx *ctx = new ec_ctx();
    ctx->device = device;
    ctx->stream = (hipStream_t)stream;  // NULL: the device's default stream
    ctx->own_stream = false;
    ctx->red_cap = 4096;
    if (hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ctx->ncu < 1)
        ctx->ncu = 256;
    if (hipMalloc(&ctx->red_partial, sizeof(double) * ctx->red_cap) != hipSuccess ||
        hipHostMalloc(&ctx->red_host, sizeof(double) * 4) != hipSuccess) {
        delete ctx;
        return EC_E_OOM;
    }
    *out = ctx;
    return EC_OK;
}

extern "C" int ec_destroy(ec_ctx *ctx) {
    if (!ctx) return EC_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto &r : ctx->prof_recs) {
        (void)r;
    }
    for (auto e : ctx->event_pool) hipEventDestroy(e);
    if (ctx->red_partial) hipFree(ctx->red_partial);
    if (ctx->ws) hipFree(ctx->ws);
    if (ctx->ws2) hipFree(ctx->ws2);
    if (ctx->ws3) hipFree(ctx->ws3);
    if (ctx->red_host) hipHostFree(ctx->red_host);
    if (ctx->capturing) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->cap_stream, &g);
        if (g) hipGraphDestroy(g);
        ctx->stream = ctx->user_stream;
    }
    if (ctx->order_event) hipEventDestroy(ctx->order_event);
    if (ctx->cap_stream) hipStreamDestroy(ctx->cap_stream);
    if (ctx->own_stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return EC_OK;
}

extern "C" const char *ec_last_error(ec_ctx *ctx) { return ctx ? ctx->err : "null context"; }

// ---------------------------------------------------------------- graph capture
// The context's calls between ec_capture_begin and ec_capture_end are recorded
// on a private stream (the legacy default stream cannot be captured) instead
// of being run; the instantiated graph then replays them on the context's
// stream, ordered with its other work.
extern "C" int ec_capture_begin(ec_ctx *ctx) {
    EC_REQUIRE(ctx, ctx, "null context");
    EC_REQUIRE(ctx, !ctx->capturing, "ec_capture_begin: already capturing");
    EC_REQUIRE(ctx, !ctx->prof, "ec_capture_begin: per-launch profiling is on");
    if (!ctx->cap_stream) EC_HIP(ctx, hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking));
    // the captured work reads what the context's stream has queued so far
    EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    EC_HIP(ctx, hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeRelaxed));
    ctx->user_stream = ctx->stream;
    ctx->stream = ctx->cap_stream;
    ctx->capturing = true;
    return EC_OK;
}

extern "C" int ec_capture_end(ec_ctx *ctx, void **graph) {
    EC_REQUIRE(ctx, ctx && graph, "ec_capture_end: null argument");
    EC_REQUIRE(ctx, ctx->capturing, "ec_capture_end: not capturing");
    *graph = nullptr;
    ctx->stream = ctx->user_stream;
    ctx->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->cap_stream, &g);
    if (e != hipSuccess || !g) return ec_fail(ctx, EC_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return ec_fail(ctx, EC_E_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
    *graph = ge;
    return EC_OK;
}

extern "C" int ec_graph_launch(ec_ctx *ctx, void *graph) {
    EC_REQUIRE(ctx, ctx && graph, "ec_graph_launch: null argument");
    EC_REQUIRE(ctx, !ctx->capturing, "ec_graph_launch: the context is capturing");
    EC_HIP(ctx, hipGraphLaunch((hipGraphExec_t)graph, ctx->stream));
    return EC_OK;
}

extern "C" int ec_graph_destroy(ec_ctx *ctx, void *graph) {
    EC_REQUIRE(ctx, ctx, "null context");
    if (graph) EC_HIP(ctx, hipGraphExecDestroy((hipGraphExec_t)graph));
    return EC_OK;
}

extern "C" int ec_stream_after(ec_ctx *ctx, ec_ctx *other) {
    EC_REQUIRE(ctx, ctx && other && other != ctx, "ec_stream_after: needs two different contexts");
    EC_REQUIRE(ctx, other->device == ctx->device, "ec_stream_after: contexts of devices %d and %d", ctx->device,
               other->device);
    EC_REQUIRE(ctx, other->stream != ctx->stream, "ec_stream_after: both contexts order on one stream");
    if (!other->order_event) EC_HIP(ctx, hipEventCreateWithFlags(&other->order_event, hipEventDisableTiming));
    EC_HIP(ctx, hipEventRecord(other->order_event, other->stream));
    EC_HIP(ctx, hipStreamWaitEvent(ctx->stream, other->order_event, 0));
    return EC_OK;
}

extern "C" int ec_sync(ec_ctx *ctx) {
    EC_REQUIRE(ctx, ctx, "null context");
    EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EC_OK;
}

// ---------------------------------------------------------------- memory
extern "C" int ec_alloc(ec_ctx *ctx, size_t bytes, void **ptr) {
    EC_REQUIRE(ctx, ctx && ptr, "ec_alloc: null argument");
    *ptr = nullptr;
    if (bytes == 0) return EC_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) return ec_fail(ctx, EC_E_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return EC_OK;
}

extern "C" int ec_free(ec_ctx *ctx, void *ptr) {
    EC_REQUIRE(ctx, ctx, "null context");
    if (ptr) EC_HIP(ctx, hipFree(ptr));
    return EC_OK;
}

extern "C" int ec_h2d(ec_ctx *ctx, void *dst, const void *src, size_t bytes) {
    EC_REQUIRE(ctx, ctx && (bytes == 0 || (dst && src)), "ec_h2d: null argument");
    EC_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EC_OK;
}

extern "C" int ec_d2h(ec_ctx *ctx, void *dst, const void *src, size_t bytes) {
    EC_REQUIRE(ctx, ctx && (bytes == 0 || (dst && src)), "ec_d2h: null argument");
    EC_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EC_OK;
}

extern "C" int ec_d2d(ec_ctx *ctx, void *dst, const void *src, size_t bytes) {
    EC_REQUIRE(ctx, ctx && (bytes == 0 || (dst && src)), "ec_d2d: null argument");
    EC_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return EC_OK;
}

// Zero fill through a kernel (graph-capture safe, unlike a memset node).
__global__ void k_zero(uint32_t *p, size_t nwords, uint8_t *tail, size_t ntail) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < nwords; i += stride) p[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

extern "C" int ec_zero(ec_ctx *ctx, void *dst, size_t bytes) {
    EC_REQUIRE(ctx, ctx && (bytes == 0 || dst), "ec_zero: null argument");
    if (bytes == 0) return EC_OK;
    EC_REQUIRE(ctx, ((uintptr_t)dst & 3) == 0, "ec_zero: pointer must be 4-byte aligned");
    size_t nwords = bytes / 4, ntail = bytes % 4;
    int blocks = (int)std::min<size_t>((nwords + 255) / 256 + 1, 65536);
    k_zero<<<blocks, 256, 0, ctx->stream>>>((uint32_t *)dst, nwords, (uint8_t *)dst + nwords * 4, ntail);
    EC_CHECK_LAUNCH(ctx, "k_zero");
    return EC_OK;
}

template <typename T>
__global__ void k_axpy(size_t count, T a, const T *__restrict__ x, T *__restrict__ y) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) y[i] += a * x[i];
}

extern "C" int ec_axpy(ec_ctx *ctx, int dtype, size_t count, double a, const void *x, void *y) {
    EC_REQUIRE(ctx, ctx && (count == 0 || (x && y)), "ec_axpy: null argument");
    if (count == 0) return EC_OK;
    int blocks = (int)std::min<size_t>((count + 255) / 256, 1 << 20);
    if (dtype == EC_F64)
        k_axpy<double><<<blocks, 256, 0, ctx->stream>>>(count, a, (const double *)x, (double *)y);
    else
        k_axpy<float><<<blocks, 256, 0, ctx->stream>>>(count, (float)a, (const float *)x, (float *)y);
    EC_CHECK_LAUNCH(ctx, "k_axpy");
    return EC_OK;
}

template <typename T>
__global__ void k_axpby(size_t count, T a, const T *__restrict__ x, T b, T *__restrict__ y) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) y[i] = a * x[i] + b * y[i];
}

extern "C" int ec_axpby(ec_ctx *ctx, int dtype, size_t count, double a, const void *x, double b, void *y) {
    EC_REQUIRE(ctx, ctx && (count == 0 || (x && y)), "ec_axpby: null argument");
    if (count == 0) return EC_OK;
    const int blocks = (int)std::min<size_t>((count + 255) / 256, 1 << 20);
    if (dtype == EC_F64)
        k_axpby<double><<<blocks, 256, 0, ctx->stream>>>(count, a, (const double *)x, b, (double *)y);
    else
        k_axpby<float><<<blocks, 256, 0, ctx->stream>>>(count, (float)a, (const float *)x, (float)b, (float *)y);
    EC_CHECK_LAUNCH(ctx, "k_axpby");
    return EC_OK;
}

template <typename Y, typename X>
__global__ void k_axpy2(size_t count, double a, const X *__restrict__ x, Y *__restrict__ y) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) y[i] = (Y)((double)y[i] + a * (double)x[i]);
}

extern "C" int ec_axpy2(ec_ctx *ctx, int ydtype, size_t count, double a, int xdtype, const void *x, void *y) {
    EC_REQUIRE(ctx, ctx && (count == 0 || (x && y)), "ec_axpy2: null argument");
    if (count == 0) return EC_OK;
    const int blocks = (int)std::min<size_t>((count + 255) / 256, 1 << 20);
    if (ydtype == EC_F32 && xdtype == EC_F64)
        k_axpy2<float, double><<<blocks, 256, 0, ctx->stream>>>(count, a, (const double *)x, (float *)y);
    else if (ydtype == EC_F64 && xdtype == EC_F64)
        k_axpy2<double, double><<<blocks, 256, 0, ctx->stream>>>(count, a, (const double *)x, (double *)y);
    else if (ydtype == EC_F64 && xdtype == EC_F32)
        k_axpy2<double, float><<<blocks, 256, 0, ctx->stream>>>(count, a, (const float *)x, (double *)y);
    else
        k_axpy2<float, float><<<blocks, 256, 0, ctx->stream>>>(count, a, (const float *)x, (float *)y);
    EC_CHECK_LAUNCH(ctx, "k_axpy2");
    return EC_OK;
}

template <typename D, typename S>
__global__ void k_convert(size_t count, D *__restrict__ dst, const S *__restrict__ src) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) dst[i] = (D)src[i];
}

extern "C" int ec_convert(ec_ctx *ctx, int dst_dtype, void *dst, int src_dtype, const void *src,
                          size_t count) {
    EC_REQUIRE(ctx, ctx && (count == 0 || (dst && src)), "ec_convert: null argument");
    if (count == 0) return EC_OK;
    int blocks = (int)std::min<size_t>((count + 255) / 256, 1 << 20);
    if (dst_dtype == EC_F32 && src_dtype == EC_F64)
        k_convert<float, double><<<blocks, 256, 0, ctx->stream>>>(count, (float *)dst, (const double *)src);
    else if (dst_dtype == EC_F64 && src_dtype == EC_F32)
        k_convert<double, float><<<blocks, 256, 0, ctx->stream>>>(count, (double *)dst, (const float *)src);
    else if (dst_dtype == EC_F64 && src_dtype == EC_F64)
        EC_HIP(ctx, hipMemcpyAsync(dst, src, count * 8, hipMemcpyDeviceToDevice, ctx->stream));
    else if (dst_dtype == EC_F32 && src_dtype == EC_F32)
        EC_HIP(ctx, hipMemcpyAsync(dst, src, count * 4, hipMemcpyDeviceToDevice, ctx->stream));
    else
        return ec_fail(ctx, EC_E_INVALID, "ec_convert: bad dtype");
    EC_CHECK_LAUNCH(ctx, "k_convert");
    return EC_OK;
}

// ---------------------------------------------------------------- timing
int ec_prof_start(ec_ctx *ctx, const char *name, double bytes) {
    if (!ctx || !ctx->prof) return -1;
    while (ctx->event_pool.size() < ctx->event_next + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ctx->event_pool.push_back(e);
    }
    ec_prof_rec r;
    r.name = name;
    r.bytes = bytes;
    r.start = ctx->event_pool[ctx->event_next++];
    r.stop = ctx->event_pool[ctx->event_next++];
    hipEventRecord(r.start, ctx->stream);
    ctx->prof_recs.push_back(r);
    return (int)ctx->prof_recs.size() - 1;
}

void ec_prof_stop(ec_ctx *ctx, int rec) {
    if (!ctx || rec < 0 || !ctx->prof) return;
    hipEventRecord(ctx->prof_recs[rec].stop, ctx->stream);
}

extern "C" int ec_profile_begin(ec_ctx *ctx) {
    EC_REQUIRE(ctx, ctx, "null context");
    ctx->prof = true;
    ctx->prof_recs.clear();
    ctx->event_next = 0;
    return EC_OK;
}

extern "C" int ec_profile_end(ec_ctx *ctx, const char **names, double *ms, double *bytes, int cap,
                              int *count) {
    EC_REQUIRE(ctx, ctx, "null context");
    EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int k = 0;
    for (auto &r : ctx->prof_recs) {
        if (k >= cap) break;
        float t = 0.f;
        EC_HIP(ctx, hipEventElapsedTime(&t, r.start, r.stop));
        if (names) names[k] = r.name;
        if (ms) ms[k] = t;
        if (bytes) bytes[k] = r.bytes;
        k++;
    }
    if (count) *count = (int)ctx->prof_recs.size();
    ctx->prof = false;
    ctx->prof_recs.clear();
    ctx->event_next = 0;
    return EC_OK;
}

// ---------------------------------------------------------------- reductions
// FP64 accumulation; wave-level butterfly over 64 lanes then an LDS pass.
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

template <int OP>  // 0 sum, 1 max
__device__ double block_reduce(double v) {
    __shared__ double part[16];
    v = OP == 0 ? wave_sum(v) : wave_max(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) part[wid] = v;
    __syncthreads();
    const int nw = blockDim.x >> 6;
    v = (threadIdx.x < nw) ? part[threadIdx.x] : 0.0;
    if (wid == 0) v = OP == 0 ? wave_sum(v) : wave_max(v);
    return v;
}

// MODE 0: sum x^2, 1: max |x|, 2: sum (x - y)^2
template <typename T, int MODE>
__global__ void k_reduce(size_t count, const T *__restrict__ x, const T *__restrict__ y,
                         double *__restrict__ partial) {
    double acc = 0.0;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) {
        double v = (double)x[i];
        if (MODE == 0) acc += v * v;
        else if (MODE == 1) acc = fmax(acc, fabs(v));
        else {
            double d = v - (double)y[i];
            acc += d * d;
        }
    }
    acc = block_reduce<MODE == 1 ? 1 : 0>(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

template <int OP>
__global__ void k_reduce_final(int nparts, const double *__restrict__ partial, double *__restrict__ out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc = OP == 0 ? acc + partial[i] : fmax(acc, partial[i]);
    acc = block_reduce<OP>(acc);
    if (threadIdx.x == 0) out[0] = acc;
}

template <typename T, int MODE>
static int reduce_launch(ec_ctx *ctx, size_t count, const void *x, const void *y, double *out) {
    const int threads = 256;
    int blocks = (int)std::min<size_t>((count + threads - 1) / threads, (size_t)2048);
    if (blocks < 1) blocks = 1;
    k_reduce<T, MODE><<<blocks, threads, 0, ctx->stream>>>(count, (const T *)x, (const T *)y, ctx->red_partial);
    EC_CHECK_LAUNCH(ctx, "k_reduce");
    k_reduce_final<MODE == 1 ? 1 : 0><<<1, 1024, 0, ctx->stream>>>(blocks, ctx->red_partial, ctx->red_partial + 2048);
    EC_CHECK_LAUNCH(ctx, "k_reduce_final");
    EC_HIP(ctx, hipMemcpyAsync(ctx->red_host, ctx->red_partial + 2048, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    EC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = ctx->red_host[0];
    return EC_OK;
}

extern "C" int ec_reduce_maxabs(ec_ctx *ctx, int dtype, size_t count, const void *x, double *out) {
    EC_REQUIRE(ctx, ctx && out && (count == 0 || x), "ec_reduce_maxabs: null argument");
    if (count == 0) { *out = 0.0; return EC_OK; }
    return dtype == EC_F64 ? reduce_launch<double, 1>(ctx, count, x, nullptr, out)
                           : reduce_launch<float, 1>(ctx, count, x, nullptr, out);
}

extern "C" int ec_reduce_sumsq(ec_ctx *ctx, int dtype, size_t count, const void *x, double *out) {
    EC_REQUIRE(ctx, ctx && out && (count == 0 || x), "ec_reduce_sumsq: null argument");
    if (count == 0) { *out = 0.0; return EC_OK; }
    return dtype == EC_F64 ? reduce_launch<double, 0>(ctx, count, x, nullptr, out)
                           : reduce_launch<float, 0>(ctx, count, x, nullptr, out);
}

extern "C" int ec_reduce_diffsq(ec_ctx *ctx, int dtype, size_t count, const void *x, const void *y,
                                double *out) {
    EC_REQUIRE(ctx, ctx && out && (count == 0 || (x && y)), "ec_reduce_diffsq: null argument");
    if (count == 0) { *out = 0.0; return EC_OK; }
    return dtype == EC_F64 ? reduce_launch<double, 2>(ctx, count, x, y, out)
                           : reduce_launch<float, 2>(ctx, count, x, y, out);
}

// ---------------------------------------------------------------- random numbers
// Counter-based normals: splitmix64 hashing + Box-Muller, the same integer
// recipe the FP64 CPU oracle uses, so both produce the same variates up to
// the last-ulp differences of log/cos.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

template <typename T>
__global__ void k_randn(size_t count, uint64_t seed, uint64_t offset, T *__restrict__ out) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) {
        const uint64_t ctr = offset + i;
        const uint64_t h1 = splitmix64(seed ^ splitmix64(2 * ctr));
        const uint64_t h2 = splitmix64(seed ^ splitmix64(2 * ctr + 1));
        const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;
        const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
        out[i] = (T)(sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2));
    }
}

extern "C" int ec_randn(ec_ctx *ctx, int dtype, size_t count, uint64_t seed, uint64_t offset, void *out) {
    EC_REQUIRE(ctx, ctx && (count == 0 || out), "ec_randn: null argument");
    if (count == 0) return EC_OK;
    int blocks = (int)std::min<size_t>((count + 255) / 256, 1 << 16);
    if (dtype == EC_F64)
        k_randn<double><<<blocks, 256, 0, ctx->stream>>>(count, seed, offset, (double *)out);
    else
        k_randn<float><<<blocks, 256, 0, ctx->stream>>>(count, seed, offset, (float *)out);
    EC_CHECK_LAUNCH(ctx, "k_randn");
    return EC_OK;
}
