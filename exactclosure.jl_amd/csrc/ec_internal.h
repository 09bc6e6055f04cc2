// Internal definitions shared by the HIP translation units of libexactclosure.
// gfx950 (CDNA4) only: 64-wide wavefronts, 160 KiB LDS per CU.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/exactclosure.h"

struct ec_prof_rec {
    const char *name;
    hipEvent_t start, stop;
    double bytes;  // algorithmic bytes moved by this launch
};

struct ec_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    char err[1024] = {0};
    // Reduction workspace (device partials + pinned host result).
    double *red_partial = nullptr;
    double *red_host = nullptr;
    int red_cap = 0;
    // Per-launch timing, enabled by ec_profile_begin.
    bool prof = false;
    std::vector<ec_prof_rec> prof_recs;
    std::vector<hipEvent_t> event_pool;
    size_t event_next = 0;
    // Launch-shape choices of this context (set by the autotuner or by hand):
    // z planes per block of the velocity filter (256 is fastest at 1024^3) and
    // the x/y/z block targets of the flux-form closure kernel.
    int fv_tz = 256;
    // compute units of the device (persistent launches size their grid by it)
    int ncu = 0;
    // column passes of large FP32 spectra: persistent blocks that prefetch
    // their next tile (1) or one block per tile (0); the same bits either way
    int fft_cols_persistent = 1;
    // the z pass of FP32 Poisson solves: tridiagonal solves per column (1) or
    // the forward transform, scale and inverse transform (0)
    int z_poisson_tri = 1;
    int cl_tx = 16, cl_ty = 16, cl_tz = 64;
    // Workspace of the separable filter passes (grown on demand, FP64).
    void *ws = nullptr;
    size_t ws_doubles = 0;
    void *ws2 = nullptr;
    size_t ws2_doubles = 0;
    // The split-window partials of the velocity filter (FP64, grown on demand).
    void *ws3 = nullptr;
    size_t ws3_doubles = 0;
    // Graph capture (ec_capture_begin/end): while capturing, `stream` is the
    // private capture stream and `user_stream` the context's own.
    hipStream_t cap_stream = nullptr;
    hipStream_t user_stream = nullptr;
    bool capturing = false;
    // recorded on this context's stream by ec_stream_after(other, this)
    hipEvent_t order_event = nullptr;
};

// ---- error handling: no C++ exception crosses the C ABI ----
int ec_fail(ec_ctx *ctx, int code, const char *fmt, ...);

#define EC_HIP(ctx, call)                                                                     \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            (void)hipGetLastError(); /* do not leave the error for the next launch check */   \
            return ec_fail((ctx), EC_E_HIP, "%s failed: %s (%s:%d)", #call,                   \
                           hipGetErrorString(e_), __FILE__, __LINE__);                        \
        }                                                                                     \
    } while (0)

#define EC_CHECK_LAUNCH(ctx, name)                                                            \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess)                                                                 \
            return ec_fail((ctx), EC_E_HIP, "launch of %s failed: %s", (name),                \
                           hipGetErrorString(e_));                                            \
    } while (0)

#define EC_REQUIRE(ctx, cond, ...)                                                            \
    do {                                                                                      \
        if (!(cond)) return ec_fail((ctx), EC_E_INVALID, __VA_ARGS__);                        \
    } while (0)

// ---- launch timing (HIP events on the context stream) ----
// Returns the record index or -1 when profiling is off.
int ec_prof_start(ec_ctx *ctx, const char *name, double bytes);
void ec_prof_stop(ec_ctx *ctx, int rec);

struct ec_prof_scope {
    ec_ctx *ctx;
    int rec;
    ec_prof_scope(ec_ctx *c, const char *name, double bytes) : ctx(c), rec(ec_prof_start(c, name, bytes)) {}
    ~ec_prof_scope() { ec_prof_stop(ctx, rec); }
};

// ---- complex helpers ----
// Aligned to its size, so that a complex value moves as one 8- (FP32) or
// 16-byte (FP64) access; LDS accesses then take 16-bit immediate offsets.
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T x, y;
};

// Streaming (non-temporal) stores for outputs a kernel writes once and does
// not read back: they do not displace the lines its loads share in L2.
template <typename T>
__device__ __forceinline__ void ec_nt_store(T *p, T v) {
    __builtin_nontemporal_store(v, p);
}
// Streaming loads for inputs a kernel reads exactly once: they leave L2 to
// the lines that neighbouring blocks share.
template <typename T>
__device__ __forceinline__ T ec_nt_load(const T *p) {
    return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ cplx<T> ec_nt_load(const cplx<T> *p) {  // one 8- or 16-byte access
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(p));
    return cplx<T>{v.x, v.y};
}
template <typename VT>
__device__ __forceinline__ VT ec_nt_load16(const VT *p) {  // a 16-byte vector (float4, double2)
    static_assert(sizeof(VT) == 16, "16-byte vector");
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const u4 w = __builtin_nontemporal_load(reinterpret_cast<const u4 *>(p));
    VT v;
    __builtin_memcpy(&v, &w, 16);
    return v;
}
// a value or a 16-byte vector, read once
template <typename VT>
__device__ __forceinline__ VT ec_nt_ld(const VT *p) {
    if constexpr (sizeof(VT) == 16) return ec_nt_load16(p);
    else return ec_nt_load(p);
}
template <typename VT>
__device__ __forceinline__ void ec_nt_store16(VT *p, const VT &v) {  // a 16-byte vector (float4, double2)
    static_assert(sizeof(VT) == 16, "16-byte vector");
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    u4 w;
    __builtin_memcpy(&w, &v, 16);
    __builtin_nontemporal_store(w, reinterpret_cast<u4 *>(p));
}
// RK4 in stored stage states (ec_ns_rk4_update_states): the states
// u_{s+1} = u0 + c_s dt k_s (c = 1/2, 1/2, 1) carry k1, k2, k3, so the new
// state u0 + dt/6 (k1 + 2 k2 + 2 k3 + k4) of rk4_timestep! is
// u0 + ((u1 - u0) + 2 (u2 - u0) + (u3 - u0)) / 3 + dt/6 k4.  The differences
// of neighbouring states are formed first (exact where they are within a
// factor 2 of each other), so the sum meets u0 once, as u0 + dt/6 (acc + k4)
// does.
template <typename T>
__device__ __forceinline__ T ec_rk4_stage_coef(int stage, T dt) {
    return stage == 2 ? dt : dt / 2;
}
template <typename T>
__device__ __forceinline__ T ec_rk4_combine(T u0, T u1, T u2, T u3, T k, T dt) {
    const T d = (u1 - u0) + 2 * (u2 - u0) + (u3 - u0);
    return u0 + (d * (T)(1.0 / 3.0) + (dt / 6) * k);
}
// Non-temporal stores through a raw buffer descriptor of a uniform base: the
// address is the descriptor's base plus a 32-bit byte offset, so a kernel that
// writes many fields at the same offsets (a plane of a tile) forms no per-lane
// 64-bit addresses.  Word 3 is the gfx9 raw-buffer format (32-bit data);
// num_records only bounds the offsets, which stay below 2^31.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t ec_rsrc(void *base) {
    return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ void ec_nt_store_at(__amdgpu_buffer_rsrc_t r, unsigned byte_off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, byte_off, 0, 2 /* nt */);
}
__device__ __forceinline__ void ec_nt_store_at(__amdgpu_buffer_rsrc_t r, unsigned byte_off, double v) {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, byte_off, 0, 2 /* nt */);
}
template <typename T>
__device__ __forceinline__ void ec_nt_store(cplx<T> *p, cplx<T> v) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<v2 *>(p));
}

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
    return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
    return {a.x - b.x, a.y - b.y};
}

// periodic neighbours without division (|shift| == 1)
__device__ __forceinline__ int ec_up(int i, int n) { return i + 1 == n ? 0 : i + 1; }
__device__ __forceinline__ int ec_dn(int i, int n) { return i == 0 ? n - 1 : i - 1; }
// periodic wrap of an index within one period of [0, n): i in [-n, 2n)
__device__ __forceinline__ int ec_wrap_any(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// XCD-aware tile order for a 1D grid.  gfx950 dispatches workgroup b to XCD
// b % 8, and each XCD has its own L2; this maps the workgroups of one XCD to
// one contiguous range of logical tiles, so tiles whose halos overlap are
// fetched into the same L2 at about the same time.  A bijection on
// [0, total) for any total.
__device__ __forceinline__ unsigned ec_xcd_tile(unsigned b, unsigned total) {
    constexpr unsigned NX = 8;
    const unsigned q = total / NX, r = total % NX, x = b % NX;
    return x * q + (x < r ? x : r) + b / NX;
}

static inline size_t ec_cube(int n) { return (size_t)n * (size_t)n * (size_t)n; }

// Poisson solver state (FFT plan, spectrum and pressure buffers).
struct ec_poisson {
    int n = 0;
    int dtype = EC_F64;
    double l = 1.0;
    int nh = 0;
    int pitch = 0;            // row pitch of spec in complex elements: nh rounded up to 16
    void *spec = nullptr;     // (pitch, n, n) complex of dtype, columns kx < nh used
    void *p = nullptr;        // n^3 pressure of dtype
    void *twiddle = nullptr;  // exp(-2 pi i t / n), t in [0, n), complex of dtype
    void *eig = nullptr;      // sin^2(pi k / n), k in [0, n), real of dtype
    int nrad = 0;
    int radix[32] = {0};
    bool own_p = true;
};

// Internal entry points shared between translation units.



int ec_closure_tiled(ec_ctx *ctx, int dtype, int n, double l, double visc, void *const u[3], const void *p,
                     int nsets, const ec_closure_set *sets);
