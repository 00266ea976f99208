// Building blocks shared by the linearisation kernels (k_assemble.hip,
// k_assemble_frames.hip): LDS-only barriers, the register butterfly that sums
// the 27 pair-sum entries over a wavefront, MFMA wrappers, refined
// reciprocals, the robust weight and the measurement record.
#pragma once
#include "spgo_device.h"

namespace spgo {

namespace {

constexpr int NWAVE = WG / 64;

// workgroup barrier that orders LDS only: the kernel never reads back what it
// stores to global memory, so it does not wait for those stores
// (s_waitcnt lgkmcnt(0) with vmcnt/expcnt at their maximum)
__device__ __forceinline__ void lds_barrier() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// ---- sums over the wavefront of the 27 pair-sum entries --------------------
// A transposing butterfly: at every exchange a lane keeps one half of its
// entries and adds its partner's copy of that half, 27 -> 14 -> 7 -> 4 -> 2 -> 1
// entries, all in registers.  Partners differ in lane bit 5 and 4
// (v_permlane32_swap / v_permlane16_swap), then are mirrored within rows,
// half rows and quads (lane ^ 15, ^ 7, ^ 3: DPP), the lane with the bit clear
// keeping the first half; a last quad swap (lane ^ 1) adds the two copies.
// The masks 32, 16, 15, 7, 3, 1 are independent, so every total covers each
// of the 64 lanes once.

// Lower half: x + the upper partner's x; upper half: y + the lower partner's y.
// v_permlane32_swap(x, y) exchanges x of the upper half with y of the lower.
__device__ __forceinline__ float keep_sum32(float x, float y) {
    const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}
__device__ __forceinline__ double keep_sum32(double x, double y) {
    const unsigned long long bx = (unsigned long long)__double_as_longlong(x);
    const unsigned long long by = (unsigned long long)__double_as_longlong(y);
    const auto pl = __builtin_amdgcn_permlane32_swap((unsigned)bx, (unsigned)by, false, false);
    const auto ph = __builtin_amdgcn_permlane32_swap((unsigned)(bx >> 32), (unsigned)(by >> 32), false, false);
    const double p0 = __longlong_as_double((long long)(((unsigned long long)ph[0] << 32) | pl[0]));
    const double p1 = __longlong_as_double((long long)(((unsigned long long)ph[1] << 32) | pl[1]));
    return p0 + p1;
}
// The same between the rows (16 lanes) of each row pair: even rows keep x.
__device__ __forceinline__ float keep_sum16(float x, float y) {
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}
__device__ __forceinline__ double keep_sum16(double x, double y) {
    const unsigned long long bx = (unsigned long long)__double_as_longlong(x);
    const unsigned long long by = (unsigned long long)__double_as_longlong(y);
    const auto pl = __builtin_amdgcn_permlane16_swap((unsigned)bx, (unsigned)by, false, false);
    const auto ph = __builtin_amdgcn_permlane16_swap((unsigned)(bx >> 32), (unsigned)(by >> 32), false, false);
    const double p0 = __longlong_as_double((long long)(((unsigned long long)ph[0] << 32) | pl[0]));
    const double p1 = __longlong_as_double((long long)(((unsigned long long)ph[1] << 32) | pl[1]));
    return p0 + p1;
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// DPP exchange with the partner selected by CTRL, which differs in lane bit BIT
template <int CTRL, int BIT, typename T>
__device__ __forceinline__ T keep_sum_dpp(T x, T y, int lane) {
    const bool up = (lane >> BIT) & 1;
    return (up ? y : x) + dpp_move<CTRL>(up ? x : y);
}
constexpr int DPP_ROW_MIRROR = 0x140, DPP_ROW_HALF_MIRROR = 0x141;
constexpr int DPP_QUAD_REVERSE = 0x1B, DPP_QUAD_SWAP_PAIRS = 0xB1;   // quad_perm [3,2,1,0], [1,0,3,2]

// Returns the wavefront total of one entry in lanes 2i and 2i + 1 (which one:
// wave_sum27_entry).  Every lane of the wavefront must be active.
template <typename T>
__device__ __forceinline__ T wave_sum27(T (&v)[27], int lane) {
    T a[14], b[7], c[4], d[2];
#pragma unroll
    for (int i = 0; i < 14; ++i) a[i] = keep_sum32(v[i], i + 14 < 27 ? v[i + 14] : (T)0);
#pragma unroll
    for (int i = 0; i < 7; ++i) b[i] = keep_sum16(a[i], a[i + 7]);
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] = keep_sum_dpp<DPP_ROW_MIRROR, 3>(b[i], i + 4 < 7 ? b[i + 4] : (T)0, lane);
#pragma unroll
    for (int i = 0; i < 2; ++i) d[i] = keep_sum_dpp<DPP_ROW_HALF_MIRROR, 2>(c[i], c[i + 2], lane);
    T e = keep_sum_dpp<DPP_QUAD_REVERSE, 1>(d[0], d[1], lane);
    e += dpp_move<DPP_QUAD_SWAP_PAIRS>(e);
    return e;
}

// The entry whose total wave_sum27 leaves in this lane, -1 for padding: the
// same exchanges applied to the codes q + 1 give 64 (q + 1), exactly.
template <typename T>
__device__ __forceinline__ int wave_sum27_entry(int lane) {
    T code[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) code[i] = (T)(i + 1);
    return (int)wave_sum27(code, lane) / 64 - 1;
}

constexpr int PRE_SEG = 2;   // slot segments per wavefront whose measurements are loaded up front

template <typename T> struct Vec4;
template <> struct Vec4<double> { typedef double __attribute__((ext_vector_type(4))) type; };
template <> struct Vec4<float> { typedef float __attribute__((ext_vector_type(4))) type; };

__device__ __forceinline__ Vec4<double>::type mfma16x16x4(double a, double b, Vec4<double>::type c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ Vec4<float>::type mfma16x16x4(float a, float b, Vec4<float>::type c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// 1/x: the hardware estimate refined by two Newton steps (FP64); plain
// division in FP32
__device__ __forceinline__ double rcp_refined(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}
// FP32: the hardware reciprocal and reciprocal square root (1 ulp), where an
// IEEE division takes about ten instructions
__device__ __forceinline__ float rcp_refined(float x) { return __builtin_amdgcn_rcpf(x); }

// 1/sqrt(x): the hardware estimate refined by one Halley step (FP64)
__device__ __forceinline__ double rsqrt_refined(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(0.375, e, 0.5), y);
}
__device__ __forceinline__ float rsqrt_refined(float x) { return __builtin_amdgcn_rsqf(x); }

// square root and quotient: the hardware's 1-ulp forms in FP32, IEEE in FP64
__device__ __forceinline__ float root(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double root(double x) { return sqrt(x); }
__device__ __forceinline__ float quot(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double quot(double a, double b) { return a / b; }

// Robust weight w = sqrt(kappa * rho'(s)) and cost 1/2 kappa rho(s) of the
// squared residual norm s.  inv_param = 1 / param.  Cauchy:
// rho = c^2 log(1 + s / c^2), w = sqrt(kappa) / sqrt(1 + s / c^2).
template <typename T>
__device__ __forceinline__ void robust_eval(const Robust& rb, T inv_param, T s, T kappa, T& w, double& cost) {
    T rho, drho;
    if (rb.kind == 0) {
        const T c2 = (T)rb.param;
        const T x = s * inv_param;
        rho = c2 * log1p(x);
        w = (kappa == (T)16 ? (T)4 : (T)1) * rsqrt_refined((T)1 + x);
        cost += 0.5 * (double)kappa * (double)rho;
        return;
    } else if (rb.kind == 1) {
        const T d = (T)rb.param;
        const T rs = root(s);
        if (rs <= d) {
            rho = s;
            drho = (T)1;
        } else {
            rho = (T)2 * d * rs - d * d;
            drho = quot(d, rs);
        }
    } else {
        rho = s;
        drho = (T)1;
    }
    w = root(kappa * drho);
    cost += 0.5 * (double)kappa * (double)rho;
}

// One wavefront's measurement: ids and pixels, loaded ahead of use.
struct MeasIn {
    int t;
    uint32_t meta;
    float4 uv;
};

__device__ __forceinline__ MeasIn meas_load(const Layout& lay, int ob0, int o, bool valid) {
    MeasIn m;
    const int oc = valid ? ob0 + o : ob0;
    m.t = lay.ob_t[oc];
    m.meta = lay.ob_meta[oc];
    m.uv = lay.ob_uv[oc];
    return m;
}

}  // namespace

}  // namespace spgo
