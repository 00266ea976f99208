// Device-side SE(3) helpers and wavefront reductions for gfx950.
// Poses are 3x4 row-major [R | t] (the top three rows of the reference's
// row-major 4x4 matrices).
#pragma once
#include <hip/hip_runtime.h>

#include "spgo_internal.h"

namespace spgo {

// Rotation-angle threshold of the reference's small-angle branches
// (include/lie.h:4).
constexpr double SMALL_ANGLE = 1e-4;

template <typename T>
__device__ __forceinline__ void pose_load(const double* __restrict__ p, T* out) {
#pragma unroll
    for (int i = 0; i < 12; ++i) out[i] = (T)p[i];
}

// C = A * B for affine 3x4 poses.
template <typename T>
__device__ __forceinline__ void pose_mul(const T* A, const T* B, T* C) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            T s = A[4 * r + 0] * B[c] + A[4 * r + 1] * B[4 + c] + A[4 * r + 2] * B[8 + c];
            if (c == 3) s += A[4 * r + 3];
            C[4 * r + c] = s;
        }
    }
}

// Rigid inverse [R^T | -R^T t] (src/LieUtils/utils.cu:38-69).
template <typename T>
__device__ __forceinline__ void pose_inv(const T* A, T* B) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) B[4 * r + c] = A[4 * c + r];
#pragma unroll
    for (int r = 0; r < 3; ++r)
        B[4 * r + 3] = -(B[4 * r + 0] * A[3] + B[4 * r + 1] * A[7] + B[4 * r + 2] * A[11]);
}

// Exp of the twist (psi, rho): R = Rodrigues(psi), t = Jl(psi) rho, with the
// reference's second-order branch below |psi| = 1e-4
// (SE3.cu:3-38, SO3.cu:5-32, SO3.cu:82-107).
__device__ __forceinline__ void exp_se3(const double* xi, double* T) {
    const double p0 = xi[0], p1 = xi[1], p2 = xi[2];
    const double th = sqrt(p0 * p0 + p1 * p1 + p2 * p2);
    double S[9] = {0.0, -p2, p1, p2, 0.0, -p0, -p1, p0, 0.0};
    double S2[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            S2[3 * r + c] = S[3 * r] * S[c] + S[3 * r + 1] * S[3 + c] + S[3 * r + 2] * S[6 + c];
    double a, b, c1, c2;
    if (th > SMALL_ANGLE) {
        double s, co;
        sincos(th, &s, &co);
        a = s / th;
        b = (1.0 - co) / (th * th);
        c1 = b;
        c2 = (th - s) / (th * th * th);
    } else {
        a = 1.0;
        b = 0.5;
        c1 = 0.5;
        c2 = 1.0 / 6.0;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double t = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double id = (r == c) ? 1.0 : 0.0;
            T[4 * r + c] = id + a * S[3 * r + c] + b * S2[3 * r + c];
            t += (id + c1 * S[3 * r + c] + c2 * S2[3 * r + c]) * xi[3 + c];
        }
        T[4 * r + 3] = t;
    }
}

// out = Ad(G) x with Ad = [[R, 0], [[t]x R, R]] acting on (psi, rho).
__device__ __forceinline__ void adjoint_apply(const double* G, const double* x, double* out) {
    double Rp[3], Rr[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        Rp[r] = G[4 * r] * x[0] + G[4 * r + 1] * x[1] + G[4 * r + 2] * x[2];
        Rr[r] = G[4 * r] * x[3] + G[4 * r + 1] * x[4] + G[4 * r + 2] * x[5];
    }
    const double t0 = G[3], t1 = G[7], t2 = G[11];
    out[0] = Rp[0];
    out[1] = Rp[1];
    out[2] = Rp[2];
    out[3] = t1 * Rp[2] - t2 * Rp[1] + Rr[0];
    out[4] = t2 * Rp[0] - t0 * Rp[2] + Rr[1];
    out[5] = t0 * Rp[1] - t1 * Rp[0] + Rr[2];
}

// Entry (r, c) of Ad(G) (6x6).
__device__ __forceinline__ double adjoint_entry(const double* G, int r, int c) {
    if (r < 3) return c < 3 ? G[4 * r + c] : 0.0;
    if (c >= 3) return G[4 * (r - 3) + (c - 3)];
    // ([t]x R)[r-3][c]
    const int i = r - 3;
    const double t0 = G[3], t1 = G[7], t2 = G[11];
    const double R0 = G[c], R1 = G[4 + c], R2 = G[8 + c];
    if (i == 0) return t1 * R2 - t2 * R1;
    if (i == 1) return t2 * R0 - t0 * R2;
    return t0 * R1 - t1 * R0;
}

// ---------------------------------------------------------------------------
// wavefront (64-lane) reductions
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the four 16-lane rows of a wavefront (lanes l, l^16, l^32, l^48),
// the same in every lane, without LDS: v_permlane16_swap(v, v) returns the
// values of rows (2i, 2i+1) of the lane's row pair, v_permlane32_swap(v, v)
// those of the lower and upper half.  The additions are (r0 + r1) + (r2 + r3),
// as with exchanges over 16 then 32 lanes.
__device__ __forceinline__ float row_sum4(float v) {
    const unsigned u = __float_as_uint(v);
    const auto p = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    const float s = __uint_as_float(p[0]) + __uint_as_float(p[1]);
    const unsigned w = __float_as_uint(s);
    const auto q = __builtin_amdgcn_permlane32_swap(w, w, false, false);
    return __uint_as_float(q[0]) + __uint_as_float(q[1]);
}

__device__ __forceinline__ double row_sum4(double v) {
    auto mk = [](unsigned lo, unsigned hi) { return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo)); };
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    auto pl = __builtin_amdgcn_permlane16_swap((unsigned)b, (unsigned)b, false, false);
    auto ph = __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
    const double s = mk(pl[0], ph[0]) + mk(pl[1], ph[1]);
    b = (unsigned long long)__double_as_longlong(s);
    pl = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
    ph = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
    return mk(pl[0], ph[0]) + mk(pl[1], ph[1]);
}

// Row of the 16x16 MFMA accumulator held in register v of lane `lane`
// (column = lane & 15).  Measured on gfx950: the f32 16x16x4 result keeps
// four consecutive rows per lane group, the f64 16x16x4 result interleaves them.
__device__ __forceinline__ int mfma16_row(float, int lane, int v) { return 4 * (lane >> 4) + v; }
__device__ __forceinline__ int mfma16_row(double, int lane, int v) { return (lane >> 4) + 4 * v; }

// Index of (p, q), p <= q < 6, in the row-major upper triangle (21 entries).
__device__ __forceinline__ int sym6(int p, int q) {
    if (p > q) {
        const int t = p;
        p = q;
        q = t;
    }
    return p * 6 - (p * (p - 1)) / 2 + (q - p);
}

}  // namespace spgo
