// Learned preconditioner for CG on the reduced pose system.
//
// No reference counterpart: the reference factors the reduced system with
// torch's Cholesky (linear_solvers.cu:195-218).  Here a small message-passing
// network over the pose covisibility band predicts a factored approximate
// inverse P = G^T G of the block-banded reduced system H:
//   G_ii = (I + tril(out_i[0:21])) C_i,   G_{i,i-1} = out_i[21:57] C_{i-1},
// C_i the inverse of the lower Cholesky factor of H_ii.  With the network's
// head at zero P is block Jacobi.  Systems are stored as block bands,
// Hb[i][d] = H_{i,i-d} (6x6 row-major, d = 0..W), nodes of all systems
// concatenated.
//
// Kernels: features (C_i, normalised neighbour blocks and their norms), the
// network (embedding, three message-passing layers, head) forward and
// backward in FP32, the training loss in FP64 -- Kaporin's condition measure
// of P^1/2 H P^1/2 without its constant,
//   L = log(sum_i t_i / 6n) - (2 / 6n) sum_i sum_k log |(A_i)_kk|,
//   t_i = tr(A_i A_i^T) + 2 tr(A_i N_i B_i^T) + tr(B_i B_i^T),
// with A_i = I + tril(out_i[0:21]), B_i = out_i[21:57], N_i = C_i H_{i,i-1}
// C_{i-1}^T -- Adam, and PCG (one workgroup per system, FP64).
#include <hip/hip_cooperative_groups.h>

#include "spgo_device.h"
#include "spgo_internal.h"
#include "spgo_lpcg.h"

namespace spgo {

namespace {

__device__ __forceinline__ bool chol6(const double* A, double* L) {
    // lower Cholesky factor of the SPD 6x6 A (row-major); false if not positive
    for (int i = 0; i < 36; ++i) L[i] = 0.0;
    for (int j = 0; j < 6; ++j) {
        double d = A[j * 6 + j];
        for (int k = 0; k < j; ++k) d -= L[j * 6 + k] * L[j * 6 + k];
        if (!(d > 0.0)) return false;
        const double ljj = sqrt(d);
        L[j * 6 + j] = ljj;
        for (int i = j + 1; i < 6; ++i) {
            double s = A[i * 6 + j];
            for (int k = 0; k < j; ++k) s -= L[i * 6 + k] * L[j * 6 + k];
            L[i * 6 + j] = s / ljj;
        }
    }
    return true;
}

__device__ __forceinline__ void inv_lower6(const double* L, double* C) {
    // C = L^-1 (lower triangular)
    for (int i = 0; i < 36; ++i) C[i] = 0.0;
    for (int j = 0; j < 6; ++j) {
        C[j * 6 + j] = 1.0 / L[j * 6 + j];
        for (int i = j + 1; i < 6; ++i) {
            double s = 0.0;
            for (int k = j; k < i; ++k) s += L[i * 6 + k] * C[k * 6 + j];
            C[i * 6 + j] = -s / L[i * 6 + i];
        }
    }
}

__device__ __forceinline__ int sys_lo(const int* sys_begin, const int* node_sys, int i) {
    return sys_begin[node_sys[i]];
}
__device__ __forceinline__ int sys_hi(const int* sys_begin, const int* node_sys, int i) {
    return sys_begin[node_sys[i] + 1];
}

}  // namespace

// ---- features -------------------------------------------------------------
__global__ void k_lp_chol(int T, int W, const double* Hb, double* C, int* bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    double A[36], L[36], Ci[36];
    const double* h = Hb + (long long)i * (W + 1) * 36;
    for (int k = 0; k < 36; ++k) A[k] = h[k];
    if (!chol6(A, L)) {
        atomicOr(bad, 1);
        for (int k = 0; k < 36; ++k) C[(long long)i * 36 + k] = (k % 7 == 0) ? 1.0 : 0.0;
        return;
    }
    inv_lower6(L, Ci);
    for (int k = 0; k < 36; ++k) C[(long long)i * 36 + k] = Ci[k];
}

// thread per (node, d), d = 1..W: s[i][d] = ||C_i H_{i,i-d} C_{i-d}^T||_F and,
// for d = 1, the normalised block itself into the node input
__global__ void k_lp_feat1(int T, int W, const double* Hb, const double* C, const int* node_sys,
                           const int* sys_begin, double* s, float* X, double* N1) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T * W) return;
    const int i = t / W, d = t - i * W + 1;
    const int lo = sys_lo(sys_begin, node_sys, i);
    double nrm = 0.0;
    double Nb[36];
    if (i - d >= lo) {
        const double* Hd = Hb + ((long long)i * (W + 1) + d) * 36;
        const double* Ci = C + (long long)i * 36;
        const double* Cj = C + (long long)(i - d) * 36;
        double Tm[36];
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                double a = 0.0;
                for (int k = 0; k <= r; ++k) a += Ci[r * 6 + k] * Hd[k * 6 + c];
                Tm[r * 6 + c] = a;
            }
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) {
                double a = 0.0;
                for (int k = 0; k <= c; ++k) a += Tm[r * 6 + k] * Cj[c * 6 + k];
                Nb[r * 6 + c] = a;
                nrm += a * a;
            }
    } else {
        for (int k = 0; k < 36; ++k) Nb[k] = 0.0;
    }
    s[(long long)i * (W + 1) + d] = sqrt(nrm);
    if (d == 1)
        for (int k = 0; k < 36; ++k) {
            X[(long long)i * LP_NIN + k] = (float)Nb[k];
            N1[(long long)i * 36 + k] = Nb[k];
        }
}

__global__ void k_lp_feat2(int T, int W, const double* s, const int* node_sys, const int* sys_begin, float* X,
                           double* wsum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    const int lo = sys_lo(sys_begin, node_sys, i), hi = sys_hi(sys_begin, node_sys, i);
    double tot = 0.0;
    for (int d = 1; d <= W; ++d) {
        if (i - d >= lo) tot += s[(long long)i * (W + 1) + d];
        if (i + d < hi) tot += s[(long long)(i + d) * (W + 1) + d];
    }
    float* x = X + (long long)i * LP_NIN;
    x[36] = W >= 1 ? (float)s[(long long)i * (W + 1) + 1] : 0.f;
    x[37] = (W >= 1 && i + 1 < hi) ? (float)s[(long long)(i + 1) * (W + 1) + 1] : 0.f;
    x[38] = (float)(tot / (W > 0 ? W : 1));
    x[39] = 1.f;
    wsum[i] = 1.0 + tot;
}

// ---- network --------------------------------------------------------------
// The dense layers are products over the nodes on MFMA
// (v_mfma_f32_16x16x16_bf16): activations and weights enter as bfloat16
// (rounded to nearest even), products accumulate in float32, the master
// weights, biases and stored activations stay float32.
typedef short __attribute__((ext_vector_type(4))) bf16x4;
typedef float __attribute__((ext_vector_type(4))) f32x4;

__device__ __forceinline__ short to_bf16(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

constexpr int MM_NODES = 64;           // nodes per workgroup of k_lp_mm (16 per wavefront)
constexpr int MM_KP = 128 + 8;         // LDS row stride (bf16) of the staged operands
constexpr int WG_NODES = 256;          // nodes per workgroup of k_lp_wgrad

// Z[n][c] = bias[c] + sum_k A[n][k] B[c][k], A = [in1 | in2] (T x (I1 + I2)),
// B = [W1 | W2] (C x (I1 + I2)), or with trans the transpose of W1 (I1 x C:
// B[c][k] = W1[k][c], no second input).  C <= 64, I1 + I2 <= 128.  With hrelu
// also max(Z, 0); with accumulate Z is added to.  Wavefront w computes nodes
// 16w .. 16w + 15 of the workgroup's 64 against all 64 columns (four 16x16 tiles).
template <bool TRANS>
__global__ __launch_bounds__(256) void k_lp_mm(int T, int C, int I1, const float* in1, const float* W1, int I2,
                                               const float* in2, const float* W2, const float* bias, float* z,
                                               float* hrelu, int accumulate) {
    __shared__ short As[MM_NODES * MM_KP];
    __shared__ short Bs[64 * MM_KP];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int n0 = blockIdx.x * MM_NODES;
    const int K = I1 + I2, KP = (K + 15) & ~15;
    // both operands staged in batches of eight loads in flight: thread t
    // stages entries 32 (t & 3) .. + 31 of row t >> 2 of each (a loop with a
    // run-time trip count waited for each load in turn)
    const int sr = tid >> 2, kb = 32 * (tid & 3);
    if (kb < KP) {
        const int n = n0 + sr;
        const float* w1 = TRANS ? W1 + sr + (long long)kb * C : W1 + sr * I1 + kb;
        const float* w2 = TRANS ? W1 : W2 + sr * I2 + kb - I1;
        const float* a1 = in1 + (long long)n * I1 + kb;
        const float* a2 = in2 + (long long)n * I2 + kb - I1;
#pragma unroll 1
        for (int jb = 0; jb < 32; jb += 8) {
            float vb[8], va[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = kb + jb + j;
                vb[j] = 0.f;
                va[j] = 0.f;
                if (sr < C && k < K) vb[j] = TRANS ? w1[(long long)(jb + j) * C] : (k < I1 ? w1[jb + j] : w2[jb + j]);
                if (n < T && k < K) va[j] = k < I1 ? a1[jb + j] : a2[jb + j];
            }
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (kb + jb + j < KP) {
                    Bs[sr * MM_KP + kb + jb + j] = to_bf16(vb[j]);
                    As[sr * MM_KP + kb + jb + j] = to_bf16(va[j]);
                }
        }
    }
    __syncthreads();
    f32x4 acc[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) acc[ct] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const short* ar = As + (16 * wv + (lane & 15)) * MM_KP + 4 * (lane >> 4);
    for (int kk = 0; kk < KP; kk += 16) {
        const bf16x4 a = *(const bf16x4*)(ar + kk);
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            const bf16x4 bv = *(const bf16x4*)(Bs + (16 * ct + (lane & 15)) * MM_KP + 4 * (lane >> 4) + kk);
            acc[ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[ct], 0, 0, 0);
        }
    }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const int c = 16 * ct + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int n = n0 + 16 * wv + 4 * (lane >> 4) + v;
            if (n < T && c < C) {
                const long long o = (long long)n * C + c;
                float val = acc[ct][v] + (bias ? bias[c] : 0.f);
                if (accumulate) val += z[o];
                z[o] = val;
                if (hrelu) hrelu[o] = val > 0.f ? val : 0.f;
            }
        }
    }
}

// Partial weight gradients over WG_NODES nodes: part[blk][o I + i] =
// sum_n dy[n][o] x[n][i] on MFMA (the node index is the contraction), and
// part[blk][O I + o] = sum_n dy[n][o].  O, I <= 64.
__global__ __launch_bounds__(256) void k_lp_wgrad(int T, int O, int I, const float* dy, const float* x, float* part) {
    __shared__ short Ys[64 * (64 + 8)];   // [o][node]
    __shared__ short Xs[64 * (64 + 8)];   // [i][node]
    constexpr int S = 64 + 8;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nb = blockIdx.x * WG_NODES;
    f32x4 acc[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) acc[ct] = (f32x4){0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    for (int sub = 0; sub < WG_NODES; sub += 64) {
        __syncthreads();
        for (int e = tid; e < 64 * 64; e += 256) {
            const int k = e >> 6, c = e & 63, n = nb + sub + k;
            const bool in = n < T;
            Ys[c * S + k] = to_bf16(in && c < O ? dy[(long long)n * O + c] : 0.f);
            Xs[c * S + k] = to_bf16(in && c < I ? x[(long long)n * I + c] : 0.f);
        }
        if (tid < O)
            for (int k = 0; k < 64; ++k) {
                const int n = nb + sub + k;
                if (n < T) bsum += dy[(long long)n * O + tid];
            }
        __syncthreads();
        const short* ar = Ys + (16 * wv + (lane & 15)) * S + 4 * (lane >> 4);
#pragma unroll
        for (int kk = 0; kk < 64; kk += 16) {
            const bf16x4 a = *(const bf16x4*)(ar + kk);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const bf16x4 bv = *(const bf16x4*)(Xs + (16 * ct + (lane & 15)) * S + 4 * (lane >> 4) + kk);
                acc[ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[ct], 0, 0, 0);
            }
        }
    }
    float* out = part + (long long)blockIdx.x * (O * I + O);
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const int i = 16 * ct + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int o = 16 * wv + 4 * (lane >> 4) + v;
            if (o < O && i < I) out[o * I + i] = acc[ct][v];
        }
    }
    if (tid < O) out[O * I + tid] = bsum;
}

// out[t] = sum over blocks of part[blk][t], blocks in order (deterministic)
__global__ void k_lp_reduce_parts(int nblk, int n, const float* part, float* gW, float* gb, int nW) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float a0 = 0.f, a1 = 0.f;
    int b = 0;
    for (; b + 1 < nblk; b += 2) {
        a0 += part[(long long)b * n + t];
        a1 += part[(long long)(b + 1) * n + t];
    }
    if (b < nblk) a0 += part[(long long)b * n + t];
    const float s = a0 + a1;
    if (t < nW) gW[t] = s;
    else if (gb) gb[t - nW] = s;
}

// m_i = sum_j s_ij h_j / wsum_i over the band neighbours j of i (same system)
__global__ void k_lp_agg(int T, int W, const double* s, const double* wsum, const int* node_sys,
                         const int* sys_begin, const float* hin, float* m) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T * LP_HID) return;
    const int i = t / LP_HID, k = t - i * LP_HID;
    const int lo = sys_lo(sys_begin, node_sys, i), hi = sys_hi(sys_begin, node_sys, i);
    // every load is issued (outside the system's range the indices fall back to
    // node i and the term is not added): a load under a condition waits for its
    // data before the next one starts
    double a = 0.0;
#pragma unroll 4
    for (int d = 1; d <= W; ++d) {
        const bool dn = i - d >= lo, up = i + d < hi;
        const double sl = s[(long long)i * (W + 1) + d];
        const double su = s[(long long)(up ? i + d : i) * (W + 1) + d];
        const float hl = hin[(long long)(dn ? i - d : i) * LP_HID + k];
        const float hu = hin[(long long)(up ? i + d : i) * LP_HID + k];
        a = dn ? a + sl * hl : a;
        a = up ? a + su * hu : a;
    }
    m[t] = (float)(a / wsum[i]);
}

// transpose of k_lp_agg, added: dh_j += sum_i s_ij / wsum_i dm_i
__global__ void k_lp_aggT(int T, int W, const double* s, const double* wsum, const int* node_sys,
                          const int* sys_begin, const float* dm, float* dh) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T * LP_HID) return;
    const int j = t / LP_HID, k = t - j * LP_HID;
    const int lo = sys_lo(sys_begin, node_sys, j), hi = sys_hi(sys_begin, node_sys, j);
    double a = 0.0;
    for (int d = 1; d <= W; ++d) {
        if (j + d < hi) a += s[(long long)(j + d) * (W + 1) + d] * dm[(long long)(j + d) * LP_HID + k] / wsum[j + d];
        if (j - d >= lo) a += s[(long long)j * (W + 1) + d] * dm[(long long)(j - d) * LP_HID + k] / wsum[j - d];
    }
    dh[t] += (float)a;
}

__global__ void k_lp_reluback(int n, const float* dh, const float* z, float* dz) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dz[t] = z[t] > 0.f ? dh[t] : 0.f;
}

// ---- preconditioner blocks and the loss -------------------------------------
__device__ __forceinline__ void tril_index(int k, int& r, int& c) {
    r = 0;
    while ((r + 1) * (r + 2) / 2 <= k) ++r;
    c = k - r * (r + 1) / 2;
}

// G_ii = (I + tril(out[0:21])) C_i, G_{i,i-1} = out[21:57] C_{i-1}; out == nullptr: block Jacobi.
// Six threads per node, one row r of both blocks each; row r of the unit lower
// factor held in registers (compile-time indices), each entry summed over k
// from its column up, as G_ii(r, c) = sum_{k=c..r} A(r, k) C_i(k, c).
__global__ void k_lp_buildG(int T, const float* out, const double* C, const int* node_sys, const int* sys_begin,
                            double* Gd, double* Go) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / 6, r = t - 6 * (t / 6);
    if (i >= T) return;
    const int lo = sys_lo(sys_begin, node_sys, i);
    const float* oi = out ? out + (long long)i * LP_NOUT : nullptr;
    double A[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        A[k] = k == r ? 1.0 : 0.0;
        if (oi && k <= r) A[k] += (double)oi[r * (r + 1) / 2 + k];
    }
    const double* Ci = C + (long long)i * 36;
    double* gd = Gd + (long long)i * 36 + 6 * r;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        double a = 0.0;
#pragma unroll
        for (int k = c; k < 6; ++k)
            if (k <= r) a += A[k] * Ci[k * 6 + c];
        gd[c] = a;
    }
    double* go = Go + (long long)i * 36 + 6 * r;
    if (oi && i > lo) {
        const double* Cj = C + (long long)(i - 1) * 36;
        double B[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) B[k] = (double)oi[21 + r * 6 + k];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double a = 0.0;
#pragma unroll
            for (int k = c; k < 6; ++k) a += B[k] * Cj[k * 6 + c];
            go[c] = a;
        }
    } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) go[c] = 0.0;
    }
}

__device__ __forceinline__ void band_row(int i, int lo, int hi, int W, const double* Hb, const double* x,
                                         double* y) {
    for (int a = 0; a < 6; ++a) y[a] = 0.0;
    for (int d = 0; d <= W; ++d) {
        if (i - d >= lo) {
            const double* B = Hb + ((long long)i * (W + 1) + d) * 36;
            const double* xv = x + (long long)(i - d) * 6;
            for (int a = 0; a < 6; ++a)
                for (int b = 0; b < 6; ++b) y[a] += B[a * 6 + b] * xv[b];
        }
        if (d > 0 && i + d < hi) {
            const double* B = Hb + ((long long)(i + d) * (W + 1) + d) * 36;
            const double* xv = x + (long long)(i + d) * 6;
            for (int a = 0; a < 6; ++a)
                for (int b = 0; b < 6; ++b) y[a] += B[b * 6 + a] * xv[b];
        }
    }
}

// per node: t_i and sum_k log |(A_i)_kk|
__device__ __forceinline__ void lp_AB(const float* out, int i, bool has_b, double* A, double* B) {
    for (int k = 0; k < 36; ++k) A[k] = (k % 7 == 0) ? 1.0 : 0.0;
    for (int k = 0; k < 21; ++k) {
        int r, c;
        tril_index(k, r, c);
        A[r * 6 + c] += (double)out[(long long)i * LP_NOUT + k];
    }
    for (int k = 0; k < 36; ++k) B[k] = has_b ? (double)out[(long long)i * LP_NOUT + 21 + k] : 0.0;
}

__global__ void k_lp_kap_node(int T, const float* out, const double* N1, const int* node_sys, const int* sys_begin,
                              double* tnode, double* ldnode) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    double A[36], B[36];
    lp_AB(out, i, i > sys_lo(sys_begin, node_sys, i), A, B);
    const double* N = N1 + (long long)i * 36;
    double t = 0.0, ld = 0.0;
    for (int k = 0; k < 36; ++k) t += A[k] * A[k] + B[k] * B[k];
    // 2 tr(A N B^T) = 2 sum_rc (A N)_rc B_rc
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
            double an = 0.0;
            for (int k = 0; k <= r; ++k) an += A[r * 6 + k] * N[k * 6 + c];
            t += 2.0 * an * B[r * 6 + c];
        }
    for (int k = 0; k < 6; ++k) ld += log(fabs(A[k * 7]));
    tnode[i] = t;
    ldnode[i] = ld;
}

// per system (one workgroup): Tsys = sum t_i, loss = log(Tsys / 6n) - 2 sum ld_i / 6n
__global__ __launch_bounds__(256) void k_lp_kap_sys(const int* sys_begin, const double* tnode, const double* ldnode,
                                                    double* Tsys, double* loss) {
    __shared__ double r0[256], r1[256];
    const int sidx = blockIdx.x, tid = threadIdx.x;
    const int lo = sys_begin[sidx], hi = sys_begin[sidx + 1];
    double a = 0.0, b = 0.0;
    for (int i = lo + tid; i < hi; i += 256) {
        a += tnode[i];
        b += ldnode[i];
    }
    r0[tid] = a;
    r1[tid] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            r0[tid] += r0[tid + w];
            r1[tid] += r1[tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double n6 = 6.0 * (hi - lo);
        Tsys[sidx] = r0[0];
        loss[sidx] = log(r0[0] / n6) - 2.0 * r1[0] / n6;
    }
}

// dL/dout, mean over the S systems: dA = (2 A + 2 B N^T) / T_s - (2 / 6n) diag(1 / a_kk)
// (lower triangle), dB = (2 A N + 2 B) / T_s
__global__ void k_lp_kap_dout(int T, int S, const float* out, const double* N1, const double* Tsys,
                              const int* node_sys, const int* sys_begin, float* dout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    const int sidx = node_sys[i];
    const int lo = sys_begin[sidx], hi = sys_begin[sidx + 1];
    const bool has_b = i > lo;
    double A[36], B[36];
    lp_AB(out, i, has_b, A, B);
    const double* N = N1 + (long long)i * 36;
    const double iT = 1.0 / Tsys[sidx], n6 = 6.0 * (hi - lo), sc = 1.0 / S;
    for (int k = 0; k < 21; ++k) {
        int r, c;
        tril_index(k, r, c);
        double bn = 0.0;   // (B N^T)_rc
        for (int q = 0; q < 6; ++q) bn += B[r * 6 + q] * N[c * 6 + q];
        double d = (2.0 * A[r * 6 + c] + 2.0 * bn) * iT;
        if (r == c) d -= 2.0 / n6 / A[r * 7];
        dout[(long long)i * LP_NOUT + k] = (float)(d * sc);
    }
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
            double d = 0.0;
            if (has_b) {
                double an = 0.0;
                for (int k = 0; k <= r; ++k) an += A[r * 6 + k] * N[k * 6 + c];
                d = (2.0 * an + 2.0 * B[r * 6 + c]) * iT;
            }
            dout[(long long)i * LP_NOUT + 21 + r * 6 + c] = (float)(d * sc);
        }
}

__global__ void k_lp_adam(int n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
                          float eps, float c1, float c2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gg = g[t];
    const float mm = b1 * m[t] + (1.f - b1) * gg;
    const float vv = b2 * v[t] + (1.f - b2) * gg * gg;
    m[t] = mm;
    v[t] = vv;
    p[t] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// ---- PCG: one workgroup per system ------------------------------------------
// z = G^T G r with the blocks Gd (Go == nullptr: block diagonal); x starts at 0.
__global__ __launch_bounds__(256) void k_lp_pcg(const int* sys_begin, int W, const double* Hb, const double* Gd,
                                                const double* Go, const double* g, double* x, double* r, double* z,
                                                double* p, double* q, double* tmp, double tol, int maxit,
                                                int* iters, double* relres, Control* ctl, const int* bad) {
    __shared__ double red[256];
    const int sidx = blockIdx.x, tid = threadIdx.x;
    const int lo = sys_begin[sidx], hi = sys_begin[sidx + 1];
    auto sum = [&](double v) {
        red[tid] = v;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) red[tid] += red[tid + w];
            __syncthreads();
        }
        const double s = red[0];
        __syncthreads();
        return s;
    };
    auto prec = [&]() {
        // tmp = G r, then z = G^T tmp
        for (int i = lo + tid; i < hi; i += 256)
            for (int a = 0; a < 6; ++a) {
                double s = 0.0;
                for (int b = 0; b < 6; ++b) s += Gd[(long long)i * 36 + a * 6 + b] * r[(long long)i * 6 + b];
                if (Go && i > lo)
                    for (int b = 0; b < 6; ++b) s += Go[(long long)i * 36 + a * 6 + b] * r[(long long)(i - 1) * 6 + b];
                tmp[(long long)i * 6 + a] = s;
            }
        __syncthreads();
        for (int i = lo + tid; i < hi; i += 256)
            for (int a = 0; a < 6; ++a) {
                double s = 0.0;
                for (int b = 0; b < 6; ++b) s += Gd[(long long)i * 36 + b * 6 + a] * tmp[(long long)i * 6 + b];
                if (Go && i + 1 < hi)
                    for (int b = 0; b < 6; ++b)
                        s += Go[(long long)(i + 1) * 36 + b * 6 + a] * tmp[(long long)(i + 1) * 6 + b];
                z[(long long)i * 6 + a] = s;
            }
        __syncthreads();
    };
    double gg = 0.0;
    for (long long e = 6LL * lo + tid; e < 6LL * hi; e += 256) {
        x[e] = 0.0;
        r[e] = g[e];
        gg += g[e] * g[e];
    }
    __syncthreads();
    const double g0 = sqrt(sum(gg));
    prec();
    double rz = 0.0;
    for (long long e = 6LL * lo + tid; e < 6LL * hi; e += 256) {
        p[e] = z[e];
        rz += r[e] * z[e];
    }
    rz = sum(rz);
    int it = 0;
    double rn = g0;
    bool broke = bad && *bad;   // a diagonal block that is not positive definite: no CG
    while (!broke && it < maxit && rn > tol * g0 && g0 > 0.0) {
        double pq = 0.0;
        for (int i = lo + tid; i < hi; i += 256) {
            double yy[6];
            band_row(i, lo, hi, W, Hb, p, yy);
            for (int a = 0; a < 6; ++a) {
                q[(long long)i * 6 + a] = yy[a];
                pq += p[(long long)i * 6 + a] * yy[a];
            }
        }
        pq = sum(pq);
        if (!(pq > 0.0) || !(rz > 0.0)) {   // not positive definite (or not finite)
            broke = true;
            break;
        }
        const double al = rz / pq;
        double rr = 0.0;
        for (long long e = 6LL * lo + tid; e < 6LL * hi; e += 256) {
            x[e] += al * p[e];
            r[e] -= al * q[e];
            rr += r[e] * r[e];
        }
        rn = sqrt(sum(rr));
        ++it;
        if (!(rn > tol * g0)) break;
        prec();
        double rz2 = 0.0;
        for (long long e = 6LL * lo + tid; e < 6LL * hi; e += 256) rz2 += r[e] * z[e];
        rz2 = sum(rz2);
        const double be = rz2 / rz;
        rz = rz2;
        for (long long e = 6LL * lo + tid; e < 6LL * hi; e += 256) p[e] = z[e] + be * p[e];
        __syncthreads();
    }
    if (tid == 0) {
        const bool ok = !broke && (rn <= tol * g0 || g0 == 0.0);
        if (iters) iters[sidx] = it;
        if (relres) relres[sidx] = broke ? -1.0 : (g0 > 0.0 ? rn / g0 : 0.0);
        if (ctl) {
            ctl[sidx].pcg_iters = it;
            ctl[sidx].pcg_fail = ok ? 0 : 1;
        }
    }
}

// Hb[i][d] = H_{i,i-d} from the lower triangle of a dense row-major Pp x Pp H
// (the diagonal block symmetrised); zero where i - d < 0
__global__ void k_lp_band_from_dense(int N, int W, const double* H, int Pp, double* Hb) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)N * (W + 1) * 36) return;
    const int e = (int)(t % 36);
    const long long bd = t / 36;
    const int d = (int)(bd % (W + 1)), i = (int)(bd / (W + 1));
    const int r = e / 6, c = e - r * 6;
    double v = 0.0;
    if (i - d >= 0) {
        int row = 6 * i + r, col = 6 * (i - d) + c;
        if (col > row) {
            const int tmp = row;
            row = col;
            col = tmp;
        }
        v = H[(long long)row * Pp + col];
    }
    Hb[t] = v;
}

// ---- PCG on one large system: every workgroup of a cooperative launch --------
// Work item e = 6 i + a (node i, row a) in grid-stride loops; global sums are
// block partials in part[slot][block], summed by every block in the same order
// after a grid barrier, so all blocks take the same decisions.  The pair sum
// uses slots 0 and 1, the p.q sum slots 2 and 3: a block writes a slot again
// only after passing the next barrier, which every reader of it has passed.
__global__ __launch_bounds__(256) void k_lp_pcg_coop(int T, int W, const double* Hb, const double* Gd,
                                                     const double* Go, const double* g, double* x, double* r,
                                                     double* z, double* p, double* q, double* tmp, double* part,
                                                     double tol, int maxit, Control* ctl, const int* bad,
                                                     const int* far_beg, const int* far_col, const double* far_val) {
    // Two grid barriers per iteration.  After p.q gives alpha, every thread
    // forms the new residual of the block rows it needs (i - 1 .. i + 1) from
    // the old one and q, and with it z = G^T G r locally: no barrier inside the
    // preconditioner.  One sum gives r.r and r.z.  Then p = z + beta p and
    // q = H z + beta q (= H p): z was complete before that sum's barrier, so
    // the next product needs none of its own.  r is double-buffered (the other
    // copy lives in tmp) because neighbours read the old residual while its
    // owner writes the new one.
    namespace cgs = cooperative_groups;
    cgs::grid_group grid = cgs::this_grid();
    __shared__ double red[2][4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nblk = gridDim.x;
    const int e0 = blockIdx.x * 256 + tid, es = nblk * 256, E = 6 * T;
    // two block sums, the same in every thread
    auto block_sum2 = [&](double& a, double& b) {
        a = wave_sum(a);
        b = wave_sum(b);
        if (lane == 0) {
            red[0][wv] = a;
            red[1][wv] = b;
        }
        __syncthreads();
        a = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        b = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        __syncthreads();
    };
    // grid sums: block partials to slots s, s + 1 of part, one grid barrier,
    // then every block adds all partials in the same order (so all blocks
    // hold the same alpha and beta)
    auto grid_sum2 = [&](double& a, double& b, int slot) {
        block_sum2(a, b);
        if (tid == 0) {
            part[slot * nblk + blockIdx.x] = a;
            part[(slot + 1) * nblk + blockIdx.x] = b;
        }
        grid.sync();
        a = 0.0;
        b = 0.0;
        for (int k = tid; k < nblk; k += 256) {
            a += part[slot * nblk + k];
            b += part[(slot + 1) * nblk + k];
        }
        block_sum2(a, b);
    };
    // y_i[a] = (H v)_i[a] over the band, four block distances at a time with
    // every load issued before the products (a load under a condition waits
    // for its data before the next one starts): indices past the band are
    // clamped into it and their terms weighted by zero
    auto hrow = [&](int e, const double* v) {
        const int i = e / 6, a = e - 6 * (e / 6);
        double y = 0.0;
        for (int d0 = 0; d0 <= W; d0 += 4) {
            double bl[4][6], vl[4][6], bu[4][6], vu[4][6], wl[4], wu[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int d = d0 + k;
                const bool lo = d <= W && i - d >= 0, up = d > 0 && d <= W && i + d < T;
                wl[k] = lo ? 1.0 : 0.0;
                wu[k] = up ? 1.0 : 0.0;
                const int dl = lo ? d : 0, du = up ? d : 0;
                const double* Bl = Hb + ((long long)i * (W + 1) + dl) * 36 + a * 6;
                const double* Vl = v + 6LL * (i - dl);
                const double* Bu = Hb + ((long long)(i + du) * (W + 1) + du) * 36 + a;
                const double* Vu = v + 6LL * (i + du);
#pragma unroll
                for (int b = 0; b < 6; ++b) {
                    bl[k][b] = Bl[b];
                    vl[k][b] = Vl[b];
                    bu[k][b] = Bu[6 * b];
                    vu[k][b] = Vu[b];
                }
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                double sl = 0.0, su = 0.0;
#pragma unroll
                for (int b = 0; b < 6; ++b) {
                    sl += bl[k][b] * vl[k][b];
                    su += bu[k][b] * vu[k][b];
                }
                y += wl[k] * sl + wu[k] * su;
            }
        }
        if (far_beg) {
            // blocks outside the band (both triangles listed)
            for (int k = far_beg[i]; k < far_beg[i + 1]; ++k) {
                const double* B = far_val + 36LL * k + 6 * a;
                const double* V = v + 6LL * far_col[k];
                double s = 0.0;
#pragma unroll
                for (int b = 0; b < 6; ++b) s += B[b] * V[b];
                y += s;
            }
        }
        return y;
    };
    // block row i of G v, v given element-wise by vv(j, b)
    auto gblock = [&](int i, auto&& vv, double* t) {
        double vi[6], vm[6];
        const bool off = Go != nullptr && i > 0;
        const int im = off ? i - 1 : i;
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            vi[b] = vv(i, b);
            vm[b] = vv(im, b);
        }
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            double s = 0.0;
#pragma unroll
            for (int b = 0; b < 6; ++b) s += Gd[(long long)i * 36 + a * 6 + b] * vi[b];
            if (off) {
#pragma unroll
                for (int b = 0; b < 6; ++b) s += Go[(long long)i * 36 + a * 6 + b] * vm[b];
            }
            t[a] = s;
        }
    };
    // element (i, a) of G^T G v (the second block row's loads are issued with
    // the first: at the last node it repeats row i, weighted by zero)
    auto mprec = [&](int i, int a, auto&& vv) {
        const bool nx = Go != nullptr && i + 1 < T;
        const int in = nx ? i + 1 : i;
        double t[6], tn[6];
        gblock(i, vv, t);
        gblock(in, vv, tn);
        double s = 0.0, sn = 0.0;
#pragma unroll
        for (int b = 0; b < 6; ++b) s += Gd[(long long)i * 36 + b * 6 + a] * t[b];
        if (Go != nullptr) {
#pragma unroll
            for (int b = 0; b < 6; ++b) sn += Go[(long long)in * 36 + b * 6 + a] * tn[b];
        }
        return nx ? s + sn : s;
    };
    double* rA = r;     // residual read this iteration
    double* rB = tmp;   // residual written this iteration
    // r = g, z = M g, p = z
    double gg = 0.0, rz = 0.0;
    for (int e = e0; e < E; e += es) {
        const int i = e / 6, a = e - 6 * i;
        const double zv = mprec(i, a, [&](int j, int b) { return g[6 * j + b]; });
        const double gv = g[e];
        x[e] = 0.0;
        rA[e] = gv;
        z[e] = zv;
        p[e] = zv;
        gg += gv * gv;
        rz += gv * zv;
    }
    grid_sum2(gg, rz, 0);
    const double g0 = sqrt(gg);
    bool broke = bad && *bad;
    int it = 0;
    double rn = g0;
    if (!broke && g0 > 0.0) {
        // q = H p with p = z, complete after the barrier
        double pq = 0.0, unused = 0.0;
        for (int e = e0; e < E; e += es) {
            const double y = hrow(e, z);
            q[e] = y;
            pq += z[e] * y;
        }
        grid_sum2(pq, unused, 2);
        while (it < maxit) {
            if (!(pq > 0.0) || !(rz > 0.0)) {
                broke = true;
                break;
            }
            const double al = rz / pq;
            const double* ro = rA;
            auto rnew = [&](int j, int b) { return fma(-al, q[6 * j + b], ro[6 * j + b]); };
            double rr = 0.0, rz2 = 0.0;
            for (int e = e0; e < E; e += es) {
                const int i = e / 6, a = e - 6 * i;
                const double rv = rnew(i, a);
                const double zv = mprec(i, a, rnew);
                x[e] += al * p[e];
                rB[e] = rv;
                z[e] = zv;
                rr += rv * rv;
                rz2 += rv * zv;
            }
            grid_sum2(rr, rz2, 0);
            ++it;
            rn = sqrt(rr);
            double* t = rA;
            rA = rB;
            rB = t;
            if (!(rn > tol * g0)) break;
            const double be = rz2 / rz;
            rz = rz2;
            pq = 0.0;
            unused = 0.0;
            for (int e = e0; e < E; e += es) {
                const double pv = z[e] + be * p[e];
                const double qv = hrow(e, z) + be * q[e];
                p[e] = pv;
                q[e] = qv;
                pq += pv * qv;
            }
            grid_sum2(pq, unused, 2);
        }
    }
    if (blockIdx.x == 0 && tid == 0) {
        const bool ok = !broke && (rn <= tol * g0 || g0 == 0.0);
        ctl->pcg_iters = it;
        ctl->pcg_fail = ok ? 0 : 1;
    }
}

__global__ void k_lp_zero_flag(int* p) { *p = 0; }

// ---- launchers ------------------------------------------------------------------
namespace {
inline unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }
}

void lp_launch_features(const LpSet& s, double* C, double* sfeat, float* X, double* N1, double* wsum, int* bad,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_lp_chol, dim3(nblk(s.T)), dim3(256), 0, st, s.T, s.W, s.Hb, C, bad);
    if (s.W > 0)
        hipLaunchKernelGGL(k_lp_feat1, dim3(nblk((long long)s.T * s.W)), dim3(256), 0, st, s.T, s.W, s.Hb, C,
                           s.node_sys, s.sys_begin, sfeat, X, N1);
    hipLaunchKernelGGL(k_lp_feat2, dim3(nblk(s.T)), dim3(256), 0, st, s.T, s.W, sfeat, s.node_sys, s.sys_begin, X,
                       wsum);
}

void lp_launch_linear(int T, int O, int I1, const float* in1, const float* W1, int I2, const float* in2,
                      const float* W2, const float* b, float* z, float* h, hipStream_t st) {
    hipLaunchKernelGGL(k_lp_mm<false>, dim3(nblk(T, MM_NODES)), dim3(256), 0, st, T, O, I1, in1, W1, I2, in2, W2, b,
                       z, h, 0);
}

void lp_launch_linearT(int T, int O, int I, const float* dy, const float* Wt, float* dx, int acc, hipStream_t st) {
    // dx (T x I) = dy (T x O) W (O x I): B[c][k] = W[k][c]
    hipLaunchKernelGGL(k_lp_mm<true>, dim3(nblk(T, MM_NODES)), dim3(256), 0, st, T, I, O, dy, Wt, 0,
                       (const float*)nullptr, (const float*)nullptr, (const float*)nullptr, dx, (float*)nullptr, acc);
}

int lp_wgrad_blocks(int T) { return (T + WG_NODES - 1) / WG_NODES; }

void lp_launch_wgrad(int T, int O, int I, const float* dy, const float* x, float* gW, float* gb, float* part,
                     hipStream_t st) {
    const int nb = lp_wgrad_blocks(T);
    hipLaunchKernelGGL(k_lp_wgrad, dim3(nb), dim3(256), 0, st, T, O, I, dy, x, part);
    const int n = O * I + O;
    hipLaunchKernelGGL(k_lp_reduce_parts, dim3(nblk(n)), dim3(256), 0, st, nb, n, part, gW, gb, O * I);
}

void lp_launch_agg(const LpSet& s, const double* sfeat, const double* wsum, const float* hin, float* m, int transpose,
                   hipStream_t st) {
    if (transpose)
        hipLaunchKernelGGL(k_lp_aggT, dim3(nblk((long long)s.T * LP_HID)), dim3(256), 0, st, s.T, s.W, sfeat, wsum,
                           s.node_sys, s.sys_begin, hin, m);
    else
        hipLaunchKernelGGL(k_lp_agg, dim3(nblk((long long)s.T * LP_HID)), dim3(256), 0, st, s.T, s.W, sfeat, wsum,
                           s.node_sys, s.sys_begin, hin, m);
}

void lp_launch_reluback(int n, const float* dh, const float* z, float* dz, hipStream_t st) {
    hipLaunchKernelGGL(k_lp_reluback, dim3(nblk(n)), dim3(256), 0, st, n, dh, z, dz);
}

void lp_launch_buildG(const LpSet& s, const float* out, const double* C, double* Gd, double* Go, hipStream_t st) {
    hipLaunchKernelGGL(k_lp_buildG, dim3(nblk(6 * s.T)), dim3(256), 0, st, s.T, out, C, s.node_sys, s.sys_begin, Gd, Go);
}

void lp_launch_kap(const LpSet& s, const float* out, const double* N1, double* tnode, double* ldnode, double* Tsys,
                   double* loss, hipStream_t st) {
    hipLaunchKernelGGL(k_lp_kap_node, dim3(nblk(s.T)), dim3(256), 0, st, s.T, out, N1, s.node_sys, s.sys_begin, tnode,
                       ldnode);
    hipLaunchKernelGGL(k_lp_kap_sys, dim3(s.S), dim3(256), 0, st, s.sys_begin, tnode, ldnode, Tsys, loss);
}

void lp_launch_kap_dout(const LpSet& s, const float* out, const double* N1, const double* Tsys, float* dout,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_lp_kap_dout, dim3(nblk(s.T)), dim3(256), 0, st, s.T, s.S, out, N1, Tsys, s.node_sys,
                       s.sys_begin, dout);
}

void lp_launch_adam(int n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                    int step, hipStream_t st) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_lp_adam, dim3(nblk(n)), dim3(256), 0, st, n, p, g, m, v, lr, b1, b2, eps, c1, c2);
}

void lp_launch_pcg(const LpSet& s, const double* Gd, const double* Go, double* x, double* r, double* z, double* p,
                   double* q, double* tmp, double tol, int maxit, int* iters, double* relres, hipStream_t st,
                   Control* ctl, const int* bad) {
    hipLaunchKernelGGL(k_lp_pcg, dim3(s.S), dim3(256), 0, st, s.sys_begin, s.W, s.Hb, Gd, Go, s.g, x, r, z, p, q, tmp,
                       tol, maxit, iters, relres, ctl, bad);
}

int lp_coop_max_blocks() {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_lp_pcg_coop, 256, 0) != hipSuccess) return 0;
    return per_cu * cus;
}

int lp_launch_pcg_coop(int T, int W, const double* Hb, const double* Gd, const double* Go, const double* g, double* x,
                       double* r, double* z, double* p, double* q, double* tmp, double* part, int blocks, double tol,
                       int maxit, Control* ctl, const int* bad, const int* far_beg, const int* far_col,
                       const double* far_val, hipStream_t st) {
    void* args[] = {&T, &W, &Hb, &Gd, &Go, &g, &x, &r, &z, &p, &q, &tmp, &part, &tol, &maxit, &ctl, &bad,
                    &far_beg, &far_col, &far_val};
    return (int)hipLaunchCooperativeKernel((const void*)k_lp_pcg_coop, dim3(blocks), dim3(256), args, 0, st);
}

// far_val[k] = H_{i,j} (i = far_row[k], j = far_col[k]) from the lower triangle
__global__ void k_lp_far_from_dense(int nfar, const int* far_row, const int* far_col, const double* H, int Pp,
                                    double* far_val) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nfar * 36) return;
    const int k = (int)(t / 36), e = (int)(t % 36);
    int row = 6 * far_row[k] + e / 6, col = 6 * far_col[k] + e % 6;
    if (col > row) {
        const int tmp = row;
        row = col;
        col = tmp;
    }
    far_val[t] = H[(long long)row * Pp + col];
}

void lp_launch_far_from_dense(int nfar, const int* far_row, const int* far_col, const double* H, int Pp,
                              double* far_val, hipStream_t st) {
    hipLaunchKernelGGL(k_lp_far_from_dense, dim3(nblk((long long)nfar * 36)), dim3(256), 0, st, nfar, far_row,
                       far_col, H, Pp, far_val);
}

void lp_launch_zero_flag(int* p, hipStream_t st) { hipLaunchKernelGGL(k_lp_zero_flag, dim3(1), dim3(1), 0, st, p); }

void lp_launch_band_from_dense(int N, int W, const double* H, int Pp, double* Hb, hipStream_t st) {
    hipLaunchKernelGGL(k_lp_band_from_dense, dim3(nblk((long long)N * (W + 1) * 36)), dim3(256), 0, st, N, W, H, Pp,
                       Hb);
}

}  // namespace spgo
