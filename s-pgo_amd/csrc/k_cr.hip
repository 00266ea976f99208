// Block cyclic reduction of the banded reduced pose system: the factorisation
// and both solves of H delta = g (the reference's dense torch Cholesky and
// cholesky_solve, linear_solvers.cu:195-218), for profiles at most four
// 16-row blocks deep.
//
// A profile D blocks of 16 deep makes H block tridiagonal in super-rows of
// S = 16 D rows: super-row I couples only to I - 1 and I + 1.  Level l
// eliminates, all at once, the super-rows i = 2^l (mod 2^(l+1)) from the
// system of the super-rows that are multiples of 2^l.  With the diagonal block
// D_i = L L^T, W = L^-1, its neighbours p = i - 2^l and q = i + 2^l, E_i the
// coupling of i to p and E_q that of q to i:
//     U_i = W E_i,  V_i = W E_q^T,  y_i = W b_i
//     D_p -= U_i^T U_i,  D_q -= V_i^T V_i,  b_p -= U_i^T y_i,  b_q -= V_i^T y_i
//     new coupling of q to p:  -V_i^T U_i
// This is the Cholesky factorisation of H with its super-rows reordered
// (odd-even), so it needs H positive definite and nothing more; every pivot is
// checked.  After the last level super-row 0 is alone; the solution then comes
// back level by level, x_i = W^T (y_i - U_i x_p - V_i x_q).
//
// The sequential part is one S x S factorisation per level (log2 of the
// number of super-rows levels) instead of a column chain over all of H.
// Per level two launches: k_cr_factor (one workgroup per eliminated
// super-row: D_i = L L^T, W = L^-1, and U, V, y as 16 x 16 FP64 MFMA tiles)
// and k_cr_update (the survivors' new D, coupling and b as MFMA tiles).  Once
// at most three super-rows are left, k_cr_tail solves them in one workgroup as
// a block-tridiagonal Cholesky, and one k_cr_back launch brings the solution
// back down through the levels (workgroups handing solutions over by flags).  Level 0 reads H and g in place (the lower triangle);
// later levels read the workspace.  Rows past the padded size are identity.
//
// k_cr_factor factors D right-looking over 16-wide block columns.  The
// diagonal block is factored by one wavefront, one row per lane (the four
// 16-lane rows of the wavefront repeat it): the column entries reach the
// other lanes by DPP row broadcasts (row_newbcast), which ride on a register
// move, so a column costs no LDS round trip or lane read.  The same loop
// inverts the unit lower factor (forward substitution of the identity, rows
// broadcast the same way); W_kk = diag(1/L_jj) (unit L)^-1.  The panel blocks
// below are A_bk W_kk^T and the trailing update L_bk L_ck^T, both on MFMA;
// W's blocks below its diagonal, W_bc = -W_bb sum_m L_bm W_mc, are formed on
// MFMA by the other wavefronts while the next diagonal block is factored.
#include <utility>

#include "spgo_device.h"
#include "spgo_kernels.h"
#include "spgo_pose.h"

namespace spgo {

namespace {
typedef double __attribute__((ext_vector_type(4))) d4;

// lane J of each 16-lane row, to every lane of that row
// (row_newbcast writes every lane, so the value the instruction would keep
// in unwritten lanes is v itself and needs no register of its own)
template <int J>
__device__ __forceinline__ double row_bcast(double v) {
    return __builtin_amdgcn_update_dpp(v, v, 0x150 + J, 0xf, 0xf, true);
}

// The diagonal-block factorisation is a fixed instruction schedule, one
// inline-asm block per column (ALU only), generated at compile time.  A
// column's pivot chain (broadcast, rsqrt, one Newton-Halley step, scaling) has
// the previous column's updates of the inverse interleaved with it, and the
// update that feeds the next pivot is issued first.  The hazards the compiler
// cannot see inside the block are handled in it: a DPP read needs two wait
// states after a VALU write of its source (each block starts with s_nop 1, as
// the compiler may place a register copy right before it, and inside a block
// no DPP source is written fewer than two instructions earlier), and a VALU
// read of a transcendental (v_rsq_f64) result needs one.
struct AsmText {
    char buf[6144] = {};
    unsigned long n = 0;
    constexpr void put(const char* p) {
        while (*p) buf[n++] = *p++;
    }
    constexpr void num(unsigned v) {
        char t[12] = {};
        int k = 0;
        do {
            t[k++] = char('0' + v % 10);
            v /= 10;
        } while (v);
        while (k) buf[n++] = t[--k];
    }
    constexpr void hex(unsigned v) {
        put("0x");
        for (int sh = 28; sh >= 0; sh -= 4) buf[n++] = "0123456789abcdef"[(v >> sh) & 15];
    }
    constexpr void op(int i) {
        buf[n++] = '%';
        num((unsigned)i);
    }
    constexpr const char* data() const { return buf; }
    constexpr unsigned long size() const { return n; }
};

// Operands of a column block: 0-15 the row (av), 16-19 this 16-lane row's four
// columns of the inverse (yv), 20 m, 21 d, 22 y0, 23 t, 24 e, 25 inv, 26 mi,
// 27 ivr; inputs 28 m of the previous column, 29 the constant 0.375, 30 the
// lane's factor (r > J), 31 the lane's factor (r == J).
enum { OP_Y = 16, OP_M = 20, OP_D, OP_Y0, OP_T, OP_E, OP_INV, OP_MI, OP_IVR, OP_MPREV, OP_C375, OP_GT, OP_EQ };

constexpr void asm_op3(AsmText& a, const char* ins, int o0, const char* s1, int o1, int o2) {
    a.put(ins);
    a.op(o0);
    a.put(s1);
    a.op(o1);
    a.put(", ");
    a.op(o2);
}
// The work column J defers into its chain's gaps: the updates of column J - 1
// that no pivot before column J + 1 needs, item k of the list
//   k < 4:   yv[k] -= (yv[k] of lane J - 1) m_(J-1)        (the inverse)
//   k >= 4:  av[q] -= L_q,J-1 L_r,J-1, q = J + k - 3        (the trailing update)
constexpr int column_pool(int J) { return J >= 1 ? 4 + (15 - J) : 0; }
constexpr void column_pool_item(AsmText& a, int J, int k) {
    if (k < 4) {
        asm_op3(a, "v_fmac_f64_dpp ", OP_Y + k, ", -", OP_Y + k, OP_MPREV);
        a.put(" row_newbcast:");
        a.num((unsigned)(J - 1));
    } else {
        const int q = J + k - 3;
        asm_op3(a, "v_fmac_f64_dpp ", q, ", -", J - 1, J - 1);
        a.put(" row_newbcast:");
        a.num((unsigned)q);
    }
    a.put(" row_mask:0xf bank_mask:0xf\n");
}
// pool items [k0, k1) of column J
constexpr void column_pool_range(AsmText& a, int J, int k0, int k1) {
    const int n = column_pool(J);
    for (int k = k0; k < k1 && k < n; ++k) column_pool_item(a, J, k);
}

// Column J: pivot d = a_JJ (lane J), y0 = rsqrt(d), one Halley step to inv,
// L_rJ = a_rJ inv (into av[J]), m = L_rJ inv on lanes r > J (0 elsewhere),
// ivr += inv on lane J, and the update of av[J + 1] (the next pivot's entry);
// column J - 1's other updates fill the gaps of this chain.  VALU only: a
// VALU result read by a scalar instruction (a pivot test) or a scalar write of
// EXEC stalls the wavefront's issue.
template <int J>
constexpr AsmText column_asm() {
    AsmText a;
    a.put("s_nop 1\nv_mov_b64_dpp ");
    a.op(OP_D);
    a.put(", ");
    a.op(J);
    a.put(" row_newbcast:");
    a.num(J);
    a.put(" row_mask:0xf bank_mask:0xf\nv_rsq_f64 ");
    a.op(OP_Y0);
    a.put(", ");
    a.op(OP_D);
    a.put("\n");
    // a VALU read of a transcendental result needs one wait state: the pool
    // items give it, column 0 (no pool) an s_nop
    if (column_pool(J) == 0) a.put("s_nop 0\n");
    column_pool_range(a, J, 0, 4);
    asm_op3(a, "v_mul_f64 ", OP_T, ", ", OP_D, OP_Y0);   // t = d y0
    a.put("\n");
    column_pool_range(a, J, 4, 6);
    asm_op3(a, "v_fma_f64 ", OP_E, ", -", OP_T, OP_Y0);   // e = 1 - t y0
    a.put(", 1.0\n");
    column_pool_range(a, J, 6, 8);
    asm_op3(a, "v_mul_f64 ", OP_T, ", ", OP_Y0, OP_E);   // t = y0 e
    a.put("\n");
    asm_op3(a, "v_fma_f64 ", OP_E, ", ", OP_E, OP_C375);   // e = 0.375 e + 0.5
    a.put(", 0.5\n");
    column_pool_range(a, J, 8, 10);
    asm_op3(a, "v_fma_f64 ", OP_INV, ", ", OP_T, OP_E);   // inv = t e + y0
    a.put(", ");
    a.op(OP_Y0);
    a.put("\n");
    column_pool_range(a, J, 10, 12);
    asm_op3(a, "v_mul_f64 ", J, ", ", J, OP_INV);   // l = a_rJ inv
    a.put("\n");
    asm_op3(a, "v_mul_f64 ", OP_MI, ", ", OP_INV, OP_GT);
    a.put("\n");
    asm_op3(a, "v_fma_f64 ", OP_IVR, ", ", OP_INV, OP_EQ);
    a.put(", ");
    a.op(OP_IVR);
    a.put("\n");
    asm_op3(a, "v_mul_f64 ", OP_M, ", ", J, OP_MI);
    a.put("\n");
    column_pool_range(a, J, 12, 32);
    if (J + 1 < 16) {
        // a_r,J+1 -= L_J+1,J L_rJ: the next pivot's entry
        asm_op3(a, "v_fmac_f64_dpp ", J + 1, ", -", J, J);
        a.put(" row_newbcast:");
        a.num((unsigned)(J + 1));
        a.put(" row_mask:0xf bank_mask:0xf\n");
    }
    return a;
}
// after column 15 nothing is left: it has no trailing columns, and its
// inverse update has no rows below it (m = 0 on every lane)
template <int J> struct ColumnAsm {
    static constexpr AsmText text = column_asm<J>();
};

// keep the compiler from reordering the LDS accesses of one wavefront (the
// hardware executes a wavefront's LDS operations in order)
__device__ __forceinline__ void wave_lds_order() { __atomic_signal_fence(__ATOMIC_SEQ_CST); }

template <class F, int... I>
__device__ __forceinline__ void static_for_seq(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_seq(f, std::make_integer_sequence<int, N>{});
}

// 1/sqrt(x): the hardware estimate refined by one Halley step
__device__ __forceinline__ double rsqrt_refined(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(0.375, e, 0.5), y);
}

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

struct CrArgs {
    const double* H;   // reduced system, Pp x Pp, lower triangle
    const double* g;   // right-hand side (Pp)
    double* x;         // solution (Pp)
    double* D;         // per super-row: current diagonal block (S x S, lower tiles)
    double* E;         // per super-row: current coupling to its left neighbour (S x S)
    double* W;         // per eliminated super-row: L^-1 (S x S)
    double* U;         // per eliminated super-row: W E_i
    double* V;         // per eliminated super-row: W E_q^T
    double* b;         // per super-row: current right-hand side (S)
    double* y;         // per eliminated super-row: W b_i (S)
    int* sync;         // [ns + 1], [ns + 4]: k_cr_tail3's flags; [ns + 3]: k_cr_back's finished workgroups
    Control* ctl;
    int Pp, ns;
    int gate;          // 1: a CG solve may have solved this iteration (return when ctl->pcg_fail is 0)
};

// Blocks, solutions and ready flags pass between workgroups of one launch by
// relaxed device-scope atomics, coherent across the XCDs' L2 caches without an
// acquire or release (which would write back or invalidate a whole L2): a
// writer waits until its stores have completed before it raises a flag, and a
// reader loads the data after it has seen the flag.
__device__ __forceinline__ double xget(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void xput(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
constexpr int CR_SPIN_MAX = 1 << 20;
// wait for a ready flag; past the bound (or once the Cholesky flag is down) the
// solution is discarded: the flag goes down and nothing waits any longer
__device__ __forceinline__ void cr_wait(const int* flag, Control* ctl) {
    int n = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if (__hip_atomic_load(&ctl->chol_ok, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
        if (++n > CR_SPIN_MAX) {
            __hip_atomic_store(&ctl->chol_ok, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// The back-substitution hands solutions over through the solution vector
// itself: the tail kernel sets every row it does not solve to NaN, and a row's
// entry is ready once it is no longer NaN (a NaN from a failed factorisation
// runs into the bounded wait, which marks the factorisation failed).
__device__ __forceinline__ double cr_poll(const double* p, Control* ctl) {
    int n = 0;
    double v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (__builtin_isnan(v)) {
        if (++n > CR_SPIN_MAX) {
            __hip_atomic_store(&ctl->chol_ok, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return v;
}

// Element (r, c) of super-block (I, J) of H, identity past the padded size.
// Branch-free: the address is clamped into H and the value selected after the
// load, so a thread's loads are all issued before the first one is waited for
// (a load under a per-lane branch is waited for before the branch closes).
template <int S>
__device__ __forceinline__ double h_ld(const CrArgs& a, int I, int J, int r, int c) {
    const int gr = S * I + r, gc = S * J + c;
    const bool in = gr < a.Pp && gc < a.Pp;
    const double v = a.H[(long long)min(gr, a.Pp - 1) * a.Pp + min(gc, a.Pp - 1)];
    return in ? v : (gr == gc ? 1.0 : 0.0);
}

// g[S I + t], zero past the padded size (branch-free as h_ld)
template <int S>
__device__ __forceinline__ double g_ld(const CrArgs& a, int I, int t) {
    const int gr = S * I + t;
    const double v = a.g[min(gr, a.Pp - 1)];
    return gr < a.Pp ? v : 0.0;
}
}  // namespace

// sum_{m=c}^{b-1} L_bm W_mc (a 16 x 16 MFMA tile) for W's block (b, c), b > c:
// every operand of the at most P - 1 block products is read from LDS first
// (the addresses are valid for every m < P), then the products are taken
template <int P>
__device__ __forceinline__ d4 w_row_sum(double (*A)[16 * P + 1], double (*Wl)[16 * P + 1], int bb, int c,
                                        int lane) {
    double la[P][4], wb[P][4];
#pragma unroll
    for (int m = 0; m < P; ++m)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            la[m][s] = A[16 * bb + (lane & 15)][16 * m + 4 * s + (lane >> 4)];
            wb[m][s] = Wl[16 * m + 4 * s + (lane >> 4)][16 * c + (lane & 15)];
        }
    d4 t = 0.0;
#pragma unroll
    for (int m = 0; m < P; ++m)
        if (m >= c && m < bb) {
#pragma unroll
            for (int s = 0; s < 4; ++s) t = mfma(la[m][s], wb[m][s], t);
        }
    return t;
}

// Factor the S x S block in A (lower blocks, the diagonal ones whole) as
// A = L L^T with L in A's lower triangle and W = L^-1 in Wl (zeroed by the
// caller), by the whole workgroup; the caller has synchronised after filling
// A and Wl and synchronises before reading them.  Returns, on wavefront 0,
// whether a pivot was not positive and finite.
template <int P>
__device__ __forceinline__ bool cr_factor(double (*A)[16 * P + 1], double (*Wl)[16 * P + 1]) {
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    (void)tid;
    // the diagonal factor's per-lane factors (r > j) and (r == j), formed once
    double gtv[16], eqv[16];
    static_for<16>([&](auto J) {
        constexpr int j = decltype(J)::value;
        gtv[j] = (lane & 15) > j ? 1.0 : 0.0;
        eqv[j] = (lane & 15) == j ? 1.0 : 0.0;
    });
    bool bad = false;
    // Step k: wavefront 0 factors the diagonal block k while the others form W's
    // block row k - 1 and the trailing update of step k - 1 (all but the next
    // diagonal tile, which wavefront 0 updated itself); after a barrier the
    // panel blocks of column k, wavefront 0 taking block row k + 1 and then the
    // tile (k + 1, k + 1), so it can start the next diagonal block at once.
    for (int k = 0; k < P; ++k) {
        if (w == 0) {
            const int r = lane & 15, k0 = 16 * k;
            double av[16], yv[4];
            const int R = lane >> 4;   // this 16-lane row keeps columns 4R .. 4R + 3 of the inverse
            static_for<16>([&](auto J) {
                constexpr int j = decltype(J)::value;
                av[j] = A[k0 + r][k0 + j];
            });
            static_for<4>([&](auto I) {
                constexpr int c = decltype(I)::value;
                yv[c] = (4 * R + c == r) ? 1.0 : 0.0;
            });
            // column j: pivot d = a_jj, 1/sqrt(d) refined by one Halley step,
            // L_rj = a_rj / sqrt(d) (into av[j]), then a_rq -= L_rj L_qj for
            // q > j (q = j + 1 at once, the others during column j + 1).  The
            // unit factor's inverse: rows r > j take yv -= (L_rj / L_jj) yv_j
            // (m), applied during column j + 1.
            double ivr = 0.0, mprev = 0.0;
            const double c375 = 0.375;
            static_for<16>([&](auto J) {
                constexpr int j = decltype(J)::value;
                double m, d, y0, t, e, inv, mi;
                asm volatile((ColumnAsm<j>::text)
                             : "+v"(av[0]), "+v"(av[1]), "+v"(av[2]), "+v"(av[3]), "+v"(av[4]), "+v"(av[5]),
                               "+v"(av[6]), "+v"(av[7]), "+v"(av[8]), "+v"(av[9]), "+v"(av[10]), "+v"(av[11]),
                               "+v"(av[12]), "+v"(av[13]), "+v"(av[14]), "+v"(av[15]), "+v"(yv[0]), "+v"(yv[1]),
                               "+v"(yv[2]), "+v"(yv[3]), "=&v"(m), "=&v"(d), "=&v"(y0), "=&v"(t), "=&v"(e),
                               "=&v"(inv), "=&v"(mi), "+v"(ivr)
                             : "v"(mprev), "v"(c375), "v"(gtv[j]), "v"(eqv[j]));
                mprev = m;
            });
            // every pivot positive and finite: 1/sqrt of it is too (ivr of lane r
            // is that of pivot r; the four 16-lane rows agree)
            bad = bad || __builtin_amdgcn_ballot_w64(!(ivr > 0.0 && ivr < __builtin_inf())) != 0;
            static_for<4>([&](auto I) {
                constexpr int j = decltype(I)::value;
                const int c = 4 * R + j;
                Wl[k0 + r][k0 + c] = (c <= r) ? yv[j] * ivr : 0.0;
            });
            if (lane < 16) {
                static_for<16>([&](auto J) {
                    constexpr int j = decltype(J)::value;
                    A[k0 + r][k0 + j] = (j <= r) ? av[j] : 0.0;
                });
            }
        } else {
            if (w < k) {
                // W block row k - 1, column block c = w - 1 < k - 1:
                // W_bc = -W_bb sum_{m=c}^{b-1} L_bm W_mc, b = k - 1
                const int bb = k - 1, c = w - 1;
                const d4 t = w_row_sum<P>(A, Wl, bb, c, lane);
                d4 u = 0.0;
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    u = mfma(Wl[16 * bb + (lane & 15)][16 * bb + 4 * s + (lane >> 4)], t[s], u);
#pragma unroll
                for (int v = 0; v < 4; ++v) Wl[16 * bb + (lane >> 4) + 4 * v][16 * c + (lane & 15)] = -u[v];
            }
            if (k >= 1) {
                // trailing update of step k - 1, A_bc -= L_b,k-1 L_c,k-1^T for
                // k <= c <= b < P, except (k, k): tiles 1 .. n(n+1)/2 - 1 of the
                // lower triangle of the n x n trailing blocks
                const int n = P - k, T = n * (n + 1) / 2, k0 = 16 * (k - 1);
                for (int t = w; t < T; t += 3) {
                    int bq = 0;
                    while ((bq + 1) * (bq + 2) / 2 <= t) ++bq;
                    const int bb = k + bq, cc = k + (t - bq * (bq + 1) / 2);
                    d4 acc = 0.0;
#pragma unroll
                    for (int s = 0; s < 4; ++s)
                        acc = mfma(A[16 * bb + (lane & 15)][k0 + 4 * s + (lane >> 4)],
                                   A[16 * cc + (lane & 15)][k0 + 4 * s + (lane >> 4)], acc);
#pragma unroll
                    for (int v = 0; v < 4; ++v) A[16 * bb + (lane >> 4) + 4 * v][16 * cc + (lane & 15)] -= acc[v];
                }
            }
        }
        __syncthreads();
        if (k + 1 < P) {
            // panel: L_bk = A_bk W_kk^T, one block row per wavefront
            const int k0 = 16 * k;
            if (w < P - 1 - k) {
                const int bb = k + 1 + w;
                d4 acc = 0.0;
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    acc = mfma(A[16 * bb + (lane & 15)][k0 + 4 * s + (lane >> 4)],
                               Wl[k0 + (lane & 15)][k0 + 4 * s + (lane >> 4)], acc);
#pragma unroll
                for (int v = 0; v < 4; ++v) A[16 * bb + (lane >> 4) + 4 * v][k0 + (lane & 15)] = acc[v];
            }
            if (w == 0) {
                // the next diagonal tile from this wavefront's own panel block
                wave_lds_order();
                const int b1 = 16 * (k + 1);
                d4 acc = 0.0;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const double x = A[b1 + (lane & 15)][k0 + 4 * s + (lane >> 4)];
                    acc = mfma(x, x, acc);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) A[b1 + (lane >> 4) + 4 * v][b1 + (lane & 15)] -= acc[v];
            }
            __syncthreads();
        }
    }
    if (P > 1) {
        // the last block row of W
        if (w < P - 1) {
            const int bb = P - 1, c = w;
            const d4 t = w_row_sum<P>(A, Wl, bb, c, lane);
            d4 u = 0.0;
#pragma unroll
            for (int s = 0; s < 4; ++s) u = mfma(Wl[16 * bb + (lane & 15)][16 * bb + 4 * s + (lane >> 4)], t[s], u);
#pragma unroll
            for (int v = 0; v < 4; ++v) Wl[16 * bb + (lane >> 4) + 4 * v][16 * c + (lane & 15)] = -u[v];
        }
    }
    return bad;
}

// Load super-row i's diagonal block into A (lower blocks, the diagonal ones
// whole: the diagonal factor's lanes also carry the entries above the
// diagonal, which the elimination reduces to rounding noise only if they start
// symmetric) and zero Wl; all loads of a thread are issued before its stores.
template <int P>
__device__ __forceinline__ void cr_load_block(const CrArgs& a, int i, bool fromH, double (*A)[16 * P + 1],
                                              double (*Wl)[16 * P + 1]) {
    constexpr int S = 16 * P, NE = S * S / 256;
    const int tid = threadIdx.x;
    double ld[NE];
    if (fromH) {
#pragma unroll
        for (int t = 0; t < NE; ++t) {
            const int e = tid + 256 * t, r = e / S, c = e - (e / S) * S;
            ld[t] = h_ld<S>(a, i, i, max(r, c), min(r, c));
        }
    } else {
#pragma unroll
        for (int t = 0; t < NE; ++t) ld[t] = a.D[(long long)i * S * S + tid + 256 * t];
    }
#pragma unroll
    for (int t = 0; t < NE; ++t) {
        const int e = tid + 256 * t, r = e / S, c = e - (e / S) * S;
        A[r][c] = (c >> 4) <= (r >> 4) ? ld[t] : 0.0;
        Wl[r][c] = 0.0;
    }
}

// Rows the back-substitution solves (those not a multiple of 2^L) to NaN:
// not ready (see cr_poll).
template <int S>
__device__ __forceinline__ void cr_mark_unsolved(const CrArgs& a, int L, int tid) {
    const double nan = __builtin_nan("");
    for (int e = tid; e < a.Pp; e += 256)
        if (((e / S) & ((1 << L) - 1)) != 0) a.x[e] = nan;
}

// One workgroup per super-row i eliminated at this level: D_i = L L^T, W = L^-1,
// then U_i = W E_i, V_i = W E_q^T (MFMA tiles, wavefront w taking tile column w
// of both) and y_i = W b_i; W, U, V and y are stored for the survivors' update
// and the back-substitution.
template <int P>
__global__ __launch_bounds__(256) void k_cr_factor(CrArgs a, int level) {
    constexpr int S = 16 * P, SP = S + 1, NE = S * S / 256, Q = S / 4;
    __shared__ double A[S][SP];    // D, then its factor L (lower blocks)
    __shared__ double Wl[S][SP];   // W = L^-1
    __shared__ double Ei[S][SP];   // coupling of i to p = i - 2^l
    __shared__ double Eq[S][SP];   // coupling of q = i + 2^l to i
    __shared__ double vb[S];
    if (a.gate && a.ctl->pcg_fail == 0) return;   // this iteration's CG solve converged
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int st = 1 << level, i = (2 * blockIdx.x + 1) << level, q = i + st;
    const bool fromH = level == 0, has_q = q < a.ns;
    const long long SS = (long long)S * S;
    // D_i and b_i first; the coupling blocks E_i, E_q are requested behind them
    // and stay in registers while D_i is factored, so their loads overlap the
    // factorisation instead of preceding it
    double le[NE], lq[NE];
    {
        constexpr int NB = NE;
        double ld[NB];
        if (fromH) {
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                const int e = tid + 256 * t, r = e / S, c = e - (e / S) * S;
                ld[t] = h_ld<S>(a, i, i, max(r, c), min(r, c));
            }
        } else {
#pragma unroll
            for (int t = 0; t < NB; ++t) ld[t] = a.D[(long long)i * SS + tid + 256 * t];
        }
        const double bv = tid < S ? (fromH ? g_ld<S>(a, i, tid) : a.b[(long long)i * S + tid]) : 0.0;
        if (fromH) {
#pragma unroll
            for (int u = 0; u < NE; ++u) {
                const int e = tid + 256 * u, r = e / S, c = e - (e / S) * S;
                le[u] = h_ld<S>(a, i, i - 1, r, c);
                lq[u] = has_q ? h_ld<S>(a, q, i, r, c) : 0.0;
            }
        } else {
#pragma unroll
            for (int u = 0; u < NE; ++u) {
                le[u] = a.E[i * SS + tid + 256 * u];
                lq[u] = has_q ? a.E[q * SS + tid + 256 * u] : 0.0;
            }
        }
#pragma unroll
        for (int t = 0; t < NB; ++t) {
            const int e = tid + 256 * t, r = e / S, c = e - (e / S) * S;
            A[r][c] = (c >> 4) <= (r >> 4) ? ld[t] : 0.0;
            Wl[r][c] = 0.0;
        }
        if (tid < S) vb[tid] = bv;
    }
    __syncthreads();
    const bool bad = cr_factor<P>(A, Wl);
    if (bad && tid == 0) a.ctl->chol_ok = 0;
#pragma unroll
    for (int u = 0; u < NE; ++u) {
        const int e = tid + 256 * u, r = e / S, c = e - (e / S) * S;
        Ei[r][c] = le[u];
        Eq[r][c] = lq[u];
    }
    __syncthreads();
    if (w < P) {
        // tile column c = w of U = W E_i and of V = W E_q^T; W is lower triangular
        const int c = w, n = 16 * c + (lane & 15);
        double* Uo = a.U + i * SS;
        double* Vo = a.V + i * SS;
        static_for<P>([&](auto Rr) {
            constexpr int r = decltype(Rr)::value;
            // the tile row's operands first, then its MFMAs
            double wa[r + 1][4], eu[r + 1][4], ev[r + 1][4];
#pragma unroll
            for (int m = 0; m <= r; ++m)
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) {
                    const int kk = 16 * m + 4 * s4 + (lane >> 4);
                    wa[m][s4] = Wl[16 * r + (lane & 15)][kk];
                    eu[m][s4] = Ei[kk][n];
                    ev[m][s4] = Eq[n][kk];
                }
            d4 u = 0.0, v = 0.0;
#pragma unroll
            for (int m = 0; m <= r; ++m)
#pragma unroll
                for (int s4 = 0; s4 < 4; ++s4) {
                    u = mfma(wa[m][s4], eu[m][s4], u);
                    v = mfma(wa[m][s4], ev[m][s4], v);
                }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                Uo[(16 * r + (lane >> 4) + 4 * e) * S + n] = u[e];
                if (has_q) Vo[(16 * r + (lane >> 4) + 4 * e) * S + n] = v[e];
            }
        });
    }
    // y = W b, four threads per row
    {
        const int row = tid >> 2, pt = tid & 3;
        if (row < S) {
            double sacc = 0.0;
            for (int c = Q * pt; c < Q * pt + Q; ++c) sacc = fma(Wl[row][c], vb[c], sacc);
            sacc += __shfl_xor(sacc, 1);
            sacc += __shfl_xor(sacc, 2);
            if (pt == 0) a.y[(long long)i * S + row] = sacc;
        }
    }
    double* Wo = a.W + i * SS;
    for (int e = tid; e < S * S; e += 256) {
        const int r = e / S, c = e - (e / S) * S;
        Wo[e] = Wl[r][c];
    }
}

// The super-rows left after the last level, q_k = k 2^L (k < t), form a block
// tridiagonal system (D_k, coupling E_k of q_k to q_(k-1), b_k).  One
// workgroup solves it by the block Cholesky factorisation, k in order:
//     M_k = E_k W_(k-1)^T,  L_k L_k^T = D_k - M_k M_k^T,  W_k = L_k^-1,
//     y_k = W_k (b_k - M_k y_(k-1)),
// then x_(t-1) = W_(t-1)^T y_(t-1), x_k = W_k^T (y_k - M_(k+1)^T x_(k+1)).
// W_k, M_k and y_k wait in the rows' workspace for the backward sweep.
template <int P>
__global__ __launch_bounds__(256) void k_cr_tail(CrArgs a, int L) {
    constexpr int S = 16 * P, SP = S + 1, Q = S / 4;
    __shared__ double A[S][SP];     // D_k, then L_k
    __shared__ double W0[S][SP];    // W of the even steps
    __shared__ double W1[S][SP];    // W of the odd steps
    __shared__ double Mb[S][SP];    // E_k, then M_k
    __shared__ double vb[S], vy[S], vx[S];
    if (a.gate && a.ctl->pcg_fail == 0) return;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int t = (a.ns + (1 << L) - 1) >> L;
    const bool fromH = L == 0;
    const long long SS = (long long)S * S;
    const int row = tid >> 2, part = tid & 3;
    bool bad = false;
    for (int k = 0; k < t; ++k) {
        const int q = k << L;
        double(*Wc)[SP] = (k & 1) ? W1 : W0;
        double(*Wp)[SP] = (k & 1) ? W0 : W1;
        cr_load_block<P>(a, q, fromH, A, Wc);
        for (int e = tid; e < S; e += 256) vb[e] = fromH ? g_ld<S>(a, q, e) : a.b[(long long)q * S + e];
        if (k >= 1) {
            constexpr int NE = S * S / 256;
            double ld[NE];
            if (fromH) {
#pragma unroll
                for (int u = 0; u < NE; ++u) {
                    const int e = tid + 256 * u, r = e / S, c = e - (e / S) * S;
                    ld[u] = h_ld<S>(a, q, q - 1, r, c);
                }
            } else {
#pragma unroll
                for (int u = 0; u < NE; ++u) ld[u] = a.E[q * SS + tid + 256 * u];
            }
#pragma unroll
            for (int u = 0; u < NE; ++u) {
                const int e = tid + 256 * u, r = e / S, c = e - (e / S) * S;
                Mb[r][c] = ld[u];
            }
        }
        __syncthreads();
        if (k >= 1) {
            // M = E W_prev^T, wavefront w forms block row w (all its E reads
            // before its writes, so in place); W_prev is lower triangular
            if (w < P) {
                d4 acc[P];
                double ea[P][4];
#pragma unroll
                for (int m = 0; m < P; ++m)
#pragma unroll
                    for (int s4 = 0; s4 < 4; ++s4) ea[m][s4] = Mb[16 * w + (lane & 15)][16 * m + 4 * s4 + (lane >> 4)];
                static_for<P>([&](auto Cc) {
                    constexpr int c = decltype(Cc)::value;
                    double wb[c + 1][4];
#pragma unroll
                    for (int m = 0; m <= c; ++m)
#pragma unroll
                        for (int s4 = 0; s4 < 4; ++s4) wb[m][s4] = Wp[16 * c + (lane & 15)][16 * m + 4 * s4 + (lane >> 4)];
                    acc[c] = 0.0;
#pragma unroll
                    for (int m = 0; m <= c; ++m)
#pragma unroll
                        for (int s4 = 0; s4 < 4; ++s4) acc[c] = mfma(ea[m][s4], wb[m][s4], acc[c]);
                });
                wave_lds_order();
#pragma unroll
                for (int c = 0; c < P; ++c)
#pragma unroll
                    for (int v = 0; v < 4; ++v) Mb[16 * w + (lane >> 4) + 4 * v][16 * c + (lane & 15)] = acc[c][v];
            }
            __syncthreads();
            // D -= M M^T (lower tiles, the diagonal ones whole); b -= M y_prev
            for (int tt = w; tt < P * (P + 1) / 2; tt += 4) {
                int r = 0;
                while ((r + 1) * (r + 2) / 2 <= tt) ++r;
                const int c = tt - r * (r + 1) / 2;
                double ma[S / 4], mc[S / 4];
#pragma unroll
                for (int s4 = 0; s4 < S / 4; ++s4) {
                    ma[s4] = Mb[16 * r + (lane & 15)][4 * s4 + (lane >> 4)];
                    mc[s4] = Mb[16 * c + (lane & 15)][4 * s4 + (lane >> 4)];
                }
                d4 acc = 0.0;
#pragma unroll
                for (int s4 = 0; s4 < S / 4; ++s4) acc = mfma(ma[s4], mc[s4], acc);
#pragma unroll
                for (int v = 0; v < 4; ++v) A[16 * r + (lane >> 4) + 4 * v][16 * c + (lane & 15)] -= acc[v];
            }
            if (row < S) {
                double s = 0.0;
                for (int c = Q * part; c < Q * part + Q; ++c) s = fma(Mb[row][c], vy[c], s);
                s += __shfl_xor(s, 1);
                s += __shfl_xor(s, 2);
                if (part == 0) vb[row] -= s;
            }
            // M_k for the backward sweep
            for (int e = tid; e < S * S; e += 256) {
                const int r = e / S, c = e - (e / S) * S;
                a.U[q * SS + e] = Mb[r][c];
            }
            __syncthreads();
        }
        bad = cr_factor<P>(A, Wc) || bad;
        __syncthreads();
        // y_k = W_k b
        if (row < S) {
            double s = 0.0;
            for (int c = Q * part; c < Q * part + Q; ++c) s = fma(Wc[row][c], vb[c], s);
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            if (part == 0) {
                vy[row] = s;
                a.y[(long long)q * S + row] = s;
            }
        }
        if (k + 1 < t)
            for (int e = tid; e < S * S; e += 256) {
                const int r = e / S, c = e - (e / S) * S;
                a.W[q * SS + e] = Wc[r][c];
            }
        __syncthreads();
    }
    if (bad && tid == 0) a.ctl->chol_ok = 0;
    // backward: x_(t-1) = W^T y with W_(t-1) still in LDS
    {
        double(*Wc)[SP] = ((t - 1) & 1) ? W1 : W0;
        if (row < S) {
            double s = 0.0;
            for (int r = Q * part; r < Q * part + Q; ++r) s = fma(Wc[r][row], vy[r], s);
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            if (part == 0) {
                vx[row] = s;
                const int gr = S * ((t - 1) << L) + row;
                if (gr < a.Pp) a.x[gr] = s;
            }
        }
    }
    for (int k = t - 2; k >= 0; --k) {
        const int q = k << L, q1 = (k + 1) << L;
        __syncthreads();
        // z = y_k - M_(k+1)^T x_(k+1) (M into Mb, W_k into A)
        {
            constexpr int NE = S * S / 256;
            double lm[NE], lw[NE];
#pragma unroll
            for (int u = 0; u < NE; ++u) {
                lm[u] = a.U[q1 * SS + tid + 256 * u];
                lw[u] = a.W[q * SS + tid + 256 * u];
            }
            if (tid < S) vb[tid] = a.y[(long long)q * S + tid];
#pragma unroll
            for (int u = 0; u < NE; ++u) {
                const int e = tid + 256 * u, r = e / S, c = e - (e / S) * S;
                Mb[r][c] = lm[u];
                A[r][c] = lw[u];
            }
        }
        __syncthreads();
        if (row < S) {
            double s = 0.0;
            for (int r = Q * part; r < Q * part + Q; ++r) s = fma(Mb[r][row], vx[r], s);
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            if (part == 0) vb[row] -= s;
        }
        __syncthreads();
        if (row < S) {
            double s = 0.0;
            for (int r = Q * part; r < Q * part + Q; ++r) s = fma(A[r][row], vb[r], s);
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            if (part == 0) {
                vx[row] = s;
                const int gr = S * q + row;
                if (gr < a.Pp) a.x[gr] = s;
            }
        }
    }
    // the rows the back-substitution solves start unsolved
    if (L > 0) cr_mark_unsolved<S>(a, L, tid);
}

// Survivor q of this level: D_q -= V_im^T V_im + U_ip^T U_ip (blockIdx.y < P: a
// block row of lower tiles), its new coupling -V_im^T U_im (< 2P: a block row),
// and b_q -= V_im^T y_im + U_ip^T y_ip (= 2P); im = q - 2^l, ip = q + 2^l.
template <int P>
__global__ __launch_bounds__(256) void k_cr_update(CrArgs a, int level) {
    constexpr int S = 16 * P, K4 = S / 4;
    if (a.gate && a.ctl->pcg_fail == 0) return;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int st = 1 << level, q = (2 * blockIdx.x) << level;
    const int im = q - st, ip = q + st;
    const bool has_m = q > 0, has_p = ip < a.ns;
    const bool fromH = level == 0;
    const long long SS = (long long)S * S;
    const int part = blockIdx.y;
    if (blockIdx.x == 0 && part == 0 && tid < 4) a.sync[a.ns + 1 + tid] = 0;   // k_cr_tail3's flags, k_cr_back's count
    if (part < 2 * P) {
        const bool isE = part >= P;
        const int r = isE ? part - P : part, c = w;
        if (c >= P || (!isE && c > r) || (isE && !has_m)) return;
        const int R = 16 * r + (lane & 15), C = 16 * c + (lane & 15);
        // every load first: both products' operands and the current block
        double va[K4], vb[K4], ua[K4], ub[K4], base[4];
        const bool use_p = !isE && has_p;
        if (has_m) {
            const double* Vm = a.V + im * SS;
            const double* Bm = isE ? a.U + im * SS : Vm;
#pragma unroll
            for (int s = 0; s < K4; ++s) {
                const int kk = 4 * s + (lane >> 4);
                va[s] = Vm[kk * S + R];
                vb[s] = Bm[kk * S + C];
            }
        }
        if (use_p) {
            const double* Up = a.U + ip * SS;
#pragma unroll
            for (int s = 0; s < K4; ++s) {
                const int kk = 4 * s + (lane >> 4);
                ua[s] = Up[kk * S + R];
                ub[s] = Up[kk * S + C];
            }
        }
        if (!isE) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rr = 16 * r + (lane >> 4) + 4 * v, cc = 16 * c + (lane & 15);
                base[v] = fromH ? h_ld<S>(a, q, q, max(rr, cc), min(rr, cc)) : a.D[q * SS + rr * S + cc];
            }
        }
        d4 acc = 0.0;
        if (has_m) {
#pragma unroll
            for (int s = 0; s < K4; ++s) acc = mfma(va[s], vb[s], acc);
        }
        if (use_p) {
#pragma unroll
            for (int s = 0; s < K4; ++s) acc = mfma(ua[s], ub[s], acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int rr = 16 * r + (lane >> 4) + 4 * v, cc = 16 * c + (lane & 15);
            if (isE) a.E[q * SS + rr * S + cc] = -acc[v];
            else a.D[q * SS + rr * S + cc] = base[v] - acc[v];
        }
        return;
    }
    // b: thread (t, h) sums the half h of k for entry t of both products;
    // 2S <= 256 threads, halves combined through LDS
    __shared__ double half[S];
    const int t = tid % S, h = tid / S;
    double s = 0.0;
    if (h < 2) {
        constexpr int KH = S / 2;
        double vv[KH], yv[KH], uu[KH], yu[KH];
        if (has_m) {
            const double* Vm = a.V + im * SS;
            const double* ym = a.y + (long long)im * S;
#pragma unroll
            for (int k = 0; k < KH; ++k) {
                vv[k] = Vm[(KH * h + k) * S + t];
                yv[k] = ym[KH * h + k];
            }
        }
        if (has_p) {
            const double* Up = a.U + ip * SS;
            const double* yp = a.y + (long long)ip * S;
#pragma unroll
            for (int k = 0; k < KH; ++k) {
                uu[k] = Up[(KH * h + k) * S + t];
                yu[k] = yp[KH * h + k];
            }
        }
        double b0 = 0.0;
        if (h == 0) b0 = fromH ? g_ld<S>(a, q, t) : a.b[(long long)q * S + t];
        s = b0;
        if (has_m) {
#pragma unroll
            for (int k = 0; k < KH; ++k) s = fma(-vv[k], yv[k], s);
        }
        if (has_p) {
#pragma unroll
            for (int k = 0; k < KH; ++k) s = fma(-uu[k], yu[k], s);
        }
        if (h == 1) half[t] = s;
    }
    __syncthreads();
    if (h == 0) a.b[(long long)q * S + t] = s + half[t];
}

// x_i = W_i^T (y_i - U_i x_p - V_i x_q) for every super-row eliminated by the
// levels, one workgroup each, in one launch.  The rows of the highest level
// have the lowest workgroup indices, so a row waits only for rows of
// workgroups dispatched before it (the launch's one-per-CU workgroups are
// resident together on any device with as many CUs; every wait is bounded
// regardless).  A workgroup stages its U, V, W (coalesced loads into LDS) and y,
// then reads its two neighbours' solutions (one level up, or rows of the tail)
// entry by entry as they become ready (cr_poll), and stores its own.
template <int P>
__global__ __launch_bounds__(256) void k_cr_back(CrArgs a, int L, PoseStep ps, int fuse) {
    constexpr int S = 16 * P, SP = S + 1, Q = S / 4, NE = S * S / 256;
    __shared__ double Us[S][SP], Vs[S][SP], Ws[S][SP];
    __shared__ double xp[S], xq[S], z[S], yv[S];
    __shared__ int last;
    if (a.gate && a.ctl->pcg_fail == 0) return;
    const int tid = threadIdx.x;
    // the rows of the highest level first: a row waits only for rows of higher
    // levels (and the tail's), which have lower workgroup indices
    int tk = blockIdx.x, l = L - 1;
    for (; l > 0; --l) {
        const int stl = 1 << l, n = (a.ns - stl + 2 * stl - 1) / (2 * stl);
        if (tk < n) break;
        tk -= n;
    }
    const int i = (2 * tk + 1) << l, st = 1 << l, p = i - st, q = i + st;
    const long long SS = (long long)S * S;
    const bool has_q = q < a.ns;
    {
        double lu[NE], lv[NE], lw[NE];
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            lu[u] = a.U[i * SS + tid + 256 * u];
            lv[u] = has_q ? a.V[i * SS + tid + 256 * u] : 0.0;
            lw[u] = a.W[i * SS + tid + 256 * u];
        }
        if (tid < S) yv[tid] = a.y[(long long)i * S + tid];
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int e = tid + 256 * u, r = e / S, c = e - (e / S) * S;
            Us[r][c] = lu[u];
            Vs[r][c] = lv[u];
            Ws[r][c] = lw[u];
        }
    }
    // the neighbours' solutions, entry by entry as they become ready (wavefront
    // 0 row p, wavefront 1 row q)
    if (tid < S) xp[tid] = cr_poll(a.x + S * p + tid, a.ctl);   // rows of p lie below S i <= Pp
    if (tid >= 64 && tid < 64 + S) {
        const int gq = S * q + tid - 64;
        xq[tid - 64] = (has_q && gq < a.Pp) ? cr_poll(a.x + gq, a.ctl) : 0.0;
    }
    __syncthreads();
    const int row = tid >> 2, pt = tid & 3;
    if (row < S) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < Q; ++j) {
            const int c = Q * pt + j;
            s = fma(Us[row][c], xp[c], s);
            s = fma(Vs[row][c], xq[c], s);
        }
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        if (pt == 0) z[row] = yv[row] - s;
    }
    __syncthreads();
    if (row < S) {
        // W is lower triangular: W(r, row) = 0 for r < row
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < Q; ++j) s = fma(Ws[Q * pt + j][row], z[Q * pt + j], s);
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        const int gr = S * i + row;
        if (pt == 0 && gr < a.Pp) xput(a.x + gr, s);
    }
    if (fuse) {
        // the last workgroup to finish (its count taken after its own solution
        // stores completed) finds every row solved and runs the pose update (it
        // reads the solution with device-coherent loads); it leaves the counter
        // at zero for the next solve
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __syncthreads();
        if (tid == 0)
            last = __hip_atomic_fetch_add(a.sync + a.ns + 3, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                   (int)gridDim.x - 1;
        __syncthreads();
        if (last) {
            if (tid == 0) __hip_atomic_store(a.sync + a.ns + 3, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pose_update_block<true>(ps.pd, ps.frame_mode, ps.cam, ps.b, a.ctl, a.x, ps.step, 0);
        }
    }
}

// Helpers of the three-row tail (super-block tiles in LDS, 16 x 16 MFMA tiles).
// C = X^T X (lower tiles, the diagonal ones whole), wavefront w taking tiles
// w, w + 4, ...
template <int P>
__device__ __forceinline__ void cr_gram(double (*X)[16 * P + 1], double (*C)[16 * P + 1], int w, int lane) {
    constexpr int S = 16 * P;
    for (int tt = w; tt < P * (P + 1) / 2; tt += 4) {
        int r = 0;
        while ((r + 1) * (r + 2) / 2 <= tt) ++r;
        const int c = tt - r * (r + 1) / 2;
        double xa[S / 4], xc[S / 4];
#pragma unroll
        for (int s4 = 0; s4 < S / 4; ++s4) {
            xa[s4] = X[4 * s4 + (lane >> 4)][16 * r + (lane & 15)];
            xc[s4] = X[4 * s4 + (lane >> 4)][16 * c + (lane & 15)];
        }
        d4 acc = 0.0;
#pragma unroll
        for (int s4 = 0; s4 < S / 4; ++s4) acc = mfma(xa[s4], xc[s4], acc);
#pragma unroll
        for (int v = 0; v < 4; ++v) C[16 * r + (lane >> 4) + 4 * v][16 * c + (lane & 15)] = acc[v];
    }
}
// X = W E (tr = false) or W E^T (tr = true) in place of E in Eb, wavefront w
// taking tile column w (its reads of Eb's column block, or row block, come
// before its writes of column block w); W lower triangular
template <int P>
__device__ __forceinline__ void cr_times_w(double (*Wl)[16 * P + 1], double (*Eb)[16 * P + 1], bool tr, int w,
                                           int lane) {
    if (w >= P) return;
    const int c = w, n = 16 * c + (lane & 15);
    d4 acc[P];
    static_for<P>([&](auto Rr) {
        constexpr int r = decltype(Rr)::value;
        double wa[r + 1][4], eb[r + 1][4];
#pragma unroll
        for (int m = 0; m <= r; ++m)
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const int kk = 16 * m + 4 * s4 + (lane >> 4);
                wa[m][s4] = Wl[16 * r + (lane & 15)][kk];
                eb[m][s4] = tr ? Eb[n][kk] : Eb[kk][n];
            }
        acc[r] = 0.0;
#pragma unroll
        for (int m = 0; m <= r; ++m)
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) acc[r] = mfma(wa[m][s4], eb[m][s4], acc[r]);
    });
    if (tr) return;   // W E^T: the caller takes it from registers (rows of Eb are still read)
    wave_lds_order();
#pragma unroll
    for (int r = 0; r < P; ++r)
#pragma unroll
        for (int v = 0; v < 4; ++v) Eb[16 * r + (lane >> 4) + 4 * v][n] = acc[r][v];
}
// out = M^T v (tr) or M v, four threads per entry
template <int P>
__device__ __forceinline__ double cr_matvec(double (*M)[16 * P + 1], const double* v, bool tr, int tid) {
    constexpr int S = 16 * P, Q = S / 4;
    const int row = tid >> 2, pt = tid & 3;
    double s = 0.0;
    if (row < S) {
#pragma unroll
        for (int j = 0; j < Q; ++j) {
            const int k = Q * pt + j;
            s = fma(tr ? M[k][row] : M[row][k], v[k], s);
        }
    }
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    return s;
}

// Three super-rows left (r0 = 0, r1 = 2^L, r2 = 2^(L+1)): the two ends are
// eliminated concurrently, then the middle one, so only two factorisations are
// sequential.  Two workgroups:
//   workgroup 0: D_0 = L_0 L_0^T, W_0, V_0 = W_0 E_1^T, y_0 = W_0 b_0, and the
//             middle row's share C_0 = V_0^T V_0, c_0 = V_0^T y_0, published
//             with a ready flag; then, with W_0, V_0 and y_0 still in LDS, it
//             waits for x_1 and solves x_0 = W_0^T (y_0 - V_0 x_1);
//   workgroup 1: D_2 = L_2 L_2^T, W_2, U_2 = W_2 E_2, y_2, C_2 = U_2^T U_2,
//             c_2 = U_2^T y_2; waits for workgroup 0's share; factors
//             D_1 - C_0 - C_2, solves x_1 = W_1^T W_1 (b_1 - c_0 - c_2),
//             publishes it with a second flag, then x_2 = W_2^T (y_2 - U_2 x_1)
//             while workgroup 0 forms x_0.
// The two waits rely on the grid's two workgroups being resident together
// (two workgroups on an otherwise idle device); like every wait here they are
// bounded, and a wait that runs out marks the factorisation failed.
// What the workgroups hand over moves by device-coherent atomics (xput / xget).
// sync[ns + 1] and sync[ns + 4] (the flags) are cleared by k_cr_update.
template <int P>
__global__ __launch_bounds__(256) void k_cr_tail3(CrArgs a, int L) {
    constexpr int S = 16 * P, SP = S + 1, NE = S * S / 256;
    __shared__ double A[S][SP], Wl[S][SP], Mb[S][SP], W2s[S][SP];
    __shared__ double vb[S], vy[S], vc[S], vx[S], vz[S];
    if (a.gate && a.ctl->pcg_fail == 0) return;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const long long SS = (long long)S * S;
    const int r0 = 0, r1 = 1 << L, r2 = 2 << L;
    int* flag = a.sync + a.ns + 1;    // workgroup 0's share of the middle row published
    int* flagx = a.sync + a.ns + 4;   // x_1 published
    const int tk = blockIdx.x;                    // 0: the left end; 1: the right end, then the middle row
    const int e = tk == 0 ? r0 : r2;              // this workgroup's end row
    const int en = tk == 0 ? r1 : r2;             // the coupling block: E of r1 (to r0) or of r2 (to r1)
    // D_e, the coupling, b_e; workgroup 1 also requests D_1 and b_1, which it
    // holds in registers until the left end's share has arrived
    double ld1[NE];
    double b1v = 0.0;
    {
        double le[NE];
#pragma unroll
        for (int u = 0; u < NE; ++u) le[u] = a.E[en * SS + tid + 256 * u];
        cr_load_block<P>(a, e, false, A, Wl);
        if (tid < S) vb[tid] = a.b[(long long)e * S + tid];
#pragma unroll
        for (int u = 0; u < NE; ++u) ld1[u] = tk == 1 ? a.D[r1 * SS + tid + 256 * u] : 0.0;
        if (tk == 1 && tid < S) b1v = a.b[(long long)r1 * S + tid];
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int q = tid + 256 * u, r = q / S, c = q - (q / S) * S;
            Mb[r][c] = le[u];
        }
    }
    __syncthreads();
    bool bad = cr_factor<P>(A, Wl);
    __syncthreads();
    {
        const double y = cr_matvec<P>(Wl, vb, false, tid);
        if ((tid & 3) == 0 && (tid >> 2) < S) vy[tid >> 2] = y;
    }
    if (tk == 0) {
        // V_0 = W_0 E_1^T from registers into A (L_0 is no longer needed)
        if (w < P) {
            const int c = w, n = 16 * c + (lane & 15);
            static_for<P>([&](auto Rr) {
                constexpr int r = decltype(Rr)::value;
                double wa[r + 1][4], eb[r + 1][4];
#pragma unroll
                for (int m = 0; m <= r; ++m)
#pragma unroll
                    for (int s4 = 0; s4 < 4; ++s4) {
                        const int kk = 16 * m + 4 * s4 + (lane >> 4);
                        wa[m][s4] = Wl[16 * r + (lane & 15)][kk];
                        eb[m][s4] = Mb[n][kk];
                    }
                d4 acc = 0.0;
#pragma unroll
                for (int m = 0; m <= r; ++m)
#pragma unroll
                    for (int s4 = 0; s4 < 4; ++s4) acc = mfma(wa[m][s4], eb[m][s4], acc);
#pragma unroll
                for (int v = 0; v < 4; ++v) A[16 * r + (lane >> 4) + 4 * v][n] = acc[v];
            });
        }
        __syncthreads();
        cr_gram<P>(A, Mb, w, lane);   // C_0 into Mb
        {
            const double c0 = cr_matvec<P>(A, vy, true, tid);
            if ((tid & 3) == 0 && (tid >> 2) < S) vc[tid >> 2] = c0;
        }
        __syncthreads();
        for (int q = tid; q < S * S; q += 256) {
            const int r = q / S, c = q - (q / S) * S;
            xput(a.U + r0 * SS + q, Mb[r][c]);
        }
        if (tid < S) xput(a.b + (long long)r0 * S + tid, vc[tid]);
        if (bad && tid == 0) a.ctl->chol_ok = 0;
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __syncthreads();
        if (tid == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // x_0 = W_0^T (y_0 - V_0 x_1): W_0 in Wl, V_0 in A, y_0 in vy
        if (tid == 0) cr_wait(flagx, a.ctl);
        __syncthreads();
        if (tid < S) vx[tid] = xget(a.x + S * r1 + tid);   // rows of r1 lie below S r2 < Pp
        __syncthreads();
        {
            const double vx1 = cr_matvec<P>(A, vx, false, tid);
            if ((tid & 3) == 0 && (tid >> 2) < S) vz[tid >> 2] = vy[tid >> 2] - vx1;
        }
        __syncthreads();
        {
            const double x0 = cr_matvec<P>(Wl, vz, true, tid);
            const int gr = S * r0 + (tid >> 2);
            if ((tid & 3) == 0 && (tid >> 2) < S && gr < a.Pp) a.x[gr] = x0;
        }
        return;
    }
    // workgroup 1: U_2 = W_2 E_2 in place in Mb, then C_2 = U_2^T U_2 into A, c_2
    cr_times_w<P>(Wl, Mb, false, w, lane);
    __syncthreads();
    cr_gram<P>(Mb, A, w, lane);
    {
        const double c2 = cr_matvec<P>(Mb, vy, true, tid);
        if ((tid & 3) == 0 && (tid >> 2) < S) vc[tid >> 2] = c2;
    }
    // W_2 waits in LDS; the rows the back-substitution solves start unsolved
    for (int q = tid; q < S * S; q += 256) {
        const int r = q / S, c = q - (q / S) * S;
        W2s[r][c] = Wl[r][c];
    }
    cr_mark_unsolved<S>(a, L, tid);
    if (tid < S) vz[tid] = vy[tid];   // y_2
    if (tid == 0) cr_wait(flag, a.ctl);
    __syncthreads();
    // D_1 - C_0 - C_2 into A, b_1 - c_0 - c_2 into vb
    {
        double lc[NE];
#pragma unroll
        for (int u = 0; u < NE; ++u) lc[u] = xget(a.U + r0 * SS + tid + 256 * u);
        if (tid < S) vb[tid] = b1v - xget(a.b + (long long)r0 * S + tid) - vc[tid];
#pragma unroll
        for (int u = 0; u < NE; ++u) {
            const int q = tid + 256 * u, r = q / S, c = q - (q / S) * S;
            A[r][c] = (c >> 4) <= (r >> 4) ? ld1[u] - lc[u] - A[r][c] : 0.0;
            Wl[r][c] = 0.0;
        }
    }
    __syncthreads();
    bad = cr_factor<P>(A, Wl) || bad;
    __syncthreads();
    {
        const double y1 = cr_matvec<P>(Wl, vb, false, tid);
        if ((tid & 3) == 0 && (tid >> 2) < S) vy[tid >> 2] = y1;
    }
    __syncthreads();
    {
        // x_1, published to workgroup 0 with the second flag
        const double x1 = cr_matvec<P>(Wl, vy, true, tid);
        if ((tid & 3) == 0 && (tid >> 2) < S) {
            vx[tid >> 2] = x1;
            xput(a.x + S * r1 + (tid >> 2), x1);   // rows of r1 lie below S r2 < Pp
        }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (tid == 0) __hip_atomic_store(flagx, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // x_2 = W_2^T (y_2 - U_2 x_1) (U_2 is still in Mb)
    {
        const double ux = cr_matvec<P>(Mb, vx, false, tid);
        if ((tid & 3) == 0 && (tid >> 2) < S) vz[tid >> 2] -= ux;
    }
    __syncthreads();
    {
        const double x2 = cr_matvec<P>(W2s, vz, true, tid);
        const int gr = S * r2 + (tid >> 2);
        if ((tid & 3) == 0 && (tid >> 2) < S && gr < a.Pp) a.x[gr] = x2;
    }
    if (bad && tid == 0) a.ctl->chol_ok = 0;
}

int cr_max_depth() { return 4; }

size_t cr_work_doubles(int Pp, int D) {
    const size_t S = 16 * (size_t)D, ns = (Pp + S - 1) / S;
    return ns * (5 * S * S + 2 * S) + (ns + 7) / 2 + 1;   // the blocks, then the flags and tickets
}

// super-rows left to the sequential tail: each level of the reduction costs
// about as much as two or three tail steps, so the levels stop at this many
constexpr int CR_TAIL_ROWS = 3;

template <int P>
static bool launch_cr_p(const CrArgs& a, const PoseStep* ps, hipStream_t s) {
    int L = 0;
    while (((a.ns + (1 << L) - 1) >> L) > CR_TAIL_ROWS) ++L;
    for (int l = 0; l < L; ++l) {
        const int st = 1 << l;
        const int n_elim = (a.ns - st + 2 * st - 1) / (2 * st);
        const int n_surv = (a.ns + 2 * st - 1) / (2 * st);
        hipLaunchKernelGGL(k_cr_factor<P>, dim3(n_elim), dim3(256), 0, s, a, l);
        hipLaunchKernelGGL(k_cr_update<P>, dim3(n_surv, 2 * P + 1), dim3(256), 0, s, a, l);
    }
    if (L >= 1 && ((a.ns + (1 << L) - 1) >> L) == 3)
        hipLaunchKernelGGL(k_cr_tail3<P>, dim3(2), dim3(256), 0, s, a, L);
    else
        hipLaunchKernelGGL(k_cr_tail<P>, dim3(1), dim3(256), 0, s, a, L);
    int n_back = 0;
    for (int l = 0; l < L; ++l) {
        const int st = 1 << l;
        n_back += (a.ns - st + 2 * st - 1) / (2 * st);
    }
    if (n_back == 0) return false;
    hipLaunchKernelGGL(k_cr_back<P>, dim3(n_back), dim3(256), 0, s, a, L, ps ? *ps : PoseStep{}, ps ? 1 : 0);
    return ps != nullptr;
}

bool launch_cr(const double* H, const double* g, double* x, int Pp, int D, double* work, Control* ctl, int gated,
               const PoseStep* ps, hipStream_t s) {
    CrArgs a;
    const size_t S = 16 * (size_t)D;
    a.ns = (int)((Pp + S - 1) / S);
    const size_t ns = a.ns;
    a.H = H;
    a.g = g;
    a.x = x;
    a.Pp = Pp;
    a.ctl = ctl;
    a.gate = gated;
    a.D = work;
    a.E = a.D + ns * S * S;
    a.W = a.E + ns * S * S;
    a.U = a.W + ns * S * S;
    a.V = a.U + ns * S * S;
    a.b = a.V + ns * S * S;
    a.y = a.b + ns * S;
    a.sync = reinterpret_cast<int*>(a.y + ns * S);
    switch (D) {
    case 1: return launch_cr_p<1>(a, ps, s);
    case 2: return launch_cr_p<2>(a, ps, s);
    case 3: return launch_cr_p<3>(a, ps, s);
    default: return launch_cr_p<4>(a, ps, s);
    }
}

}  // namespace spgo
