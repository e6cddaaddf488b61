// Kernels of the Bermudan-swaption exposure path (gfx950):
//   berm_sim_kernel        exact HW1F step + trapezoid log-numeraire per path (FP64, Philox or
//                          host draws), writing R, LB, BN = exp(LB) and the standardised rate X.
//   berm_policy_*          one backward step of the Longstaff-Schwartz policy for every trade at
//                          once: moments of the ITM paths, 4x4 solve, exercise update.
//   berm_apply_kernel      forward application of the fitted policy to fresh paths.
//   berm_judge_*           per (trade, date) cubic regression of the pathwise value over the
//                          paths still alive, and its exposure.
//   mlp_*                  the continuation network: BF16 fragment packing, forward+backward
//                          training step, Adam, and the exposure / error evaluation.
// Every reduction runs in a fixed order (wave butterflies, then waves in index order, then chunks
// in index order), so results do not depend on scheduling.
#include "eqhip_berm_kernels.hpp"
#include "eqhip_device.hpp"

namespace eqhip {
namespace berm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef short short4v __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ float group16_sum(float v) {  // over the 16 lanes of one lane group
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Block sum of M per-thread values (256 threads), written to out[0..M).
template <int M>
__device__ __forceinline__ void block_sum_store(double (&acc)[M], double* out) {
    __shared__ double red[kWaves][M];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < M; ++q) acc[q] = wave_sum(acc[q]);
    if (l == 0) {
#pragma unroll
        for (int q = 0; q < M; ++q) red[w][q] = acc[q];
    }
    __syncthreads();
    if (threadIdx.x < M) {
        double s = red[0][threadIdx.x];
        s += red[1][threadIdx.x];
        s += red[2][threadIdx.x];
        s += red[3][threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// N [(1 - P(T, T+rem)) - K sum_{m=1..rem} P(T, T+m)], bonds from the A(m), B(m) table.
__device__ __forceinline__ double exercise_value(const double2* AB, double r, int rem, double N, double K) {
#pragma clang fp contract(off)
    double ann = 0.0, pl = 0.0;
    for (int m = 1; m <= rem; ++m) {
        const double2 v = AB[m];
        pl = v.x * exp(-(v.y * r));
        ann += pl;
    }
    return N * ((1.0 - pl) - K * ann);
}

// Gaussian elimination with partial pivoting on the 4x4 normal equations; 0 if singular.
__device__ int solve4(double M[4][4], const double v[4], double c[4]) {
#pragma clang fp contract(off)
    double A[4][5];
    for (int i = 0; i < 4; ++i) {
        for (int j = 0; j < 4; ++j) A[i][j] = M[i][j];
        A[i][4] = v[i];
    }
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int i = col + 1; i < 4; ++i)
            if (fabs(A[i][col]) > fabs(A[piv][col])) piv = i;
        if (fabs(A[piv][col]) < 1e-300) return 0;
        for (int j = 0; j < 5; ++j) {
            double t = A[col][j];
            A[col][j] = A[piv][j];
            A[piv][j] = t;
        }
        for (int i = col + 1; i < 4; ++i) {
            double f = A[i][col] / A[col][col];
            for (int j = col; j < 5; ++j) A[i][j] -= f * A[col][j];
        }
    }
    for (int i = 3; i >= 0; --i) {
        double s = A[i][4];
        for (int j = i + 1; j < 4; ++j) s -= A[i][j] * c[j];
        c[i] = s / A[i][i];
    }
    return 1;
}

// Moment sums [count, x, x^2, x^3, xx, x x2, x x3, x2 x2, x2 x3, x3 x3, y, x y, x2 y, x3 y] of one
// observation; the oracle accumulates phi_i phi_k with phi = (1, x, x*x, (x*x)*x).
__device__ __forceinline__ void add_moments(double (&acc)[14], double x, double yy) {
#pragma clang fp contract(off)
    const double x2 = x * x, x3 = x2 * x;
    acc[0] += 1.0;
    acc[1] += x;
    acc[2] += x2;
    acc[3] += x3;
    acc[4] += x * x;
    acc[5] += x * x2;
    acc[6] += x * x3;
    acc[7] += x2 * x2;
    acc[8] += x2 * x3;
    acc[9] += x3 * x3;
    acc[10] += yy;
    acc[11] += x * yy;
    acc[12] += x2 * yy;
    acc[13] += x3 * yy;
}

__device__ int solve_moments(const double* s, double c[4]) {
    double M[4][4] = {{s[0], s[1], s[2], s[3]}, {s[1], s[4], s[5], s[6]}, {s[2], s[5], s[7], s[8]},
                      {s[3], s[6], s[8], s[9]}};
    const double v[4] = {s[10], s[11], s[12], s[13]};
    return solve4(M, v, c);
}

__device__ __forceinline__ double cubic(const double* c, double x) {
#pragma clang fp contract(off)
    return c[0] + x * (c[1] + x * (c[2] + x * c[3]));
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Simulation: r' = (r e + b(1-e)) + sd z; LB' = LB + (r + r') dt/2 (hull_white.rs:114-120 uses the
// same trapezoid for the path discount factor); X = (r - m_t) / s_t with the exact moments of r_t.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) berm_sim_kernel(SimArgsB g) {
#pragma clang fp contract(off)
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n_paths) return;
    const uint64_t n = g.n_paths;
    double r = g.r0, acc = 0.0, z_next = 0.0;
    for (uint32_t s = 0;; ++s) {
        const uint64_t o = (uint64_t)s * n + p;
        const double t = (double)s / g.spy;
        const double m = g.b + (g.r0 - g.b) * exp(-g.a * t);
        const double sd = g.sigma * sqrt((1.0 - exp(-2.0 * g.a * t)) / (2.0 * g.a));
        g.R[o] = r;
        g.LB[o] = acc;
        g.BN[o] = exp(acc);
        g.X[o] = (s == 0) ? 0.0 : (r - m) / sd;
        if (s == g.n_steps) break;
        double z;
        if (g.Z) {
            z = g.Z[(uint64_t)s * n + p];
        } else if ((s & 1u) == 0u) {
            philox_normal_pair(g.seed, p, s >> 1, z, z_next);
        } else {
            z = z_next;
        }
        const double rn = (r * g.decay + g.mean_c) + g.sd * z;
        acc += (r + rn) * g.half_dt;
        r = rn;
    }
}

// A(m), B(m) of the m-year bond (models/hull_white.rs:50-80), m = 1..max_ex.
__global__ void berm_ab_kernel(int max_ex, double a, double b, double sigma, double2* out) {
#pragma clang fp contract(off)
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > max_ex) return;
    if (m == 0) {
        out[0] = make_double2(1.0, 0.0);
        return;
    }
    const double tau = (double)m;
    const double bf = hw_b(a, tau);
    out[m] = make_double2(hw_a(a, b, sigma, tau, bf), bf);
}

// ---------------------------------------------------------------------------------------------
// Longstaff-Schwartz policy, backward step bi (exercise index e = n_ex - 1 - bi of every trade
// with n_ex > bi, swap years left = bi + 1). Grid (chunks, trades).
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) berm_policy_moments(Book bk, PathSet ps, int bi, double* part) {
#pragma clang fp contract(off)
    const int j = blockIdx.y, c = blockIdx.x;
    const int L = bk.n_ex[j];
    if (bi >= L) return;  // uniform over the block
    const int T = bk.first_ex[j] + (L - 1 - bi);
    const uint64_t n = ps.n, row = (uint64_t)T * bk.spy * n;
    const double N = bk.notional[j], K = bk.strike[j];
    double acc[14];
#pragma unroll
    for (int q = 0; q < 14; ++q) acc[q] = 0.0;
    const uint64_t p0 = (uint64_t)c * kChunk;
    const uint64_t p1 = (p0 + kChunk < n) ? p0 + kChunk : n;
    for (uint64_t p = p0 + threadIdx.x; p < p1; p += kThreads) {
        const double u = exercise_value(bk.AB, ps.R[row + p], bi + 1, N, K);
        if (!(u > 0.0)) continue;
        add_moments(acc, ps.X[row + p], ps.y0[(uint64_t)j * n + p] * ps.BN[row + p]);
    }
    block_sum_store<14>(acc, part + ((uint64_t)j * gridDim.x + c) * kMoments);
}

__global__ void berm_policy_solve(Book bk, int bi, int n_chunks, const double* part, double* ex_coef) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= bk.n_trades) return;
    const int L = bk.n_ex[j];
    if (bi >= L) return;
    const int e = L - 1 - bi;
    double s[14];
    for (int q = 0; q < 14; ++q) s[q] = 0.0;
    for (int c = 0; c < n_chunks; ++c)
        for (int q = 0; q < 14; ++q) s[q] += part[((uint64_t)j * n_chunks + c) * kMoments + q];
    double co[4] = {0.0, 0.0, 0.0, 0.0};
    const int fit = (e < L - 1) && s[0] >= 8.0 && solve_moments(s, co);
    for (int i = 0; i < 4; ++i) ex_coef[((uint64_t)j * bk.max_ex + e) * 4 + i] = fit ? co[i] : 0.0;
}

__global__ void __launch_bounds__(kThreads) berm_policy_update(Book bk, PathSet ps, int bi, const double* ex_coef) {
#pragma clang fp contract(off)
    const int j = blockIdx.y, c = blockIdx.x;
    const int L = bk.n_ex[j];
    if (bi >= L) return;
    const int e = L - 1 - bi;
    const int T = bk.first_ex[j] + e;
    const uint64_t n = ps.n, row = (uint64_t)T * bk.spy * n;
    const double N = bk.notional[j], K = bk.strike[j];
    const double* co = ex_coef + ((uint64_t)j * bk.max_ex + e) * 4;
    const double c0 = co[0], c1 = co[1], c2 = co[2], c3 = co[3];
    const uint64_t p0 = (uint64_t)c * kChunk;
    const uint64_t p1 = (p0 + kChunk < n) ? p0 + kChunk : n;
    for (uint64_t p = p0 + threadIdx.x; p < p1; p += kThreads) {
        const double u = exercise_value(bk.AB, ps.R[row + p], bi + 1, N, K);
        if (!(u > 0.0)) continue;
        const double x = ps.X[row + p];
        const double cont = c0 + x * (c1 + x * (c2 + x * c3));
        if (u > cont) {
            ps.tau[(uint64_t)j * n + p] = (int8_t)T;
            ps.y0[(uint64_t)j * n + p] = u * exp(-ps.LB[row + p]);
        }
    }
}

// Fitted policy on other paths: first exercise date where U > 0 and U > C(x).
__global__ void __launch_bounds__(kThreads) berm_apply_kernel(Book bk, PathSet ps, const double* ex_coef) {
#pragma clang fp contract(off)
    const int j = blockIdx.y;
    const uint64_t n = ps.n;
    const uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    if (p >= n) return;
    const int E = bk.first_ex[j], L = bk.n_ex[j];
    const double N = bk.notional[j], K = bk.strike[j];
    int t_ex = -1;
    double yy = 0.0;
    for (int e = 0; e < L; ++e) {
        const int T = E + e;
        const uint64_t o = (uint64_t)T * bk.spy * n + p;
        const double u = exercise_value(bk.AB, ps.R[o], L - e, N, K);
        if (!(u > 0.0)) continue;
        const double cont = cubic(ex_coef + ((uint64_t)j * bk.max_ex + e) * 4, ps.X[o]);
        if (u > cont) {
            t_ex = T;
            yy = u * exp(-ps.LB[o]);
            break;
        }
    }
    ps.tau[(uint64_t)j * n + p] = (int8_t)t_ex;
    ps.y0[(uint64_t)j * n + p] = yy;
}

// Chunk sums of y0 (the price is their total over n). Grid (chunks, trades).
__global__ void __launch_bounds__(kThreads) berm_sum_y0(Book bk, PathSet ps, double* part) {
    const int j = blockIdx.y, c = blockIdx.x;
    const uint64_t n = ps.n;
    const uint64_t p0 = (uint64_t)c * kChunk;
    const uint64_t p1 = (p0 + kChunk < n) ? p0 + kChunk : n;
    double acc[1] = {0.0};
    for (uint64_t p = p0 + threadIdx.x; p < p1; p += kThreads) acc[0] += ps.y0[(uint64_t)j * n + p];
    block_sum_store<1>(acc, part + ((uint64_t)j * gridDim.x + c));
}

// ---------------------------------------------------------------------------------------------
// Judge: continuation regression per (trade, date k) over alive paths (not exercised before t_k),
// target y0 * BN; then EE = (1/n) sum_alive max(C(x), 0). Grid (chunks, trades), dates looped.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) berm_judge_moments(Book bk, PathSet ps, double* part) {
#pragma clang fp contract(off)
    const int j = blockIdx.y, c = blockIdx.x;
    const uint64_t n = ps.n;
    const uint64_t p0 = (uint64_t)c * kChunk;
    const uint64_t p1 = (p0 + kChunk < n) ? p0 + kChunk : n;
    const int K = bk.kmax[j];
    for (int k = 1; k <= K; ++k) {
        double acc[14];
#pragma unroll
        for (int q = 0; q < 14; ++q) acc[q] = 0.0;
        const uint64_t row = (uint64_t)k * n;
        for (uint64_t p = p0 + threadIdx.x; p < p1; p += kThreads) {
            const int tv = ps.tau[(uint64_t)j * n + p];
            if (tv >= 0 && tv * bk.spy < k) continue;
            add_moments(acc, ps.X[row + p], ps.y0[(uint64_t)j * n + p] * ps.BN[row + p]);
        }
        block_sum_store<14>(acc, part + (((uint64_t)j * bk.n_steps + (k - 1)) * gridDim.x + c) * kMoments);
    }
}

__global__ void berm_judge_solve(Book bk, int n_chunks, const double* part, double* coef) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= bk.n_trades * bk.n_steps) return;
    const int j = row / bk.n_steps, k = row % bk.n_steps + 1;
    double co[4] = {0.0, 0.0, 0.0, 0.0};
    if (k <= bk.kmax[j]) {
        double s[14];
        for (int q = 0; q < 14; ++q) s[q] = 0.0;
        for (int c = 0; c < n_chunks; ++c)
            for (int q = 0; q < 14; ++q) s[q] += part[((uint64_t)row * n_chunks + c) * kMoments + q];
        if (!(s[0] >= 8.0 && solve_moments(s, co))) co[0] = co[1] = co[2] = co[3] = 0.0;
    }
    for (int i = 0; i < 4; ++i) coef[(uint64_t)row * 4 + i] = co[i];
}

__global__ void __launch_bounds__(kThreads) berm_judge_ee(Book bk, PathSet ps, const double* coef, double* part) {
#pragma clang fp contract(off)
    const int j = blockIdx.y, c = blockIdx.x;
    const uint64_t n = ps.n;
    const uint64_t p0 = (uint64_t)c * kChunk;
    const uint64_t p1 = (p0 + kChunk < n) ? p0 + kChunk : n;
    const int K = bk.kmax[j];
    for (int k = 1; k <= K; ++k) {
        const double* co = coef + ((uint64_t)j * bk.n_steps + (k - 1)) * 4;
        const double c0 = co[0], c1 = co[1], c2 = co[2], c3 = co[3];
        double acc[1] = {0.0};
        const uint64_t row = (uint64_t)k * n;
        for (uint64_t p = p0 + threadIdx.x; p < p1; p += kThreads) {
            const int tv = ps.tau[(uint64_t)j * n + p];
            if (tv >= 0 && tv * bk.spy < k) continue;
            const double x = ps.X[row + p];
            const double v = c0 + x * (c1 + x * (c2 + x * c3));
            acc[0] += (v > 0.0) ? v : 0.0;
        }
        block_sum_store<1>(acc, part + ((uint64_t)j * bk.n_steps + (k - 1)) * gridDim.x + c);
    }
}

// out[row] = (sum_c part[row][c][0]) / n for the valid rows of [n_trades][n_rows_per_trade]
// (kmax null: every row valid), times row_scale[trade] when given.
__global__ void berm_finalize(int n_rows, int n_chunks, const double* part, int stride, double n_div,
                              const double* trade_scale, int rows_per_trade, const int* kmax, double* out) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n_rows) return;
    const int j = row / rows_per_trade, k = row % rows_per_trade + 1;
    if (kmax && k > kmax[j]) {
        out[row] = 0.0;
        return;
    }
    double s = 0.0;
    for (int c = 0; c < n_chunks; ++c) s += part[((uint64_t)row * n_chunks + c) * stride];
    double v = s / n_div;
    if (trade_scale) v = v * trade_scale[j];
    out[row] = v;
}

// ---------------------------------------------------------------------------------------------
// Continuation network.
// ---------------------------------------------------------------------------------------------
namespace {

// r = 1 / (1 + 2^zs) for zs = 2 log2(e) z, so that tanh(z) = 1 - 2r: the forward operands carry the
// factor 2 log2(e), and the next layer absorbs the affine map r -> 1 - 2r (weights -2W, bias
// b + W 1), leaving the hardware exp2, one add and the reciprocal per activation. Saturates to
// r = 0 / 1 (tanh = +-1) when 2^zs overflows / underflows.
__device__ __forceinline__ float act_r(float zs) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(zs));
}

__device__ __forceinline__ bf16x8 as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ bf16x8 pack_b(const f32x4& lo, const f32x4& hi) {
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[i] = (__bf16)lo[i];
        v[4 + i] = (__bf16)hi[i];
    }
    return v;
}

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(short4v, a),
                                                    __builtin_bit_cast(short4v, b), c, 0, 0, 0);
}

// Forward operands of one network held in registers by every lane of a wave.
// Forward operands of one network held in registers by every lane of a wave: the scaled BF16
// weight fragments and the folded FP32 vectors (b2', b3', w4' = -2 w4, b4').
struct FwdW {
    bf16x8 w1[4], w2[8], w3[8];
    f32x4 b2[4], b3[4], w4[4];
    float b4;
};

__device__ __forceinline__ void load_fwd(FwdW& W, const uint4* fr, const float* fv, int l) {
    const int g = l >> 4;
#pragma unroll
    for (int f = 0; f < 4; ++f) W.w1[f] = as_bf8(fr[(kFragW1 + f) * 64 + l]);
#pragma unroll
    for (int f = 0; f < 8; ++f) {
        W.w2[f] = as_bf8(fr[(kFragW2 + f) * 64 + l]);
        W.w3[f] = as_bf8(fr[(kFragW3 + f) * 64 + l]);
    }
#pragma unroll
    for (int R = 0; R < 4; ++R)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int f = 16 * R + 4 * g + i;
            W.b2[R][i] = fv[FV_B2 + f];
            W.b3[R][i] = fv[FV_B3 + f];
            W.w4[R][i] = fv[FV_W4 + f];
        }
    W.b4 = fv[FV_B4];
}

// One 16-sample tile: bx is layer 1's B operand (lane group 0 holds u0, u1, u2, 1 of sample l%16).
// Returns y of sample l%16 in every lane; r1..r3 are the FP32 activations r = (1 - tanh) / 2 in
// accumulator layout.
__device__ __forceinline__ float mlp_fwd(const FwdW& W, bf16x8 bx, f32x4 (&r1)[4], f32x4 (&r2)[4],
                                         f32x4 (&r3)[4]) {
#pragma unroll
    for (int R = 0; R < 4; ++R) {
        f32x4 z = {0.f, 0.f, 0.f, 0.f};
        z = mfma32(W.w1[R], bx, z);
#pragma unroll
        for (int i = 0; i < 4; ++i) r1[R][i] = act_r(z[i]);
    }
    bf16x8 q0 = pack_b(r1[0], r1[1]), q1 = pack_b(r1[2], r1[3]);
#pragma unroll
    for (int R = 0; R < 4; ++R) {
        f32x4 z = W.b2[R];
        z = mfma32(W.w2[2 * R], q0, z);
        z = mfma32(W.w2[2 * R + 1], q1, z);
#pragma unroll
        for (int i = 0; i < 4; ++i) r2[R][i] = act_r(z[i]);
    }
    q0 = pack_b(r2[0], r2[1]);
    q1 = pack_b(r2[2], r2[3]);
#pragma unroll
    for (int R = 0; R < 4; ++R) {
        f32x4 z = W.b3[R];
        z = mfma32(W.w3[2 * R], q0, z);
        z = mfma32(W.w3[2 * R + 1], q1, z);
#pragma unroll
        for (int i = 0; i < 4; ++i) r3[R][i] = act_r(z[i]);
    }
    float part = 0.f;
#pragma unroll
    for (int R = 0; R < 4; ++R)
#pragma unroll
        for (int i = 0; i < 4; ++i) part += W.w4[R][i] * r3[R][i];
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    return part + W.b4;
}

__device__ __forceinline__ bf16x8 input_operand(int g, float u0, float u1, float u2) {
    bf16x8 bx;
#pragma unroll
    for (int i = 0; i < 8; ++i) bx[i] = (__bf16)0.0f;
    if (g == 0) {
        bx[0] = (__bf16)u0;
        bx[1] = (__bf16)u1;
        bx[2] = (__bf16)u2;
        bx[3] = (__bf16)1.0f;
    }
    return bx;
}

// acc[Ro*4+Ci] += dz[out block Ro] (x) hp[in block Ci] over the tile's 16 samples, through a
// wave-private LDS transpose (features x samples, BF16) and v_mfma_f32_16x16x16_bf16.
template <int NCI>
__device__ __forceinline__ void outer_acc(const f32x4 (&dz)[4], const f32x4* hp, __bf16* sZ, const __bf16* sH,
                                          f32x4 (&acc)[4 * NCI], int l, bool write_h) {
    const int g = l >> 4, n = l & 15;
#pragma unroll
    for (int R = 0; R < 4; ++R)
#pragma unroll
        for (int i = 0; i < 4; ++i) sZ[(16 * R + 4 * g + i) * 16 + n] = (__bf16)dz[R][i];
    if (write_h) {
        __bf16* sHw = const_cast<__bf16*>(sH);
#pragma unroll
        for (int R = 0; R < 4; ++R)
#pragma unroll
            for (int i = 0; i < 4; ++i) sHw[(16 * R + 4 * g + i) * 16 + n] = (__bf16)hp[R][i];
    }
    __syncthreads();
    bf16x4 A[4], B[NCI];
#pragma unroll
    for (int R = 0; R < 4; ++R) A[R] = *reinterpret_cast<const bf16x4*>(sZ + (16 * R + n) * 16 + 4 * g);
#pragma unroll
    for (int C = 0; C < NCI; ++C) B[C] = *reinterpret_cast<const bf16x4*>(sH + (16 * C + n) * 16 + 4 * g);
#pragma unroll
    for (int R = 0; R < 4; ++R)
#pragma unroll
        for (int C = 0; C < NCI; ++C) acc[R * NCI + C] = mfma16(A[R], B[C], acc[R * NCI + C]);
    __syncthreads();
}

}  // namespace

// Seeded Glorot-uniform initial weights, zero biases (the layer-1 bias column is drawn like W1).
__global__ void mlp_init_kernel(int n_nets, uint64_t seed, float* master) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n_nets * kParamStride) return;
    const int q = (int)(i % kParamStride);
    const uint64_t net = i / kParamStride;
    uint32_t o[4];
    philox4x32_10((uint32_t)q, (uint32_t)net, 0x5EEDu, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), o);
    const float u = ((float)(o[0] >> 8) + 0.5f) * (1.0f / 16777216.0f) * 2.0f - 1.0f;  // (-1, 1)
    float v = 0.0f;
    if (q < OFF_W2) v = u * 1.2f;                                   // fan-in 3 on inputs of scale ~1
    else if (q < OFF_B2) v = u * 0.2165f;                           // sqrt(6 / 128)
    else if (q >= OFF_W3 && q < OFF_B3) v = u * 0.2165f;
    else if (q >= OFF_W4 && q < OFF_B4) v = u * 0.3038f;            // sqrt(6 / 65)
    master[i] = (q < kNParams) ? v : 0.0f;
}

// BF16 operand fragments of every network; grid (kNFrag, n_nets), 64 threads.
__global__ void mlp_pack_kernel(const float* master, uint4* frags) {
    const int net = blockIdx.y, f = blockIdx.x, l = threadIdx.x;
    const int m = l & 15, g = l >> 4;
    const float* th = master + (size_t)net * kParamStride;
    bf16x8 v;
    if (f < kFragW2) {  // layer 1 (incl. its bias column), scaled by 2 log2(e)
#pragma unroll
        for (int i = 0; i < 8; ++i)
            v[i] = (__bf16)((g == 0 && i < 4) ? kActScale * th[OFF_W1 + (16 * f + m) * 4 + i] : 0.0f);
    } else {
        const int idx = f - kFragW2;
        const int mat = idx >> 3, R = (idx & 7) >> 1, Q = idx & 1;
        const int off = (mat == 0 || mat == 2) ? OFF_W2 : OFF_W3;
        const bool trans = mat >= 2;  // backward operands W^T are unscaled
        const float scale = trans ? 1.0f : -2.0f * kActScale;
        const int row = 16 * R + m;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int kf = 32 * Q + (i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4));
            v[i] = (__bf16)(scale * (trans ? th[off + kf * kH + row] : th[off + row * kH + kf]));
        }
    }
    frags[((size_t)net * kNFrag + f) * 64 + l] = __builtin_bit_cast(uint4, v);
}

// Folded FP32 forward vectors of every network: b' = 2 log2(e) (b + W 1) for layers 2 and 3,
// w4' = -2 w4, b4' = b4 + sum(w4). Grid n_nets, 256 threads.
__global__ void mlp_fold_kernel(const float* master, float* fwdvec) {
    const int net = blockIdx.x, t = threadIdx.x;
    const float* th = master + (size_t)net * kParamStride;
    float* fv = fwdvec + (size_t)net * kFwdStride;
    if (t < 128) {
        const int f = t & 63;
        const int ow = (t < 64) ? OFF_W2 : OFF_W3, ob = (t < 64) ? OFF_B2 : OFF_B3;
        float rs = 0.0f;
        for (int k = 0; k < kH; ++k) rs += th[ow + f * kH + k];
        fv[(t < 64 ? FV_B2 : FV_B3) + f] = kActScale * (th[ob + f] + rs);
    } else if (t < 192) {
        fv[FV_W4 + (t - 128)] = -2.0f * th[OFF_W4 + (t - 128)];
    } else if (t == 192) {
        float s = 0.0f;
        for (int k = 0; k < kH; ++k) s += th[OFF_W4 + k];
        fv[FV_B4] = th[OFF_B4] + s;
    }
}

// One training step of every network (one block each): forward, backward and the weight-gradient
// partials of each wave. SRC 0 draws (path, date) samples of the training set with Philox
// (counter = sample, step, net), SRC 1 reads explicit samples.
template <int SRC>
__global__ void __launch_bounds__(kThreads) mlp_train_kernel(Book bk, PathSet ps, Norm nm, const uint4* frags,
                                                             const float* fwdvec, int tiles_per_wave, uint32_t step,
                                                             uint64_t seed, const float* xu, const float* xt,
                                                             const float* xm, float* grad_part, float* loss_part) {
    const int j = blockIdx.x;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, n16 = l & 15;
    __shared__ uint4 sWT[16 * 64];                        // W2^T then W3^T fragments
    // 16-byte aligned: the transposed tiles are read back as 8-byte vectors
    __shared__ __attribute__((aligned(16))) __bf16 sZ[kWaves][kH * 16];
    __shared__ __attribute__((aligned(16))) __bf16 sH[kWaves][kH * 16];
    __shared__ __attribute__((aligned(16))) __bf16 sU[kWaves][16 * 16];
    const uint4* fr = frags + (size_t)j * kNFrag * 64;
    for (int q = threadIdx.x; q < 16 * 64; q += kThreads) sWT[q] = fr[kFragW2T * 64 + q];
    for (int q = l; q < 16 * 16; q += 64) sU[w][q] = (__bf16)0.0f;
    FwdW W;
    load_fwd(W, fr, fwdvec + (size_t)j * kFwdStride, l);
    f32x4 dW1[4], dW2[16], dW3[16], gb2[4], gb3[4], gw4[4];
    const f32x4 zero4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 16; ++q) dW2[q] = dW3[q] = zero4;
#pragma unroll
    for (int q = 0; q < 4; ++q) dW1[q] = gb2[q] = gb3[q] = gw4[q] = zero4;
    float gb4 = 0.f, lsum = 0.f;
    const int batch = tiles_per_wave * 16 * kWaves;
    const float inv_b2 = 2.0f / (float)batch;
    const uint64_t n = ps.n;
    // SRC 0: a sample is a Philox-drawn (path, date) of the training set, gathered from HBM one tile
    // ahead, so that the gathers of tile tt + 1 are in flight during tile tt's forward and backward pass
    // (a block has one wave per SIMD, with no other wave to hide their latency)
    struct Gather {
        double r, lb, y0, bn;
        int tv, k;
    };
    auto gather = [&](int tt) {
        Gather s;
        const int i = (w * tiles_per_wave + tt) * 16 + n16;
        uint32_t o[4];
        philox4x32_10((uint32_t)i, step, (uint32_t)j, 0x7A11u, (uint32_t)seed, (uint32_t)(seed >> 32), o);
        const uint64_t p = __umul64hi(((uint64_t)o[0] << 32) | o[1], n);
        s.k = 1 + (int)__umulhi(o[2], (uint32_t)bk.kmax[j]);
        const uint64_t ro = (uint64_t)s.k * n + p;
        s.r = ps.R[ro];
        s.lb = ps.LB[ro];
        s.bn = ps.BN[ro];
        s.tv = ps.tau[(uint64_t)j * n + p];
        s.y0 = ps.y0[(uint64_t)j * n + p];
        return s;
    };
    Gather next{};
    if (SRC == 0 && tiles_per_wave > 0) next = gather(0);
    __syncthreads();
    for (int tt = 0; tt < tiles_per_wave; ++tt) {
        const int i = (w * tiles_per_wave + tt) * 16 + n16;
        float u0, u1, u2, tgt, msk;
        if (SRC == 1) {
            const float* q = xu + ((size_t)j * batch + i) * 3;
            u0 = q[0];
            u1 = q[1];
            u2 = q[2];
            tgt = xt[(size_t)j * batch + i];
            msk = xm[(size_t)j * batch + i];
        } else {
#pragma clang fp contract(off)
            const Gather c = next;
            if (tt + 1 < tiles_per_wave) next = gather(tt + 1);
            u0 = (float)((c.r - nm.r_shift) * nm.r_scale);
            u1 = (float)(c.lb * nm.lb_scale + nm.lb_shift);
            u2 = (float)((double)c.k / bk.spy * nm.t_scale + nm.t_shift);
            tgt = (float)(c.y0 * c.bn / bk.notional[j] * nm.y_scale);
            msk = (c.tv < 0 || c.tv * bk.spy >= c.k) ? 1.0f : 0.0f;
        }
        if (g == 0) {
            sU[w][0 * 16 + n16] = (__bf16)u0;
            sU[w][1 * 16 + n16] = (__bf16)u1;
            sU[w][2 * 16 + n16] = (__bf16)u2;
            sU[w][3 * 16 + n16] = (__bf16)1.0f;
        }
        f32x4 h1[4], h2[4], h3[4];
        const float y = mlp_fwd(W, input_operand(g, u0, u1, u2), h1, h2, h3);
#pragma unroll
        for (int R = 0; R < 4; ++R)
#pragma unroll
            for (int q = 0; q < 4; ++q) {  // r -> tanh = 1 - 2r
                h1[R][q] = fmaf(-2.0f, h1[R][q], 1.0f);
                h2[R][q] = fmaf(-2.0f, h2[R][q], 1.0f);
                h3[R][q] = fmaf(-2.0f, h3[R][q], 1.0f);
            }
        const float res = (y - tgt) * msk;
        const float dy = inv_b2 * res;
        if (g == 0) {
            lsum += res * (y - tgt);
            gb4 += dy;
        }
        f32x4 dz[4];
#pragma unroll
        for (int R = 0; R < 4; ++R)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                gw4[R][q] += h3[R][q] * dy;
                dz[R][q] = (-0.5f * W.w4[R][q]) * dy * (1.0f - h3[R][q] * h3[R][q]);  // w4 = -w4'/2
                gb3[R][q] += dz[R][q];
            }
        outer_acc<4>(dz, h2, sZ[w], sH[w], dW3, l, true);
        // dh2 = W3^T dz3, dz2 = dh2 (1 - h2^2)
        {
            const bf16x8 q0 = pack_b(dz[0], dz[1]), q1 = pack_b(dz[2], dz[3]);
            f32x4 nz[4];
#pragma unroll
            for (int R = 0; R < 4; ++R) {
                f32x4 z = zero4;
                z = mfma32(as_bf8(sWT[(8 + 2 * R) * 64 + l]), q0, z);
                z = mfma32(as_bf8(sWT[(8 + 2 * R + 1) * 64 + l]), q1, z);
#pragma unroll
                for (int q = 0; q < 4; ++q) nz[R][q] = z[q] * (1.0f - h2[R][q] * h2[R][q]);
            }
#pragma unroll
            for (int R = 0; R < 4; ++R) {
                dz[R] = nz[R];
                gb2[R] += nz[R];
            }
        }
        outer_acc<4>(dz, h1, sZ[w], sH[w], dW2, l, true);
        // dh1 = W2^T dz2, dz1 = dh1 (1 - h1^2)
        {
            const bf16x8 q0 = pack_b(dz[0], dz[1]), q1 = pack_b(dz[2], dz[3]);
#pragma unroll
            for (int R = 0; R < 4; ++R) {
                f32x4 z = zero4;
                z = mfma32(as_bf8(sWT[(2 * R) * 64 + l]), q0, z);
                z = mfma32(as_bf8(sWT[(2 * R + 1) * 64 + l]), q1, z);
#pragma unroll
                for (int q = 0; q < 4; ++q) dz[R][q] = z[q] * (1.0f - h1[R][q] * h1[R][q]);
            }
        }
        outer_acc<1>(dz, nullptr, sZ[w], sU[w], dW1, l, false);
    }
    // this wave's partial gradient
    float* gp = grad_part + ((size_t)j * kWaves + w) * kParamStride;
#pragma unroll
    for (int R = 0; R < 4; ++R)
#pragma unroll
        for (int C = 0; C < 4; ++C)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int o = (16 * R + 4 * g + q) * kH + 16 * C + n16;
                gp[OFF_W2 + o] = dW2[R * 4 + C][q];
                gp[OFF_W3 + o] = dW3[R * 4 + C][q];
            }
    if (n16 < 4) {
#pragma unroll
        for (int R = 0; R < 4; ++R)
#pragma unroll
            for (int q = 0; q < 4; ++q) gp[OFF_W1 + (16 * R + 4 * g + q) * 4 + n16] = dW1[R][q];
    }
#pragma unroll
    for (int R = 0; R < 4; ++R)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float s2 = group16_sum(gb2[R][q]);
            const float s3 = group16_sum(gb3[R][q]);
            const float s4 = group16_sum(gw4[R][q]);
            if (n16 == 0) {
                const int f = 16 * R + 4 * g + q;
                gp[OFF_B2 + f] = s2;
                gp[OFF_B3 + f] = s3;
                gp[OFF_W4 + f] = s4;
            }
        }
    const float s4b = group16_sum(gb4);
    const float sl = group16_sum(lsum);
    if (l == 0) {
        gp[OFF_B4] = s4b;
        loss_part[(size_t)j * kWaves + w] = sl / (float)batch;
    }
}

template __global__ void mlp_train_kernel<0>(Book, PathSet, Norm, const uint4*, const float*, int, uint32_t, uint64_t,
                                             const float*, const float*, const float*, float*, float*);
template __global__ void mlp_train_kernel<1>(Book, PathSet, Norm, const uint4*, const float*, int, uint32_t, uint64_t,
                                             const float*, const float*, const float*, float*, float*);

// Sums the wave partials in wave order; Adam (Kingma & Ba 2015) on the float32 master weights.
__global__ void mlp_adam_kernel(int n_nets, float* master, float* m1, float* v1, const float* grad_part, float lr,
                                float beta1, float beta2, float eps, float bc1, float bc2, int do_update,
                                float* grad_out, const float* loss_part, float* loss_out) {
#pragma clang fp contract(off)
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n_nets * kParamStride) return;
    const int q = (int)(i % kParamStride);
    const uint64_t net = i / kParamStride;
    if (q == 0 && loss_out) {
        const float* lp = loss_part + net * kWaves;
        loss_out[net] = ((lp[0] + lp[1]) + lp[2]) + lp[3];
    }
    if (q >= kNParams) return;
    const float* gp = grad_part + net * kWaves * kParamStride + q;
    const float gr = ((gp[0] + gp[kParamStride]) + gp[2 * kParamStride]) + gp[3 * kParamStride];
    if (grad_out) grad_out[net * kNParams + q] = gr;
    if (!do_update) return;
    const float m = beta1 * m1[i] + (1.0f - beta1) * gr;
    const float v = beta2 * v1[i] + (1.0f - beta2) * gr * gr;
    m1[i] = m;
    v1[i] = v;
    const float mh = m / bc1, vh = v / bc2;
    master[i] = master[i] - lr * mh / (sqrtf(vh) + eps);
}

// MODE 0: chunk sums of alive * max(y, 0) per (trade, date). MODE 1: squared errors of the network
// and of the polynomial `poly` against the pathwise value over alive paths, and their count.
// Grid (chunks, trades); 4 waves per block, each a quarter of the chunk.
template <int MODE>
__global__ void __launch_bounds__(kThreads) mlp_eval_kernel(Book bk, PathSet ps, Norm nm, const uint4* frags,
                                                            const float* fwdvec, const double* poly, double* part) {
    const int j = blockIdx.y, c = blockIdx.x;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, n16 = l & 15;
    constexpr int NQ = (MODE == 0) ? 1 : 3;
    __shared__ double red[kWaves][NQ];
    FwdW W;
    load_fwd(W, frags + (size_t)j * kNFrag * 64, fwdvec + (size_t)j * kFwdStride, l);
    const uint64_t n = ps.n;
    const uint64_t q0 = (uint64_t)c * kChunk + (uint64_t)w * (kChunk / kWaves);
    const uint64_t q1 = (q0 + kChunk / kWaves < n) ? q0 + kChunk / kWaves : n;
    const int K = bk.kmax[j];
    const double inv_scale = nm.y_scale / bk.notional[j];
    for (int k = 1; k <= K; ++k) {
        double acc[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
        float u2;
        {
#pragma clang fp contract(off)
            u2 = (float)((double)k / bk.spy * nm.t_scale + nm.t_shift);
        }
        double c0 = 0, c1 = 0, c2 = 0, c3 = 0;
        if (MODE == 1) {
            const double* co = poly + ((uint64_t)j * bk.n_steps + (k - 1)) * 4;
            c0 = co[0];
            c1 = co[1];
            c2 = co[2];
            c3 = co[3];
        }
        const uint64_t row = (uint64_t)k * n;
        for (uint64_t base = q0; base < q1; base += 64) {
            const uint64_t p = base + l;
            const bool valid = p < q1;
            int tv = -1;
            double r = nm.r_shift, lb = 0.0;
            if (valid) {
                r = ps.R[row + p];
                lb = ps.LB[row + p];
                tv = ps.tau[(uint64_t)j * n + p];
            }
            const bool alive = valid && (tv < 0 || tv * bk.spy >= k);
            if (!__any(alive)) continue;
            float u0, u1;
            {
#pragma clang fp contract(off)
                u0 = (float)((r - nm.r_shift) * nm.r_scale);
                u1 = (float)(lb * nm.lb_scale + nm.lb_shift);
            }
            float my_y = 0.f;
            if constexpr (MODE == 0) {
                // only the sum over alive paths is needed: move their inputs to the first n_alive lanes
                // (a bijective forward permute, dead lanes behind) and run ceil(n_alive / 16) tiles
                const uint64_t amask = __ballot(alive);
                const int n_alive = __popcll(amask);
                const int rank_alive = __builtin_amdgcn_mbcnt_hi((uint32_t)(amask >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)amask, 0u));
                const int pos = alive ? rank_alive : n_alive + (l - rank_alive);
                const float v0 = __int_as_float(__builtin_amdgcn_ds_permute(pos << 2, __float_as_int(u0)));
                const float v1 = __int_as_float(__builtin_amdgcn_ds_permute(pos << 2, __float_as_int(u1)));
                const int n_sub = (n_alive + 15) >> 4;
                for (int s = 0; s < n_sub; ++s) {
                    const float a0 = __shfl(v0, 16 * s + n16), a1 = __shfl(v1, 16 * s + n16);
                    f32x4 h1[4], h2[4], h3[4];
                    const float y = mlp_fwd(W, input_operand(g, a0, a1, u2), h1, h2, h3);
                    if (g == s) my_y = y;  // lane l holds compact sample l
                }
                if (l < n_alive) acc[0] += (double)fmaxf(my_y, 0.0f);
            } else {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const float a0 = __shfl(u0, 16 * s + n16), a1 = __shfl(u1, 16 * s + n16);
                    f32x4 h1[4], h2[4], h3[4];
                    const float y = mlp_fwd(W, input_operand(g, a0, a1, u2), h1, h2, h3);
                    if (g == s) my_y = y;  // lane l holds sample 16 (l/16) + l%16 = l of this 64-path tile
                }
                if (alive) {
#pragma clang fp contract(off)
                    const double tgt = ps.y0[(uint64_t)j * n + p] * ps.BN[row + p] * inv_scale;
                    const double x = ps.X[row + p];
                    const double cp = (c0 + x * (c1 + x * (c2 + x * c3))) * inv_scale;
                    const double e1 = (double)my_y - tgt, e2 = cp - tgt;
                    acc[0] += e1 * e1;
                    acc[1] += e2 * e2;
                    acc[2] += 1.0;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] = wave_sum(acc[q]);
        if (l == 0)
#pragma unroll
            for (int q = 0; q < NQ; ++q) red[w][q] = acc[q];
        __syncthreads();
        if (threadIdx.x < NQ) {
            const int q = threadIdx.x;
            part[(((uint64_t)j * bk.n_steps + (k - 1)) * gridDim.x + c) * NQ + q] =
                ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
        }
        __syncthreads();
    }
}

template __global__ void mlp_eval_kernel<0>(Book, PathSet, Norm, const uint4*, const float*, const double*, double*);
template __global__ void mlp_eval_kernel<1>(Book, PathSet, Norm, const uint4*, const float*, const double*, double*);

// Outputs on explicit inputs xu [net][n][3] -> y [net][n]; grid (ceil(n/64), nets), one wave per block.
__global__ void __launch_bounds__(64) mlp_forward_kernel(int n, const uint4* frags, const float* fwdvec,
                                                         const float* xu, float* y) {
    const int j = blockIdx.y, l = threadIdx.x, g = l >> 4, n16 = l & 15;
    FwdW W;
    load_fwd(W, frags + (size_t)j * kNFrag * 64, fwdvec + (size_t)j * kFwdStride, l);
    const int p = blockIdx.x * 64 + l;
    float u0 = 0.f, u1 = 0.f, u2 = 0.f;
    if (p < n) {
        const float* q = xu + ((size_t)j * n + p) * 3;
        u0 = q[0];
        u1 = q[1];
        u2 = q[2];
    }
    float my_y = 0.f;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const float a0 = __shfl(u0, 16 * s + n16), a1 = __shfl(u1, 16 * s + n16), a2 = __shfl(u2, 16 * s + n16);
        f32x4 h1[4], h2[4], h3[4];
        const float yy = mlp_fwd(W, input_operand(g, a0, a1, a2), h1, h2, h3);
        if (g == s) my_y = yy;
    }
    if (p < n) y[(size_t)j * n + p] = my_y;
}

// ---------------------------------------------------------------------------------------------
// Lattice evaluation of the continuation networks. At date k every path's network inputs (r, LB)
// lie in the box [r_lo, r_hi] x [lb_lo, lb_hi] of the set's paths at that date. Each network is
// evaluated by MFMA at nr x nl nodes spanning the box (node (i, m) at r_lo + i dr, lb_lo + m dl),
// and a path's output is the bilinear interpolation of the four nodes around its (r, LB): the
// network's work per date is nr nl samples instead of one per alive path.
// ---------------------------------------------------------------------------------------------

// Per-date bounds of R and LB over the set's paths: (r_lo, r_hi, lb_lo, lb_hi) of date k = 1..n_steps.
__global__ void __launch_bounds__(kThreads) berm_bounds_kernel(PathSet ps, double* bounds) {
    const int k = blockIdx.x + 1;
    const uint64_t row = (uint64_t)k * ps.n;
    double v[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (uint64_t p = threadIdx.x; p < ps.n; p += kThreads) {
        const double r = ps.R[row + p], lb = ps.LB[row + p];
        v[0] = fmin(v[0], r);
        v[1] = fmax(v[1], r);
        v[2] = fmin(v[2], lb);
        v[3] = fmax(v[3], lb);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        v[0] = fmin(v[0], __shfl_xor(v[0], o));
        v[1] = fmax(v[1], __shfl_xor(v[1], o));
        v[2] = fmin(v[2], __shfl_xor(v[2], o));
        v[3] = fmax(v[3], __shfl_xor(v[3], o));
    }
    __shared__ double red[kWaves][4];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q) red[w][q] = v[q];
    __syncthreads();
    if (threadIdx.x < 4) {
        const int q = threadIdx.x;
        double m = red[0][q];
        for (int i = 1; i < kWaves; ++i) m = (q & 1) ? fmax(m, red[i][q]) : fmin(m, red[i][q]);
        bounds[(size_t)(k - 1) * 4 + q] = m;
    }
}

namespace {

// Lattice node spacing of one axis (zero for a degenerate box, whose nodes then coincide).
__device__ __forceinline__ double node_step(double lo, double hi, int n) { return (hi - lo) / (double)(n - 1); }

// Two 16-sample tiles through the network, layer by layer, so that one tile's matrix products and
// the other's activations are independent instructions the scheduler can overlap.
__device__ __forceinline__ void mlp_fwd2(const FwdW& W, bf16x8 bxa, bf16x8 bxb, float& ya, float& yb) {
    f32x4 ra[4], rb[4];
#pragma unroll
    for (int R = 0; R < 4; ++R) {
        f32x4 za = {0.f, 0.f, 0.f, 0.f}, zb = {0.f, 0.f, 0.f, 0.f};
        za = mfma32(W.w1[R], bxa, za);
        zb = mfma32(W.w1[R], bxb, zb);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ra[R][i] = act_r(za[i]);
            rb[R][i] = act_r(zb[i]);
        }
    }
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        const bf16x8* Wl = layer == 0 ? W.w2 : W.w3;
        const f32x4* bl = layer == 0 ? W.b2 : W.b3;
        const bf16x8 a0 = pack_b(ra[0], ra[1]), a1 = pack_b(ra[2], ra[3]);
        const bf16x8 b0 = pack_b(rb[0], rb[1]), b1 = pack_b(rb[2], rb[3]);
#pragma unroll
        for (int R = 0; R < 4; ++R) {
            f32x4 za = bl[R], zb = bl[R];
            za = mfma32(Wl[2 * R], a0, za);
            zb = mfma32(Wl[2 * R], b0, zb);
            za = mfma32(Wl[2 * R + 1], a1, za);
            zb = mfma32(Wl[2 * R + 1], b1, zb);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ra[R][i] = act_r(za[i]);
                rb[R][i] = act_r(zb[i]);
            }
        }
    }
    float pa = 0.f, pb = 0.f;
#pragma unroll
    for (int R = 0; R < 4; ++R)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            pa += W.w4[R][i] * ra[R][i];
            pb += W.w4[R][i] * rb[R][i];
        }
    pa += __shfl_xor(pa, 16);
    pb += __shfl_xor(pb, 16);
    pa += __shfl_xor(pa, 32);
    pb += __shfl_xor(pb, 32);
    ya = pa + W.b4;
    yb = pb + W.b4;
}

}  // namespace

// Node values of every (trade, date k <= kmax) network: table[trade][k-1][m nr + i] (grid n_steps x
// n_trades, 4 waves, two tiles per wave at a time; nr nl a multiple of 32). The inputs are formed as
// the per-path evaluation forms them (FP64 affine maps rounded to float32).
__global__ void __launch_bounds__(kThreads) mlp_lattice_kernel(Book bk, Norm nm, const uint4* frags,
                                                               const float* fwdvec, const double* bounds, int nr,
                                                               int nl, float* table) {
    const int k = blockIdx.x + 1, j = blockIdx.y;
    if (k > bk.kmax[j]) return;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, n16 = l & 15;
    FwdW W;
    load_fwd(W, frags + (size_t)j * kNFrag * 64, fwdvec + (size_t)j * kFwdStride, l);
    const double* bd = bounds + (size_t)(k - 1) * 4;
    const double r_lo = bd[0], dr = node_step(bd[0], bd[1], nr), lb_lo = bd[2], dl = node_step(bd[2], bd[3], nl);
    float u2;
    {
#pragma clang fp contract(off)
        u2 = (float)((double)k / bk.spy * nm.t_scale + nm.t_shift);
    }
    const int ns = nr * nl, n_tiles = ns / 16;
    float* out = table + ((size_t)j * bk.n_steps + (k - 1)) * ns;
    auto inputs = [&](int s, float& a0, float& a1) {
#pragma clang fp contract(off)
        const int m = s / nr, i = s - m * nr;
        const double r = r_lo + (double)i * dr, lb = lb_lo + (double)m * dl;
        a0 = (float)((r - nm.r_shift) * nm.r_scale);
        a1 = (float)(lb * nm.lb_scale + nm.lb_shift);
    };
    for (int t = 2 * w; t < n_tiles; t += 2 * kWaves) {
        const int sa = 16 * t + n16, sb = sa + 16;
        float a0, a1, b0, b1, ya, yb;
        inputs(sa, a0, a1);
        inputs(sb, b0, b1);
        mlp_fwd2(W, input_operand(g, a0, a1, u2), input_operand(g, b0, b1, u2), ya, yb);
        if (g == 0) {
            out[sa] = ya;
            out[sb] = yb;
        }
    }
}

// Exposure sums from the node tables: per (trade, date) the chunk sum of max(y, 0) over the paths
// alive at the date, y interpolated at the path's (r, LB). Grid (chunks, trade groups of
// kLatticeGroup), kLatticeThreads threads; the group's node tables of the date in LDS ([group][nr nl]
// floats, dynamic; nr nl a power of two). A thread takes kLatticeUnroll paths at a time and forms their
// lattice coordinates once for the group's trades. Each thread sums its paths in path order (FP32), then the
// waves (FP64) and the block's waves in index order.
__global__ void __launch_bounds__(kLatticeThreads) mlp_lattice_ee_kernel(Book bk, PathSet ps, const double* bounds,
                                                                         const float* table, int nr, int nl,
                                                                         double* part) {
    constexpr int TG = kLatticeGroup, NW = kLatticeThreads / 64;
    extern __shared__ float sT[];
    __shared__ double red[NW][TG];
    const int c = blockIdx.x, j0 = blockIdx.y * TG;
    const int nt = min(TG, bk.n_trades - j0);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int ns = nr * nl;
    int kend = 0;
    for (int q = 0; q < nt; ++q) kend = max(kend, bk.kmax[j0 + q]);
    const uint64_t n = ps.n;
    const uint64_t p0 = (uint64_t)c * kChunk;
    const uint64_t p1 = (p0 + kChunk < n) ? p0 + kChunk : n;
    for (int k = 1; k <= kend; ++k) {
        __syncthreads();  // the previous date's readers of sT and red are done
        {  // float4 copies; ns is a power of two, so a trade's table is i >> sh
            const int sh = __builtin_ctz((unsigned)ns) - 2;
            float4* d4 = reinterpret_cast<float4*>(sT);
            for (int i = threadIdx.x; i < (nt * ns) >> 2; i += kLatticeThreads) {
                const int q = i >> sh;
                const float4* src =
                    reinterpret_cast<const float4*>(table + ((size_t)(j0 + q) * bk.n_steps + (k - 1)) * ns);
                d4[i] = (k <= bk.kmax[j0 + q]) ? src[i - (q << sh)] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        __syncthreads();
        const double* bd = bounds + (size_t)(k - 1) * 4;
        const double r_lo = bd[0], dr = node_step(bd[0], bd[1], nr), lb_lo = bd[2], dl = node_step(bd[2], bd[3], nl);
        const double ir_s = dr > 0.0 ? 1.0 / dr : 0.0, il_s = dl > 0.0 ? 1.0 / dl : 0.0;
        float acc[TG];
#pragma unroll
        for (int q = 0; q < TG; ++q) acc[q] = 0.0f;
        const uint64_t row = (uint64_t)k * n;
        // kLatticeUnroll paths per thread at a time, their loads issued together (invalid ones clamp to
        // the chunk's first path and add nothing)
        constexpr int U = kLatticeUnroll;
        for (uint64_t pb = p0 + threadIdx.x; pb < p1; pb += (uint64_t)U * kLatticeThreads) {
            double rv[U], lv[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t p = pb + (uint64_t)u * kLatticeThreads;
                ok[u] = p < p1;
                const uint64_t pc = ok[u] ? p : p0;
                rv[u] = ps.R[row + pc];
                lv[u] = ps.LB[row + pc];
            }
            int base[U];
            float wr[U], wl[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const double fr = fmin(fmax((rv[u] - r_lo) * ir_s, 0.0), (double)(nr - 1));
                const double fl = fmin(fmax((lv[u] - lb_lo) * il_s, 0.0), (double)(nl - 1));
                const int ir = min((int)fr, nr - 2), il = min((int)fl, nl - 2);
                wr[u] = (float)(fr - (double)ir);
                wl[u] = (float)(fl - (double)il);
                base[u] = il * nr + ir;
            }
#pragma unroll
            for (int q = 0; q < TG; ++q) {
                if (q >= nt || k > bk.kmax[j0 + q]) continue;  // uniform
                int tv[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    tv[u] = ok[u] ? ps.tau[(uint64_t)(j0 + q) * n + pb + (uint64_t)u * kLatticeThreads] : 0;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (!ok[u] || (tv[u] >= 0 && tv[u] * bk.spy < k)) continue;
                    const float* T = sT + q * ns + base[u];
                    const float y0 = T[0] + wr[u] * (T[1] - T[0]);
                    const float y1 = T[nr] + wr[u] * (T[nr + 1] - T[nr]);
                    const float y = y0 + wl[u] * (y1 - y0);
                    acc[q] += fmaxf(y, 0.0f);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < TG; ++q) {
            const double s = wave_sum((double)acc[q]);
            if (l == 0) red[w][q] = s;
        }
        __syncthreads();
        if (threadIdx.x < nt && k <= bk.kmax[j0 + threadIdx.x]) {
            const int q = threadIdx.x;
            double s = red[0][q];
            for (int i = 1; i < NW; ++i) s += red[i][q];
            part[((uint64_t)(j0 + q) * bk.n_steps + (k - 1)) * gridDim.x + c] = s;
        }
    }
}

}  // namespace berm
}  // namespace eqhip
