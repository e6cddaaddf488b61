// Bermudan payer swaptions with a learned continuation value: device-side declarations.
//
// Paths live in HBM step-major ([grid step][path], coalesced across a wave): the short rate R,
// the trapezoid log-numeraire LB = int_0^t r ds, the numeraire BN = exp(LB) and the standardised
// short rate X = (r - E[r_t]) / sd(r_t). Per (trade, path): the exercise year tau (int8, -1 =
// never) and the pathwise cashflow discounted to 0, y0.
//
// The continuation network is 3 -> 64 -> 64 -> 64 -> 1 with tanh, one per trade, computed with
// v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32 accumulation). Activations are kept transposed
// (features x samples) so that a layer's accumulator tile is already the next layer's B operand:
// lane l holds features 4(l/16)+i of sample l%16 in each 16-row block, and the weights' K slots
// are permuted to match (slot i of lane group g in K-block Q is feature 32Q+4g+i for i<4 and
// 32Q+16+4g+(i-4) for i>=4), so no activation ever leaves the registers between layers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace eqhip {
namespace berm {

constexpr int kH = 64;
constexpr int kNParams = 8641;
constexpr int kParamStride = 8704;
constexpr int OFF_W1 = 0, OFF_W2 = 256, OFF_B2 = 4352, OFF_W3 = 4416, OFF_B3 = 8512, OFF_W4 = 8576,
              OFF_B4 = 8640;
// BF16 operand fragments per network, 64 lanes x 16 bytes each.
constexpr int kFragW1 = 0, kFragW2 = 4, kFragW3 = 12, kFragW2T = 20, kFragW3T = 28, kNFrag = 36;
// Forward activation: r = 1 / (1 + 2^zs), tanh = 1 - 2r, with zs = kActScale z (kActScale = 2 log2 e)
// carried by the forward weights; folded FP32 forward vectors per network (stride kFwdStride).
constexpr float kActScale = 2.8853900817779268f;
constexpr int kFwdStride = 256, FV_B2 = 0, FV_B3 = 64, FV_W4 = 128, FV_B4 = 192;

constexpr int kThreads = 256;      // every Bermudan kernel runs 4 waves per block
constexpr int kWaves = 4;
constexpr int kChunk = 8192;       // paths per block in the path-parallel kernels
constexpr int kMoments = 16;       // 10 distinct entries of the 4x4 normal matrix, 4 of the rhs, count, pad

// Per-set path store.
struct PathSet {
    uint64_t n;            // paths
    int n_chunks;          // ceil(n / kChunk)
    const double* R;       // [n_steps+1][n]
    const double* LB;
    const double* BN;
    const double* X;
    int8_t* tau;           // [n_trades][n]
    double* y0;            // [n_trades][n]
};

struct Book {
    int n_trades;
    int spy;                  // grid steps per year
    int n_steps;              // grid steps (monitoring dates k = 1..n_steps)
    int max_ex;
    const double* notional;
    const double* strike;
    const int* first_ex;
    const int* n_ex;
    const int* kmax;          // last monitoring date with a possibly non-zero exposure
    const double2* AB;        // [max_ex + 1]: A(m), B(m) of the m-year bond, m = 1..max_ex
    double a, b, sigma, r0;
};

struct Norm {
    double r_shift, r_scale, lb_scale, lb_shift, t_scale, t_shift, y_scale;
};

struct SimArgsB {
    double r0, decay, mean_c, sd, half_dt;
    double a, b, sigma;
    double spy;         // grid steps per year (t = s / spy)
    uint32_t n_steps;
    uint64_t n_paths, seed;
    const double* Z;    // host draws [n_steps][n] or null
    double *R, *LB, *BN, *X;
};

__global__ void berm_sim_kernel(SimArgsB g);
__global__ void berm_ab_kernel(int max_ex, double a, double b, double sigma, double2* out);
__global__ void berm_policy_moments(Book bk, PathSet ps, int bi, double* part);
__global__ void berm_policy_solve(Book bk, int bi, int n_chunks, const double* part, double* ex_coef);
__global__ void berm_policy_update(Book bk, PathSet ps, int bi, const double* ex_coef);
__global__ void berm_apply_kernel(Book bk, PathSet ps, const double* ex_coef);
__global__ void berm_sum_y0(Book bk, PathSet ps, double* part);
__global__ void berm_judge_moments(Book bk, PathSet ps, double* part);
__global__ void berm_judge_solve(Book bk, int n_chunks, const double* part, double* coef);
__global__ void berm_judge_ee(Book bk, PathSet ps, const double* coef, double* part);
__global__ void berm_finalize(int n_rows, int n_chunks, const double* part, int stride, double n_div,
                              const double* trade_scale, int rows_per_trade, const int* kmax, double* out);

// network
__global__ void mlp_init_kernel(int n_nets, uint64_t seed, float* master);
__global__ void mlp_pack_kernel(const float* master, uint4* frags);
__global__ void mlp_fold_kernel(const float* master, float* fwdvec);
template <int SRC>
__global__ void mlp_train_kernel(Book bk, PathSet ps, Norm nm, const uint4* frags, const float* fwdvec,
                                 int tiles_per_wave, uint32_t step, uint64_t seed, const float* xu,
                                 const float* xt, const float* xm, float* grad_part, float* loss_part);
__global__ void mlp_adam_kernel(int n_nets, float* master, float* m1, float* v1, const float* grad_part,
                                float lr, float beta1, float beta2, float eps, float bc1, float bc2,
                                int do_update, float* grad_out, const float* loss_part, float* loss_out);
template <int MODE>
__global__ void mlp_eval_kernel(Book bk, PathSet ps, Norm nm, const uint4* frags, const float* fwdvec,
                                const double* poly, double* part);
__global__ void mlp_forward_kernel(int n, const uint4* frags, const float* fwdvec, const float* xu, float* y);
// lattice evaluation of the networks (per-date node tables, bilinear interpolation per path)
constexpr int kLatticeGroup = 16;      // trades per block of the interpolation kernel
constexpr int kLatticeThreads = 512;   // threads per block of the interpolation kernel
constexpr int kLatticeUnroll = 8;      // paths per thread in flight in the interpolation kernel
constexpr int kLatticeMaxNodes = 1024; // nr nl: the group's tables fill at most 64 KB of LDS
__global__ void berm_bounds_kernel(PathSet ps, double* bounds);
__global__ void mlp_lattice_kernel(Book bk, Norm nm, const uint4* frags, const float* fwdvec, const double* bounds,
                                   int nr, int nl, float* table);
__global__ void mlp_lattice_ee_kernel(Book bk, PathSet ps, const double* bounds, const float* table, int nr, int nl,
                                      double* part);

}  // namespace berm
}  // namespace eqhip
