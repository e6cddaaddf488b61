// Host side of the Bermudan-swaption exposure path: resident path sets, the Longstaff-Schwartz
// policy, the polynomial judge and the continuation networks, behind the eqhip_berm_* C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "eqhip.h"
#include "eqhip_berm_kernels.hpp"
#include "eqhip_host.hpp"

using namespace eqhip_host;
using namespace eqhip::berm;

namespace {

struct HostSet {
    uint64_t n = 0;
    int n_chunks = 0;
    DevBuf R, LB, BN, X, tau, y0;
    DevBuf judge_coef;          // [n_trades][n_steps][4] of the last polynomial fit on this set
    DevBuf bounds;              // [n_steps][4] per-date (r_lo, r_hi, lb_lo, lb_hi) over the paths
    bool have_policy = false, have_judge = false, have_bounds = false;

    PathSet view() const {
        PathSet ps;
        ps.n = n;
        ps.n_chunks = n_chunks;
        ps.R = R.as<const double>();
        ps.LB = LB.as<const double>();
        ps.BN = BN.as<const double>();
        ps.X = X.as<const double>();
        ps.tau = tau.as<int8_t>();
        ps.y0 = y0.as<double>();
        return ps;
    }
};

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct eqhip_berm {
    eqhip_context* ctx = nullptr;
    int device = 0;  // kept so that destroying the book never touches its context
    eqhip_model_hw1f model{};
    int n_trades = 0, max_ex = 0, spy = 0, n_steps = 0;
    std::vector<double> notional, strike;
    std::vector<int> first_ex, n_ex, kmax;
    DevBuf d_notional, d_strike, d_first, d_nex, d_kmax, d_ab, d_coef, d_part, d_inv_scale;
    HostSet sets[2];  // 0 = training, 1 = inference
    bool have_policy = false;
    // networks
    eqhip_berm_mlp_cfg cfg{};
    bool have_net = false;
    uint32_t adam_t = 0;
    DevBuf master, m1, v1, frags, fwdvec, grad_part, loss_part, loss_step, grad_out, xs_u, xs_t, xs_m;
    DevBuf lattice;  // node values [n_trades][n_steps][nr nl] of the last lattice evaluation
    eqhip_berm_timing timing{};

    Book book() const {
        Book b;
        b.n_trades = n_trades;
        b.spy = spy;
        b.n_steps = n_steps;
        b.max_ex = max_ex;
        b.notional = d_notional.as<const double>();
        b.strike = d_strike.as<const double>();
        b.first_ex = d_first.as<const int>();
        b.n_ex = d_nex.as<const int>();
        b.kmax = d_kmax.as<const int>();
        b.AB = d_ab.as<const double2>();
        b.a = model.a;
        b.b = model.b;
        b.sigma = model.sigma;
        b.r0 = model.r0;
        return b;
    }
    Norm norm() const {
        return Norm{cfg.r_shift, cfg.r_scale, cfg.lb_scale, cfg.lb_shift, cfg.t_scale, cfg.t_shift, cfg.y_scale};
    }
    hipStream_t st() const { return ctx->stream; }
    void sync() { HIPCHK(hipStreamSynchronize(st())); }
    size_t part_bytes(int set, int per) const {
        return (size_t)n_trades * std::max(1, n_steps) * sets[set].n_chunks * per * sizeof(double);
    }
};

namespace {

void simulate_set(eqhip_berm& B, int set, uint64_t n, uint64_t seed, const double* z) {
    HostSet& S = B.sets[set];
    S.n = n;
    S.n_chunks = (int)((n + kChunk - 1) / kChunk);
    const size_t cells = (size_t)(B.n_steps + 1) * n;
    S.R.alloc(cells * 8);
    S.LB.alloc(cells * 8);
    S.BN.alloc(cells * 8);
    S.X.alloc(cells * 8);
    S.tau.alloc((size_t)B.n_trades * n);
    S.y0.alloc((size_t)B.n_trades * n * 8);
    DevBuf dz;
    if (z) {
        const size_t bytes = (size_t)B.n_steps * n * 8;
        dz.alloc(bytes);
        HIPCHK(hipMemcpy(dz.p, z, bytes, hipMemcpyHostToDevice));
    }
    const double a = B.model.a, b = B.model.b, sigma = B.model.sigma;
    const double dt = 1.0 / (double)B.spy;
    // exact-scheme constants computed as the oracle's simulate_one does
    const double e = std::exp(-(a * dt));
    SimArgsB g;
    g.r0 = B.model.r0;
    g.decay = e;
    g.mean_c = b * (1.0 - e);
    g.sd = sigma * std::sqrt((1.0 - std::exp(-(2.0 * a * dt))) / (2.0 * a));
    g.half_dt = 0.5 * dt;
    g.a = a;
    g.b = b;
    g.sigma = sigma;
    g.spy = (double)B.spy;
    g.n_steps = (uint32_t)B.n_steps;
    g.n_paths = n;
    g.seed = seed;
    g.Z = z ? dz.as<const double>() : nullptr;
    g.R = S.R.as<double>();
    g.LB = S.LB.as<double>();
    g.BN = S.BN.as<double>();
    g.X = S.X.as<double>();
    hipLaunchKernelGGL(berm_sim_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, B.st(), g);
    HIPCHK(hipGetLastError());
    if (z) B.sync();  // dz dies at scope exit
}

void ensure_part(eqhip_berm& B, size_t bytes) {
    if (B.d_part.bytes < bytes) B.d_part.alloc(bytes);
}

void finalize(eqhip_berm& B, int n_rows, int n_chunks, const double* part, int stride, double n_div,
              const double* trade_scale, int rows_per_trade, const int* kmax, double* d_out) {
    hipLaunchKernelGGL(berm_finalize, dim3((unsigned)((n_rows + 127) / 128)), dim3(128), 0, B.st(), n_rows, n_chunks,
                       part, stride, n_div, trade_scale, rows_per_trade, kmax, d_out);
    HIPCHK(hipGetLastError());
}

// Price (mean of y0) of every trade on one set, into host out[n_trades].
void set_price(eqhip_berm& B, int set, double* out) {
    HostSet& S = B.sets[set];
    ensure_part(B, (size_t)B.n_trades * S.n_chunks * 8);
    hipLaunchKernelGGL(berm_sum_y0, dim3(S.n_chunks, B.n_trades), dim3(kThreads), 0, B.st(), B.book(), S.view(),
                       B.d_part.as<double>());
    HIPCHK(hipGetLastError());
    DevBuf d;
    d.alloc((size_t)B.n_trades * 8);
    finalize(B, B.n_trades, S.n_chunks, B.d_part.as<const double>(), 1, (double)S.n, nullptr, 1, nullptr,
             d.as<double>());
    HIPCHK(hipMemcpyAsync(out, d.p, (size_t)B.n_trades * 8, hipMemcpyDeviceToHost, B.st()));
    B.sync();
}

void fit_policy(eqhip_berm& B) {
    HostSet& T = B.sets[0];
    const PathSet ps = T.view();
    const Book bk = B.book();
    HIPCHK(hipMemsetAsync(T.tau.p, 0xFF, T.tau.bytes, B.st()));  // -1: never exercised
    HIPCHK(hipMemsetAsync(T.y0.p, 0, T.y0.bytes, B.st()));
    HIPCHK(hipMemsetAsync(B.d_coef.p, 0, B.d_coef.bytes, B.st()));
    ensure_part(B, (size_t)B.n_trades * T.n_chunks * kMoments * 8);
    for (int bi = 0; bi < B.max_ex; ++bi) {
        hipLaunchKernelGGL(berm_policy_moments, dim3(T.n_chunks, B.n_trades), dim3(kThreads), 0, B.st(), bk, ps, bi,
                           B.d_part.as<double>());
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(berm_policy_solve, dim3((B.n_trades + 63) / 64), dim3(64), 0, B.st(), bk, bi, T.n_chunks,
                           B.d_part.as<const double>(), B.d_coef.as<double>());
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(berm_policy_update, dim3(T.n_chunks, B.n_trades), dim3(kThreads), 0, B.st(), bk, ps, bi,
                           B.d_coef.as<const double>());
        HIPCHK(hipGetLastError());
    }
    T.have_policy = true;
}

void apply_policy(eqhip_berm& B, int set) {
    HostSet& S = B.sets[set];
    hipLaunchKernelGGL(berm_apply_kernel, dim3((unsigned)((S.n + kThreads - 1) / kThreads), B.n_trades), dim3(kThreads),
                       0, B.st(), B.book(), S.view(), B.d_coef.as<const double>());
    HIPCHK(hipGetLastError());
    S.have_policy = true;
}

void require_set(const eqhip_berm& B, int set) {
    if (set != 0 && set != 1) raise(EQHIP_E_INVALID, "set must be 0 (training) or 1 (inference)");
    if (!B.have_policy) raise(EQHIP_E_INVALID, "the exercise policy has not been fitted (eqhip_berm_fit_policy)");
    (void)B;
}

void judge(eqhip_berm& B, int set, double* ee, double* coef) {
    HostSet& S = B.sets[set];
    const Book bk = B.book();
    const PathSet ps = S.view();
    const int rows = B.n_trades * B.n_steps;
    ensure_part(B, B.part_bytes(set, kMoments));
    S.judge_coef.alloc((size_t)rows * 4 * 8);
    hipLaunchKernelGGL(berm_judge_moments, dim3(S.n_chunks, B.n_trades), dim3(kThreads), 0, B.st(), bk, ps,
                       B.d_part.as<double>());
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(berm_judge_solve, dim3((rows + 127) / 128), dim3(128), 0, B.st(), bk, S.n_chunks,
                       B.d_part.as<const double>(), S.judge_coef.as<double>());
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(berm_judge_ee, dim3(S.n_chunks, B.n_trades), dim3(kThreads), 0, B.st(), bk, ps,
                       S.judge_coef.as<const double>(), B.d_part.as<double>());
    HIPCHK(hipGetLastError());
    DevBuf d;
    d.alloc((size_t)rows * 8);
    finalize(B, rows, S.n_chunks, B.d_part.as<const double>(), 1, (double)S.n, nullptr, B.n_steps,
             B.d_kmax.as<const int>(), d.as<double>());
    if (ee) HIPCHK(hipMemcpyAsync(ee, d.p, (size_t)rows * 8, hipMemcpyDeviceToHost, B.st()));
    if (coef) HIPCHK(hipMemcpyAsync(coef, S.judge_coef.p, (size_t)rows * 32, hipMemcpyDeviceToHost, B.st()));
    B.sync();
    S.have_judge = true;
}

void alloc_net(eqhip_berm& B) {
    const size_t np = (size_t)B.n_trades * kParamStride;
    B.master.alloc(np * 4);
    B.m1.alloc(np * 4);
    B.v1.alloc(np * 4);
    B.frags.alloc((size_t)B.n_trades * kNFrag * 64 * 16);
    B.fwdvec.alloc((size_t)B.n_trades * kFwdStride * 4);
    B.grad_part.alloc(np * kWaves * 4);
    B.loss_part.alloc((size_t)B.n_trades * kWaves * 4);
    B.loss_step.alloc((size_t)B.n_trades * 4);
    B.grad_out.alloc((size_t)B.n_trades * kNParams * 4);
    HIPCHK(hipMemsetAsync(B.m1.p, 0, B.m1.bytes, B.st()));
    HIPCHK(hipMemsetAsync(B.v1.p, 0, B.v1.bytes, B.st()));
    HIPCHK(hipMemsetAsync(B.grad_part.p, 0, B.grad_part.bytes, B.st()));
    B.adam_t = 0;
}

void pack(eqhip_berm& B) {
    hipLaunchKernelGGL(mlp_pack_kernel, dim3(kNFrag, B.n_trades), dim3(64), 0, B.st(), B.master.as<const float>(),
                       B.frags.as<uint4>());
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(mlp_fold_kernel, dim3(B.n_trades), dim3(256), 0, B.st(), B.master.as<const float>(),
                       B.fwdvec.as<float>());
    HIPCHK(hipGetLastError());
}

void check_cfg(const eqhip_berm_mlp_cfg& c) {
    if (c.batch == 0 || c.batch % 64 != 0) raise(EQHIP_E_INVALID, "batch must be a positive multiple of 64");
    if (!(c.r_scale != 0.0 && c.lb_scale != 0.0 && c.t_scale != 0.0 && c.y_scale > 0.0))
        raise(EQHIP_E_INVALID, "input / output scales must be non-zero (y_scale > 0)");
}

// Adam step over the current grad_part (update = 0: only the summed gradient and loss).
void adam(eqhip_berm& B, int update, float lr, float* d_grad_out, float* d_loss_out) {
    const uint64_t total = (uint64_t)B.n_trades * kParamStride;
    float bc1 = 1.0f, bc2 = 1.0f;
    if (update) {
        ++B.adam_t;
        bc1 = (float)(1.0 - std::pow((double)B.cfg.beta1, (double)B.adam_t));
        bc2 = (float)(1.0 - std::pow((double)B.cfg.beta2, (double)B.adam_t));
    }
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, B.st(), B.n_trades,
                       B.master.as<float>(), B.m1.as<float>(), B.v1.as<float>(), B.grad_part.as<const float>(), lr,
                       B.cfg.beta1, B.cfg.beta2, B.cfg.eps, bc1, bc2, update, d_grad_out,
                       B.loss_part.as<const float>(), d_loss_out);
    HIPCHK(hipGetLastError());
    if (update) pack(B);
}

void stage_samples(eqhip_berm& B, const float* u, const float* target, const float* mask, uint32_t n) {
    if (n == 0 || n % 64 != 0) raise(EQHIP_E_INVALID, "n must be a positive multiple of 64");
    if (!u || !target || !mask) raise(EQHIP_E_INVALID, "null sample array");
    const size_t m = (size_t)B.n_trades * n;
    B.xs_u.alloc(m * 3 * 4);
    B.xs_t.alloc(m * 4);
    B.xs_m.alloc(m * 4);
    HIPCHK(hipMemcpy(B.xs_u.p, u, m * 12, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(B.xs_t.p, target, m * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(B.xs_m.p, mask, m * 4, hipMemcpyHostToDevice));
}

void train_explicit(eqhip_berm& B, uint32_t n) {
    hipLaunchKernelGGL(mlp_train_kernel<1>, dim3(B.n_trades), dim3(kThreads), 0, B.st(), B.book(), B.sets[0].view(),
                       B.norm(), B.frags.as<const uint4>(), B.fwdvec.as<const float>(), (int)(n / 64), 0u,
                       (uint64_t)0, B.xs_u.as<const float>(), B.xs_t.as<const float>(), B.xs_m.as<const float>(),
                       B.grad_part.as<float>(), B.loss_part.as<float>());
    HIPCHK(hipGetLastError());
}

void require_net(const eqhip_berm& B) {
    if (!B.have_net) raise(EQHIP_E_INVALID, "no network parameters (eqhip_berm_train or eqhip_berm_set_params)");
}

}  // namespace

extern "C" {

int eqhip_berm_create(eqhip_context* ctx, const eqhip_model_hw1f* model, const eqhip_berm_book* book,
                      const eqhip_berm_mc* mc, const double* z_train, const double* z_infer, eqhip_berm** out) {
    return guarded([&] {
        if (!ctx || !model || !book || !mc || !out) raise(EQHIP_E_INVALID, "null argument");
        *out = nullptr;
        if (!(model->a > 0.0) || !(model->sigma > 0.0)) raise(EQHIP_E_INVALID, "HW1F needs a > 0 and sigma > 0");
        if (book->n_trades <= 0 || !book->notional || !book->strike || !book->first_exercise || !book->n_exercise)
            raise(EQHIP_E_INVALID, "empty or incomplete Bermudan book");
        if (mc->steps_per_year == 0 || mc->horizon_years == 0) raise(EQHIP_E_INVALID, "empty grid");
        if (mc->n_train == 0 || mc->n_infer == 0) raise(EQHIP_E_INVALID, "n_train and n_infer must be positive");
        if (mc->rng != EQHIP_RNG_PHILOX && mc->rng != EQHIP_RNG_HOST_DRAWS) raise(EQHIP_E_INVALID, "unknown rng");
        if (mc->rng == EQHIP_RNG_HOST_DRAWS && (!z_train || !z_infer))
            raise(EQHIP_E_INVALID, "host-draw mode needs z_train and z_infer");
        if (mc->n_train >= (1ull << 32)) raise(EQHIP_E_CAPACITY, "n_train must be below 2^32");
        auto B = std::make_unique<eqhip_berm>();
        B->ctx = ctx;
        B->device = ctx->device;
        B->model = *model;
        B->n_trades = book->n_trades;
        B->spy = (int)mc->steps_per_year;
        B->n_steps = (int)(mc->steps_per_year * mc->horizon_years);
        for (int j = 0; j < book->n_trades; ++j) {
            const int E = book->first_exercise[j], L = book->n_exercise[j];
            if (E < 1 || L < 1) raise(EQHIP_E_INVALID, "first_exercise and n_exercise must be >= 1");
            if (E + L > (int)mc->horizon_years) raise(EQHIP_E_INVALID, "a maturity lies beyond the horizon");
            if (E + L > 127) raise(EQHIP_E_CAPACITY, "exercise years must stay below 127");
            if (!std::isfinite(book->notional[j]) || !std::isfinite(book->strike[j]) || book->notional[j] <= 0.0)
                raise(EQHIP_E_INVALID, "notionals must be positive and finite, strikes finite");
            B->notional.push_back(book->notional[j]);
            B->strike.push_back(book->strike[j]);
            B->first_ex.push_back(E);
            B->n_ex.push_back(L);
            B->kmax.push_back((E + L - 1) * B->spy);  // no exposure after the last exercise date
            B->max_ex = std::max(B->max_ex, L);
        }
        HIPCHK(hipSetDevice(ctx->device));
        upload(B->d_notional, B->notional, B->st());
        upload(B->d_strike, B->strike, B->st());
        upload(B->d_first, B->first_ex, B->st());
        upload(B->d_nex, B->n_ex, B->st());
        upload(B->d_kmax, B->kmax, B->st());
        std::vector<double> inv(B->n_trades);
        B->d_ab.alloc((size_t)(B->max_ex + 1) * 16);
        hipLaunchKernelGGL(berm_ab_kernel, dim3(1), dim3(128), 0, B->st(), B->max_ex, model->a, model->b, model->sigma,
                           B->d_ab.as<double2>());
        HIPCHK(hipGetLastError());
        B->d_coef.alloc((size_t)B->n_trades * B->max_ex * 4 * 8);
        const auto t0 = std::chrono::steady_clock::now();
        const bool host = mc->rng == EQHIP_RNG_HOST_DRAWS;
        simulate_set(*B, 0, mc->n_train, mc->seed_train, host ? z_train : nullptr);
        simulate_set(*B, 1, mc->n_infer, mc->seed_infer, host ? z_infer : nullptr);
        B->sync();
        B->timing.simulate_ms = ms_since(t0);
        *out = B.release();
    });
}

int eqhip_berm_fit_policy(eqhip_berm* B, double* ex_coef, double* price_train, double* price_infer) {
    return guarded([&] {
        if (!B) raise(EQHIP_E_INVALID, "null handle");
        HIPCHK(hipSetDevice(B->ctx->device));
        auto t0 = std::chrono::steady_clock::now();
        fit_policy(*B);
        B->sync();
        B->timing.policy_ms = ms_since(t0);
        t0 = std::chrono::steady_clock::now();
        apply_policy(*B, 1);
        B->sync();
        B->timing.apply_ms = ms_since(t0);
        B->have_policy = true;
        if (ex_coef)
            HIPCHK(hipMemcpyAsync(ex_coef, B->d_coef.p, B->d_coef.bytes, hipMemcpyDeviceToHost, B->st()));
        B->sync();
        if (price_train) set_price(*B, 0, price_train);
        if (price_infer) set_price(*B, 1, price_infer);
    });
}

int eqhip_berm_pathwise(eqhip_berm* B, int32_t set, int32_t* tau, double* y0) {
    return guarded([&] {
        if (!B) raise(EQHIP_E_INVALID, "null handle");
        require_set(*B, set);
        HIPCHK(hipSetDevice(B->ctx->device));
        const HostSet& S = B->sets[set];
        const size_t m = (size_t)B->n_trades * S.n;
        B->sync();
        if (tau) {
            std::vector<int8_t> t8(m);
            HIPCHK(hipMemcpy(t8.data(), S.tau.p, m, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < m; ++i) tau[i] = t8[i];
        }
        if (y0) HIPCHK(hipMemcpy(y0, S.y0.p, m * 8, hipMemcpyDeviceToHost));
    });
}

int eqhip_berm_exposure_lsm(eqhip_berm* B, int32_t set, double* ee, double* coef) {
    return guarded([&] {
        if (!B) raise(EQHIP_E_INVALID, "null handle");
        require_set(*B, set);
        HIPCHK(hipSetDevice(B->ctx->device));
        const auto t0 = std::chrono::steady_clock::now();
        judge(*B, set, ee, coef);
        B->timing.judge_ms = ms_since(t0);
    });
}

int eqhip_berm_train(eqhip_berm* B, const eqhip_berm_mlp_cfg* cfg, float* loss_curve) {
    return guarded([&] {
        if (!B || !cfg) raise(EQHIP_E_INVALID, "null argument");
        require_set(*B, 0);
        check_cfg(*cfg);
        HIPCHK(hipSetDevice(B->ctx->device));
        B->cfg = *cfg;
        const auto t0 = std::chrono::steady_clock::now();
        alloc_net(*B);
        const uint64_t total = (uint64_t)B->n_trades * kParamStride;
        hipLaunchKernelGGL(mlp_init_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, B->st(), B->n_trades,
                           cfg->seed, B->master.as<float>());
        HIPCHK(hipGetLastError());
        pack(*B);
        DevBuf hist;
        hist.alloc((size_t)std::max(1u, cfg->steps) * B->n_trades * 4);
        const int tpw = (int)(cfg->batch / 64);
        for (uint32_t s = 0; s < cfg->steps; ++s) {
            hipLaunchKernelGGL(mlp_train_kernel<0>, dim3(B->n_trades), dim3(kThreads), 0, B->st(), B->book(),
                               B->sets[0].view(), B->norm(), B->frags.as<const uint4>(), B->fwdvec.as<const float>(),
                               tpw, s, cfg->seed, nullptr, nullptr, nullptr, B->grad_part.as<float>(),
                               B->loss_part.as<float>());
            HIPCHK(hipGetLastError());
            const double frac = cfg->steps > 1 ? (double)s / (double)(cfg->steps - 1) : 1.0;
            const float lr = (float)(cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1.0 + std::cos(M_PI * frac)));
            adam(*B, 1, lr, nullptr, hist.as<float>() + (size_t)s * B->n_trades);
        }
        B->have_net = true;
        B->sync();  // the history is written by the stream's last kernels
        if (loss_curve && cfg->steps) {
            std::vector<float> h((size_t)cfg->steps * B->n_trades);
            HIPCHK(hipMemcpy(h.data(), hist.p, h.size() * 4, hipMemcpyDeviceToHost));
            for (uint32_t s = 0; s < cfg->steps; ++s) {
                double acc = 0.0;
                for (int j = 0; j < B->n_trades; ++j) acc += h[(size_t)s * B->n_trades + j];
                loss_curve[s] = (float)(acc / B->n_trades);
            }
        }
        B->sync();
        B->timing.train_ms = ms_since(t0);
    });
}

int eqhip_berm_exposure_mlp(eqhip_berm* B, int32_t set, double* ee) {
    return guarded([&] {
        if (!B || !ee) raise(EQHIP_E_INVALID, "null argument");
        require_set(*B, set);
        require_net(*B);
        HIPCHK(hipSetDevice(B->ctx->device));
        const auto t0 = std::chrono::steady_clock::now();
        HostSet& S = B->sets[set];
        const int rows = B->n_trades * B->n_steps;
        ensure_part(*B, B->part_bytes(set, 1));
        hipLaunchKernelGGL(mlp_eval_kernel<0>, dim3(S.n_chunks, B->n_trades), dim3(kThreads), 0, B->st(), B->book(),
                           S.view(), B->norm(), B->frags.as<const uint4>(), B->fwdvec.as<const float>(), nullptr,
                           B->d_part.as<double>());
        HIPCHK(hipGetLastError());
        std::vector<double> scale(B->n_trades);
        for (int j = 0; j < B->n_trades; ++j) scale[j] = B->notional[j] / B->cfg.y_scale;
        upload(B->d_inv_scale, scale, B->st());
        DevBuf d;
        d.alloc((size_t)rows * 8);
        finalize(*B, rows, S.n_chunks, B->d_part.as<const double>(), 1, (double)S.n, B->d_inv_scale.as<const double>(),
                 B->n_steps, B->d_kmax.as<const int>(), d.as<double>());
        HIPCHK(hipMemcpyAsync(ee, d.p, (size_t)rows * 8, hipMemcpyDeviceToHost, B->st()));
        B->sync();
        B->timing.infer_ms = ms_since(t0);
    });
}

int eqhip_berm_exposure_mlp_lattice(eqhip_berm* B, int32_t set, int32_t nr, int32_t nl, double* ee) {
    return guarded([&] {
        if (!B || !ee) raise(EQHIP_E_INVALID, "null argument");
        require_set(*B, set);
        require_net(*B);
        auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
        if (nr < 2 || nl < 2 || !pow2(nr) || !pow2(nl) || nr * nl < 32 || nr * nl > kLatticeMaxNodes)
            raise(EQHIP_E_INVALID, "lattice: nr, nl powers of two >= 2 with 32 <= nr nl <= " +
                                       std::to_string(kLatticeMaxNodes));
        HIPCHK(hipSetDevice(B->ctx->device));
        const auto t0 = std::chrono::steady_clock::now();
        HostSet& S = B->sets[set];
        const int rows = B->n_trades * B->n_steps;
        if (!S.have_bounds) {  // the paths of a set do not change: their per-date box once
            S.bounds.alloc((size_t)B->n_steps * 4 * 8);
            hipLaunchKernelGGL(berm_bounds_kernel, dim3(B->n_steps), dim3(kThreads), 0, B->st(), S.view(),
                               S.bounds.as<double>());
            HIPCHK(hipGetLastError());
            S.have_bounds = true;
        }
        const size_t ns = (size_t)nr * nl;
        if (B->lattice.bytes < (size_t)rows * ns * 4) B->lattice.alloc((size_t)rows * ns * 4);
        hipEvent_t ev[3];
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventRecord(ev[0], B->st()));
        hipLaunchKernelGGL(mlp_lattice_kernel, dim3(B->n_steps, B->n_trades), dim3(kThreads), 0, B->st(), B->book(),
                           B->norm(), B->frags.as<const uint4>(), B->fwdvec.as<const float>(),
                           S.bounds.as<const double>(), nr, nl, B->lattice.as<float>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ev[1], B->st()));
        ensure_part(*B, B->part_bytes(set, 1));
        const size_t lds = (size_t)kLatticeGroup * ns * 4;
        HIPCHK(hipFuncSetAttribute((const void*)mlp_lattice_ee_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
        hipLaunchKernelGGL(mlp_lattice_ee_kernel, dim3(S.n_chunks, (B->n_trades + kLatticeGroup - 1) / kLatticeGroup),
                           dim3(kLatticeThreads), lds, B->st(), B->book(), S.view(), S.bounds.as<const double>(),
                           B->lattice.as<const float>(), nr, nl, B->d_part.as<double>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ev[2], B->st()));
        std::vector<double> scale(B->n_trades);
        for (int j = 0; j < B->n_trades; ++j) scale[j] = B->notional[j] / B->cfg.y_scale;
        upload(B->d_inv_scale, scale, B->st());
        DevBuf d;
        d.alloc((size_t)rows * 8);
        finalize(*B, rows, S.n_chunks, B->d_part.as<const double>(), 1, (double)S.n, B->d_inv_scale.as<const double>(),
                 B->n_steps, B->d_kmax.as<const int>(), d.as<double>());
        HIPCHK(hipMemcpyAsync(ee, d.p, (size_t)rows * 8, hipMemcpyDeviceToHost, B->st()));
        B->sync();
        float m0 = 0.f, m1 = 0.f;
        HIPCHK(hipEventElapsedTime(&m0, ev[0], ev[1]));
        HIPCHK(hipEventElapsedTime(&m1, ev[1], ev[2]));
        for (auto& e : ev) (void)hipEventDestroy(e);
        B->timing.lattice_ms = ms_since(t0);
        B->timing.lattice_net_ms = m0;
        B->timing.lattice_interp_ms = m1;
    });
}

int eqhip_berm_mse(eqhip_berm* B, double* mse_mlp, double* mse_poly) {
    return guarded([&] {
        if (!B || !mse_mlp || !mse_poly) raise(EQHIP_E_INVALID, "null argument");
        require_set(*B, 1);
        require_net(*B);
        const int poly_set = B->sets[0].have_judge ? 0 : 1;
        if (!B->sets[poly_set].have_judge) raise(EQHIP_E_INVALID, "no polynomial fit yet (eqhip_berm_exposure_lsm)");
        HIPCHK(hipSetDevice(B->ctx->device));
        HostSet& S = B->sets[1];
        const int rows = B->n_trades * B->n_steps;
        ensure_part(*B, B->part_bytes(1, 3));
        hipLaunchKernelGGL(mlp_eval_kernel<1>, dim3(S.n_chunks, B->n_trades), dim3(kThreads), 0, B->st(), B->book(),
                           S.view(), B->norm(), B->frags.as<const uint4>(), B->fwdvec.as<const float>(),
                           B->sets[poly_set].judge_coef.as<const double>(), B->d_part.as<double>());
        HIPCHK(hipGetLastError());
        DevBuf d;
        d.alloc((size_t)rows * 3 * 8);
        for (int q = 0; q < 3; ++q)
            finalize(*B, rows, S.n_chunks, B->d_part.as<const double>() + q, 3, 1.0, nullptr, B->n_steps,
                     B->d_kmax.as<const int>(), d.as<double>() + (size_t)q * rows);
        std::vector<double> h((size_t)rows * 3);
        HIPCHK(hipMemcpyAsync(h.data(), d.p, h.size() * 8, hipMemcpyDeviceToHost, B->st()));
        B->sync();
        const double ys2 = B->cfg.y_scale * B->cfg.y_scale;
        for (int j = 0; j < B->n_trades; ++j) {
            double e1 = 0.0, e2 = 0.0, cnt = 0.0;
            for (int k = 0; k < B->n_steps; ++k) {
                e1 += h[(size_t)j * B->n_steps + k];
                e2 += h[(size_t)rows + (size_t)j * B->n_steps + k];
                cnt += h[(size_t)2 * rows + (size_t)j * B->n_steps + k];
            }
            mse_mlp[j] = cnt > 0 ? e1 / cnt / ys2 : 0.0;
            mse_poly[j] = cnt > 0 ? e2 / cnt / ys2 : 0.0;
        }
    });
}

int eqhip_berm_get_params(eqhip_berm* B, float* params) {
    return guarded([&] {
        if (!B || !params) raise(EQHIP_E_INVALID, "null argument");
        require_net(*B);
        HIPCHK(hipSetDevice(B->ctx->device));
        B->sync();
        HIPCHK(hipMemcpy2D(params, kNParams * 4, B->master.p, kParamStride * 4, kNParams * 4, B->n_trades,
                           hipMemcpyDeviceToHost));
    });
}

int eqhip_berm_set_params(eqhip_berm* B, const eqhip_berm_mlp_cfg* cfg, const float* params) {
    return guarded([&] {
        if (!B || !cfg || !params) raise(EQHIP_E_INVALID, "null argument");
        check_cfg(*cfg);
        HIPCHK(hipSetDevice(B->ctx->device));
        B->cfg = *cfg;
        alloc_net(*B);
        B->sync();  // the copies below run on the null stream
        HIPCHK(hipMemset(B->master.p, 0, B->master.bytes));
        HIPCHK(hipMemcpy2D(B->master.p, kParamStride * 4, params, kNParams * 4, kNParams * 4, B->n_trades,
                           hipMemcpyHostToDevice));
        pack(*B);
        B->sync();
        B->have_net = true;
    });
}

int eqhip_berm_mlp_grad(eqhip_berm* B, const float* u, const float* target, const float* mask, uint32_t n,
                        float* loss, float* grad) {
    return guarded([&] {
        if (!B || !loss || !grad) raise(EQHIP_E_INVALID, "null argument");
        require_net(*B);
        HIPCHK(hipSetDevice(B->ctx->device));
        stage_samples(*B, u, target, mask, n);
        train_explicit(*B, n);
        adam(*B, 0, 0.0f, B->grad_out.as<float>(), B->loss_step.as<float>());
        HIPCHK(hipMemcpyAsync(grad, B->grad_out.p, (size_t)B->n_trades * kNParams * 4, hipMemcpyDeviceToHost, B->st()));
        HIPCHK(hipMemcpyAsync(loss, B->loss_step.p, (size_t)B->n_trades * 4, hipMemcpyDeviceToHost, B->st()));
        B->sync();
    });
}

int eqhip_berm_adam_step(eqhip_berm* B, const float* u, const float* target, const float* mask, uint32_t n,
                         uint32_t step, float lr) {
    return guarded([&] {
        if (!B) raise(EQHIP_E_INVALID, "null handle");
        require_net(*B);
        if (step != B->adam_t + 1) raise(EQHIP_E_INVALID, "Adam steps must be consecutive from 1");
        HIPCHK(hipSetDevice(B->ctx->device));
        stage_samples(*B, u, target, mask, n);
        train_explicit(*B, n);
        adam(*B, 1, lr, nullptr, nullptr);
        B->sync();
    });
}

int eqhip_berm_mlp_forward(eqhip_berm* B, const float* u, uint32_t n, float* y) {
    return guarded([&] {
        if (!B || !u || !y) raise(EQHIP_E_INVALID, "null argument");
        if (n == 0) raise(EQHIP_E_INVALID, "n must be positive");
        require_net(*B);
        HIPCHK(hipSetDevice(B->ctx->device));
        const size_t m = (size_t)B->n_trades * n;
        DevBuf du, dy;
        du.alloc(m * 12);
        dy.alloc(m * 4);
        HIPCHK(hipMemcpy(du.p, u, m * 12, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + 63) / 64, B->n_trades), dim3(64), 0, B->st(), (int)n,
                           B->frags.as<const uint4>(), B->fwdvec.as<const float>(), du.as<const float>(),
                           dy.as<float>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(y, dy.p, m * 4, hipMemcpyDeviceToHost, B->st()));
        B->sync();
    });
}

int eqhip_berm_get_timing(const eqhip_berm* B, eqhip_berm_timing* out) {
    return guarded([&] {
        if (!B || !out) raise(EQHIP_E_INVALID, "null argument");
        *out = B->timing;
    });
}

int eqhip_berm_destroy(eqhip_berm* B) {
    return guarded([&] {
        if (!B) return;
        (void)hipSetDevice(B->device);
        delete B;  // hipFree waits for the device
    });
}

}  // extern "C"
