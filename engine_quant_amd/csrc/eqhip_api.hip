// Host side of libeqhip: contexts, the exposure plan (schedule classes, cash-flow date table,
// per-date descriptors, device buffers), kernel launches and the C ABI of include/eqhip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "eqhip.h"
#include "eqhip_kernels.hpp"

using namespace eqhip;

namespace {

thread_local std::string g_last_error;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void raise(int code, const std::string& msg) { throw Error(code, msg); }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) raise(EQHIP_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHK(x) hip_check((x), #x)

template <class F>
int guarded(F&& f) {
    try {
        f();
        return EQHIP_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return EQHIP_E_CAPACITY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return EQHIP_E_INVALID;
    }
}

// Device allocation owned by a plan; freed when the plan dies.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t n) {
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
        bytes = n;
        if (n) HIPCHK(hipMalloc(&p, n));
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Synchronous copy: the source vector may be a temporary of the caller.
template <class T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t) {
    b.alloc(std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty()) HIPCHK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

// Host twin of the affine ZCB (models/hull_white.rs:50-80), used only for the par rate of
// use_par_rate swaps (api.rs:53-57, irs.rs:126-145), a deterministic set-up scalar.
double host_zcb(double a, double b, double sigma, double r, double t, double maturity) {
    double tau = maturity - t;
    double bf = ((-std::exp(a * (-tau))) + 1.0) / a;
    double a2 = a * a, s2 = sigma * sigma;
    double term1 = ((bf + (-tau)) * (a2 * b - s2 * 0.5)) / a2;
    double term2 = ((s2 * bf) * bf) / (a * 4.0);
    return std::exp(term1 - term2) * std::exp(-(bf * r));
}

double host_par_rate(double a, double b, double sigma, double r, double start, const double* pay,
                     const double* acc, int m) {
    double num = host_zcb(a, b, sigma, r, start, start) - host_zcb(a, b, sigma, r, start, pay[m - 1]);
    double den = 0.0;
    for (int i = 0; i < m; ++i) {
        double leg = host_zcb(a, b, sigma, r, start, pay[i]) * acc[i];
        den = (i == 0) ? leg : den + leg;
    }
    return num / den;
}

bool is_reset_date(double start, const double* pay, int m, double t) {
    if (std::fabs(t - start) < kDateTolHost) return true;
    for (int i = 0; i + 1 < m; ++i)
        if (std::fabs(pay[i] - t) < kDateTolHost) return true;
    return false;
}

}  // namespace

struct eqhip_context {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t lds_limit = 65536;
};

struct eqhip_plan {
    eqhip_context* ctx = nullptr;
    eqhip_model_hw1f model{};
    eqhip_mc mc{};
    int n_dates = 0, n_trades = 0, S = 0, C = 0, NF = 0, D = 0, Fmax = 0, n_slots = 0;
    bool det = false;
    double dt = 1.0;
    int tiles_per_block = 1, n_chunks = 1;
    uint64_t n_tiles = 0;
    int want_trade = 0;
    int n_items = 0;
    int set_mode = 0;
    std::vector<double> dates;
    size_t lds_bytes = 0;

    DevBuf slot_of_step, R, Z, AB, date_slot, date_dfirst, date_det, fdesc, fam_off, flow_d, flow_tau,
        flow_cls, fix_off, fix_slot, fix_AB, set_off, set_cls, set_alpha, set_beta, X, items, cls_mat, tr_K,
        tr_absN, tr_N, tr_NK, tr_orig, partial, trade_ee, trade_cva, set_ee, set_pfe, set_cva, neff, qrank,
        d_dates, vplan, units;
    ValuePlan vp{};  // host copy of the device descriptor (per-run pointers patched at launch)
    std::vector<int4> unit_list;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    eqhip_timing timing{};

    ~eqhip_plan() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

namespace {

struct ClassKey {
    double start;
    std::vector<double> pay, acc;
    bool operator<(const ClassKey& o) const {
        return std::tie(start, pay, acc) < std::tie(o.start, o.pay, o.acc);
    }
};

void validate_portfolio(const eqhip_portfolio* pf) {
    if (!pf) raise(EQHIP_E_INVALID, "portfolio is null");
    if (pf->n_trades < 1) raise(EQHIP_E_INVALID, "portfolio has no trades");
    if (pf->n_sets < 1) raise(EQHIP_E_INVALID, "portfolio has no netting sets");
    if (!pf->set_id || !pf->notional || !pf->fixed_rate || !pf->start || !pf->flow_offset ||
        !pf->payment_times || !pf->accruals)
        raise(EQHIP_E_INVALID, "portfolio has a null array");
    if (pf->flow_offset[0] != 0) raise(EQHIP_E_INVALID, "flow_offset[0] must be 0");
    for (int j = 0; j < pf->n_trades; ++j) {
        int s = pf->set_id[j];
        if (s < 0 || s >= pf->n_sets) raise(EQHIP_E_INVALID, "set_id out of range for trade " + std::to_string(j));
        int f0 = pf->flow_offset[j], f1 = pf->flow_offset[j + 1];
        if (f1 <= f0)
            raise(EQHIP_E_INVALID, "trade " + std::to_string(j) + ": a swap needs at least one period");
        for (int f = f0; f < f1; ++f) {
            double t = pf->payment_times[f];
            double prev = (f == f0) ? pf->start[j] : pf->payment_times[f - 1];
            if (!(t > prev))
                raise(EQHIP_E_INVALID, "trade " + std::to_string(j) +
                                           ": payment times must increase strictly after start");
        }
    }
}

void choose_launch(eqhip_plan& P) {
    P.set_mode = (P.S <= kMaxSplitSets) ? 0 : 1;
    size_t lds = (size_t)8 * kPathsPerTile * (1 + P.D + P.Fmax) + (size_t)16 * kPathsPerTile * P.C +
                 (P.set_mode == 0 ? (size_t)8 * kPathsPerTile * kValueWaves * P.S : 0);
    if (lds > P.ctx->lds_limit)
        raise(EQHIP_E_CAPACITY, "portfolio too diverse for one plan: " + std::to_string(P.D) +
                                    " cash-flow dates and " + std::to_string(P.C) + " schedule classes need " +
                                    std::to_string(lds) + " B of LDS (limit " + std::to_string(P.ctx->lds_limit) +
                                    "); split it into several plans");
    P.lds_bytes = lds;
    P.n_tiles = (P.mc.n_paths + kPathsPerTile - 1) / kPathsPerTile;
    // enough units for ~8 blocks per CU (one 512-thread block per CU at this LDS footprint)
    uint64_t target_chunks = std::max<uint64_t>(1, (2048 + P.n_dates - 1) / P.n_dates);
    uint64_t tpb = std::max<uint64_t>(1, (P.n_tiles + target_chunks - 1) / target_chunks);
    P.tiles_per_block = (int)tpb;
    P.n_chunks = (int)((P.n_tiles + tpb - 1) / tpb);
}

void build_value_plan(eqhip_plan& P);

void build_plan(eqhip_plan& P, const eqhip_model_hw1f* model, const eqhip_portfolio* pf,
                const double* dates, int n_dates, const eqhip_mc* mc, int want_trade) {
    if (!model || !mc) raise(EQHIP_E_INVALID, "model / mc is null");
    if (!(model->a != 0.0)) raise(EQHIP_E_INVALID, "HullWhite1F: a must be non-zero");
    validate_portfolio(pf);
    if (n_dates < 1 || !dates) raise(EQHIP_E_INVALID, "monitoring_times vacío: at least one date is required");
    if (mc->n_paths < 1) raise(EQHIP_E_INVALID, "PricingContext: n_paths must be > 0");
    if (mc->n_steps < 1) raise(EQHIP_E_INVALID, "PricingContext: n_steps must be > 0");
    if (mc->rng != EQHIP_RNG_HOST_DRAWS && mc->rng != EQHIP_RNG_PHILOX) raise(EQHIP_E_INVALID, "unknown rng");
    if (mc->disc != EQHIP_DISC_EULER && mc->disc != EQHIP_DISC_EXACT) raise(EQHIP_E_INVALID, "unknown discretisation");
    if (mc->n_paths > (uint64_t)1 << 40) raise(EQHIP_E_CAPACITY, "n_paths too large");

    P.model = *model;
    P.mc = *mc;
    P.n_dates = n_dates;
    P.dates.assign(dates, dates + n_dates);
    P.n_trades = pf->n_trades;
    P.S = pf->n_sets;
    P.want_trade = want_trade ? 1 : 0;
    hipStream_t st = P.ctx->stream;

    // time grid (exposure.rs:91-116): dt = t_max / n_steps, idx = round(t / dt)
    double t_max = -INFINITY;
    for (double t : P.dates) t_max = std::max(t_max, t);
    P.det = !(t_max > 0.0);
    P.dt = P.det ? 1.0 : t_max / (double)mc->n_steps;
    auto grid_of = [&](double t) -> int64_t {
        if (P.det) return 0;
        double g = std::round(t / P.dt);
        return g < 0.0 ? 0 : (int64_t)g;
    };
    const uint32_t steps = P.det ? 0u : mc->n_steps;

    // schedule classes
    std::map<ClassKey, int> cls_of;
    std::vector<ClassKey> classes;
    std::vector<int> trade_cls(pf->n_trades);
    for (int j = 0; j < pf->n_trades; ++j) {
        int f0 = pf->flow_offset[j], f1 = pf->flow_offset[j + 1];
        ClassKey key{pf->start[j], std::vector<double>(pf->payment_times + f0, pf->payment_times + f1),
                     std::vector<double>(pf->accruals + f0, pf->accruals + f1)};
        auto it = cls_of.find(key);
        if (it == cls_of.end()) {
            it = cls_of.emplace(key, (int)classes.size()).first;
            classes.push_back(key);
        }
        trade_cls[j] = it->second;
    }
    P.C = (int)classes.size();

    // cash-flow date table
    std::vector<double> T;
    for (auto& c : classes) {
        T.push_back(c.start);
        T.insert(T.end(), c.pay.begin(), c.pay.end());
    }
    std::sort(T.begin(), T.end());
    T.erase(std::unique(T.begin(), T.end()), T.end());
    P.D = (int)T.size();
    auto didx = [&](double t) { return (int)(std::lower_bound(T.begin(), T.end(), t) - T.begin()); };

    // schedule families: classes whose (start, payment, accrual) sequences are prefixes of one
    // representative schedule share one running annuity sum
    std::vector<int> cls_order(P.C);
    for (int c = 0; c < P.C; ++c) cls_order[c] = c;
    std::stable_sort(cls_order.begin(), cls_order.end(),
                     [&](int x, int y) { return classes[x].pay.size() > classes[y].pay.size(); });
    std::vector<int> fam_rep;               // representative class of each family
    std::vector<int> cls_fam(P.C), cls_len(P.C);
    for (int c : cls_order) {
        const ClassKey& K = classes[c];
        const size_t m = K.pay.size();
        int fam = -1;
        for (size_t F = 0; F < fam_rep.size() && fam < 0; ++F) {
            const ClassKey& R = classes[fam_rep[F]];
            if (R.start == K.start && std::equal(K.pay.begin(), K.pay.end(), R.pay.begin()) &&
                std::equal(K.acc.begin(), K.acc.end(), R.acc.begin()))
                fam = (int)F;
        }
        if (fam < 0) {
            fam = (int)fam_rep.size();
            fam_rep.push_back(c);
        }
        cls_fam[c] = fam;
        cls_len[c] = (int)m;
    }
    P.NF = (int)fam_rep.size();
    std::vector<int> fam_off(P.NF + 1, 0), flow_d, flow_cls;
    std::vector<double> flow_tau;
    for (int F = 0; F < P.NF; ++F) {
        const ClassKey& R = classes[fam_rep[F]];
        for (size_t i = 0; i < R.pay.size(); ++i) {
            flow_d.push_back(didx(R.pay[i]));
            flow_tau.push_back(R.acc[i]);
            flow_cls.push_back(-1);
        }
        fam_off[F + 1] = (int)flow_d.size();
    }
    for (int c = 0; c < P.C; ++c) flow_cls[fam_off[cls_fam[c]] + cls_len[c] - 1] = c;
    std::vector<double> cls_mat(P.C);
    for (int c = 0; c < P.C; ++c) cls_mat[c] = classes[c].pay.back();

    // slots: grid indices at which r is stored
    std::map<int64_t, int> slot_of;
    auto slot = [&](int64_t g) {
        auto it = slot_of.find(g);
        if (it == slot_of.end()) it = slot_of.emplace(g, (int)slot_of.size()).first;
        return it->second;
    };
    std::vector<int> date_slot(n_dates), date_dfirst(n_dates), date_det(n_dates);
    std::vector<int4> fdesc((size_t)n_dates * P.NF);
    std::vector<int> fix_off(n_dates + 1, 0), fix_slot;
    std::vector<double> fix_tau;
    std::vector<double> ab_tau((size_t)n_dates * P.D, -1.0);
    for (int k = 0; k < n_dates; ++k) {
        const double t = P.dates[k];
        const int64_t g = grid_of(t);
        date_slot[k] = slot(g);
        date_det[k] = (g == 0) ? 1 : 0;
        int dfirst = (int)(std::upper_bound(T.begin(), T.end(), t + kDateTolHost) - T.begin());
        date_dfirst[k] = dfirst;
        for (int d = dfirst; d < P.D; ++d) ab_tau[(size_t)k * P.D + d] = T[d] - t;
        std::map<std::pair<int, double>, int> fix_here;
        // a family is alive while its representative (longest) schedule is; every live member
        // class shares the representative's period state at t (same start, same earlier flows)
        for (int F = 0; F < P.NF; ++F) {
            const ClassKey& K = classes[fam_rep[F]];
            const int m = (int)K.pay.size();
            int4 fd = make_int4(0, 0, 0, 0);
            if (t < K.pay[m - 1] - kDateTolHost) {
                int first = 0;
                while (first < m && !(K.pay[first] > t + kDateTolHost)) ++first;
                fd.y = first;
                if (std::fabs(t - K.start) < kDateTolHost) {
                    fd.x = 1;
                } else if (t < K.start) {
                    fd.x = 2;
                    fd.z = didx(K.start);
                } else {
                    double prev = (first == 0) ? K.start : K.pay[first - 1];
                    if (std::fabs(t - prev) < kDateTolHost) {
                        fd.x = 1;
                    } else {
                        fd.x = 3;
                        fd.z = didx(K.pay[first]);
                        int fs = slot(grid_of(prev));
                        double ftau = K.pay[first] - prev;
                        auto key = std::make_pair(fs, ftau);
                        auto it = fix_here.find(key);
                        if (it == fix_here.end()) {
                            it = fix_here.emplace(key, (int)fix_here.size()).first;
                            fix_slot.push_back(fs);
                            fix_tau.push_back(ftau);
                        }
                        fd.w = it->second;
                    }
                }
            }
            fdesc[(size_t)k * P.NF + F] = fd;
        }
        fix_off[k + 1] = (int)fix_slot.size();
        P.Fmax = std::max(P.Fmax, (int)fix_here.size());
    }
    P.n_slots = (int)slot_of.size();
    std::vector<int> slot_of_step(steps + 1, -1);
    for (auto& kv : slot_of) slot_of_step[(size_t)kv.first] = kv.second;

    // netting sets: per (set, class) sums of N and N*K in trade order
    std::vector<std::map<int, std::pair<double, double>>> set_cls_sum(P.S);
    for (int j = 0; j < pf->n_trades; ++j) {
        auto& e = set_cls_sum[pf->set_id[j]][trade_cls[j]];
        double n = pf->notional[j];
        e.first += n;
        e.second += n * pf->fixed_rate[j];
    }
    std::vector<int> set_off(P.S + 1, 0), set_cls;
    std::vector<double> set_alpha, set_beta;
    for (int s = 0; s < P.S; ++s) {
        for (auto& kv : set_cls_sum[s]) {
            set_cls.push_back(kv.first);
            set_alpha.push_back(kv.second.first);
            set_beta.push_back(kv.second.second);
        }
        set_off[s + 1] = (int)set_cls.size();
    }

    // per-trade work items: trades sorted by (class, sign), up to kTradesPerItem per item
    std::vector<int4> items;
    std::vector<int> tr_orig;
    std::vector<double> tr_K, tr_absN, tr_N, tr_NK;
    if (P.want_trade) {
        auto sgn = [&](int j) { return pf->notional[j] < 0.0 ? -1 : 1; };
        std::vector<int> order(pf->n_trades);
        for (int j = 0; j < pf->n_trades; ++j) order[j] = j;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return std::make_pair(trade_cls[x], sgn(x)) < std::make_pair(trade_cls[y], sgn(y));
        });
        for (int i = 0; i < pf->n_trades;) {
            const int c = trade_cls[order[i]], sg = sgn(order[i]);
            int4 item = make_int4(c, i, 0, sg);
            while (i < pf->n_trades && trade_cls[order[i]] == c && sgn(order[i]) == sg && item.z < kTradesPerItem) {
                const int j = order[i];
                tr_orig.push_back(j);
                tr_K.push_back(pf->fixed_rate[j]);
                tr_absN.push_back(std::fabs(pf->notional[j]));
                tr_N.push_back(pf->notional[j]);
                tr_NK.push_back(pf->notional[j] * pf->fixed_rate[j]);
                ++i;
                ++item.z;
            }
            items.push_back(item);
        }
        P.n_items = (int)items.size();
        if (P.n_items > kValueThreads * kItemsPerThread)
            raise(EQHIP_E_CAPACITY, "per-trade EE supports at most " +
                                        std::to_string(kValueThreads * kItemsPerThread) +
                                        " work items per plan (got " + std::to_string(P.n_items) + ")");
    }

    choose_launch(P);

    // effective path counts and PFE ranks (exposure.rs:42-49; grid index 0 has one row)
    std::vector<uint64_t> neff(n_dates), qrank(n_dates);
    for (int k = 0; k < n_dates; ++k) {
        uint64_t n = date_det[k] ? 1 : mc->n_paths;
        uint64_t q = (uint64_t)std::ceil(0.95 * (double)n);
        q = (q == 0) ? 0 : q - 1;
        if (q > n - 1) q = n - 1;
        neff[k] = n;
        qrank[k] = q;
    }

    // memory check for the large buffers
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    double need = 8.0 * ((double)P.n_slots * mc->n_paths + (double)P.S * n_dates * mc->n_paths +
                         (P.want_trade ? (double)P.n_trades * n_dates * P.n_chunks : 0.0) +
                         (mc->rng == EQHIP_RNG_HOST_DRAWS ? (double)mc->n_steps * mc->n_paths : 0.0));
    if (need > 0.9 * (double)free_b)
        raise(EQHIP_E_CAPACITY, "exposure plan needs " + std::to_string(need / 1e9) + " GB of device memory, " +
                                    std::to_string(free_b / 1e9) + " GB free");

    // uploads
    upload(P.slot_of_step, slot_of_step, st);
    upload(P.date_slot, date_slot, st);
    upload(P.date_dfirst, date_dfirst, st);
    upload(P.date_det, date_det, st);
    upload(P.fdesc, fdesc, st);
    upload(P.fam_off, fam_off, st);
    upload(P.flow_d, flow_d, st);
    upload(P.flow_tau, flow_tau, st);
    upload(P.flow_cls, flow_cls, st);
    upload(P.cls_mat, cls_mat, st);
    upload(P.fix_off, fix_off, st);
    upload(P.fix_slot, fix_slot, st);
    upload(P.set_off, set_off, st);
    upload(P.set_cls, set_cls, st);
    upload(P.set_alpha, set_alpha, st);
    upload(P.set_beta, set_beta, st);
    upload(P.neff, neff, st);
    upload(P.qrank, qrank, st);
    upload(P.d_dates, P.dates, st);
    if (P.want_trade) {
        upload(P.items, items, st);
        upload(P.tr_K, tr_K, st);
        upload(P.tr_absN, tr_absN, st);
        upload(P.tr_N, tr_N, st);
        upload(P.tr_NK, tr_NK, st);
        upload(P.tr_orig, tr_orig, st);
        P.partial.alloc((size_t)P.n_trades * n_dates * P.n_chunks * 8);
        P.trade_ee.alloc((size_t)P.n_trades * n_dates * 8);
        P.trade_cva.alloc((size_t)P.n_trades * 8);
    }
    // ZCB factor tables, computed on the device
    {
        DevBuf tau;
        upload(tau, ab_tau, st);
        P.AB.alloc(ab_tau.size() * sizeof(double2));
        int n = (int)ab_tau.size();
        hipLaunchKernelGGL(ab_kernel, dim3((n + 255) / 256), dim3(256), 0, st, tau.as<double>(), n,
                           P.model.a, P.model.b, P.model.sigma, P.AB.as<double2>());
        HIPCHK(hipGetLastError());
        DevBuf ftau;
        upload(ftau, fix_tau, st);
        P.fix_AB.alloc(std::max<size_t>(1, fix_tau.size()) * sizeof(double2));
        int nf = (int)fix_tau.size();
        if (nf > 0) {
            hipLaunchKernelGGL(ab_kernel, dim3((nf + 255) / 256), dim3(256), 0, st, ftau.as<double>(), nf,
                               P.model.a, P.model.b, P.model.sigma, P.fix_AB.as<double2>());
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(st));  // tau buffers die here
    }
    P.R.alloc((size_t)P.n_slots * mc->n_paths * 8);
    P.X.alloc((size_t)P.S * n_dates * mc->n_paths * 8);
    P.set_ee.alloc((size_t)P.S * n_dates * 8);
    P.set_pfe.alloc((size_t)P.S * n_dates * 8);
    P.set_cva.alloc((size_t)P.S * 8);
    if (mc->rng == EQHIP_RNG_HOST_DRAWS) P.Z.alloc((size_t)mc->n_steps * mc->n_paths * 8);
    build_value_plan(P);
    for (auto& e : P.ev) HIPCHK(hipEventCreate(&e));
}

// Device descriptor of a plan (pointers into the plan's buffers) and its work units.
void build_value_plan(eqhip_plan& P) {
    ValuePlan& g = P.vp;
    g = ValuePlan{};
    g.R = P.R.as<double>();
    g.n_paths = P.mc.n_paths;
    g.n_tiles = P.n_tiles;
    g.n_dates = P.n_dates;
    g.D = P.D;
    g.C = P.C;
    g.NF = P.NF;
    g.S = P.S;
    g.Fmax = P.Fmax;
    g.n_chunks = P.n_chunks;
    g.tiles_per_block = P.tiles_per_block;
    g.set_mode = P.set_mode;
    g.dates = P.d_dates.as<double>();
    g.AB = P.AB.as<double2>();
    g.date_slot = P.date_slot.as<int>();
    g.date_dfirst = P.date_dfirst.as<int>();
    g.date_det = P.date_det.as<int>();
    g.fdesc = P.fdesc.as<int4>();
    g.fam_off = P.fam_off.as<int>();
    g.flow_d = P.flow_d.as<int>();
    g.flow_tau = P.flow_tau.as<double>();
    g.flow_cls = P.flow_cls.as<int>();
    g.fix_off = P.fix_off.as<int>();
    g.fix_slot = P.fix_slot.as<int>();
    g.fix_AB = P.fix_AB.as<double2>();
    g.set_off = P.set_off.as<int>();
    g.set_cls = P.set_cls.as<int>();
    g.set_alpha = P.set_alpha.as<double>();
    g.set_beta = P.set_beta.as<double>();
    g.X = P.X.as<double>();
    g.want_trade = P.want_trade;
    g.n_items = P.n_items;
    g.n_trades = P.n_trades;
    g.items = P.items.as<int4>();
    g.cls_mat = P.cls_mat.as<double>();
    g.tr_K = P.tr_K.as<double>();
    g.tr_absN = P.tr_absN.as<double>();
    g.tr_N = P.tr_N.as<double>();
    g.tr_NK = P.tr_NK.as<double>();
    g.tr_orig = P.tr_orig.as<int>();
    g.partial = P.partial.as<double>();
    P.vplan.alloc(sizeof(ValuePlan));
    HIPCHK(hipMemcpy(P.vplan.p, &g, sizeof(ValuePlan), hipMemcpyHostToDevice));
    P.unit_list.clear();
    for (int k = 0; k < P.n_dates; ++k)
        for (int c = 0; c < P.n_chunks; ++c) P.unit_list.push_back(make_int4(0, k, c, 0));
    upload(P.units, P.unit_list, P.ctx->stream);
}

void launch_value(eqhip_plan& P, double* path_mtm, double* set_mtm) {
    hipStream_t st = P.ctx->stream;
    const bool patched = path_mtm || set_mtm || P.vp.path_mtm || P.vp.set_mtm;
    if (patched) {  // per-run debug outputs live in the descriptor
        P.vp.path_mtm = path_mtm;
        P.vp.set_mtm = set_mtm;
        HIPCHK(hipMemcpy(P.vplan.p, &P.vp, sizeof(ValuePlan), hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(value_kernel, dim3((unsigned)P.unit_list.size()), dim3(kValueThreads), P.lds_bytes, st,
                       P.vplan.as<const ValuePlan>(), P.units.as<const int4>());
    HIPCHK(hipGetLastError());
}

void execute_plan(eqhip_plan& P, const double* host_draws, const eqhip_credit* credit,
                  eqhip_outputs* out) {
    if (!out || !out->set_ee || !out->set_pfe) raise(EQHIP_E_INVALID, "outputs set_ee / set_pfe are required");
    if ((out->trade_ee || out->trade_cva) && !P.want_trade)
        raise(EQHIP_E_INVALID, "per-trade outputs need a plan created with want_trade_ee");
    if ((out->set_cva || out->trade_cva) && !credit) raise(EQHIP_E_INVALID, "CVA needs credit data");
    hipStream_t st = P.ctx->stream;
    HIPCHK(hipSetDevice(P.ctx->device));
    const uint64_t np = P.mc.n_paths;
    if (P.mc.rng == EQHIP_RNG_HOST_DRAWS) {
        if (!host_draws) raise(EQHIP_E_INVALID, "host_draws required in EQHIP_RNG_HOST_DRAWS mode");
        HIPCHK(hipMemcpyAsync(P.Z.p, host_draws, (size_t)P.mc.n_steps * np * 8, hipMemcpyHostToDevice, st));
    }
    DevBuf dpath, dset;
    if (out->path_mtm) {
        if (!P.want_trade) raise(EQHIP_E_INVALID, "path_mtm needs a plan created with want_trade_ee");
        dpath.alloc((size_t)P.n_dates * P.n_trades * np * 8);
    }
    if (out->set_mtm) dset.alloc((size_t)P.n_dates * P.S * np * 8);

    HIPCHK(hipEventRecord(P.ev[0], st));
    SimArgs s{};
    s.a = P.model.a;
    s.b = P.model.b;
    s.sigma = P.model.sigma;
    s.r0 = P.model.r0;
    s.dt = P.dt;
    s.sqrt_dt = std::sqrt(P.dt);
    s.decay = std::exp(-(P.model.a * P.dt));
    s.mean_c = P.model.b * (1.0 - s.decay);
    s.sd = P.model.sigma * std::sqrt((1.0 - std::exp(-(2.0 * P.model.a * P.dt))) / (2.0 * P.model.a));
    s.n_paths = np;
    s.n_steps = P.det ? 0u : P.mc.n_steps;
    s.seed = P.mc.seed;
    s.slot_of_step = P.slot_of_step.as<int>();
    s.Z = P.Z.as<double>();
    s.R = P.R.as<double>();
    dim3 sgrid((unsigned)((np + 255) / 256));
    int key = P.mc.disc * 2 + P.mc.rng;
    switch (key) {
        case 0: hipLaunchKernelGGL((simulate_kernel<0, 0>), sgrid, dim3(256), 0, st, s); break;
        case 1: hipLaunchKernelGGL((simulate_kernel<0, 1>), sgrid, dim3(256), 0, st, s); break;
        case 2: hipLaunchKernelGGL((simulate_kernel<1, 0>), sgrid, dim3(256), 0, st, s); break;
        default: hipLaunchKernelGGL((simulate_kernel<1, 1>), sgrid, dim3(256), 0, st, s); break;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(P.ev[1], st));
    launch_value(P, dpath.as<double>(), dset.as<double>());
    HIPCHK(hipEventRecord(P.ev[2], st));
    hipLaunchKernelGGL(select_kernel, dim3((unsigned)(P.S * P.n_dates)), dim3(kSelectThreads), 0, st,
                       P.X.as<double>(), P.n_dates, np, P.neff.as<uint64_t>(), P.qrank.as<uint64_t>(),
                       P.set_ee.as<double>(), P.set_pfe.as<double>());
    HIPCHK(hipGetLastError());
    if (P.want_trade) {
        int n = P.n_trades * P.n_dates;
        hipLaunchKernelGGL(trade_ee_kernel, dim3((n + 255) / 256), dim3(256), 0, st, P.partial.as<double>(),
                           P.n_trades, P.n_dates, P.n_chunks, P.neff.as<uint64_t>(), P.trade_ee.as<double>());
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(P.ev[3], st));
    double h = credit ? credit->hazard_rate : 0.0, rr = credit ? credit->recovery_rate : 0.0;
    if (out->set_cva) {
        hipLaunchKernelGGL(cva_kernel, dim3((P.S + 63) / 64), dim3(64), 0, st, P.set_ee.as<double>(), P.S,
                           P.n_dates, P.d_dates.as<double>(), P.model.a, P.model.b, P.model.sigma,
                           P.model.r0, h, rr, P.set_cva.as<double>());
        HIPCHK(hipGetLastError());
    }
    if (out->trade_cva) {
        hipLaunchKernelGGL(cva_kernel, dim3((P.n_trades + 63) / 64), dim3(64), 0, st, P.trade_ee.as<double>(),
                           P.n_trades, P.n_dates, P.d_dates.as<double>(), P.model.a, P.model.b,
                           P.model.sigma, P.model.r0, h, rr, P.trade_cva.as<double>());
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(P.ev[4], st));

    auto d2h = [&](void* dst, const DevBuf& src, size_t bytes) {
        if (dst) HIPCHK(hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToHost, st));
    };
    d2h(out->set_ee, P.set_ee, (size_t)P.S * P.n_dates * 8);
    d2h(out->set_pfe, P.set_pfe, (size_t)P.S * P.n_dates * 8);
    d2h(out->set_cva, P.set_cva, (size_t)P.S * 8);
    d2h(out->trade_ee, P.trade_ee, (size_t)P.n_trades * P.n_dates * 8);
    d2h(out->trade_cva, P.trade_cva, (size_t)P.n_trades * 8);
    d2h(out->path_mtm, dpath, (size_t)P.n_dates * P.n_trades * np * 8);
    d2h(out->set_mtm, dset, (size_t)P.n_dates * P.S * np * 8);
    HIPCHK(hipStreamSynchronize(st));

    float ms[4];
    for (int i = 0; i < 4; ++i) HIPCHK(hipEventElapsedTime(&ms[i], P.ev[i], P.ev[i + 1]));
    P.timing.simulate_ms = ms[0];
    P.timing.value_ms = ms[1];
    P.timing.reduce_ms = ms[2];
    P.timing.cva_ms = ms[3];
    P.timing.total_ms = (double)ms[0] + ms[1] + ms[2] + ms[3];
    P.timing.n_value_blocks = (uint64_t)P.n_dates * P.n_chunks;
    P.timing.paths_per_tile = (uint32_t)kPathsPerTile;
    P.timing.tiles_per_block = (uint32_t)P.tiles_per_block;
}

eqhip_context* check_ctx(eqhip_context* ctx) {
    if (!ctx) raise(EQHIP_E_INVALID, "context is null");
    HIPCHK(hipSetDevice(ctx->device));
    return ctx;
}

// Shared body of the two reference-shaped profile entry points: singleton netting sets over a
// shared schedule.
void profile_batch(eqhip_context* ctx, double a, double b, double sigma, double r0,
                   const double* notionals, const double* fixed_rates, int n_trades, double start,
                   const double* pay, const double* acc, int m, const double* times, int n_dates,
                   uint32_t n_steps, uint64_t n_paths, uint64_t seed, int rng, int disc,
                   const double* host_draws, int allow_any_date, double* out_times, double* out_ee,
                   double* out_pfe, int32_t* n_out) {
    check_ctx(ctx);
    if (n_dates < 1 || !times) raise(EQHIP_E_INVALID, "monitoring_times vacío");
    if (n_steps < 1) raise(EQHIP_E_INVALID, "n_steps debe ser al menos 1");
    if (m < 1 || !pay || !acc) raise(EQHIP_E_INVALID, "un swap necesita al menos un periodo");
    if (!allow_any_date)
        for (int k = 0; k < n_dates; ++k)
            if (!is_reset_date(start, pay, m, times[k]))
                raise(EQHIP_E_INVALID, "t=" + std::to_string(times[k]) + " no es una fecha de reseteo válida del swap");
    double t_max = -INFINITY;
    for (int k = 0; k < n_dates; ++k) t_max = std::max(t_max, times[k]);
    int nd = (t_max <= 0.0) ? 1 : n_dates;  // exposure.rs:98-108 returns only the first date
    std::vector<int32_t> set_id(n_trades), off(n_trades + 1);
    std::vector<double> st(n_trades, start), P_(n_trades * (size_t)m), A_(n_trades * (size_t)m);
    for (int j = 0; j < n_trades; ++j) {
        set_id[j] = j;
        off[j] = j * m;
        std::copy(pay, pay + m, P_.begin() + (size_t)j * m);
        std::copy(acc, acc + m, A_.begin() + (size_t)j * m);
    }
    off[n_trades] = n_trades * m;
    eqhip_portfolio pf{n_trades, n_trades, set_id.data(), notionals, fixed_rates, st.data(),
                       off.data(), P_.data(), A_.data()};
    eqhip_model_hw1f model{a, b, sigma, r0};
    eqhip_mc mc{n_paths, n_steps, seed, rng, disc};
    eqhip_plan P;
    P.ctx = ctx;
    build_plan(P, &model, &pf, times, nd, &mc, 0);
    std::vector<double> ee((size_t)n_trades * nd), pfe((size_t)n_trades * nd);
    eqhip_outputs out{};
    out.set_ee = ee.data();
    out.set_pfe = pfe.data();
    execute_plan(P, host_draws, nullptr, &out);
    for (int j = 0; j < n_trades; ++j)
        for (int k = 0; k < nd; ++k) {
            out_ee[(size_t)j * n_dates + k] = ee[(size_t)j * nd + k];
            out_pfe[(size_t)j * n_dates + k] = pfe[(size_t)j * nd + k];
        }
    for (int k = 0; k < nd; ++k) out_times[k] = times[k];
    if (n_out) *n_out = nd;
}

}  // namespace

extern "C" {

int eqhip_is_gpu_backend_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        return 0;
    }
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
            return 1;
    }
    return 0;
}

size_t eqhip_last_error(char* buf, size_t len) {
    if (buf && len) {
        size_t n = std::min(len - 1, g_last_error.size());
        std::memcpy(buf, g_last_error.data(), n);
        buf[n] = '\0';
    }
    return g_last_error.size();
}

int eqhip_context_create(int device, eqhip_context** out) {
    return guarded([&] {
        if (!out) raise(EQHIP_E_INVALID, "out is null");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
            (void)hipGetLastError();
            raise(EQHIP_E_NO_DEVICE, "backend 'gpu' requested but no HIP device is visible");
        }
        if (device < 0 || device >= n) raise(EQHIP_E_NO_DEVICE, "device index out of range");
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            raise(EQHIP_E_NO_DEVICE, std::string("this build targets gfx950, device is ") + prop.gcnArchName);
        auto ctx = std::make_unique<eqhip_context>();
        ctx->device = device;
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->lds_limit = prop.sharedMemPerBlock;  // 160 KB per workgroup on gfx950
        if (hipFuncSetAttribute((const void*)value_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)ctx->lds_limit) != hipSuccess) {
            (void)hipGetLastError();
            ctx->lds_limit = std::min<size_t>(ctx->lds_limit, 65536);
        }
        *out = ctx.release();
    });
}

int eqhip_context_destroy(eqhip_context* ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int eqhip_plan_create(eqhip_context* ctx, const eqhip_model_hw1f* model, const eqhip_portfolio* portfolio,
                      const double* dates, int32_t n_dates, const eqhip_mc* mc, int32_t want_trade_ee,
                      eqhip_plan** out) {
    return guarded([&] {
        if (!out) raise(EQHIP_E_INVALID, "out is null");
        *out = nullptr;
        check_ctx(ctx);
        auto P = std::make_unique<eqhip_plan>();
        P->ctx = ctx;
        build_plan(*P, model, portfolio, dates, n_dates, mc, want_trade_ee);
        *out = P.release();
    });
}

int eqhip_plan_execute(eqhip_plan* plan, const double* host_draws, const eqhip_credit* credit,
                       eqhip_outputs* out) {
    return guarded([&] {
        if (!plan) raise(EQHIP_E_INVALID, "plan is null");
        execute_plan(*plan, host_draws, credit, out);
    });
}

int eqhip_plan_timing(const eqhip_plan* plan, eqhip_timing* out) {
    return guarded([&] {
        if (!plan || !out) raise(EQHIP_E_INVALID, "null argument");
        *out = plan->timing;
    });
}

int eqhip_plan_destroy(eqhip_plan* plan) {
    return guarded([&] {
        if (!plan) return;
        (void)hipSetDevice(plan->ctx->device);
        delete plan;
    });
}

int eqhip_exposure_run(eqhip_context* ctx, const eqhip_model_hw1f* model, const eqhip_portfolio* portfolio,
                       const double* dates, int32_t n_dates, const eqhip_mc* mc, const double* host_draws,
                       const eqhip_credit* credit, eqhip_outputs* out) {
    return guarded([&] {
        check_ctx(ctx);
        eqhip_plan P;
        P.ctx = ctx;
        int want_trade = out && (out->trade_ee || out->trade_cva || out->path_mtm);
        build_plan(P, model, portfolio, dates, n_dates, mc, want_trade);
        execute_plan(P, host_draws, credit, out);
    });
}

int eqhip_irs_hull_white_exposure_profile(
    eqhip_context* ctx, double a, double b, double sigma, double r0, double notional, double fixed_rate,
    int use_par_rate, double start, const double* payment_times, const double* accruals, int32_t n_flows,
    const double* monitoring_times, int32_t n_dates, uint32_t n_steps, uint64_t n_paths, uint64_t seed,
    int32_t rng, int32_t disc, const double* host_draws, int32_t allow_any_date, double* out_times,
    double* out_ee, double* out_pfe, int32_t* n_out) {
    return guarded([&] {
        if (n_flows < 1 || !payment_times || !accruals) raise(EQHIP_E_INVALID, "un swap necesita al menos un periodo");
        double k = use_par_rate ? host_par_rate(a, b, sigma, r0, start, payment_times, accruals, n_flows) : fixed_rate;
        profile_batch(ctx, a, b, sigma, r0, &notional, &k, 1, start, payment_times, accruals, n_flows,
                      monitoring_times, n_dates, n_steps, n_paths, seed, rng, disc, host_draws, allow_any_date,
                      out_times, out_ee, out_pfe, n_out);
    });
}

int eqhip_irs_hull_white_exposure_profile_batch(
    eqhip_context* ctx, double a, double b, double sigma, double r0, const double* notionals,
    const double* fixed_rates, int32_t n_trades, double start, const double* payment_times,
    const double* accruals, int32_t n_flows, const double* monitoring_times, int32_t n_dates, uint32_t n_steps,
    uint64_t n_paths, uint64_t seed, int32_t rng, int32_t disc, const double* host_draws, int32_t allow_any_date,
    double* out_times, double* out_ee, double* out_pfe, int32_t* n_out) {
    return guarded([&] {
        if (n_trades < 1 || !notionals || !fixed_rates)
            raise(EQHIP_E_INVALID, "el lote requiere al menos un trade con notional y fixed_rate");
        profile_batch(ctx, a, b, sigma, r0, notionals, fixed_rates, n_trades, start, payment_times, accruals,
                      n_flows, monitoring_times, n_dates, n_steps, n_paths, seed, rng, disc, host_draws,
                      allow_any_date, out_times, out_ee, out_pfe, n_out);
    });
}

int eqhip_unilateral_cva_from_exposure_batch(eqhip_context* ctx, double a, double b, double sigma, double r0,
                                             const double* times, const double* ee, int32_t n_profiles,
                                             int32_t n_dates, double hazard_rate, double recovery_rate,
                                             double* out_cva) {
    return guarded([&] {
        check_ctx(ctx);
        if (n_profiles < 0 || n_dates < 0 || (n_dates > 0 && (!times || !ee)) || !out_cva)
            raise(EQHIP_E_INVALID, "invalid CVA arguments");
        if (n_profiles == 0) return;
        hipStream_t st = ctx->stream;
        DevBuf dt, de, dc;
        upload(dt, std::vector<double>(times, times + n_dates), st);
        upload(de, std::vector<double>(ee, ee + (size_t)n_profiles * n_dates), st);
        dc.alloc((size_t)n_profiles * 8);
        hipLaunchKernelGGL(cva_kernel, dim3((n_profiles + 63) / 64), dim3(64), 0, st, de.as<double>(), n_profiles,
                           n_dates, dt.as<double>(), a, b, sigma, r0, hazard_rate, recovery_rate, dc.as<double>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out_cva, dc.p, (size_t)n_profiles * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    });
}

int eqhip_unilateral_cva_from_exposure(eqhip_context* ctx, double a, double b, double sigma, double r0,
                                       const double* times, const double* ee, int32_t n_dates,
                                       double hazard_rate, double recovery_rate, double* out_cva) {
    return eqhip_unilateral_cva_from_exposure_batch(ctx, a, b, sigma, r0, times, ee, 1, n_dates, hazard_rate,
                                                    recovery_rate, out_cva);
}

int eqhip_pv_dv01(eqhip_context* ctx, const double* pillars, const double* zero_rates, int32_t n_pillars,
                  const eqhip_portfolio* pf, double bump, double* out_pv, double* out_dv01, double* out_bucketed) {
    return guarded([&] {
        check_ctx(ctx);
        validate_portfolio(pf);
        if (n_pillars < 1 || !pillars || !zero_rates) raise(EQHIP_E_INVALID, "Curve: necesita al menos un pillar");
        for (int i = 1; i < n_pillars; ++i)
            if (!(pillars[i] > pillars[i - 1])) raise(EQHIP_E_INVALID, "Curve: pillars debe ser estrictamente creciente");
        if (!out_pv || !out_dv01) raise(EQHIP_E_INVALID, "out_pv / out_dv01 are required");
        hipStream_t st = ctx->stream;
        int n = pf->n_trades, nf = pf->flow_offset[n];
        DevBuf dp, dz, doff, dpay, dacc, dn, dk, ds, dpv, ddv, dbk;
        upload(dp, std::vector<double>(pillars, pillars + n_pillars), st);
        upload(dz, std::vector<double>(zero_rates, zero_rates + n_pillars), st);
        upload(doff, std::vector<int>(pf->flow_offset, pf->flow_offset + n + 1), st);
        upload(dpay, std::vector<double>(pf->payment_times, pf->payment_times + nf), st);
        upload(dacc, std::vector<double>(pf->accruals, pf->accruals + nf), st);
        upload(dn, std::vector<double>(pf->notional, pf->notional + n), st);
        upload(dk, std::vector<double>(pf->fixed_rate, pf->fixed_rate + n), st);
        upload(ds, std::vector<double>(pf->start, pf->start + n), st);
        dpv.alloc((size_t)n * 8);
        ddv.alloc((size_t)n * 8);
        if (out_bucketed) dbk.alloc((size_t)n * n_pillars * 8);
        PvArgs g{};
        g.pillars = dp.as<double>();
        g.zeros = dz.as<double>();
        g.np = n_pillars;
        g.n_trades = n;
        g.flow_off = doff.as<int>();
        g.pay = dpay.as<double>();
        g.acc = dacc.as<double>();
        g.notional = dn.as<double>();
        g.fixed = dk.as<double>();
        g.start = ds.as<double>();
        g.bump = bump;
        g.pv = dpv.as<double>();
        g.dv01 = ddv.as<double>();
        g.bucketed = out_bucketed ? dbk.as<double>() : nullptr;
        g.bucketed_on = out_bucketed ? 1 : 0;
        int total = n * (2 + n_pillars);
        hipLaunchKernelGGL(pv_dv01_kernel, dim3((total + 127) / 128), dim3(128), 0, st, g);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out_pv, dpv.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(out_dv01, ddv.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
        if (out_bucketed)
            HIPCHK(hipMemcpyAsync(out_bucketed, dbk.p, (size_t)n * n_pillars * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    });
}

}  // extern "C"
