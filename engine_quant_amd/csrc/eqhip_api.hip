// Host side of libeqhip: contexts, the exposure plan (schedule classes, cash-flow date table,
// per-date descriptors, device buffers), kernel launches and the C ABI of include/eqhip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "eqhip.h"
#include "eqhip_host.hpp"
#include "eqhip_kernels.hpp"

using namespace eqhip;

namespace eqhip_host {
std::string& last_error() {
    thread_local std::string msg;
    return msg;
}
}  // namespace eqhip_host

using namespace eqhip_host;

namespace {

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

// Host twin of the G2++ bond at the initial state (0, 0) (models/hull_white_2f.rs:66-106), for the
// par rate of use_par_rate swaps (api.rs build_irs_swap_2f, irs.rs:126-145).
double host_hw2_var(double z, double vol, double tau) {
    const double inv_z = std::pow(z, -1.0);
    const double e1 = std::exp(z * (-tau));
    const double e2 = std::exp(z * (-2.0 * tau));
    const double bracket = (((inv_z * 2.0) * e1 - (inv_z * 0.5) * e2) - inv_z * 1.5) + tau;
    return ((vol * vol) / (z * z)) * bracket;
}
double host_zcb2_at_origin(const eqhip_model_hw2f& m, double tau) {
    const double apb = m.a + m.b;
    const double ta = (std::exp(m.a * (-tau)) + (-1.0)) / m.a;
    const double tb = (std::exp(m.b * (-tau)) + (-1.0)) / m.b;
    const double tab = (std::exp(apb * (-tau)) + (-1.0)) / apb;
    const double cross = (((m.sigma * m.eta) * (2.0 * m.rho)) / (m.a * m.b)) * (((ta + tau) + tb) - tab);
    const double v = (host_hw2_var(m.a, m.sigma, tau) + host_hw2_var(m.b, m.eta, tau)) + cross;
    return std::exp(m.r0 * (-tau) + v * 0.5);
}
double host_par_rate2(const eqhip_model_hw2f& m, double start, const double* pay, const double* acc, int n) {
    const double num = host_zcb2_at_origin(m, start - start) - host_zcb2_at_origin(m, pay[n - 1] - start);
    double den = 0.0;
    for (int i = 0; i < n; ++i) {
        const double leg = host_zcb2_at_origin(m, pay[i] - start) * acc[i];
        den = (i == 0) ? leg : den + leg;
    }
    return num / den;
}

void validate_hw2f(const eqhip_model_hw2f& m) {
    if (!(m.a != 0.0) || !(m.b != 0.0) || !(m.a + m.b != 0.0))
        raise(EQHIP_E_INVALID, "HullWhite2F: a, b and a + b must be non-zero");
    if (!(std::fabs(m.rho) <= 1.0)) raise(EQHIP_E_INVALID, "HullWhite2F: rho must lie in [-1, 1]");
}

bool is_reset_date(double start, const double* pay, int m, double t) {
    if (std::fabs(t - start) < kDateTolHost) return true;
    for (int i = 0; i + 1 < m; ++i)
        if (std::fabs(pay[i] - t) < kDateTolHost) return true;
    return false;
}

}  // namespace

struct eqhip_plan {
    eqhip_context* ctx = nullptr;
    int device = 0;  // kept so that destroying the plan never touches its context
    eqhip_model_hw1f model{};
    int factors = 1;           // 2: G2++ plan (model2; state (x, y) in R, R2)
    eqhip_model_hw2f model2{};
    eqhip_mc mc{};
    int n_dates = 0, n_trades = 0, S = 0, C = 0, NF = 0, D = 0, Fmax = 0, n_slots = 0;
    bool det = false;
    double dt = 1.0;
    int tiles_per_block = 1, n_chunks = 1;
    uint64_t n_tiles = 0;
    int want_trade = 0;
    bool keep_cube = false;    // the signed netting-set values stay resident in `cube` after execute
    bool executed = false;
    int n_items = 0;
    int set_mode = 0;
    int ppl = kDefaultPpl, tpi = kDefaultTpi;  // valuation variant (paths per lane, trades per item)
    // path mode (value_path_kernel): a plan without per-trade outputs whose families and netting sets
    // fit the kernel's registers is valued one thread per path by default (path_ok: it can be)
    bool path_ok = false, path_mode = false;
    int path_ns = 0;  // the netting-set registers of the launch that serves it (a batch's largest; 0: its own)
    DevBuf pe_off, pe, pc_set, pc_ab;
    bool autotune = false;     // the first execute times the valid variants and keeps the fastest
    bool tuned = false;
    std::vector<std::string> tune_log;
    int n_items_t[3] = {0, 0, 0};             // per-trade work items of TPI 2, 4, 8
    int max_chunks = 0;                       // per-trade partial sums are sized for this many chunks
    // A portfolio too diverse for one block's LDS (or with more work items than a block has slots) is
    // valued in parts: groups of its schedule classes, each a value plan of its own over the same
    // paths (every grid step simulated into this plan's R), launched one after another into this
    // plan's netting-set rows and per-trade sums (x_mode: the first part stores its signed set
    // values, the next ones add theirs, the last one adds, takes the positive part and writes the
    // signed values out). A part points back to the plan it belongs to.
    std::vector<std::unique_ptr<eqhip_plan>> parts;
    eqhip_plan* whole = nullptr;
    int x_mode = 0;
    uint64_t target_units = 4096;   // work units the launch aims for (chunking of the path tiles)
    const double* r_shared = nullptr;  // batch-owned short-rate buffer (all grid steps), or null
    uint64_t r_stride = 0;
    std::vector<double> dates;
    size_t lds_bytes = 0;

    DevBuf lflow_off, lflow, ltau, lseg, zcls_off, zcls;
    DevBuf slot_of_step, R, Z, AB, date_slot, date_dfirst, date_det, fdesc, fam_off, flow_d, flow_tau,
        flow_cls, fix_off, fix_slot, fix_AB, set_off, set_nlive, set_cls, set_alpha, set_beta, X, cls_mat, tr_K,
        tr_absN, tr_N, tr_NK, tr_orig, partial, trade_ee, trade_cva, set_ee, set_pfe, set_cva, neff, qrank,
        items_t[3], item_first_t[3],
        d_dates, vplan, units, rows, R2, AB2, fix_B2, cube, xva_pos;
    DevBuf tables;   // storage of the host-built tables above (they are views into it)
    std::vector<uint64_t> h_neff, h_qrank;
    bool direct_out = false;  // one-shot plans: outputs go straight to the caller's arrays, and the large
                              // work buffers are views into the context's pool (under its one-shot lock)
    bool flat = false;        // homogeneous batch valued by value_flat_kernel (build_flat)
    int flat_m = 0;           // its schedule's flow count
    DevBuf flat_tables, flat_first, flat_AB, flat_acc, flat_N, flat_NK;
    ValuePlan vp{};  // host copy of the device descriptor (per-run pointers patched at launch)
    PinnedBuf h_out; // page-locked staging of the outputs
    DevBuf cva_w;    // per-date CVA factors
    // host copies of the plan tables (used by the adjoint)
    std::vector<double> h_T;
    std::vector<int4> h_fdesc;
    std::vector<int> h_fam_off, h_flow_d, h_flow_cls, h_set_off, h_set_cls, h_cls_fam, h_cls_len, h_date_dfirst,
        h_fix_off, h_fix_slot, h_date_grid, h_slot_grid;
    std::vector<double> h_flow_tau, h_set_alpha, h_set_beta, h_fix_tau;
    std::vector<int64_t> h_fix_grid;  // grid step of each fixing entry
    std::vector<int4> unit_list;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    eqhip_timing timing{};
    // per-trade PFE95 (EQHIP_PLAN_TRADE_PFE): the valuation of a group of dates writes the legs of the
    // classes alive at those dates into ccube ([date of the group][class][path], the classes of the
    // parts one after another), then one select row per (trade, date) with the trade alive
    bool want_pfe = false;
    std::vector<int> h_trade_cls;        // class of each trade of this plan's portfolio
    std::vector<double> h_cls_mat;       // last payment of each class (classes are numbered by it)
    std::vector<int> h_whole_index;      // a part: its trades' indices in the whole plan
    double2* ccube_p = nullptr;          // the whole plan's class-leg buffer (a part points into it)
    int cube_C = 0, cube_cls0 = 0;       // classes of the buffer; this value plan's first class in it
    DevBuf ccube, cube_row, cls_live0, trade_rows, trade_pfe;
    std::vector<int> grp_k;              // first date of each date group, then n_dates
    std::vector<size_t> grp_row;         // first select row of each group, then the row count
    std::vector<hipEvent_t> pfe_ev;      // before and after each group's select launch

    ~eqhip_plan() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : pfe_ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// A scenario set: the short rate at every grid step of t_max / n_steps on every path, read by the
// plans of many batch calls (their r_shared) in place of a simulation of their own.
struct eqhip_scenarios {
    int device = 0;
    eqhip_model_hw1f model{};
    eqhip_mc mc{};
    double t_max = 0.0;
    DevBuf slot_of_step;  // identity: step g is row g
    DevBuf R;             // [step][path]
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

int tpi_index(int tpi) { return tpi == 2 ? 0 : (tpi == 4 ? 1 : 2); }

// Whether the plan can run variant (ppl, tpi): its LDS fits, its per-trade work items fit the block's
// work slots, and two-factor plans take one path per lane.
bool variant_fits(const eqhip_plan& P, int ppl, int tpi) {
    if (P.factors == 2 && ppl != 1) return false;
    if (ppl == 4 && tpi != 2) return false;  // the kernel's register budget
    const bool split = P.S <= kMaxSplitSets;
    size_t lds = value_lds_bytes(P.D, P.Fmax, P.C, split ? P.S : 0, ppl);
    if (split && lds > P.ctx->lds_limit) lds = value_lds_bytes(P.D, P.Fmax, P.C, 0, ppl);
    if (P.want_trade) {
        if (P.n_items_t[tpi_index(tpi)] > kTradeSlots) return false;
        lds = std::max(lds, value_trade_scratch_bytes(tpi));
    }
    return lds <= P.ctx->lds_limit;
}

void tile_launch(eqhip_plan& P);

void choose_launch(eqhip_plan& P) {
    if (P.path_mode) {  // no LDS
        P.set_mode = 0;
        P.lds_bytes = 0;
        tile_launch(P);
        return;
    }
    const int ppl = P.ppl;
    // split mode keeps one partial sum per (set, wave) in LDS; when those do not fit, one wave per set
    P.set_mode =
        (P.S <= kMaxSplitSets && value_lds_bytes(P.D, P.Fmax, P.C, P.S, ppl) <= P.ctx->lds_limit) ? 0 : 1;
    size_t lds = value_lds_bytes(P.D, P.Fmax, P.C, P.set_mode == 0 ? P.S : 0, ppl);
    if (lds > P.ctx->lds_limit)
        raise(EQHIP_E_CAPACITY, "portfolio too diverse for one plan: " + std::to_string(P.D) +
                                    " cash-flow dates and " + std::to_string(P.C) + " schedule classes need " +
                                    std::to_string(lds) + " B of LDS (limit " + std::to_string(P.ctx->lds_limit) +
                                    "); split it into several plans");
    // per-trade sums of a unit's work slots are reduced through LDS at the end of the unit
    if (P.want_trade) lds = std::max(lds, value_trade_scratch_bytes(P.tpi));
    P.lds_bytes = lds;
    tile_launch(P);
}

// The path tiling of the plan's variant: tiles of 64 PPL paths, grouped into chunks, one work unit
// per (date, chunk).
void tile_launch(eqhip_plan& P) {
    const uint64_t tw = P.path_mode ? (uint64_t)kPathThreads * path_ppt(P.path_ns ? P.path_ns : P.S)
                                    : (uint64_t)kPathsPerTile * P.ppl;
    P.n_tiles = (P.mc.n_paths + tw - 1) / tw;
    // enough units for ~16 blocks per CU (one 1024-thread block per CU at this LDS footprint). Units run
    // date-major and blocks go to the 8 XCDs round-robin, so with a multiple of 8 chunks every unit of
    // one path chunk lands on the same XCD, whose L2 then serves the rows later dates read again (the
    // fixings of periods in progress).
    // path mode: 256-thread blocks with no LDS, several per CU: four times the units
    const uint64_t target_units = P.path_mode ? 4 * P.target_units : P.target_units;
    uint64_t target_chunks = std::max<uint64_t>(1, (target_units + P.n_dates - 1) / P.n_dates);
    target_chunks = (target_chunks + 7) / 8 * 8;
    uint64_t tpb = std::max<uint64_t>(1, (P.n_tiles + target_chunks - 1) / target_chunks);
    P.tiles_per_block = (int)tpb;
    P.n_chunks = (int)((P.n_tiles + tpb - 1) / tpb);
}

void build_value_plan(eqhip_plan& P);

// A plan's large work buffer: its own allocation, or for a one-shot plan a view into the context's pool
// buffer of that role (grown to the largest request and kept).
void work_buffer(eqhip_plan& P, DevBuf& b, DevBuf& pool, size_t bytes) {
    if (!P.direct_out) {
        b.alloc(bytes);
        return;
    }
    reserve(pool, bytes);
    b.view(pool.p, bytes);
}

void build_parts(eqhip_plan& P, const eqhip_model_hw1f* model, const eqhip_portfolio* pf, const double* dates,
                 int n_dates, const eqhip_mc* mc, const std::vector<std::vector<double>>& cls_dates,
                 const std::vector<int>& trade_cls);
void build_select_rows(eqhip_plan& P);
void build_trade_pfe(eqhip_plan& P, const eqhip_portfolio* pf);

// all_steps: r is stored at every grid step (slot = step), so plans of a batch can share one
// simulation; the batch then owns the short-rate buffer. A part of a partitioned plan (P.whole set)
// reads the whole plan's rates and writes its rows; whole_index maps the part's trades to the whole's.
void build_plan(eqhip_plan& P, const eqhip_model_hw1f* model, const eqhip_portfolio* pf,
                const double* dates, int n_dates, const eqhip_mc* mc, int want_trade, bool all_steps = false,
                const std::vector<int>* whole_index = nullptr) {
    const bool is_part = P.whole != nullptr;
    // a two-factor plan arrives with factors = 2 and model2 set, and no one-factor model
    if (!mc) raise(EQHIP_E_INVALID, "mc is null");
    if (P.factors == 2) {
        validate_hw2f(P.model2);
    } else {
        if (!model) raise(EQHIP_E_INVALID, "model is null");
        if (!(model->a != 0.0)) raise(EQHIP_E_INVALID, "HullWhite1F: a must be non-zero");
    }
    validate_portfolio(pf);
    if (n_dates < 1 || !dates) raise(EQHIP_E_INVALID, "monitoring_times vacío: at least one date is required");
    if (mc->n_paths < 1) raise(EQHIP_E_INVALID, "PricingContext: n_paths must be > 0");
    if (mc->n_steps < 1) raise(EQHIP_E_INVALID, "PricingContext: n_steps must be > 0");
    if (mc->rng != EQHIP_RNG_HOST_DRAWS && mc->rng != EQHIP_RNG_PHILOX) raise(EQHIP_E_INVALID, "unknown rng");
    if (mc->disc != EQHIP_DISC_EULER && mc->disc != EQHIP_DISC_EXACT) raise(EQHIP_E_INVALID, "unknown discretisation");
    if (mc->n_paths > (uint64_t)1 << 40) raise(EQHIP_E_CAPACITY, "n_paths too large");

    if (P.factors != 2) P.model = *model;
    P.mc = *mc;
    P.n_dates = n_dates;
    P.dates.assign(dates, dates + n_dates);
    P.n_trades = pf->n_trades;
    P.S = pf->n_sets;
    P.want_trade = (want_trade & EQHIP_PLAN_TRADE_EE) ? 1 : 0;
    P.keep_cube = (want_trade & EQHIP_PLAN_KEEP_CUBE) != 0;
    P.autotune = (want_trade & EQHIP_PLAN_AUTOTUNE) != 0;
    P.want_pfe = (want_trade & EQHIP_PLAN_TRADE_PFE) != 0;
    if (P.want_pfe && P.factors == 2) raise(EQHIP_E_INVALID, "per-trade PFE serves HullWhite1F plans only");
    if (whole_index) P.h_whole_index = *whole_index;
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
    // classes numbered by maturity (then first appearance): the maturity-sorted trade items then
    // run through consecutive class ids, and their sC columns through consecutive LDS banks
    {
        auto last_of = [&](int c) { return classes[c].pay.empty() ? classes[c].start : classes[c].pay.back(); };
        std::vector<int> by_mat(P.C);
        for (int c = 0; c < P.C; ++c) by_mat[c] = c;
        std::stable_sort(by_mat.begin(), by_mat.end(), [&](int x, int y) { return last_of(x) < last_of(y); });
        std::vector<int> new_id(P.C);
        std::vector<ClassKey> sorted;
        sorted.reserve(P.C);
        for (int c = 0; c < P.C; ++c) {
            new_id[by_mat[c]] = c;
            sorted.push_back(classes[by_mat[c]]);
        }
        classes.swap(sorted);
        for (int& c : trade_cls) c = new_id[c];
    }

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
    P.h_trade_cls = trade_cls;
    P.h_cls_mat = cls_mat;

    // slots: grid indices at which r is stored
    std::map<int64_t, int> slot_of;
    if (all_steps)
        for (int64_t g = 0; g <= (int64_t)steps; ++g) slot_of.emplace(g, (int)g);
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
        // grid index 0 has a single row under HW1F (exposure.rs:118-126: path[0] is r0 alone) but
        // n_paths equal rows under G2++ (exposure.rs:346-352: path[0] is n_paths zero states),
        // except when every date is <= 0 (one deterministic value, exposure.rs:325-337)
        date_det[k] = (g == 0 && (P.factors == 1 || P.det)) ? 1 : 0;
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
    // per-date live-flow lists (families in order), their wave segments, and the matured classes
    std::vector<int> lflow_off(n_dates + 1, 0), zcls_off(n_dates + 1, 0), zcls;
    std::vector<int4> lseg((size_t)n_dates * kValueWaves);
    std::vector<int4> lflow;
    std::vector<double> ltau;
    for (int k = 0; k < n_dates; ++k) {
        const int l0 = (int)lflow.size();
        for (int F = 0; F < P.NF; ++F) {
            const int4 fd = fdesc[(size_t)k * P.NF + F];
            const int f0 = fam_off[F], f1 = fam_off[F + 1];
            const int first = (fd.x == 0) ? f1 : f0 + fd.y;
            for (int i = f0; i < first; ++i)
                if (flow_cls[i] >= 0) zcls.push_back(flow_cls[i]);
            for (int i = first; i < f1; ++i) {
                lflow.push_back(make_int4(flow_d[i], flow_cls[i], F, 0));
                ltau.push_back(flow_tau[i]);
            }
        }
        lflow_off[k + 1] = (int)lflow.size();
        zcls_off[k + 1] = (int)zcls.size();
        const int L = (int)lflow.size() - l0;
        // segment w = [w L / waves, (w+1) L / waves); it takes the totals of the earlier segments
        // w_lo .. w-1 that end in the family its first flow belongs to (a family spans consecutive
        // segments; an empty segment in between totals zero)
        int sb[kValueWaves + 1];
        for (int w = 0; w <= kValueWaves; ++w) sb[w] = (w * L) / kValueWaves;
        auto fam_of = [&](int i) { return lflow[(size_t)l0 + i].z; };
        for (int w = 0; w < kValueWaves; ++w) {
            int w_lo = w, n_carry = 0;
            if (sb[w + 1] > sb[w]) {
                const int fam0 = fam_of(sb[w]);
                for (int w2 = w - 1; w2 >= 0; --w2) {
                    if (sb[w2 + 1] == sb[w2]) continue;  // empty
                    if (fam_of(sb[w2 + 1] - 1) != fam0) break;
                    w_lo = w2;
                }
                if (w_lo < w)
                    for (int i = sb[w]; i < sb[w + 1] && fam_of(i) == fam0; ++i) ++n_carry;
            }
            lseg[(size_t)k * kValueWaves + w] = make_int4(sb[w], sb[w + 1], w_lo, n_carry);
        }
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
    // Each set lists its classes latest maturity first. A class is alive at t while its last
    // payment is after t (cls_mat > t + tol, the rule of the matured-class lists above), so the
    // classes alive at a date are a prefix of the set's list, of length set_nlive[date][set].
    std::vector<int> set_off(P.S + 1, 0), set_cls;
    std::vector<double> set_alpha, set_beta;
    for (int s = 0; s < P.S; ++s) {
        std::vector<std::pair<int, std::pair<double, double>>> ent(set_cls_sum[s].begin(), set_cls_sum[s].end());
        std::stable_sort(ent.begin(), ent.end(),
                         [&](const auto& x, const auto& y) { return cls_mat[x.first] > cls_mat[y.first]; });
        for (auto& kv : ent) {
            set_cls.push_back(kv.first);
            set_alpha.push_back(kv.second.first);
            set_beta.push_back(kv.second.second);
        }
        set_off[s + 1] = (int)set_cls.size();
    }
    std::vector<int> set_nlive((size_t)n_dates * P.S);
    for (int k = 0; k < n_dates; ++k)
        for (int s = 0; s < P.S; ++s) {
            int n = 0;
            while (set_off[s] + n < set_off[s + 1] && cls_mat[set_cls[set_off[s] + n]] > P.dates[k] + kDateTolHost) ++n;
            set_nlive[(size_t)k * P.S + s] = n;
        }

    // Path-mode tables. Per date, its live cash-flow dates in date order, each with every family's
    // accrual at it and the netting-set entries (by family, then by set) of the classes ending at it.
    std::vector<int> pe_off(n_dates + 1, 0), pc_set;
    std::vector<PathBond> pe;
    std::vector<double2> pc_ab;
    P.path_ok = !P.want_trade && !P.want_pfe && !is_part && P.NF <= kPathFamilies && P.S <= kPathSets;
    if (P.path_ok) {
        std::vector<std::vector<int>> ents(P.C);  // set-entry indices of each class, by set
        for (int s = 0; s < P.S; ++s)
            for (int e = set_off[s]; e < set_off[s + 1]; ++e) ents[set_cls[e]].push_back(e);
        std::map<int, PathBond> at;  // cash-flow date -> its record
        std::map<int, std::array<int, kPathFamilies>> ends;  // cash-flow date -> class ending there, per family
        for (int k = 0; k < n_dates; ++k) {
            at.clear();
            ends.clear();
            for (int F = 0; F < P.NF; ++F) {
                const int4 fd = fdesc[(size_t)k * P.NF + F];
                if (fd.x == 0) continue;
                for (int i = fam_off[F] + fd.y; i < fam_off[F + 1]; ++i) {
                    const int d = flow_d[i];
                    auto it = at.find(d);
                    if (it == at.end()) {
                        PathBond b{};
                        b.d = d;
                        it = at.emplace(d, b).first;
                        std::array<int, kPathFamilies> none;
                        none.fill(-1);
                        ends.emplace(d, none);
                    }
                    it->second.tau[F] = flow_tau[i];
                    ends[d][F] = flow_cls[i];
                }
            }
            for (auto& kv : at) {
                PathBond b = kv.second;
                b.e0 = (int)pc_set.size();
                for (int F = 0; F < kPathFamilies; ++F) {
                    const int c = ends[kv.first][F];
                    if (c < 0) continue;
                    for (int e : ents[c]) {
                        pc_set.push_back((int)(std::upper_bound(set_off.begin(), set_off.end(), e) - set_off.begin()) - 1);
                        pc_ab.push_back(make_double2(set_alpha[e], set_beta[e]));
                    }
                    b.n[F] = (int)ents[c].size();
                }
                pe.push_back(b);
            }
            pe_off[k + 1] = (int)pe.size();
        }
    }

    // per-trade work items: trades sorted by (last date of the class, class, sign), up to TPI trades
    // per item, for each TPI the kernel offers (2, 4, 8). Ordering by maturity keeps the items still
    // alive at a date a contiguous suffix of the list, so the waves whose items have all matured skip
    // the date.
    std::vector<int4> items_t[3];
    std::vector<int> tr_orig, item_first_t[3];
    std::vector<double> tr_K, tr_absN, tr_N, tr_NK;
    if (P.want_trade) {
        auto sgn = [&](int j) { return pf->notional[j] < 0.0 ? -1 : 1; };
        auto last_t = [&](int j) {
            const ClassKey& K = classes[trade_cls[j]];
            return K.pay.empty() ? K.start : K.pay.back();
        };
        std::vector<int> order(pf->n_trades);
        for (int j = 0; j < pf->n_trades; ++j) order[j] = j;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return std::make_tuple(last_t(x), trade_cls[x], sgn(x)) <
                   std::make_tuple(last_t(y), trade_cls[y], sgn(y));
        });
        for (int j : order) {
            tr_orig.push_back(whole_index ? (*whole_index)[j] : j);
            tr_K.push_back(pf->fixed_rate[j]);
            tr_absN.push_back(std::fabs(pf->notional[j]));
            tr_N.push_back(pf->notional[j]);
            tr_NK.push_back(pf->notional[j] * pf->fixed_rate[j]);
        }
        const int tpis[3] = {2, 4, 8};
        for (int v = 0; v < 3; ++v) {
            std::vector<int4>& items = items_t[v];
            for (int i = 0; i < pf->n_trades;) {
                const int c = trade_cls[order[i]], sg = sgn(order[i]);
                int4 item = make_int4(c, i, 0, sg);
                while (i < pf->n_trades && trade_cls[order[i]] == c && sgn(order[i]) == sg && item.z < tpis[v]) {
                    ++i;
                    ++item.z;
                }
                items.push_back(item);
            }
            P.n_items_t[v] = (int)items.size();
            // first item alive at each date: the items alive at a date are a suffix of the list
            item_first_t[v].assign(n_dates, 0);
            for (int k = 0; k < n_dates; ++k) {
                int f = 0;
                while (f < (int)items.size() && !(cls_mat[items[f].x] > P.dates[k] + kDateTolHost)) ++f;
                item_first_t[v][k] = f;
            }
        }
    }
    // default variant: one path per lane, four trades per item (eight when four leave too many items)
    P.ppl = 1;
    P.tpi = (!P.want_trade || P.n_items_t[1] <= kTradeSlots) ? 4 : 8;
    P.n_items = P.want_trade ? P.n_items_t[tpi_index(P.tpi)] : 0;
    // path mode, whenever the plan allows it (it needs no LDS)
    P.path_mode = P.path_ok;
    // A one-factor plan that no variant fits is valued in parts (a batch job or a part must fit)
    const bool fits =
        P.path_mode || variant_fits(P, 1, 2) || variant_fits(P, 1, 4) || variant_fits(P, 1, 8);
    const bool partitioned = !fits && !is_part && !all_steps && P.factors == 1;
    if (!fits && !partitioned && P.want_trade && P.n_items_t[2] > kTradeSlots)
        raise(EQHIP_E_CAPACITY, "per-trade EE supports at most " + std::to_string(kTradeSlots) +
                                    " work items per plan (got " + std::to_string(P.n_items_t[2]) + ")");
    if (partitioned) {
        // the parts read r at their own dates and fixings: every grid step is simulated
        P.n_slots = (int)steps + 1;
        for (uint32_t g2 = 0; g2 <= steps; ++g2) slot_of_step[g2] = (int)g2;
    }

    if (partitioned)
        tile_launch(P);  // the parts' tiling (one path per lane), for the per-trade sums they share
    else
        choose_launch(P);
    P.max_chunks = P.n_chunks;  // one path per lane has the most chunks of any variant

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
    P.h_neff = neff;
    P.h_qrank = qrank;

    // memory check for the large buffers (a part has none: it writes the whole plan's)
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    double need = is_part ? 0.0 : 8.0 * ((all_steps ? 0.0 : (double)P.factors * P.n_slots * mc->n_paths) +
                         (double)(P.keep_cube ? 2 : 1) * P.S * n_dates * mc->n_paths +
                         (P.want_trade ? (double)P.n_trades * n_dates * P.n_chunks : 0.0) +
                         (mc->rng == EQHIP_RNG_HOST_DRAWS ? (double)P.factors * mc->n_steps * mc->n_paths : 0.0));
    if (need > 0.9 * (double)free_b)
        raise(EQHIP_E_CAPACITY, "exposure plan needs " + std::to_string(need / 1e9) + " GB of device memory, " +
                                    std::to_string(free_b / 1e9) + " GB free");

    // uploads: every host-built table in one allocation and one copy
    TablePack pack;
    pack.add(P.slot_of_step, slot_of_step);
    pack.add(P.date_slot, date_slot);
    pack.add(P.date_dfirst, date_dfirst);
    pack.add(P.date_det, date_det);
    pack.add(P.fdesc, fdesc);
    pack.add(P.fam_off, fam_off);
    pack.add(P.flow_d, flow_d);
    pack.add(P.flow_tau, flow_tau);
    pack.add(P.flow_cls, flow_cls);
    pack.add(P.lflow_off, lflow_off);
    pack.add(P.lflow, lflow);
    pack.add(P.ltau, ltau);
    pack.add(P.lseg, lseg);
    pack.add(P.zcls_off, zcls_off);
    pack.add(P.zcls, zcls);
    pack.add(P.cls_mat, cls_mat);
    pack.add(P.fix_off, fix_off);
    pack.add(P.fix_slot, fix_slot);
    pack.add(P.set_off, set_off);
    pack.add(P.set_nlive, set_nlive);
    pack.add(P.set_cls, set_cls);
    pack.add(P.set_alpha, set_alpha);
    pack.add(P.set_beta, set_beta);
    pack.add(P.neff, neff);
    pack.add(P.qrank, qrank);
    pack.add(P.pe_off, pe_off);
    pack.add(P.pe, pe);
    pack.add(P.pc_set, pc_set);
    pack.add(P.pc_ab, pc_ab);
    pack.add(P.d_dates, P.dates);
    if (P.want_trade) {
        for (int v = 0; v < 3; ++v) {
            pack.add(P.items_t[v], items_t[v]);
            pack.add(P.item_first_t[v], item_first_t[v]);
        }
        pack.add(P.tr_K, tr_K);
        pack.add(P.tr_absN, tr_absN);
        pack.add(P.tr_N, tr_N);
        pack.add(P.tr_NK, tr_NK);
        pack.add(P.tr_orig, tr_orig);
        if (!is_part) {
            work_buffer(P, P.partial, P.ctx->pool_partial, (size_t)P.n_trades * n_dates * P.max_chunks * 8);
            P.trade_ee.alloc((size_t)P.n_trades * n_dates * 8);
            P.trade_cva.alloc((size_t)P.n_trades * 8);
        }
    }
    DevBuf tau, ftau;  // views into the plan's tables
    pack.add(tau, ab_tau);
    pack.add(ftau, fix_tau);
    pack.flush(P.tables);
    // ZCB factor tables, computed on the device
    {
        P.AB.alloc(ab_tau.size() * sizeof(double2));
        int n = (int)ab_tau.size();
        const eqhip_model_hw2f& m2 = P.model2;
        if (P.factors == 2) {  // (c, B(a,tau)) and B(b,tau)
            P.AB2.alloc(ab_tau.size() * sizeof(double));
            hipLaunchKernelGGL(ab2_kernel, dim3((n + 255) / 256), dim3(256), 0, st, tau.as<double>(), n, m2.a, m2.b,
                               m2.sigma, m2.eta, m2.rho, m2.r0, P.AB.as<double2>(), P.AB2.as<double>());
        } else {
            hipLaunchKernelGGL(ab_kernel, dim3((n + 255) / 256), dim3(256), 0, st, tau.as<double>(), n,
                               P.model.a, P.model.b, P.model.sigma, P.AB.as<double2>());
        }
        HIPCHK(hipGetLastError());
        P.fix_AB.alloc(std::max<size_t>(1, fix_tau.size()) * sizeof(double2));
        if (P.factors == 2) P.fix_B2.alloc(std::max<size_t>(1, fix_tau.size()) * sizeof(double));
        int nf = (int)fix_tau.size();
        if (nf > 0) {
            if (P.factors == 2)
                hipLaunchKernelGGL(ab2_kernel, dim3((nf + 255) / 256), dim3(256), 0, st, ftau.as<double>(), nf, m2.a,
                                   m2.b, m2.sigma, m2.eta, m2.rho, m2.r0, P.fix_AB.as<double2>(),
                                   P.fix_B2.as<double>());
            else
                hipLaunchKernelGGL(ab_kernel, dim3((nf + 255) / 256), dim3(256), 0, st, ftau.as<double>(), nf,
                                   P.model.a, P.model.b, P.model.sigma, P.fix_AB.as<double2>());
            HIPCHK(hipGetLastError());
        }
    }
    if (!is_part) {
        if (!all_steps) work_buffer(P, P.R, P.ctx->pool_R, (size_t)P.n_slots * mc->n_paths * 8);
        if (P.factors == 2) work_buffer(P, P.R2, P.ctx->pool_R2, (size_t)P.n_slots * mc->n_paths * 8);
        work_buffer(P, P.X, P.ctx->pool_X, (size_t)P.S * n_dates * mc->n_paths * 8);
        if (P.keep_cube) P.cube.alloc((size_t)n_dates * P.S * mc->n_paths * 8);
        P.set_ee.alloc((size_t)P.S * n_dates * 8);
        P.set_pfe.alloc((size_t)P.S * n_dates * 8);
        P.set_cva.alloc((size_t)P.S * 8);
        if (mc->rng == EQHIP_RNG_HOST_DRAWS)
            work_buffer(P, P.Z, P.ctx->pool_Z, (size_t)P.factors * mc->n_steps * mc->n_paths * 8);
    }
    P.h_T = T;
    P.h_fdesc = fdesc;
    P.h_fam_off = fam_off;
    P.h_flow_d = flow_d;
    P.h_flow_cls = flow_cls;
    P.h_flow_tau = flow_tau;
    P.h_set_off = set_off;
    P.h_set_cls = set_cls;
    P.h_set_alpha = set_alpha;
    P.h_set_beta = set_beta;
    P.h_cls_fam = cls_fam;
    P.h_cls_len = cls_len;
    P.h_date_dfirst = date_dfirst;
    P.h_fix_off = fix_off;
    P.h_fix_slot = fix_slot;
    P.h_fix_tau = fix_tau;
    P.h_date_grid.resize(n_dates);
    for (int k = 0; k < n_dates; ++k) P.h_date_grid[k] = (int)grid_of(P.dates[k]);
    P.h_slot_grid.assign(P.n_slots, 0);
    for (auto& kv : slot_of) P.h_slot_grid[kv.second] = (int)kv.first;
    if (partitioned) {
        std::vector<std::vector<double>> cls_dates(P.C);
        for (int c = 0; c < P.C; ++c) {
            cls_dates[c].push_back(classes[c].start);
            cls_dates[c].insert(cls_dates[c].end(), classes[c].pay.begin(), classes[c].pay.end());
        }
        build_parts(P, model, pf, dates, n_dates, mc, cls_dates, trade_cls);
        build_select_rows(P);
    } else {
        build_value_plan(P);
    }
    if (P.want_pfe) build_trade_pfe(P, pf);
    if (!is_part)
        for (auto& e : P.ev) HIPCHK(hipEventCreate(&e));
}

// The parts of a partitioned plan: the schedule classes in maturity order (a class with many trades
// in chunks of its trades) gathered greedily into groups whose cash-flow dates, classes, fixings and
// per-trade work items fit one block, each group a part over the same paths and dates. The parts run
// in order: first, middle and last differ only in how they store the netting-set values (x_mode).
void build_parts(eqhip_plan& P, const eqhip_model_hw1f* model, const eqhip_portfolio* pf, const double* dates,
                 int n_dates, const eqhip_mc* mc, const std::vector<std::vector<double>>& cls_dates,
                 const std::vector<int>& trade_cls) {
    const size_t budget = P.ctx->lds_limit;
    const int C = (int)cls_dates.size();
    std::vector<std::vector<int>> cls_trades(C);
    for (int j = 0; j < pf->n_trades; ++j) cls_trades[trade_cls[j]].push_back(j);
    constexpr int kChunkTrades = 2048;  // at most 512 + 2 work items of four trades per class chunk
    // LDS of a group with D dates and C classes, fixings bounded by the classes (split-free set sums)
    auto group_lds = [&](size_t D, size_t nc) {
        size_t lds = value_lds_bytes((int)D, (int)nc, (int)nc, 0, 1);
        if (P.want_trade) lds = std::max(lds, value_trade_scratch_bytes(4));
        return lds;
    };
    std::vector<std::vector<int>> groups;  // trades of each group
    std::set<double> g_dates;
    size_t g_classes = 0, g_items = 0;
    for (int c = 0; c < C; ++c) {
        const auto& tr = cls_trades[c];
        for (size_t t0 = 0; t0 < tr.size(); t0 += kChunkTrades) {
            const size_t t1 = std::min(tr.size(), t0 + (size_t)kChunkTrades);
            size_t pos = 0;
            for (size_t t = t0; t < t1; ++t) pos += pf->notional[tr[t]] >= 0.0 ? 1 : 0;
            const size_t items = (pos + 3) / 4 + ((t1 - t0 - pos) + 3) / 4;
            std::set<double> nd = g_dates;
            nd.insert(cls_dates[c].begin(), cls_dates[c].end());
            const bool fits = group_lds(nd.size(), g_classes + 1) <= budget &&
                              (!P.want_trade || g_items + items <= (size_t)kTradeSlots);
            if (groups.empty() || !fits) {  // a new group, starting with this chunk
                groups.emplace_back();
                g_dates.clear();
                g_classes = g_items = 0;
                nd = std::set<double>(cls_dates[c].begin(), cls_dates[c].end());
            }
            if (group_lds(nd.size(), 1) > budget)
                raise(EQHIP_E_CAPACITY, "a schedule with " + std::to_string(cls_dates[c].size()) +
                                            " cash-flow dates needs more LDS than a block has (" +
                                            std::to_string(budget) + " B)");
            groups.back().insert(groups.back().end(), tr.begin() + (long)t0, tr.begin() + (long)t1);
            g_dates = std::move(nd);
            ++g_classes;
            g_items += items;
        }
    }
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        std::vector<int>& idx = groups[gi];
        std::sort(idx.begin(), idx.end());  // the whole portfolio's order within a part
        std::vector<int32_t> set_id, off{0};
        std::vector<double> notional, fixed, start, pay, acc;
        for (int j : idx) {
            set_id.push_back(pf->set_id[j]);
            notional.push_back(pf->notional[j]);
            fixed.push_back(pf->fixed_rate[j]);
            start.push_back(pf->start[j]);
            for (int f = pf->flow_offset[j]; f < pf->flow_offset[j + 1]; ++f) {
                pay.push_back(pf->payment_times[f]);
                acc.push_back(pf->accruals[f]);
            }
            off.push_back((int32_t)pay.size());
        }
        const eqhip_portfolio sub{(int32_t)idx.size(), P.S, set_id.data(), notional.data(), fixed.data(),
                                  start.data(), off.data(), pay.data(), acc.data()};
        auto part = std::make_unique<eqhip_plan>();
        part->ctx = P.ctx;
        part->device = P.device;
        part->whole = &P;
        part->target_units = P.target_units;
        part->x_mode = (gi == 0) ? 1 : (gi + 1 == groups.size() ? 3 : 2);
        part->r_shared = P.R.as<double>();
        part->r_stride = mc->n_paths;
        build_plan(*part, model, &sub, dates, n_dates, mc, P.want_trade ? EQHIP_PLAN_TRADE_EE : 0, true, &idx);
        if (part->n_chunks != P.n_chunks)
            raise(EQHIP_E_HIP, "internal: the parts of a plan must share its path tiling");
        P.parts.push_back(std::move(part));
    }
    P.lds_bytes = 0;
    for (auto& q : P.parts) P.lds_bytes = std::max(P.lds_bytes, q->lds_bytes);
}

void build_select_rows(eqhip_plan& P);

// Device descriptor of a plan (pointers into the plan's buffers) and its work units.
void build_value_plan(eqhip_plan& P) {
    ValuePlan& g = P.vp;
    g = ValuePlan{};
    g.R = P.r_shared ? P.r_shared : P.R.as<double>();
    g.r_stride = P.r_shared ? P.r_stride : P.mc.n_paths;
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
    g.variant = P.path_mode ? kPathVariant : kVariant(P.ppl, P.tpi);
    g.pe_off = P.pe_off.as<int>();
    g.pe = P.pe.as<PathBond>();
    g.pc_set = P.pc_set.as<int>();
    g.pc_ab = P.pc_ab.as<double2>();
    g.two_factor = (P.factors == 2) ? 1 : 0;
    g.R2 = P.R2.as<double>();
    g.AB2 = P.AB2.as<double>();
    g.fix_B2 = P.fix_B2.as<double>();
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
    g.lflow_off = P.lflow_off.as<int>();
    g.lflow = P.lflow.as<int4>();
    g.ltau = P.ltau.as<double>();
    g.lseg = P.lseg.as<int4>();
    g.zcls_off = P.zcls_off.as<int>();
    g.zcls = P.zcls.as<int>();
    g.fix_off = P.fix_off.as<int>();
    g.fix_slot = P.fix_slot.as<int>();
    g.fix_AB = P.fix_AB.as<double2>();
    g.set_off = P.set_off.as<int>();
    g.set_nlive = P.set_nlive.as<int>();
    g.set_cls = P.set_cls.as<int>();
    g.set_alpha = P.set_alpha.as<double>();
    g.set_beta = P.set_beta.as<double>();
    // a part writes the whole plan's netting-set rows and per-trade sums, indexed by its trades
    const eqhip_plan& W = P.whole ? *P.whole : P;
    g.X = W.X.as<double>();
    g.x_mode = P.x_mode;
    g.want_trade = P.want_trade;
    g.n_trades = W.n_trades;
    P.n_items = P.want_trade ? P.n_items_t[tpi_index(P.tpi)] : 0;
    g.n_items = P.n_items;
    g.item_first = P.item_first_t[tpi_index(P.tpi)].as<int>();
    g.items = P.items_t[tpi_index(P.tpi)].as<int4>();
    g.cls_mat = P.cls_mat.as<double>();
    g.tr_K = P.tr_K.as<double>();
    g.tr_absN = P.tr_absN.as<double>();
    g.tr_N = P.tr_N.as<double>();
    g.tr_NK = P.tr_NK.as<double>();
    g.tr_orig = P.tr_orig.as<int>();
    g.partial = W.partial.as<double>();
    g.ccube = P.ccube_p;
    g.cube_row = P.cube_row.as<int>();
    g.cls_live0 = P.cls_live0.as<int>();
    g.cube_C = P.cube_C;
    g.cube_cls0 = P.cube_cls0;
    P.vplan.alloc(sizeof(ValuePlan));
    HIPCHK(hipMemcpy(P.vplan.p, &g, sizeof(ValuePlan), hipMemcpyHostToDevice));
    P.unit_list.clear();
    for (int k = 0; k < P.n_dates; ++k)
        for (int c = 0; c < P.n_chunks; ++c) P.unit_list.push_back(make_int4(0, k, c, 0));
    upload(P.units, P.unit_list, P.ctx->stream);
    if (!P.whole) build_select_rows(P);
}

// The EE / PFE95 select rows of a plan's netting sets ([set][date] rows of X).
void build_select_rows(eqhip_plan& P) {
    const std::vector<uint64_t>& neff = P.h_neff;
    const std::vector<uint64_t>& qrank = P.h_qrank;
    std::vector<SelectRow> rows;
    for (int s = 0; s < P.S; ++s)
        for (int k = 0; k < P.n_dates; ++k) {
            const size_t r = (size_t)s * P.n_dates + k;
            rows.push_back(SelectRow{P.X.as<double>() + r * P.mc.n_paths, neff[k], qrank[k],
                                     P.set_ee.as<double>() + r, P.set_pfe.as<double>() + r});
        }
    upload(P.rows, rows, P.ctx->stream);
}

// Per-trade PFE95 rows of a plan created with EQHIP_PLAN_TRADE_PFE. The trades of one schedule class
// differ only in notional and fixed rate, so a trade's value on a path is N F - N K A with (F, A) the
// legs of its class: the valuation writes the legs of the classes alive at each date of a group of
// dates (16 B per class, path and date instead of 8 B per trade, path and date), and a select row per
// live (trade, date) forms the trade's positive parts from them in each of its passes. A group holds as
// many dates as a quarter of the free device memory allows (or EQHIP_TRADE_PFE_BYTES, when smaller).
void build_trade_pfe(eqhip_plan& P, const eqhip_portfolio* pf) {
    std::vector<eqhip_plan*> vps;  // the value plans: the parts, or the plan itself
    if (P.parts.empty()) vps.push_back(&P);
    for (auto& q : P.parts) vps.push_back(q.get());
    std::vector<int> cube_cls(P.n_trades, 0);
    std::vector<double> mat(P.n_trades, 0.0);
    int C_total = 0;
    for (eqhip_plan* Q : vps) {
        for (size_t i = 0; i < Q->h_trade_cls.size(); ++i) {
            const int j = Q->whole ? Q->h_whole_index[i] : (int)i;
            cube_cls[j] = C_total + Q->h_trade_cls[i];
            mat[j] = Q->h_cls_mat[Q->h_trade_cls[i]];
        }
        Q->cube_cls0 = C_total;
        C_total += Q->C;
    }
    const uint64_t np = P.mc.n_paths;
    const int nd = P.n_dates;
    const double per_date = 16.0 * (double)C_total * (double)np;
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    double budget = 0.25 * (double)free_b;
    if (const char* cap = std::getenv("EQHIP_TRADE_PFE_BYTES")) {  // a smaller buffer: more groups
        const double c = std::atof(cap);
        if (c > 0.0) budget = std::min(budget, c);
    }
    const double fit = std::floor(budget / per_date);
    if (fit < 1.0)
        raise(EQHIP_E_CAPACITY, "per-trade PFE needs " + std::to_string(per_date / 1e9) +
                                    " GB of device memory per date for the legs of " + std::to_string(C_total) +
                                    " schedule classes, " + std::to_string(free_b / 1e9) + " GB free");
    const int G = (int)std::min<double>((double)nd, fit);
    P.ccube.alloc((size_t)G * C_total * np * 16);
    P.grp_k.clear();
    for (int k0 = 0; k0 < nd; k0 += G) P.grp_k.push_back(k0);
    P.grp_k.push_back(nd);
    std::vector<int> cube_row(nd);
    for (int k = 0; k < nd; ++k) cube_row[k] = k % G;
    for (eqhip_plan* Q : vps) {
        std::vector<int> live0(nd);
        for (int k = 0; k < nd; ++k) {
            int c = 0;
            while (c < Q->C && !(Q->h_cls_mat[c] > P.dates[k] + kDateTolHost)) ++c;
            live0[k] = c;
        }
        upload(Q->cls_live0, live0, P.ctx->stream);
        upload(Q->cube_row, cube_row, P.ctx->stream);
        Q->ccube_p = P.ccube.as<double2>();
        Q->cube_C = C_total;
        build_value_plan(*Q);  // the descriptor with the cube fields
    }
    P.trade_pfe.alloc((size_t)P.n_trades * nd * 8);
    HIPCHK(hipMemsetAsync(P.trade_pfe.p, 0, (size_t)P.n_trades * nd * 8, P.ctx->stream));  // matured trades
    // Blocks reach the 8 XCDs round-robin in index order, and each XCD has its own L2: the rows of one
    // (date, class) go to one XCD's positions (x, x + 8, x + 16, ...), so that its selects, running there
    // at about the same time, read the class's legs from that L2. Each (date, class) goes to the XCD with
    // the fewest rows so far.
    constexpr int kXcds = 8;
    std::vector<TradeSelectRow> rows;
    std::vector<int> live;
    P.grp_row.assign(1, 0);
    for (size_t gi = 0; gi + 1 < P.grp_k.size(); ++gi) {
        std::vector<std::vector<TradeSelectRow>> xq(kXcds);
        for (int k = P.grp_k[gi]; k < P.grp_k[gi + 1]; ++k) {
            live.clear();
            for (int j = 0; j < P.n_trades; ++j)
                if (mat[j] > P.dates[k] + kDateTolHost) live.push_back(j);
            std::stable_sort(live.begin(), live.end(), [&](int x, int y) { return cube_cls[x] < cube_cls[y]; });
            for (size_t a = 0; a < live.size();) {
                size_t b = a;
                while (b < live.size() && cube_cls[live[b]] == cube_cls[live[a]]) ++b;
                int x = 0;
                for (int y = 1; y < kXcds; ++y)
                    if (xq[y].size() < xq[x].size()) x = y;
                for (size_t i = a; i < b; ++i) {
                    const int j = live[i];
                    const double s = pf->notional[j] < 0.0 ? -1.0 : 1.0;
                    xq[x].push_back(TradeSelectRow{
                        P.ccube.as<double2>() + ((uint64_t)cube_row[k] * C_total + cube_cls[j]) * np, P.h_neff[k],
                        P.h_qrank[k], s, s * pf->fixed_rate[j], std::fabs(pf->notional[j]),
                        P.trade_pfe.as<double>() + (size_t)j * nd + k});
                }
                a = b;
            }
        }
        size_t longest = 0;
        for (auto& q : xq) longest = std::max(longest, q.size());
        for (size_t i = 0; i < longest; ++i)
            for (int x = 0; x < kXcds; ++x)
                if (i < xq[x].size()) rows.push_back(xq[x][i]);
        P.grp_row.push_back(rows.size());
    }
    upload(P.trade_rows, rows, P.ctx->stream);
    P.pfe_ev.assign(2 * (P.grp_k.size() - 1), nullptr);
    for (auto& e : P.pfe_ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipStreamSynchronize(P.ctx->stream));
}

// The valuation of dates [k0, k1) (every date when k1 < 0); cube: write the class legs for the
// per-trade PFE rows.
// The path-mode kernel whose family and netting-set registers cover nf families and ns sets.
using ValueKernel = void (*)(const ValuePlan*, const int4*);
ValueKernel path_kernel(int factors, int nf, int ns) {
    static const ValueKernel table[2][2][4] = {
        {{value_path_kernel<1, 2, 1>, value_path_kernel<1, 2, 2>, value_path_kernel<1, 2, 4>, value_path_kernel<1, 2, 8>},
         {value_path_kernel<1, 4, 1>, value_path_kernel<1, 4, 2>, value_path_kernel<1, 4, 4>, value_path_kernel<1, 4, 8>}},
        {{value_path_kernel<2, 2, 1>, value_path_kernel<2, 2, 2>, value_path_kernel<2, 2, 4>, value_path_kernel<2, 2, 8>},
         {value_path_kernel<2, 4, 1>, value_path_kernel<2, 4, 2>, value_path_kernel<2, 4, 4>, value_path_kernel<2, 4, 8>}}};
    if (nf > kPathFamilies || ns > kPathSets) raise(EQHIP_E_HIP, "internal: plan outside path mode");
    const int nsr = path_set_regs(ns);
    const int si = nsr == 1 ? 0 : (nsr == 2 ? 1 : (nsr == 4 ? 2 : 3));
    return table[factors == 2 ? 1 : 0][nf <= 2 ? 0 : 1][si];
}

void launch_value(eqhip_plan& P, double* path_mtm, double* set_mtm, int k0 = 0, int k1 = -1, bool cube = false) {
    hipStream_t st = P.ctx->stream;
    if (!P.parts.empty()) {  // the parts in order; the last one writes the signed set values
        for (size_t i = 0; i < P.parts.size(); ++i)
            launch_value(*P.parts[i], path_mtm, i + 1 == P.parts.size() ? set_mtm : nullptr, k0, k1, cube);
        return;
    }
    if (path_mtm != P.vp.path_mtm || set_mtm != P.vp.set_mtm) {  // per-run outputs live in the descriptor
        P.vp.path_mtm = path_mtm;
        P.vp.set_mtm = set_mtm;
        HIPCHK(hipMemcpy(P.vplan.p, &P.vp, sizeof(ValuePlan), hipMemcpyHostToDevice));
    }
    if (k1 < 0) k1 = P.n_dates;
    const size_t u0 = (size_t)k0 * P.n_chunks, nu = (size_t)(k1 - k0) * P.n_chunks;
    if (nu == 0) return;
    if (cube && (P.ppl != 1 || P.factors != 1 || !P.ccube_p || P.path_mode))
        raise(EQHIP_E_HIP, "internal: class legs are written by one-factor plans at one path per lane");
    if (P.path_mode) {
        if (path_mtm) raise(EQHIP_E_HIP, "internal: path mode writes no per-trade values");
        hipLaunchKernelGGL(path_kernel(P.factors, P.NF, P.S), dim3((unsigned)nu), dim3(kPathThreads), 0, st,
                           P.vplan.as<const ValuePlan>(), P.units.as<const int4>() + u0);
        HIPCHK(hipGetLastError());
        return;
    }
    // one path per lane: the kernel of the plan's trades-per-item variant
    auto p1 = (P.tpi == 2) ? value_p1_kernel<2> : (P.tpi == 8 ? value_p1_kernel<8> : value_p1_kernel<4>);
    auto cb = (P.tpi == 2) ? value_cube_kernel<2> : (P.tpi == 8 ? value_cube_kernel<8> : value_cube_kernel<4>);
    auto v2 = (P.tpi == 2) ? value2_kernel<2> : (P.tpi == 8 ? value2_kernel<8> : value2_kernel<4>);
    auto kernel = (P.factors == 2) ? v2 : (cube ? cb : (P.ppl == 1 ? p1 : value_kernel));
    hipLaunchKernelGGL(kernel, dim3((unsigned)nu), dim3(kValueThreads), P.lds_bytes, st, P.vplan.as<const ValuePlan>(),
                       P.units.as<const int4>() + u0);
    HIPCHK(hipGetLastError());
}

void launch_value_flat(eqhip_plan& P) {
    FlatArgs g{};
    g.R = P.r_shared ? P.r_shared : P.R.as<double>();
    g.r_stride = P.r_shared ? P.r_stride : P.mc.n_paths;
    g.n_paths = P.mc.n_paths;
    g.n_dates = P.n_dates;
    g.m = P.flat_m;
    g.n_trades = P.n_trades;
    g.r0 = P.model.r0;
    g.date_slot = P.date_slot.as<int>();
    g.date_det = P.date_det.as<int>();
    g.first = P.flat_first.as<int>();
    g.AB = P.flat_AB.as<double2>();
    g.acc = P.flat_acc.as<double>();
    g.N = P.flat_N.as<double>();
    g.NK = P.flat_NK.as<double>();
    g.X = P.X.as<double>();
    dim3 grid((unsigned)((P.mc.n_paths + kFlatThreads - 1) / kFlatThreads), (unsigned)P.n_dates);
    hipLaunchKernelGGL(value_flat_kernel, grid, dim3(kFlatThreads), 0, P.ctx->stream, g);
    HIPCHK(hipGetLastError());
}

// Unilateral CVA of n_profiles EE rows ([profile][date]) on the stream: the per-date factors once
// into w (n_dates double2), then one fixed-order sum per profile.
void launch_cva_weights(hipStream_t st, const double* d_times, int n_dates, const eqhip_model_hw1f& m, double hazard,
                        double2* w) {
    hipLaunchKernelGGL(cva_weights_kernel, dim3((n_dates + 127) / 128), dim3(128), 0, st, d_times, n_dates, m.a, m.b,
                       m.sigma, m.r0, hazard, w);
    HIPCHK(hipGetLastError());
}
void launch_cva_weights2(hipStream_t st, const double* d_times, int n_dates, const eqhip_model_hw2f& m, double hazard,
                         double2* w) {
    hipLaunchKernelGGL(cva_weights2_kernel, dim3((n_dates + 127) / 128), dim3(128), 0, st, d_times, n_dates, m.a, m.b,
                       m.sigma, m.eta, m.rho, m.r0, hazard, w);
    HIPCHK(hipGetLastError());
}
void launch_cva_sum(hipStream_t st, const double* ee, int n_profiles, int n_dates, const double2* w, double recovery,
                    double* out) {
    hipLaunchKernelGGL(cva_kernel, dim3((n_profiles + 3) / 4), dim3(256), 0, st, ee, n_profiles, n_dates, w, recovery,
                       out);
    HIPCHK(hipGetLastError());
}

// HW1F paths into R[slot_of_step[step]][path] for steps 0..n_steps (Euler, kernel.rs:41-50; or the
// exact transition over dt), from host draws Z or the Philox stream of mc.seed.
void launch_simulate1(hipStream_t st, const eqhip_model_hw1f& m, const eqhip_mc& mc, double dt, uint32_t n_steps,
                      const int* slot_of_step, const double* Z, double* R) {
    const uint64_t np = mc.n_paths;
    SimArgs s{};
    s.a = m.a;
    s.b = m.b;
    s.sigma = m.sigma;
    s.r0 = m.r0;
    s.dt = dt;
    s.sqrt_dt = std::sqrt(dt);
    s.decay = std::exp(-(m.a * dt));
    s.mean_c = m.b * (1.0 - s.decay);
    s.sd = m.sigma * std::sqrt((1.0 - std::exp(-(2.0 * m.a * dt))) / (2.0 * m.a));
    s.n_paths = np;
    s.n_steps = n_steps;
    s.seed = mc.seed;
    s.slot_of_step = slot_of_step;
    s.Z = Z;
    s.R = R;
    dim3 sgrid((unsigned)((np + 255) / 256));
    dim3 pgrid((unsigned)((np + kSimPaths - 1) / kSimPaths));
    switch (mc.disc * 2 + mc.rng) {
        case 0: hipLaunchKernelGGL((simulate_kernel<0, 0>), sgrid, dim3(256), 0, st, s); break;
        case 1: hipLaunchKernelGGL((simulate_philox_kernel<0>), pgrid, dim3(256), 0, st, s); break;
        case 2: hipLaunchKernelGGL((simulate_kernel<1, 0>), sgrid, dim3(256), 0, st, s); break;
        default: hipLaunchKernelGGL((simulate_philox_kernel<1>), pgrid, dim3(256), 0, st, s); break;
    }
    HIPCHK(hipGetLastError());
}

// G2++ paths of a plan into (R, R2). Euler follows hull_white_2f.rs:116-140; the exact scheme
// draws the pair's Gaussian transition over dt (correlation c of the two increments).
void launch_simulate2(eqhip_plan& P) {
    const eqhip_model_hw2f& m = P.model2;
    const double dt = P.dt;
    Sim2Args s{};
    s.a = m.a;
    s.b = m.b;
    s.sigma = m.sigma;
    s.eta = m.eta;
    s.rho = m.rho;
    s.dt = dt;
    s.sqrt_dt = std::sqrt(dt);
    s.sqrt_1mrho2 = std::sqrt(1.0 - m.rho * m.rho);
    s.ex = std::exp(-(m.a * dt));
    s.ey = std::exp(-(m.b * dt));
    s.sdx = m.sigma * std::sqrt((1.0 - std::exp(-(2.0 * m.a * dt))) / (2.0 * m.a));
    s.sdy = m.eta * std::sqrt((1.0 - std::exp(-(2.0 * m.b * dt))) / (2.0 * m.b));
    const double cov = m.rho * m.sigma * m.eta * (1.0 - std::exp(-((m.a + m.b) * dt))) / (m.a + m.b);
    double c = (s.sdx > 0.0 && s.sdy > 0.0) ? cov / (s.sdx * s.sdy) : 0.0;
    c = std::min(1.0, std::max(-1.0, c));
    s.c_xy = c;
    s.sqrt_1mc2 = std::sqrt(1.0 - c * c);
    s.n_paths = P.mc.n_paths;
    s.n_steps = P.det ? 0u : P.mc.n_steps;
    s.seed = P.mc.seed;
    s.slot_of_step = P.slot_of_step.as<int>();
    s.Z = P.Z.as<double>();
    s.X = P.R.as<double>();
    s.Y = P.R2.as<double>();
    dim3 grid((unsigned)((P.mc.n_paths + 255) / 256));
    dim3 pgrid((unsigned)((P.mc.n_paths + kSimPaths - 1) / kSimPaths));
    hipStream_t st = P.ctx->stream;
    switch (P.mc.disc * 2 + P.mc.rng) {
        case 0: hipLaunchKernelGGL((simulate2_kernel<0, 0>), grid, dim3(256), 0, st, s); break;
        case 1: hipLaunchKernelGGL((simulate2_philox_kernel<0>), pgrid, dim3(256), 0, st, s); break;
        case 2: hipLaunchKernelGGL((simulate2_kernel<1, 0>), grid, dim3(256), 0, st, s); break;
        default: hipLaunchKernelGGL((simulate2_philox_kernel<1>), pgrid, dim3(256), 0, st, s); break;
    }
    HIPCHK(hipGetLastError());
}

// Times one valuation launch of the plan as it is laid out now (after a warm-up launch).
float time_value_launch(eqhip_plan& P) {
    hipStream_t st = P.ctx->stream;
    launch_value(P, nullptr, nullptr);
    HIPCHK(hipEventRecord(P.ev[0], st));
    launch_value(P, nullptr, nullptr);
    HIPCHK(hipEventRecord(P.ev[1], st));
    HIPCHK(hipEventSynchronize(P.ev[1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, P.ev[0], P.ev[1]));
    return ms;
}

// Autotuner of a plan: on its own simulated paths, times every (paths per lane, trades per item)
// variant that fits and keeps the fastest. The trial launches write the outputs the real launch then
// overwrites, and every variant writes the same netting-set values.
void tune_plan(eqhip_plan& P) {
    if (!P.parts.empty() || P.path_mode) {  // the parts keep their default variant; path mode has none
        P.tuned = true;
        return;
    }
    const int ppls[3] = {1, 2, 4}, tpis[3] = {2, 4, 8};
    float best = 1e30f;
    int best_ppl = P.ppl, best_tpi = P.tpi;
    for (int ppl : ppls)
        for (int tpi : tpis) {
            // without a per-trade step TPI does not matter: one variant per PPL
            if (!P.want_trade && tpi != (ppl == 4 ? 2 : 4)) continue;
            if (P.want_pfe && ppl != 1) continue;  // the class legs are written at one path per lane
            if (!variant_fits(P, ppl, tpi)) continue;
            P.ppl = ppl;
            P.tpi = tpi;
            choose_launch(P);
            build_value_plan(P);
            const float ms = time_value_launch(P);
            P.tune_log.push_back("paths/lane=" + std::to_string(ppl) + " trades/item=" + std::to_string(tpi) + ": " +
                                 std::to_string(ms) + " ms");
            if (ms < best) {
                best = ms;
                best_ppl = ppl;
                best_tpi = tpi;
            }
        }
    P.ppl = best_ppl;
    P.tpi = best_tpi;
    choose_launch(P);
    build_value_plan(P);
    P.tune_log.push_back("chosen: paths/lane=" + std::to_string(best_ppl) + " trades/item=" + std::to_string(best_tpi));
    P.tuned = true;
}

void execute_plan(eqhip_plan& P, const double* host_draws, const eqhip_credit* credit,
                  eqhip_outputs* out) {
    if (!out || !out->set_ee || !out->set_pfe) raise(EQHIP_E_INVALID, "outputs set_ee / set_pfe are required");
    if ((out->trade_ee || out->trade_cva) && !P.want_trade)
        raise(EQHIP_E_INVALID, "per-trade outputs need a plan created with want_trade_ee");
    if ((out->set_cva || out->trade_cva) && !credit) raise(EQHIP_E_INVALID, "CVA needs credit data");
    if (out->trade_pfe && !P.want_pfe)
        raise(EQHIP_E_INVALID, "trade_pfe needs a plan created with EQHIP_PLAN_TRADE_PFE");
    hipStream_t st = P.ctx->stream;
    HIPCHK(hipSetDevice(P.ctx->device));
    const uint64_t np = P.mc.n_paths;
    if (P.mc.rng == EQHIP_RNG_HOST_DRAWS) {
        if (!host_draws) raise(EQHIP_E_INVALID, "host_draws required in EQHIP_RNG_HOST_DRAWS mode");
        HIPCHK(hipMemcpyAsync(P.Z.p, host_draws, (size_t)P.factors * P.mc.n_steps * np * 8, hipMemcpyHostToDevice,
                              st));
    }
    DevBuf dpath, dset;
    if (out->path_mtm) {
        if (!P.want_trade) raise(EQHIP_E_INVALID, "path_mtm needs a plan created with want_trade_ee");
        dpath.alloc((size_t)P.n_dates * P.n_trades * np * 8);
    }
    if (out->set_mtm && !P.keep_cube) dset.alloc((size_t)P.n_dates * P.S * np * 8);
    double* set_values = P.keep_cube ? P.cube.as<double>() : dset.as<double>();

    HIPCHK(hipEventRecord(P.ev[0], st));
    if (P.r_shared) {
        // the short rates come from a scenario set simulated beforehand
    } else if (P.factors == 2) {
        launch_simulate2(P);
    } else {
        launch_simulate1(st, P.model, P.mc, P.dt, P.det ? 0u : P.mc.n_steps, P.slot_of_step.as<int>(),
                         P.Z.as<double>(), P.R.as<double>());
    }
    HIPCHK(hipEventRecord(P.ev[1], st));
    if (P.autotune && !P.tuned && !P.flat) {
        tune_plan(P);  // the paths are simulated: time the variants on them, then value with the fastest
        HIPCHK(hipEventRecord(P.ev[1], st));
    }
    if (out->trade_pfe) {  // per date group: the valuation with the class legs, then the trades' rows
        for (size_t gi = 0; gi + 1 < P.grp_k.size(); ++gi) {
            launch_value(P, dpath.as<double>(), set_values, P.grp_k[gi], P.grp_k[gi + 1], true);
            HIPCHK(hipEventRecord(P.pfe_ev[2 * gi], st));
            const size_t r0 = P.grp_row[gi], r1 = P.grp_row[gi + 1];
            if (r1 > r0)
                hipLaunchKernelGGL(trade_select_kernel, dim3((unsigned)(r1 - r0)), dim3(kSelectThreads), 0, st,
                                   P.trade_rows.as<const TradeSelectRow>() + r0);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(P.pfe_ev[2 * gi + 1], st));
        }
    } else if (P.flat && !out->path_mtm && !out->set_mtm) {
        launch_value_flat(P);
    } else {
        launch_value(P, dpath.as<double>(), set_values);
    }
    HIPCHK(hipEventRecord(P.ev[2], st));
    hipLaunchKernelGGL(select_kernel, dim3((unsigned)(P.S * P.n_dates)), dim3(kSelectThreads), 0, st,
                       P.rows.as<const SelectRow>());
    HIPCHK(hipGetLastError());
    if (P.want_trade) {
        int n = P.n_trades * P.n_dates;
        hipLaunchKernelGGL(trade_ee_kernel, dim3((n + 255) / 256), dim3(256), 0, st, P.partial.as<double>(),
                           P.n_trades, P.n_dates, P.n_chunks, P.neff.as<uint64_t>(), P.trade_ee.as<double>());
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(P.ev[3], st));
    double h = credit ? credit->hazard_rate : 0.0, rr = credit ? credit->recovery_rate : 0.0;
    if (out->set_cva || out->trade_cva) {
        reserve(P.cva_w, (size_t)P.n_dates * 16);
        if (P.factors == 2)
            launch_cva_weights2(st, P.d_dates.as<double>(), P.n_dates, P.model2, h, P.cva_w.as<double2>());
        else
            launch_cva_weights(st, P.d_dates.as<double>(), P.n_dates, P.model, h, P.cva_w.as<double2>());
    }
    if (out->set_cva)
        launch_cva_sum(st, P.set_ee.as<double>(), P.S, P.n_dates, P.cva_w.as<double2>(), rr, P.set_cva.as<double>());
    if (out->trade_cva)
        launch_cva_sum(st, P.trade_ee.as<double>(), P.n_trades, P.n_dates, P.cva_w.as<double2>(), rr,
                       P.trade_cva.as<double>());
    HIPCHK(hipEventRecord(P.ev[4], st));

    // profiles and CVAs come back through the plan's page-locked staging buffer (DMA-speed copies,
    // then host copies into the caller's arrays); the optional per-path debug outputs go directly
    struct Out {
        void* dst;
        const DevBuf* src;
        size_t bytes, off;
    };
    Out outs[6] = {{out->set_ee, &P.set_ee, (size_t)P.S * P.n_dates * 8, 0},
                   {out->set_pfe, &P.set_pfe, (size_t)P.S * P.n_dates * 8, 0},
                   {out->set_cva, &P.set_cva, (size_t)P.S * 8, 0},
                   {out->trade_ee, &P.trade_ee, (size_t)P.n_trades * P.n_dates * 8, 0},
                   {out->trade_cva, &P.trade_cva, (size_t)P.n_trades * 8, 0},
                   {out->trade_pfe, &P.trade_pfe, (size_t)P.n_trades * P.n_dates * 8, 0}};
    size_t staged = 0;
    for (Out& o : outs)
        if (o.dst) {
            o.off = staged;
            staged += (o.bytes + 255) & ~(size_t)255;
        }
    if (P.direct_out) {  // a one-shot plan: no page-locked staging to set up for a single use
        for (const Out& o : outs)
            if (o.dst) HIPCHK(hipMemcpyAsync(o.dst, o.src->p, o.bytes, hipMemcpyDeviceToHost, st));
    } else {
        reserve(P.h_out, staged);
        for (const Out& o : outs)
            if (o.dst)
                HIPCHK(hipMemcpyAsync(P.h_out.as<char>() + o.off, o.src->p, o.bytes, hipMemcpyDeviceToHost, st));
    }
    if (out->path_mtm)
        HIPCHK(hipMemcpyAsync(out->path_mtm, dpath.p, (size_t)P.n_dates * P.n_trades * np * 8, hipMemcpyDeviceToHost, st));
    if (out->set_mtm)
        HIPCHK(hipMemcpyAsync(out->set_mtm, set_values, (size_t)P.n_dates * P.S * np * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    P.executed = true;
    if (!P.direct_out)
        for (const Out& o : outs)
            if (o.dst) std::memcpy(o.dst, P.h_out.as<char>() + o.off, o.bytes);

    float ms[4];
    for (int i = 0; i < 4; ++i) HIPCHK(hipEventElapsedTime(&ms[i], P.ev[i], P.ev[i + 1]));
    double pfe_ms = 0.0;
    if (out->trade_pfe)
        for (size_t gi = 0; 2 * gi + 1 < P.pfe_ev.size(); ++gi) {
            float m = 0.0f;
            HIPCHK(hipEventElapsedTime(&m, P.pfe_ev[2 * gi], P.pfe_ev[2 * gi + 1]));
            pfe_ms += m;
        }
    P.timing.simulate_ms = ms[0];
    P.timing.value_ms = ms[1] - pfe_ms;
    P.timing.trade_pfe_ms = pfe_ms;
    P.timing.trade_pfe_groups = out->trade_pfe ? (uint32_t)(P.grp_k.size() - 1) : 0u;
    P.timing.reduce_ms = ms[2];
    P.timing.cva_ms = ms[3];
    P.timing.total_ms = (double)ms[0] + ms[1] + ms[2] + ms[3];
    P.timing.n_value_blocks = (uint64_t)P.n_dates * P.n_chunks * std::max<size_t>(1, P.parts.size());
    P.timing.n_parts = (uint32_t)std::max<size_t>(1, P.parts.size());
    P.timing.paths_per_tile =
        P.path_mode ? (uint32_t)(kPathThreads * path_ppt(P.S)) : (uint32_t)(kPathsPerTile * P.ppl);
    P.timing.tiles_per_block = (uint32_t)P.tiles_per_block;
    P.timing.trades_per_item = P.path_mode ? 0u : (uint32_t)P.tpi;
}

// ---------------------------------------------------------------------------------------------
// XVA reducer over an exposure cube (quant-engine/src/xva.rs:97-246): the row work (CSA adjustment,
// positive / negative parts, EE sums and the row quantile) on the device, then the accumulation
// over time rows on the host in the reference's loop order.
// ---------------------------------------------------------------------------------------------
struct XvaRows {
    const double* q;   // device rows of the Q cube
    uint64_t q_stride;
    const double* p;   // device rows of the P cube, or null
    uint64_t p_stride;
    uint64_t n;        // paths per row
    int n_times;
};

void validate_xva_params(const eqhip_xva_params* x) {
    if (!x) raise(EQHIP_E_INVALID, "xva params are null");
    if (!(x->confidence_level >= 0.0 && x->confidence_level <= 1.0))
        raise(EQHIP_E_INVALID, "confidence_level must be in [0,1]");
    for (const eqhip_default_spec* d : {&x->counterparty_default, &x->own_default})
        if (d->n_survival < 0 || (d->n_survival > 0 && !d->survival_probabilities))
            raise(EQHIP_E_INVALID, "survival_probabilities: null array with a positive length");
}

// xva.rs:17-23
double xva_dt(const double* times, int i) {
    return (i == 0) ? std::max(times[i], 0.0) : std::max(times[i] - times[i - 1], 0.0);
}
// xva.rs:31-46
double xva_default_increment(const eqhip_default_spec& spec, int i, const double* times) {
    if (i < spec.n_survival) {
        const double survival = spec.survival_probabilities[i];
        const double previous = (i == 0) ? 1.0 : spec.survival_probabilities[i - 1];
        return std::max(previous - survival, 0.0);
    }
    const double h = std::max(spec.hazard_rate, 0.0);
    const double start = (i == 0) ? 0.0 : times[i - 1];
    return std::exp(-(h * start)) - std::exp(-(h * std::max(times[i], start)));
}
double xva_clamp_recovery(double r) { return std::min(std::max(r, 0.0), 1.0); }

void xva_reduce(eqhip_context* ctx, const XvaRows& rows, DevBuf& pos, const eqhip_xva_params& x, const double* times,
                const double* discount, double* out_ee, double* out_pfe, double* out_measures,
                hipEvent_t* ev = nullptr, double* ms_rows = nullptr, double* ms_select = nullptr) {
#pragma clang fp contract(off)
    hipStream_t st = ctx->stream;
    const int nt = rows.n_times;
    const uint64_t n = rows.n;
    reserve(pos, (size_t)nt * n * 8);
    DevBuf sums, dev_rows;
    sums.alloc((size_t)nt * 4 * 8);  // ee, pfe, ene_q, ene_p
    double* d_sums = sums.as<double>();
    XvaRowArgs g{};
    g.q = rows.q;
    g.q_stride = rows.q_stride;
    g.p = rows.p;
    g.p_stride = rows.p_stride;
    g.n = n;
    g.pos = pos.as<double>();
    g.ene_q = d_sums + 2 * nt;
    g.ene_p = d_sums + 3 * nt;
    const eqhip_csa* csa = x.csa;
    g.csa_vm = (csa && csa->variation_margin) ? 1 : 0;
    if (csa) {
        g.threshold = std::max(csa->threshold, 0.0);
        g.mta = std::max(csa->minimum_transfer_amount, 0.0);
        g.ia = std::max(csa->independent_amount, 0.0);
    }
    // the quantile of xva.rs:66-75: sorted[round((n - 1) p)] of the positive parts
    const uint64_t rank = (uint64_t)std::llround((double)(n - 1) * x.confidence_level);
    std::vector<SelectRow> sel(nt);
    for (int k = 0; k < nt; ++k)
        sel[k] = SelectRow{g.pos + (size_t)k * n, n, std::min(rank, n - 1), d_sums + k, d_sums + nt + k};
    upload(dev_rows, sel, st);
    if (ev) HIPCHK(hipEventRecord(ev[0], st));
    hipLaunchKernelGGL(xva_rows_kernel, dim3((unsigned)nt), dim3(kSelectThreads), 0, st, g);
    HIPCHK(hipGetLastError());
    if (ev) HIPCHK(hipEventRecord(ev[1], st));
    hipLaunchKernelGGL(select_kernel, dim3((unsigned)nt), dim3(kSelectThreads), 0, st,
                       dev_rows.as<const SelectRow>());
    HIPCHK(hipGetLastError());
    if (ev) HIPCHK(hipEventRecord(ev[2], st));
    std::vector<double> h((size_t)nt * 4);
    HIPCHK(hipMemcpyAsync(h.data(), d_sums, (size_t)nt * 4 * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (ev) {
        float a = 0.f, b = 0.f;
        HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
        HIPCHK(hipEventElapsedTime(&b, ev[1], ev[2]));
        *ms_rows = a;
        *ms_select = b;
    }
    const double* ee = h.data();
    const double* pfe = h.data() + nt;
    const double* ene_q = h.data() + 2 * nt;
    const double* ene_p = h.data() + 3 * nt;
    const double inv_rc = 1.0 - xva_clamp_recovery(x.counterparty_default.recovery_rate);
    const double inv_ro = 1.0 - xva_clamp_recovery(x.own_default.recovery_rate);
    double cva = 0.0, dva = 0.0, fva = 0.0, mva = 0.0, kva = 0.0;
    for (int i = 0; i < nt; ++i) {
        const double expected_positive = ee[i];
        const double expected_negative = (rows.p ? ene_p[i] : ene_q[i]) / (double)n;
        const double d = discount ? discount[i] : std::exp(-0.0 * times[i]);
        const double yf = xva_dt(times, i);
        out_ee[i] = expected_positive;
        out_pfe[i] = pfe[i];
        cva += d * expected_positive * xva_default_increment(x.counterparty_default, i, times) * inv_rc;
        fva += d * expected_positive * x.funding_spread * yf;
        mva += d * x.initial_margin_total * x.borrowing_spread * yf;
        kva += d * expected_positive * std::max(x.capital_factor, 0.0) * yf;
        dva += d * expected_negative * xva_default_increment(x.own_default, i, times) * inv_ro;
    }
    out_measures[0] = cva;
    out_measures[1] = dva;
    out_measures[2] = fva;
    out_measures[3] = mva;
    out_measures[4] = kva;
}

// Homogeneous one-shot batches (one schedule, trade j alone in netting set j, HW1F): when every date
// is a reset date of the schedule, precedes its start or follows its maturity, the netting-set values
// come from value_flat_kernel instead of the general valuation. The date rules are those of the
// general valuation (and of irs.rs:163-187 at reset dates): flows alive at t are those paid after
// t + 1e-9; the floating leg starts at t on a reset date, at the start before it. A date inside a
// period (a fixing in progress) leaves the plan on the general kernel.
void build_flat(eqhip_plan& P, const eqhip_portfolio& pf) {
    if (P.factors != 1 || P.want_trade || P.keep_cube || P.det || P.n_dates > 65535) return;
    const int m = pf.flow_offset[1] - pf.flow_offset[0];
    const double start = pf.start[0];
    const double* pay = pf.payment_times;
    const int nd = P.n_dates;
    std::vector<int> first(nd, m);
    std::vector<double> tau((size_t)nd * (m + 1), 0.0);
    for (int k = 0; k < nd; ++k) {
        const double t = P.dates[k];
        if (t >= pay[m - 1] - kDateTolHost) continue;  // matured
        int f = 0;
        while (f < m && !(pay[f] > t + kDateTolHost)) ++f;
        double s = t;  // the floating leg's first bond date
        if (std::fabs(t - start) >= kDateTolHost) {
            if (t < start) {
                s = start;
            } else if (std::fabs(t - (f == 0 ? start : pay[f - 1])) >= kDateTolHost) {
                return;  // inside a period
            }
        }
        first[k] = f;
        for (int i = f; i < m; ++i) tau[(size_t)k * (m + 1) + i] = pay[i] - t;
        tau[(size_t)k * (m + 1) + m] = s - t;
    }
    P.flat_m = m;
    std::vector<double> acc(pf.accruals, pf.accruals + m), N(pf.n_trades), NK(pf.n_trades);
    for (int j = 0; j < pf.n_trades; ++j) {
        N[j] = pf.notional[j];
        NK[j] = pf.notional[j] * pf.fixed_rate[j];
    }
    TablePack pack;
    DevBuf d_tau;
    pack.add(P.flat_first, first);
    pack.add(d_tau, tau);
    pack.add(P.flat_acc, acc);
    pack.add(P.flat_N, N);
    pack.add(P.flat_NK, NK);
    pack.flush(P.flat_tables);
    const int n = (int)tau.size();
    P.flat_AB.alloc((size_t)n * sizeof(double2));
    hipLaunchKernelGGL(ab_kernel, dim3((n + 255) / 256), dim3(256), 0, P.ctx->stream, d_tau.as<double>(), n, P.model.a,
                       P.model.b, P.model.sigma, P.flat_AB.as<double2>());
    HIPCHK(hipGetLastError());
    P.flat = true;
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
                   double* out_pfe, int32_t* n_out, const eqhip_model_hw2f* model2 = nullptr,
                   const eqhip_scenarios* scen = nullptr) {
    check_ctx(ctx);
    if (scen) {
        if (scen->device != ctx->device) raise(EQHIP_E_INVALID, "the scenario set lives on another device");
        if (n_dates < 1 || !times) raise(EQHIP_E_INVALID, "monitoring_times vacío");
        double last = -INFINITY;
        for (int k = 0; k < n_dates; ++k) {
            if (!(times[k] >= 0.0)) raise(EQHIP_E_INVALID, "monitoring dates on a scenario set must be >= 0");
            last = std::max(last, times[k]);
        }
        if (last != scen->t_max)
            raise(EQHIP_E_INVALID, "the last monitoring date (" + std::to_string(last) +
                                       ") must be the scenario set's t_max (" + std::to_string(scen->t_max) + ")");
    }
    if (n_dates < 1 || !times) raise(EQHIP_E_INVALID, "monitoring_times vacío");
    if (n_steps < 1) raise(EQHIP_E_INVALID, "n_steps debe ser al menos 1");
    if (m < 1 || !pay || !acc) raise(EQHIP_E_INVALID, "un swap necesita al menos un periodo");
    if (!allow_any_date)
        for (int k = 0; k < n_dates; ++k)
            if (!is_reset_date(start, pay, m, times[k]))
                raise(EQHIP_E_INVALID, "t=" + std::to_string(times[k]) + " no es una fecha de reseteo válida del swap");
    double t_max = -INFINITY;
    for (int k = 0; k < n_dates; ++k) t_max = std::max(t_max, times[k]);
    int nd = (t_max <= 0.0) ? 1 : n_dates;  // exposure.rs:98-108 (and 325-337) return only the first date
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
    std::lock_guard<std::mutex> lock(ctx->oneshot_lock);
    eqhip_plan P;
    P.ctx = ctx;
    P.device = ctx->device;
    P.direct_out = true;
    if (model2) {
        P.factors = 2;
        P.model2 = *model2;
    }
    if (scen) {  // every grid step of the set's paths, nothing to simulate
        P.r_shared = scen->R.as<double>();
        P.r_stride = n_paths;
    }
    build_plan(P, model2 ? nullptr : &model, &pf, times, nd, &mc, 0, scen != nullptr);
    build_flat(P, pf);
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

// ---------------------------------------------------------------------------------------------
// Pathwise adjoint of one netting set's CVA. V_s at date k is written as
//   V_s = C0_k + sum_d nu_{k,d} P(t_k, T_d) + sum_{families mid-period} alpha_F P(t_k,T_next)/P_fix
// (reset floating legs give the constant, forward starts and maturities and annuity flows give nu);
// the adjoint kernel differentiates that per path and sweeps the short rate backwards.
// ---------------------------------------------------------------------------------------------
void sens_table(eqhip_context* ctx, const std::vector<double>& tau, const eqhip_model_hw1f& m, DevBuf& out) {
    DevBuf dt;
    upload(dt, tau, ctx->stream);
    out.alloc(std::max<size_t>(1, tau.size()) * sizeof(ZcbSens));
    int n = (int)tau.size();
    if (n > 0) {
        hipLaunchKernelGGL(zcb_sens_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, dt.as<double>(), n,
                           m.a, m.b, m.sigma, out.as<ZcbSens>());
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
}

void cva_adjoint(eqhip_plan& P, const double* host_draws, int set, const eqhip_credit* credit, double* out_cva,
                 double* out_grad, double* out_check) {
    if (set < 0 || set >= P.S) raise(EQHIP_E_INVALID, "set index out of range");
    if (!credit || !out_cva || !out_grad) raise(EQHIP_E_INVALID, "credit / out_cva / out_grad are required");
    if (P.factors != 1) raise(EQHIP_E_INVALID, "the CVA adjoint serves HullWhite1F plans only");
    if (P.det) raise(EQHIP_E_INVALID, "the adjoint needs at least one monitoring date after t = 0");
    const int nd = P.n_dates, D = P.D;
    const uint32_t steps = P.mc.n_steps;
    std::vector<int> date_of_step(steps + 1, -1);
    for (int k = 0; k < nd; ++k) {
        int g = P.h_date_grid[k];
        if (date_of_step[g] >= 0) raise(EQHIP_E_INVALID, "the adjoint needs one grid step per monitoring date");
        date_of_step[g] = k;
    }
    // forward exposure run: EE of every set (also fills R / X on the device)
    std::vector<double> ee((size_t)P.S * nd), pfe((size_t)P.S * nd), cva(P.S);
    eqhip_outputs o{};
    o.set_ee = ee.data();
    o.set_pfe = pfe.data();
    o.set_cva = cva.data();
    execute_plan(P, host_draws, credit, &o);

    // per-date linear form of V_s
    std::vector<double> C0(nd, 0.0), nu((size_t)nd * D, 0.0), mid_alpha;
    std::vector<int> mid_off(nd + 1, 0);
    std::vector<int4> mid;
    std::vector<double> fix_tau_adj;
    std::map<int64_t, int> gslot;  // fixing grid step -> G row
    std::vector<int64_t> cp_rows;  // extra R rows (fixing steps off the checkpoint grid)
    const int K = kAdjCheckpoint;
    const int n_cp = (int)(steps / K) + 1;
    auto rrow = [&](int64_t step) -> int {  // row of the forward buffer holding r at this step
        if (step % K == 0) return (int)(step / K);
        auto it = std::find(cp_rows.begin(), cp_rows.end(), step);
        if (it == cp_rows.end()) {
            cp_rows.push_back(step);
            return n_cp + (int)cp_rows.size() - 1;
        }
        return n_cp + (int)(it - cp_rows.begin());
    };
    std::vector<double> cls_alpha(P.C, 0.0), cls_beta(P.C, 0.0);
    for (int e = P.h_set_off[set]; e < P.h_set_off[set + 1]; ++e) {
        cls_alpha[P.h_set_cls[e]] = P.h_set_alpha[e];
        cls_beta[P.h_set_cls[e]] = P.h_set_beta[e];
    }
    for (int k = 0; k < nd; ++k) {
        std::map<std::pair<int, int>, int> mid_here;  // (next payment, fixing entry) -> mid entry
        for (int F = 0; F < P.NF; ++F) {
            const int4 fd = P.h_fdesc[(size_t)k * P.NF + F];
            if (fd.x == 0) continue;
            const int f0 = P.h_fam_off[F], f1 = P.h_fam_off[F + 1];
            double alpha_mid = 0.0;
            (void)f1;
            for (int c = 0; c < P.C; ++c) {
                if (P.h_cls_fam[c] != F || (cls_alpha[c] == 0.0 && cls_beta[c] == 0.0)) continue;
                const int last = P.h_cls_len[c] - 1;
                if (last < fd.y) continue;  // matured
                const double al = cls_alpha[c], be = cls_beta[c];
                nu[(size_t)k * D + P.h_flow_d[f0 + last]] -= al;
                for (int i = fd.y; i <= last; ++i) nu[(size_t)k * D + P.h_flow_d[f0 + i]] -= be * P.h_flow_tau[f0 + i];
                if (fd.x == 1) C0[k] += al;
                else if (fd.x == 2) nu[(size_t)k * D + fd.z] += al;
                else alpha_mid += al;
            }
            if (fd.x == 3 && alpha_mid != 0.0) {
                // families with the same next payment and fixing share one entry
                auto same = mid_here.find(std::make_pair(fd.z, fd.w));
                if (same != mid_here.end()) {
                    mid_alpha[(size_t)same->second] += alpha_mid;
                    continue;
                }
                const int fe = P.h_fix_off[k] + fd.w;              // fixing entry of this date
                const int64_t fstep = P.h_slot_grid[P.h_fix_slot[fe]];
                auto it = gslot.find(fstep);
                if (it == gslot.end()) it = gslot.emplace(fstep, (int)gslot.size()).first;
                mid_here.emplace(std::make_pair(fd.z, fd.w), (int)mid.size());
                mid.push_back(make_int4(fd.z, rrow(fstep), (int)fix_tau_adj.size(), it->second));
                fix_tau_adj.push_back(P.h_fix_tau[fe]);
                mid_alpha.push_back(alpha_mid);
            }
        }
        mid_off[k + 1] = (int)mid.size();
    }
    // mid entries address R rows through .y and G rows through .w; the kernel reads .y for r_fix
    std::vector<int> fixslot_of_step(steps + 1, -1);
    for (auto& kv : gslot) fixslot_of_step[(size_t)kv.first] = kv.second;
    const int n_rows = n_cp + (int)cp_rows.size();

    // weights w_k = (1-R)(S_{k-1}-S_k) P(0,t_k) / n_eff(k) and the direct CVA terms
    std::vector<double> tk(P.dates);
    DevBuf d_sens0;
    sens_table(P.ctx, tk, P.model, d_sens0);
    std::vector<ZcbSens> s0(nd);
    HIPCHK(hipMemcpy(s0.data(), d_sens0.p, nd * sizeof(ZcbSens), hipMemcpyDeviceToHost));
    const double h = credit->hazard_rate, R = credit->recovery_rate, r0 = P.model.r0;
    std::vector<double> wk(nd), neffv(nd);
    double cva_s = 0.0, dcva[6] = {0, 0, 0, 0, 0, 0};
    double prev_s = 1.0, prev_t = 0.0;
    for (int k = 0; k < nd; ++k) {
        const double t = P.dates[k];
        const double S = std::exp(-h * t);
        const double dp = prev_s - S;
        const ZcbSens z = s0[k];
        const double P0 = z.A * std::exp(-(z.B * r0));
        const double e = ee[(size_t)set * nd + k];
        const double n_eff = (P.h_date_grid[k] == 0) ? 1.0 : (double)P.mc.n_paths;
        wk[k] = (1.0 - R) * dp * P0 / n_eff;
        cva_s += (1.0 - R) * e * dp * P0;
        const double c = (1.0 - R) * dp * P0;
        dcva[0] += c * (z.lnA_a - r0 * z.B_a) * e;
        dcva[1] += c * z.lnA_b * e;
        dcva[2] += c * z.lnA_s * e;
        dcva[3] += c * (-z.B) * e;
        dcva[4] += (1.0 - R) * e * P0 * (-prev_t * prev_s + t * S);
        dcva[5] += -e * dp * P0;
        prev_s = S;
        prev_t = t;
    }

    // device tables
    std::vector<double> tau((size_t)nd * D, -1.0);
    for (int k = 0; k < nd; ++k)
        for (int d = P.h_date_dfirst[k]; d < D; ++d) tau[(size_t)k * D + d] = P.h_T[d] - P.dates[k];
    DevBuf d_zs, d_fzs, d_C0, d_nu, d_mid_off, d_mid, d_mid_alpha, d_wk, d_dos, d_fos, d_buf, d_G, d_part, d_slot;
    sens_table(P.ctx, tau, P.model, d_zs);
    sens_table(P.ctx, fix_tau_adj, P.model, d_fzs);
    upload(d_C0, C0, P.ctx->stream);
    upload(d_nu, nu, P.ctx->stream);
    upload(d_mid_off, mid_off, P.ctx->stream);
    // kernel reads r_fix from row .y and writes G row .w: pack as (next, rrow, sens, grow)
    std::vector<int4> mid_dev = mid;
    upload(d_mid, mid_dev, P.ctx->stream);
    upload(d_mid_alpha, mid_alpha, P.ctx->stream);
    upload(d_wk, wk, P.ctx->stream);
    upload(d_dos, date_of_step, P.ctx->stream);
    upload(d_fos, fixslot_of_step, P.ctx->stream);
    // forward sweep: r at checkpoints (rows 0..n_cp-1) and at off-grid fixing steps
    std::vector<int> slot_of_step(steps + 1, -1);
    for (int c = 0; c < n_cp; ++c) slot_of_step[(size_t)c * K] = c;
    for (size_t i = 0; i < cp_rows.size(); ++i) slot_of_step[(size_t)cp_rows[i]] = n_cp + (int)i;
    upload(d_slot, slot_of_step, P.ctx->stream);
    const uint64_t np = P.mc.n_paths;
    d_buf.alloc((size_t)n_rows * np * 8);
    d_G.alloc((size_t)std::max<size_t>(1, gslot.size()) * np * 8);
    hipStream_t st = P.ctx->stream;
    SimArgs sa{};
    sa.a = P.model.a;
    sa.b = P.model.b;
    sa.sigma = P.model.sigma;
    sa.r0 = P.model.r0;
    sa.dt = P.dt;
    sa.sqrt_dt = std::sqrt(P.dt);
    sa.decay = std::exp(-(P.model.a * P.dt));
    sa.mean_c = P.model.b * (1.0 - sa.decay);
    const double e2 = std::exp(-(2.0 * P.model.a * P.dt));
    const double q = (1.0 - e2) / (2.0 * P.model.a);
    sa.sd = P.model.sigma * std::sqrt(q);
    sa.n_paths = np;
    sa.n_steps = steps;
    sa.seed = P.mc.seed;
    sa.slot_of_step = d_slot.as<int>();
    sa.Z = P.Z.as<double>();
    sa.R = d_buf.as<double>();
    dim3 sgrid((unsigned)((np + 255) / 256));
    const int key = P.mc.disc * 2 + P.mc.rng;
    switch (key) {
        case 0: hipLaunchKernelGGL((simulate_kernel<0, 0>), sgrid, dim3(256), 0, st, sa); break;
        case 1: hipLaunchKernelGGL((simulate_kernel<0, 1>), sgrid, dim3(256), 0, st, sa); break;
        case 2: hipLaunchKernelGGL((simulate_kernel<1, 0>), sgrid, dim3(256), 0, st, sa); break;
        default: hipLaunchKernelGGL((simulate_kernel<1, 1>), sgrid, dim3(256), 0, st, sa); break;
    }
    HIPCHK(hipGetLastError());

    AdjArgs g{};
    g.a = sa.a;
    g.b = sa.b;
    g.sigma = sa.sigma;
    g.r0 = sa.r0;
    g.dt = sa.dt;
    g.sqrt_dt = sa.sqrt_dt;
    g.decay = sa.decay;
    g.mean_c = sa.mean_c;
    g.sd = sa.sd;
    const double a = P.model.a;
    const double dq_da = (2.0 * a * P.dt * e2 - (1.0 - e2)) / (2.0 * a * a);
    g.dsd_da = (q > 0.0) ? P.model.sigma * dq_da / (2.0 * std::sqrt(q)) : 0.0;
    g.dsd_ds = std::sqrt(q);
    g.disc = P.mc.disc;
    g.rng = P.mc.rng;
    g.seed = P.mc.seed;
    g.Z = P.Z.as<double>();
    g.n_paths = np;
    g.n_steps = steps;
    g.CP = d_buf.as<double>();
    g.RF = d_buf.as<double>();
    g.G = d_G.as<double>();
    g.n_fix_slots = (int)gslot.size();
    g.date_of_step = d_dos.as<int>();
    g.fixslot_of_step = d_fos.as<int>();
    g.D = D;
    g.date_dfirst = P.date_dfirst.as<int>();
    g.C0 = d_C0.as<double>();
    g.nu = d_nu.as<double>();
    g.zs = d_zs.as<ZcbSens>();
    g.mid_off = d_mid_off.as<int>();
    g.mid = d_mid.as<int4>();
    g.mid_alpha = d_mid_alpha.as<double>();
    g.fix_zs = d_fzs.as<ZcbSens>();
    g.wk = d_wk.as<double>();
    g.date_det = P.date_det.as<int>();
    const unsigned blocks = (unsigned)((np + kAdjThreads - 1) / kAdjThreads);
    d_part.alloc((size_t)blocks * 5 * 8);
    g.partial = d_part.as<double>();
    switch (key) {
        case 0: hipLaunchKernelGGL((adjoint_kernel<0, 0>), dim3(blocks), dim3(kAdjThreads), 0, st, g); break;
        case 1: hipLaunchKernelGGL((adjoint_kernel<0, 1>), dim3(blocks), dim3(kAdjThreads), 0, st, g); break;
        case 2: hipLaunchKernelGGL((adjoint_kernel<1, 0>), dim3(blocks), dim3(kAdjThreads), 0, st, g); break;
        default: hipLaunchKernelGGL((adjoint_kernel<1, 1>), dim3(blocks), dim3(kAdjThreads), 0, st, g); break;
    }
    HIPCHK(hipGetLastError());
    std::vector<double> part((size_t)blocks * 5);
    HIPCHK(hipMemcpyAsync(part.data(), d_part.p, part.size() * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double sum[5] = {0, 0, 0, 0, 0};
    for (unsigned b = 0; b < blocks; ++b)
        for (int c = 0; c < 5; ++c) sum[c] += part[(size_t)b * 5 + c];
    for (int c = 0; c < 4; ++c) dcva[c] += sum[c];
    *out_cva = cva_s;
    for (int c = 0; c < 6; ++c) out_grad[c] = dcva[c];
    if (out_check) *out_check = sum[4];
}

// ---------------------------------------------------------------------------------------------
// Batched netting sets: many exposure jobs (own portfolio and path count; shared model, dates,
// grid, seed, rng and discretisation) in one simulation, one valuation launch and one select
// launch. Job j's paths are the first n_j paths of the shared simulation, so every job sees the
// shocks an unbatched run would draw: set-level results are bitwise those of unbatched runs.
// ---------------------------------------------------------------------------------------------
struct eqhip_batch {
    eqhip_context* ctx = nullptr;
    int device = 0;
    eqhip_model_hw1f model{};
    std::vector<std::unique_ptr<eqhip_plan>> jobs;
    uint64_t max_paths = 0;
    uint32_t n_steps = 0;
    uint64_t seed = 0;
    int rng = 0, disc = 0;
    double dt = 1.0;
    bool det = false;
    DevBuf R, slot_of_step, plans, units, rows, counter, path_units;
    std::vector<int4> unit_list;
    std::vector<int4> path_list;      // units of the jobs in path mode, launched on their own
    int path_nf = 0, path_ns = 0;     // the most families and netting sets of those jobs
    size_t lds_bytes = 0;
    int persistent = 1;
    int best_ppl = kDefaultPpl, best_tpi = kDefaultTpi;
    uint64_t target_units_per_cu = 8;
    bool tuned = false;
    std::vector<std::string> tune_log;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    eqhip_timing timing{};
    ~eqhip_batch() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

namespace {

int device_cus(eqhip_context* ctx) {
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
    return prop.multiProcessorCount;
}

// (Re)tiles every job for the batch's chunking target and uploads the joint descriptors, unit
// list (longest units first) and select rows.
void layout_batch(eqhip_batch& B) {
    const int cus = device_cus(B.ctx);
    std::vector<ValuePlan> vps;
    std::vector<SelectRow> rows;
    std::vector<std::pair<double, int4>> costed, path_costed;
    B.lds_bytes = 0;
    B.path_nf = B.path_ns = 0;
    for (auto& J : B.jobs)  // the path-mode launch covers every path-mode job; they are tiled for it
        if (J->path_mode) {
            B.path_nf = std::max(B.path_nf, J->NF);
            B.path_ns = std::max(B.path_ns, J->S);
        }
    for (auto& J : B.jobs) J->path_ns = B.path_ns;
    uint64_t total_tiles = 0;
    for (auto& J : B.jobs) total_tiles += J->n_tiles * (uint64_t)J->n_dates;
    for (size_t j = 0; j < B.jobs.size(); ++j) {
        eqhip_plan& P = *B.jobs[j];
        // share of the batch's units proportional to the job's tile-dates
        double share = (double)(P.n_tiles * (uint64_t)P.n_dates) / (double)std::max<uint64_t>(1, total_tiles);
        P.target_units = std::max<uint64_t>(1, (uint64_t)(share * (double)(B.target_units_per_cu * cus)));
        choose_launch(P);
        if (P.want_trade) P.partial.alloc((size_t)P.n_trades * P.n_dates * P.n_chunks * 8);
        P.r_shared = B.R.as<double>();
        P.r_stride = B.max_paths;
        build_value_plan(P);
        vps.push_back(P.vp);
        B.lds_bytes = std::max(B.lds_bytes, P.lds_bytes);
        for (const int4& u : P.unit_list) {
            int4 v = u;
            v.x = (int)j;
            uint64_t t0 = (uint64_t)u.z * P.tiles_per_block;
            uint64_t tiles = std::min<uint64_t>(P.tiles_per_block, P.n_tiles - t0);
            double live = (double)P.n_trades + 1.0;  // cost ~ tiles x (trades + classes)
            (P.path_mode ? path_costed : costed).emplace_back((double)tiles * live, v);
        }
        std::vector<uint64_t> neff(P.n_dates), qrank(P.n_dates);
        HIPCHK(hipMemcpy(neff.data(), P.neff.p, P.n_dates * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(qrank.data(), P.qrank.p, P.n_dates * 8, hipMemcpyDeviceToHost));
        for (int s = 0; s < P.S; ++s)
            for (int k = 0; k < P.n_dates; ++k) {
                const size_t r = (size_t)s * P.n_dates + k;
                rows.push_back(SelectRow{P.X.as<double>() + r * P.mc.n_paths, neff[k], qrank[k],
                                         P.set_ee.as<double>() + r, P.set_pfe.as<double>() + r});
            }
    }
    auto longest_first = [](const auto& a, const auto& b) { return a.first > b.first; };
    std::stable_sort(costed.begin(), costed.end(), longest_first);
    std::stable_sort(path_costed.begin(), path_costed.end(), longest_first);
    B.unit_list.clear();
    for (auto& c : costed) B.unit_list.push_back(c.second);
    B.path_list.clear();
    for (auto& c : path_costed) B.path_list.push_back(c.second);
    upload(B.plans, vps, B.ctx->stream);
    upload(B.units, B.unit_list, B.ctx->stream);
    upload(B.path_units, B.path_list, B.ctx->stream);
    upload(B.rows, rows, B.ctx->stream);
}

void launch_batch_value(eqhip_batch& B) {
    hipStream_t st = B.ctx->stream;
    if (!B.path_list.empty()) {  // the jobs in path mode: one launch whose registers cover all of them
        hipLaunchKernelGGL(path_kernel(1, B.path_nf, B.path_ns), dim3((unsigned)B.path_list.size()),
                           dim3(kPathThreads), 0, st, B.plans.as<const ValuePlan>(), B.path_units.as<const int4>());
        HIPCHK(hipGetLastError());
    }
    const int n = (int)B.unit_list.size();
    if (n == 0) return;
    if (B.persistent) {
        hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, st, B.counter.as<unsigned int>());
        const int cus = device_cus(B.ctx);
        const int blocks_per_cu = B.lds_bytes * 2 <= B.ctx->lds_limit ? 2 : 1;
        const int grid = std::min(n, cus * blocks_per_cu);
        hipLaunchKernelGGL(value_kernel_persistent, dim3(grid), dim3(kValueThreads), B.lds_bytes, st,
                           B.plans.as<const ValuePlan>(), B.units.as<const int4>(), n, B.counter.as<unsigned int>());
    } else {
        hipLaunchKernelGGL(value_kernel, dim3(n), dim3(kValueThreads), B.lds_bytes, st, B.plans.as<const ValuePlan>(),
                           B.units.as<const int4>());
    }
    HIPCHK(hipGetLastError());
}

void simulate_batch(eqhip_batch& B) {
    SimArgs s{};
    s.a = B.model.a;
    s.b = B.model.b;
    s.sigma = B.model.sigma;
    s.r0 = B.model.r0;
    s.dt = B.dt;
    s.sqrt_dt = std::sqrt(B.dt);
    s.decay = std::exp(-(B.model.a * B.dt));
    s.mean_c = B.model.b * (1.0 - s.decay);
    s.sd = B.model.sigma * std::sqrt((1.0 - std::exp(-(2.0 * B.model.a * B.dt))) / (2.0 * B.model.a));
    s.n_paths = B.max_paths;
    s.n_steps = B.det ? 0u : B.n_steps;
    s.seed = B.seed;
    s.slot_of_step = B.slot_of_step.as<int>();
    s.Z = nullptr;
    s.R = B.R.as<double>();
    hipStream_t st = B.ctx->stream;
    const dim3 pgrid((unsigned)((s.n_paths + kSimPaths - 1) / kSimPaths));
    if (B.disc == EQHIP_DISC_EULER) hipLaunchKernelGGL((simulate_philox_kernel<0>), pgrid, dim3(256), 0, st, s);
    else hipLaunchKernelGGL((simulate_philox_kernel<1>), pgrid, dim3(256), 0, st, s);
    HIPCHK(hipGetLastError());
}

// Times one valuation launch of the batch as laid out now (after a warm-up launch).
float time_batch_launch(eqhip_batch& B, hipEvent_t e0, hipEvent_t e1) {
    hipStream_t st = B.ctx->stream;
    launch_batch_value(B);
    HIPCHK(hipEventRecord(e0, st));
    launch_batch_value(B);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    return ms;
}

// Every job's variant: the most paths per lane up to `ppl` and `tpi` trades per item where they fit,
// else the job's default.
void set_batch_variants(eqhip_batch& B, int ppl, int tpi) {
    for (auto& J : B.jobs) {
        if (J->path_mode) continue;  // no variant
        // without a per-trade step TPI does not matter: take the one each PPL allows
        auto tpi_for = [&](int p) { return J->want_trade ? tpi : (p == 4 ? 2 : 4); };
        int p = ppl;
        while (p > 1 && !variant_fits(*J, p, tpi_for(p))) p /= 2;
        if (variant_fits(*J, p, tpi_for(p))) {
            J->ppl = p;
            J->tpi = tpi_for(p);
        } else {  // the job's default
            J->ppl = kDefaultPpl;
            J->tpi = (!J->want_trade || J->n_items_t[1] <= kTradeSlots) ? 4 : 8;
        }
    }
}

// Autotuner of a batch, on the batch's own simulated paths, in two stages: the launch shape (units per
// CU, persistent grid or one block per unit) with the default variant, then at that shape the
// variant (paths per lane, trades per item) every job takes where it fits. Every trial is logged.
void autotune_batch(eqhip_batch& B) {
    simulate_batch(B);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float best = 1e30f;
    uint64_t best_t = B.target_units_per_cu;
    int best_m = B.persistent, best_ppl = kDefaultPpl, best_tpi = kDefaultTpi;
    set_batch_variants(B, kDefaultPpl, kDefaultTpi);
    bool any_tile = false;  // a job valued by the tile kernels (the rest are in path mode)
    for (auto& J : B.jobs) any_tile |= !J->path_mode;
    for (uint64_t t : {4, 8, 16}) {
        B.target_units_per_cu = t;
        layout_batch(B);
        for (int m : {1, 0}) {
            if (m == 1 && !any_tile) continue;  // the persistent grid serves tile units only
            B.persistent = m;
            const float ms = time_batch_launch(B, e0, e1);
            B.tune_log.push_back("units/CU=" + std::to_string(t) + (m ? " persistent" : " per-unit") +
                                 " paths/lane=1 trades/item=4: " + std::to_string(ms) + " ms");
            if (ms < best) {
                best = ms;
                best_t = t;
                best_m = m;
            }
        }
    }
    B.target_units_per_cu = best_t;
    B.persistent = best_m;
    bool want_trade = false;
    for (auto& J : B.jobs) want_trade |= J->want_trade != 0;
    for (int ppl : {1, 2, 4})
        for (int tpi : {2, 4, 8}) {
            if (!any_tile) break;
            if (ppl == 4 && tpi != 2) continue;
            if ((ppl == kDefaultPpl && tpi == kDefaultTpi) || (!want_trade && tpi != (ppl == 4 ? 2 : 4))) continue;
            set_batch_variants(B, ppl, tpi);
            layout_batch(B);
            const float ms = time_batch_launch(B, e0, e1);
            B.tune_log.push_back("units/CU=" + std::to_string(best_t) + (best_m ? " persistent" : " per-unit") +
                                 " paths/lane<=" + std::to_string(ppl) + " trades/item=" + std::to_string(tpi) + ": " +
                                 std::to_string(ms) + " ms");
            if (ms < best) {
                best = ms;
                best_ppl = ppl;
                best_tpi = tpi;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_batch_variants(B, best_ppl, best_tpi);
    layout_batch(B);
    B.best_ppl = best_ppl;
    B.best_tpi = best_tpi;
    int n_path = 0;
    for (auto& J : B.jobs) n_path += J->path_mode ? 1 : 0;
    B.tune_log.push_back("chosen: " + std::string(best_m ? "persistent" : "per-unit") + ", units/CU=" +
                         std::to_string(best_t) + ", paths/lane<=" + std::to_string(best_ppl) + ", trades/item=" +
                         std::to_string(best_tpi) + ", path mode: " + std::to_string(n_path) + " of " +
                         std::to_string(B.jobs.size()) + " jobs");
    B.tuned = true;
}

void execute_batch(eqhip_batch& B, const eqhip_credit* credit, eqhip_outputs* outs) {
    if (!outs) raise(EQHIP_E_INVALID, "outputs are required");
    for (size_t j = 0; j < B.jobs.size(); ++j) {
        if (!outs[j].set_ee || !outs[j].set_pfe) raise(EQHIP_E_INVALID, "outputs set_ee / set_pfe are required");
        if ((outs[j].trade_ee || outs[j].trade_cva) && !B.jobs[j]->want_trade)
            raise(EQHIP_E_INVALID, "per-trade outputs need a batch created with want_trade_ee");
        if (outs[j].path_mtm || outs[j].set_mtm) raise(EQHIP_E_INVALID, "per-path outputs are not produced by batches");
        if (outs[j].trade_pfe) raise(EQHIP_E_INVALID, "per-trade PFE is not produced by batches");
    }
    if (!B.tuned) autotune_batch(B);
    hipStream_t st = B.ctx->stream;
    HIPCHK(hipSetDevice(B.ctx->device));
    HIPCHK(hipEventRecord(B.ev[0], st));
    simulate_batch(B);
    HIPCHK(hipEventRecord(B.ev[1], st));
    launch_batch_value(B);
    HIPCHK(hipEventRecord(B.ev[2], st));
    unsigned n_rows = 0;
    for (auto& J : B.jobs) n_rows += (unsigned)(J->S * J->n_dates);
    hipLaunchKernelGGL(select_kernel, dim3(n_rows), dim3(kSelectThreads), 0, st, B.rows.as<const SelectRow>());
    HIPCHK(hipGetLastError());
    for (auto& J : B.jobs) {
        if (!J->want_trade) continue;
        int n = J->n_trades * J->n_dates;
        hipLaunchKernelGGL(trade_ee_kernel, dim3((n + 255) / 256), dim3(256), 0, st, J->partial.as<double>(),
                           J->n_trades, J->n_dates, J->n_chunks, J->neff.as<uint64_t>(), J->trade_ee.as<double>());
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(B.ev[3], st));
    double h = credit ? credit->hazard_rate : 0.0, rr = credit ? credit->recovery_rate : 0.0;
    for (size_t j = 0; j < B.jobs.size(); ++j) {
        eqhip_plan& P = *B.jobs[j];
        if (outs[j].set_cva || outs[j].trade_cva) {
            reserve(P.cva_w, (size_t)P.n_dates * 16);
            launch_cva_weights(st, P.d_dates.as<double>(), P.n_dates, P.model, h, P.cva_w.as<double2>());
        }
        if (outs[j].set_cva)
            launch_cva_sum(st, P.set_ee.as<double>(), P.S, P.n_dates, P.cva_w.as<double2>(), rr, P.set_cva.as<double>());
        if (outs[j].trade_cva)
            launch_cva_sum(st, P.trade_ee.as<double>(), P.n_trades, P.n_dates, P.cva_w.as<double2>(), rr,
                           P.trade_cva.as<double>());
    }
    HIPCHK(hipEventRecord(B.ev[4], st));
    for (size_t j = 0; j < B.jobs.size(); ++j) {
        eqhip_plan& P = *B.jobs[j];
        auto d2h = [&](void* dst, const DevBuf& src, size_t bytes) {
            if (dst) HIPCHK(hipMemcpyAsync(dst, src.p, bytes, hipMemcpyDeviceToHost, st));
        };
        d2h(outs[j].set_ee, P.set_ee, (size_t)P.S * P.n_dates * 8);
        d2h(outs[j].set_pfe, P.set_pfe, (size_t)P.S * P.n_dates * 8);
        d2h(outs[j].set_cva, P.set_cva, (size_t)P.S * 8);
        d2h(outs[j].trade_ee, P.trade_ee, (size_t)P.n_trades * P.n_dates * 8);
        d2h(outs[j].trade_cva, P.trade_cva, (size_t)P.n_trades * 8);
    }
    HIPCHK(hipStreamSynchronize(st));
    float ms[4];
    for (int i = 0; i < 4; ++i) HIPCHK(hipEventElapsedTime(&ms[i], B.ev[i], B.ev[i + 1]));
    B.timing.simulate_ms = ms[0];
    B.timing.value_ms = ms[1];
    B.timing.reduce_ms = ms[2];
    B.timing.cva_ms = ms[3];
    B.timing.total_ms = (double)ms[0] + ms[1] + ms[2] + ms[3];
    B.timing.n_value_blocks = B.unit_list.size();
    B.timing.paths_per_tile = (uint32_t)(kPathsPerTile * B.best_ppl);
    B.timing.tiles_per_block = (uint32_t)B.target_units_per_cu;
    B.timing.trades_per_item = (uint32_t)B.best_tpi;
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
        size_t n = std::min(len - 1, last_error().size());
        std::memcpy(buf, last_error().data(), n);
        buf[n] = '\0';
    }
    return last_error().size();
}

// ---------------------------------------------------------------------------------------------
// Device memory of the allocation-free CVA entry point: the per-date weights (S_{k-1} - S_k, P(0,t_k))
// of up to kCvaScratchDates dates, in a file-scope device array instead of an allocation. Every GPU the
// library runs on holds its own copy (256 KiB) until the process ends. Host code never passes the
// array's name as an address (it is a host-side stand-in): each device's address of it is resolved
// once, with that device current, when the first context on the device is created, and kept here.
// ---------------------------------------------------------------------------------------------
constexpr int kCvaScratchDates = 16384;
static __device__ __align__(256) double2 scratch[kCvaScratchDates];

namespace {
std::mutex g_scratch_lock;
std::map<int, double2*> g_scratch_addr;  // device -> its copy of `scratch`

double2* device_scratch(int device) {  // the caller has made `device` current
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    const auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) return it->second;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        raise(EQHIP_E_HIP, std::string("hipGetSymbolAddress of the CVA scratch array failed: ") + hipGetErrorString(e));
    }
    g_scratch_addr.emplace(device, static_cast<double2*>(p));
    return static_cast<double2*>(p);
}

// The caller's pointer is memory the device can read and write (device or managed memory, or
// page-locked host memory mapped into the device).
bool device_accessible(const void* p) {
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
           (attr.type == hipMemoryTypeHost && attr.devicePointer != nullptr);
}
}  // namespace

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
            raise(EQHIP_E_NO_DEVICE, std::string("the library is compiled for gfx950, device is ") + prop.gcnArchName);
        auto ctx = std::make_unique<eqhip_context>();
        ctx->device = device;
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->cva_scratch = device_scratch(device);
        ctx->lds_limit = prop.sharedMemPerBlock;  // 160 KB per workgroup on gfx950
        if (hipFuncSetAttribute((const void*)value_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)ctx->lds_limit) != hipSuccess ||
            hipFuncSetAttribute((const void*)value2_kernel<4>, hipFuncAttributeMaxDynamicSharedMemorySize,
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
        P->device = ctx->device;
        build_plan(*P, model, portfolio, dates, n_dates, mc, want_trade_ee);
        *out = P.release();
    });
}

int eqhip_plan_create_2f(eqhip_context* ctx, const eqhip_model_hw2f* model, const eqhip_portfolio* portfolio,
                         const double* dates, int32_t n_dates, const eqhip_mc* mc, int32_t want_trade_ee,
                         eqhip_plan** out) {
    return guarded([&] {
        if (!out) raise(EQHIP_E_INVALID, "out is null");
        *out = nullptr;
        check_ctx(ctx);
        if (!model) raise(EQHIP_E_INVALID, "model is null");
        auto P = std::make_unique<eqhip_plan>();
        P->ctx = ctx;
        P->device = ctx->device;
        P->factors = 2;
        P->model2 = *model;
        build_plan(*P, nullptr, portfolio, dates, n_dates, mc, want_trade_ee);
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

int eqhip_plan_set_variant(eqhip_plan* plan, int32_t paths_per_lane, int32_t trades_per_item) {
    return guarded([&] {
        if (!plan) raise(EQHIP_E_INVALID, "plan is null");
        const int p = paths_per_lane, t = trades_per_item;
        if (p == 0) {  // path mode
            if (!plan->path_ok)
                raise(EQHIP_E_CAPACITY, "path mode serves plans without per-trade outputs, with at most " +
                                            std::to_string(kPathFamilies) + " schedule families and " +
                                            std::to_string(kPathSets) + " netting sets");
            HIPCHK(hipSetDevice(plan->ctx->device));
            plan->path_mode = true;
            choose_launch(*plan);
            build_value_plan(*plan);
            plan->tuned = true;
            return;
        }
        if ((p != 1 && p != 2 && p != 4) || (t != 2 && t != 4 && t != 8))
            raise(EQHIP_E_INVALID, "paths_per_lane must be 0 (path mode), 1, 2 or 4 and trades_per_item 2, 4 or 8");
        if (!plan->parts.empty()) raise(EQHIP_E_INVALID, "a plan valued in parts keeps its parts' variant");
        if (plan->want_pfe && p != 1)
            raise(EQHIP_E_INVALID, "a plan with per-trade PFE values one path per lane");
        if (!variant_fits(*plan, p, t))
            raise(EQHIP_E_CAPACITY, "variant paths_per_lane=" + std::to_string(p) + " trades_per_item=" +
                                        std::to_string(t) + " does not fit this plan (LDS, work items or registers)");
        HIPCHK(hipSetDevice(plan->ctx->device));
        plan->path_mode = false;
        plan->ppl = p;
        plan->tpi = t;
        choose_launch(*plan);
        build_value_plan(*plan);
        plan->tuned = true;  // an explicit choice is not tuned over
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
        (void)hipSetDevice(plan->device);
        delete plan;
    });
}

int eqhip_exposure_run(eqhip_context* ctx, const eqhip_model_hw1f* model, const eqhip_portfolio* portfolio,
                       const double* dates, int32_t n_dates, const eqhip_mc* mc, const double* host_draws,
                       const eqhip_credit* credit, eqhip_outputs* out) {
    return guarded([&] {
        check_ctx(ctx);
        std::lock_guard<std::mutex> lock(ctx->oneshot_lock);
        eqhip_plan P;
        P.ctx = ctx;
        P.device = ctx->device;
        P.direct_out = true;
        int flags = (out && (out->trade_ee || out->trade_cva || out->path_mtm)) ? EQHIP_PLAN_TRADE_EE : 0;
        if (out && out->trade_pfe) flags |= EQHIP_PLAN_TRADE_PFE;
        build_plan(P, model, portfolio, dates, n_dates, mc, flags);
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

int eqhip_scenarios_create(eqhip_context* ctx, const eqhip_model_hw1f* model, double t_max, const eqhip_mc* mc,
                           eqhip_scenarios** out) {
    return guarded([&] {
        if (!out) raise(EQHIP_E_INVALID, "out is null");
        *out = nullptr;
        check_ctx(ctx);
        if (!model || !mc) raise(EQHIP_E_INVALID, "model / mc is null");
        if (!(model->a != 0.0)) raise(EQHIP_E_INVALID, "HullWhite1F: a must be non-zero");
        if (!(t_max > 0.0) || !std::isfinite(t_max)) raise(EQHIP_E_INVALID, "t_max must be positive and finite");
        if (mc->n_paths < 1) raise(EQHIP_E_INVALID, "PricingContext: n_paths must be > 0");
        if (mc->n_steps < 1) raise(EQHIP_E_INVALID, "PricingContext: n_steps must be > 0");
        if (mc->rng != EQHIP_RNG_PHILOX) raise(EQHIP_E_INVALID, "a scenario set draws from the Philox stream");
        if (mc->disc != EQHIP_DISC_EULER && mc->disc != EQHIP_DISC_EXACT)
            raise(EQHIP_E_INVALID, "unknown discretisation");
        if (mc->n_paths > (uint64_t)1 << 40) raise(EQHIP_E_CAPACITY, "n_paths too large");
        const uint32_t steps = mc->n_steps;
        const double bytes = 8.0 * (double)(steps + 1) * (double)mc->n_paths;
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        if (bytes > 0.9 * (double)free_b)
            raise(EQHIP_E_CAPACITY, "scenario set needs " + std::to_string(bytes / 1e9) + " GB of device memory, " +
                                        std::to_string(free_b / 1e9) + " GB free");
        auto S = std::make_unique<eqhip_scenarios>();
        S->device = ctx->device;
        S->model = *model;
        S->mc = *mc;
        S->t_max = t_max;
        std::vector<int> slot(steps + 1);
        for (uint32_t g = 0; g <= steps; ++g) slot[g] = (int)g;
        upload(S->slot_of_step, slot, ctx->stream);
        S->R.alloc((size_t)bytes);
        // the grid of a plan whose last date is t_max: dt = t_max / n_steps (exposure.rs:91-116)
        launch_simulate1(ctx->stream, S->model, S->mc, t_max / (double)steps, steps, S->slot_of_step.as<int>(),
                         nullptr, S->R.as<double>());
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *out = S.release();
    });
}

int eqhip_scenarios_destroy(eqhip_scenarios* scenarios) {
    return guarded([&] {
        if (!scenarios) return;
        (void)hipSetDevice(scenarios->device);
        delete scenarios;
    });
}

int eqhip_irs_hull_white_exposure_profile_batch_on(
    eqhip_context* ctx, const eqhip_scenarios* scenarios, const double* notionals, const double* fixed_rates,
    int32_t n_trades, double start, const double* payment_times, const double* accruals, int32_t n_flows,
    const double* monitoring_times, int32_t n_dates, double* out_times, double* out_ee, double* out_pfe,
    int32_t* n_out) {
    return guarded([&] {
        if (!scenarios) raise(EQHIP_E_INVALID, "scenario set is null");
        if (n_trades < 1 || !notionals || !fixed_rates)
            raise(EQHIP_E_INVALID, "el lote requiere al menos un trade con notional y fixed_rate");
        const eqhip_model_hw1f& m = scenarios->model;
        const eqhip_mc& mc = scenarios->mc;
        profile_batch(ctx, m.a, m.b, m.sigma, m.r0, notionals, fixed_rates, n_trades, start, payment_times, accruals,
                      n_flows, monitoring_times, n_dates, mc.n_steps, mc.n_paths, mc.seed, mc.rng, mc.disc, nullptr,
                      1, out_times, out_ee, out_pfe, n_out, nullptr, scenarios);
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
        DevBuf dt, de, dc, dw;
        upload(dt, std::vector<double>(times, times + n_dates), st);
        upload(de, std::vector<double>(ee, ee + (size_t)n_profiles * n_dates), st);
        dc.alloc((size_t)n_profiles * 8);
        dw.alloc(std::max<size_t>(1, (size_t)n_dates) * 16);
        const eqhip_model_hw1f m{a, b, sigma, r0};
        if (n_dates > 0) launch_cva_weights(st, dt.as<double>(), n_dates, m, hazard_rate, dw.as<double2>());
        launch_cva_sum(st, de.as<double>(), n_profiles, n_dates, dw.as<double2>(), recovery_rate, dc.as<double>());
        HIPCHK(hipMemcpyAsync(out_cva, dc.p, (size_t)n_profiles * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    });
}

int eqhip_unilateral_cva_from_exposure_batch_noalloc(eqhip_context* ctx, double a, double b, double sigma, double r0,
                                                     const double* d_times, const double* d_ee, int32_t n_profiles,
                                                     int32_t n_dates, double hazard_rate, double recovery_rate,
                                                     double* d_out_cva, void* stream) {
    return guarded([&] {
        check_ctx(ctx);
        if (n_profiles < 0 || n_dates < 0) raise(EQHIP_E_INVALID, "invalid CVA arguments");
        if (n_dates > kCvaScratchDates)
            raise(EQHIP_E_CAPACITY, "the allocation-free CVA entry point serves at most " +
                                        std::to_string(kCvaScratchDates) + " dates (got " + std::to_string(n_dates) +
                                        "); use eqhip_unilateral_cva_from_exposure_batch");
        if (n_profiles == 0) return;
        if ((n_dates > 0 && (!device_accessible(d_times) || !device_accessible(d_ee))) || !device_accessible(d_out_cva))
            raise(EQHIP_E_INVALID, "times, ee and out_cva must be device memory of the context's GPU");
        int current = -1;
        HIPCHK(hipGetDevice(&current));
        if (current != ctx->device) HIPCHK(hipSetDevice(ctx->device));
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        const eqhip_model_hw1f m{a, b, sigma, r0};
        // every weight the sum reads is written first: no clearing needed
        if (n_dates > 0) launch_cva_weights(st, d_times, n_dates, m, hazard_rate, ctx->cva_scratch);
        launch_cva_sum(st, d_ee, n_profiles, n_dates, ctx->cva_scratch, recovery_rate, d_out_cva);
        if (!stream) HIPCHK(hipStreamSynchronize(st));
    });
}

int eqhip_unilateral_cva_from_exposure(eqhip_context* ctx, double a, double b, double sigma, double r0,
                                       const double* times, const double* ee, int32_t n_dates,
                                       double hazard_rate, double recovery_rate, double* out_cva) {
    return eqhip_unilateral_cva_from_exposure_batch(ctx, a, b, sigma, r0, times, ee, 1, n_dates, hazard_rate,
                                                    recovery_rate, out_cva);
}

int eqhip_irs_hull_white_2f_exposure_profile(
    eqhip_context* ctx, double a, double b, double sigma, double eta, double rho, double r0, double notional,
    double fixed_rate, int use_par_rate, double start, const double* payment_times, const double* accruals,
    int32_t n_flows, const double* monitoring_times, int32_t n_dates, uint32_t n_steps, uint64_t n_paths,
    uint64_t seed, int32_t rng, int32_t disc, const double* host_draws, int32_t allow_any_date, double* out_times,
    double* out_ee, double* out_pfe, int32_t* n_out) {
    return guarded([&] {
        if (n_flows < 1 || !payment_times || !accruals) raise(EQHIP_E_INVALID, "un swap necesita al menos un periodo");
        const eqhip_model_hw2f m{a, b, sigma, eta, rho, r0};
        validate_hw2f(m);
        double k = use_par_rate ? host_par_rate2(m, start, payment_times, accruals, n_flows) : fixed_rate;
        profile_batch(ctx, a, b, sigma, r0, &notional, &k, 1, start, payment_times, accruals, n_flows,
                      monitoring_times, n_dates, n_steps, n_paths, seed, rng, disc, host_draws, allow_any_date,
                      out_times, out_ee, out_pfe, n_out, &m);
    });
}

int eqhip_irs_hull_white_2f_exposure_profile_batch(
    eqhip_context* ctx, double a, double b, double sigma, double eta, double rho, double r0, const double* notionals,
    const double* fixed_rates, int32_t n_trades, double start, const double* payment_times, const double* accruals,
    int32_t n_flows, const double* monitoring_times, int32_t n_dates, uint32_t n_steps, uint64_t n_paths,
    uint64_t seed, int32_t rng, int32_t disc, const double* host_draws, int32_t allow_any_date, double* out_times,
    double* out_ee, double* out_pfe, int32_t* n_out) {
    return guarded([&] {
        if (n_trades < 1 || !notionals || !fixed_rates)
            raise(EQHIP_E_INVALID, "el lote requiere al menos un trade con notional y fixed_rate");
        const eqhip_model_hw2f m{a, b, sigma, eta, rho, r0};
        profile_batch(ctx, a, b, sigma, r0, notionals, fixed_rates, n_trades, start, payment_times, accruals, n_flows,
                      monitoring_times, n_dates, n_steps, n_paths, seed, rng, disc, host_draws, allow_any_date,
                      out_times, out_ee, out_pfe, n_out, &m);
    });
}

int eqhip_unilateral_cva_from_exposure_2f_batch(eqhip_context* ctx, double a, double b, double sigma, double eta,
                                                double rho, double r0, const double* times, const double* ee,
                                                int32_t n_profiles, int32_t n_dates, double hazard_rate,
                                                double recovery_rate, double* out_cva) {
    return guarded([&] {
        check_ctx(ctx);
        if (n_profiles < 0 || n_dates < 0 || (n_dates > 0 && (!times || !ee)) || !out_cva)
            raise(EQHIP_E_INVALID, "invalid CVA arguments");
        const eqhip_model_hw2f m{a, b, sigma, eta, rho, r0};
        validate_hw2f(m);
        if (n_profiles == 0) return;
        hipStream_t st = ctx->stream;
        DevBuf dt, de, dc, dw;
        upload(dt, std::vector<double>(times, times + n_dates), st);
        upload(de, std::vector<double>(ee, ee + (size_t)n_profiles * n_dates), st);
        dc.alloc((size_t)n_profiles * 8);
        dw.alloc(std::max<size_t>(1, (size_t)n_dates) * 16);
        if (n_dates > 0) launch_cva_weights2(st, dt.as<double>(), n_dates, m, hazard_rate, dw.as<double2>());
        launch_cva_sum(st, de.as<double>(), n_profiles, n_dates, dw.as<double2>(), recovery_rate, dc.as<double>());
        HIPCHK(hipMemcpyAsync(out_cva, dc.p, (size_t)n_profiles * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    });
}

int eqhip_unilateral_cva_from_exposure_2f(eqhip_context* ctx, double a, double b, double sigma, double eta,
                                          double rho, double r0, const double* times, const double* ee,
                                          int32_t n_dates, double hazard_rate, double recovery_rate,
                                          double* out_cva) {
    return eqhip_unilateral_cva_from_exposure_2f_batch(ctx, a, b, sigma, eta, rho, r0, times, ee, 1, n_dates,
                                                       hazard_rate, recovery_rate, out_cva);
}

int eqhip_xva_calculate(eqhip_context* ctx, const eqhip_xva_params* params, const double* times, int32_t n_times,
                        uint64_t n_paths, const double* q_paths, const double* discount_factors,
                        const double* p_paths, double* out_ee, double* out_pfe, double* out_measures) {
    return guarded([&] {
        check_ctx(ctx);
        validate_xva_params(params);
        if (!out_ee || !out_pfe || !out_measures) raise(EQHIP_E_INVALID, "out_ee / out_pfe / out_measures are required");
        // ExposureCubeInput::validate, quant-domain/src/xva.rs:187-220
        if (n_times < 1 || !times || !q_paths)
            raise(EQHIP_E_INVALID, "exposure times and rows must be non-empty and have equal length");
        for (int i = 0; i < n_times; ++i)
            if (!std::isfinite(times[i]) || (i > 0 && !(times[i] > times[i - 1])) || (i == 0 && times[0] < 0.0))
                raise(EQHIP_E_INVALID, "exposure times must be finite and strictly increasing");
        if (n_paths < 1) raise(EQHIP_E_INVALID, "exposure paths must be rectangular finite rows");
        const size_t cells = (size_t)n_times * n_paths;
        for (const double* c : {q_paths, p_paths})
            if (c)
                for (size_t i = 0; i < cells; ++i)
                    if (!std::isfinite(c[i])) raise(EQHIP_E_INVALID, "exposure paths must be rectangular finite rows");
        if (discount_factors)
            for (int i = 0; i < n_times; ++i)
                if (!std::isfinite(discount_factors[i]) || discount_factors[i] < 0.0)
                    raise(EQHIP_E_INVALID, "discount_factors must match times and be non-negative");
        hipStream_t st = ctx->stream;
        DevBuf dq, dp, pos;
        dq.alloc(cells * 8);
        HIPCHK(hipMemcpyAsync(dq.p, q_paths, cells * 8, hipMemcpyHostToDevice, st));
        if (p_paths) {
            dp.alloc(cells * 8);
            HIPCHK(hipMemcpyAsync(dp.p, p_paths, cells * 8, hipMemcpyHostToDevice, st));
        }
        XvaRows rows{dq.as<double>(), n_paths, dp.as<double>(), n_paths, n_paths, n_times};
        xva_reduce(ctx, rows, pos, *params, times, discount_factors, out_ee, out_pfe, out_measures);
    });
}

int eqhip_plan_exposure_cube(eqhip_plan* plan, int32_t set, double* out_paths, double* out_discount) {
    return guarded([&] {
        if (!plan) raise(EQHIP_E_INVALID, "plan is null");
        eqhip_plan& P = *plan;
        if (!P.keep_cube) raise(EQHIP_E_INVALID, "the plan was created without EQHIP_PLAN_KEEP_CUBE");
        if (!P.executed) raise(EQHIP_E_INVALID, "the plan has not been executed");
        if (set < 0 || set >= P.S) raise(EQHIP_E_INVALID, "set index out of range");
        if (!out_paths) raise(EQHIP_E_INVALID, "out_paths is null");
        hipStream_t st = P.ctx->stream;
        HIPCHK(hipSetDevice(P.ctx->device));
        const uint64_t np = P.mc.n_paths;
        // rows (date k, set) of the [date][set][path] cube, one strided copy
        HIPCHK(hipMemcpy2DAsync(out_paths, np * 8, P.cube.as<double>() + (size_t)set * np, (size_t)P.S * np * 8, np * 8,
                                P.n_dates, hipMemcpyDeviceToHost, st));
        if (out_discount) {
            DevBuf w;
            w.alloc((size_t)P.n_dates * 16);
            if (P.factors == 2)
                launch_cva_weights2(st, P.d_dates.as<double>(), P.n_dates, P.model2, 0.0, w.as<double2>());
            else
                launch_cva_weights(st, P.d_dates.as<double>(), P.n_dates, P.model, 0.0, w.as<double2>());
            std::vector<double2> hw(P.n_dates);
            HIPCHK(hipMemcpyAsync(hw.data(), w.p, (size_t)P.n_dates * 16, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (int k = 0; k < P.n_dates; ++k) out_discount[k] = hw[k].y;
        }
        HIPCHK(hipStreamSynchronize(st));
    });
}

int eqhip_plan_xva(eqhip_plan* plan, int32_t set, const eqhip_xva_params* params, double* out_ee, double* out_pfe,
                   double* out_measures) {
    return guarded([&] {
        if (!plan) raise(EQHIP_E_INVALID, "plan is null");
        eqhip_plan& P = *plan;
        if (!P.keep_cube) raise(EQHIP_E_INVALID, "the plan was created without EQHIP_PLAN_KEEP_CUBE");
        if (!P.executed) raise(EQHIP_E_INVALID, "the plan has not been executed");
        if (set < 0 || set >= P.S) raise(EQHIP_E_INVALID, "set index out of range");
        validate_xva_params(params);
        if (!out_ee || !out_pfe || !out_measures) raise(EQHIP_E_INVALID, "out_ee / out_pfe / out_measures are required");
        for (int k = 0; k < P.n_dates; ++k)
            if (!(P.dates[k] >= 0.0) || (k > 0 && !(P.dates[k] > P.dates[k - 1])))
                raise(EQHIP_E_INVALID, "exposure times must be finite and strictly increasing");
        hipStream_t st = P.ctx->stream;
        HIPCHK(hipSetDevice(P.ctx->device));
        const uint64_t np = P.mc.n_paths;
        // the model discount P(0, t_k) of the plan's own CVA
        reserve(P.cva_w, (size_t)P.n_dates * 16);
        if (P.factors == 2)
            launch_cva_weights2(st, P.d_dates.as<double>(), P.n_dates, P.model2, 0.0, P.cva_w.as<double2>());
        else
            launch_cva_weights(st, P.d_dates.as<double>(), P.n_dates, P.model, 0.0, P.cva_w.as<double2>());
        std::vector<double2> hw(P.n_dates);
        HIPCHK(hipMemcpyAsync(hw.data(), P.cva_w.p, (size_t)P.n_dates * 16, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        std::vector<double> disc(P.n_dates);
        for (int k = 0; k < P.n_dates; ++k) disc[k] = hw[k].y;
        XvaRows rows{P.cube.as<double>() + (size_t)set * np, (uint64_t)P.S * np, nullptr, 0, np, P.n_dates};
        xva_reduce(P.ctx, rows, P.xva_pos, *params, P.dates.data(), disc.data(), out_ee, out_pfe, out_measures, P.ev,
                   &P.timing.xva_rows_ms, &P.timing.xva_select_ms);
    });
}

int eqhip_bond_exp(eqhip_context* ctx, const double* x, double* y, int32_t n) {
    return guarded([&] {
        check_ctx(ctx);
        if (n < 0 || (n > 0 && (!x || !y))) raise(EQHIP_E_INVALID, "x and y must hold n values");
        if (n == 0) return;
        HIPCHK(hipSetDevice(ctx->device));
        DevBuf buf;
        buf.alloc(2 * (size_t)n * sizeof(double));
        double* dx = buf.as<double>();
        double* dy = dx + n;
        hipStream_t st = ctx->stream;
        HIPCHK(hipMemcpyAsync(dx, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(bond_exp_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dx, dy, (int)n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(y, dy, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    });
}

int eqhip_pv_dv01(eqhip_context* ctx, const double* pillars, const double* zero_rates, int32_t n_pillars,
                  const eqhip_portfolio* pf, double bump, double* out_pv, double* out_dv01, double* out_bucketed) {
    return eqhip_pv_dv01_at(ctx, pillars, zero_rates, n_pillars, pf, bump, 0.0, out_pv, out_dv01, out_bucketed);
}

int eqhip_pv_dv01_at(eqhip_context* ctx, const double* pillars, const double* zero_rates, int32_t n_pillars,
                     const eqhip_portfolio* pf, double bump, double valuation_time, double* out_pv, double* out_dv01,
                     double* out_bucketed) {
    return guarded([&] {
        check_ctx(ctx);
        validate_portfolio(pf);
        if (!std::isfinite(valuation_time)) raise(EQHIP_E_INVALID, "valuation_time must be finite");
        for (int j = 0; j < pf->n_trades; ++j)
            if (valuation_time > pf->start[j])  // measure.cpp:166-174
                raise(EQHIP_E_INVALID, "npv_from_market: valuation_time (" + std::to_string(valuation_time) +
                                           ") posterior al start (" + std::to_string(pf->start[j]) +
                                           ") del swap -- el primer reset de la pata flotante ya habria ocurrido y "
                                           "este motor no modela fixings historicos de esa pata");
        if (n_pillars < 1 || !pillars || !zero_rates) raise(EQHIP_E_INVALID, "Curve: necesita al menos un pillar");
        for (int i = 1; i < n_pillars; ++i)
            if (!(pillars[i] > pillars[i - 1])) raise(EQHIP_E_INVALID, "Curve: pillars debe ser estrictamente creciente");
        if (!out_pv || !out_dv01) raise(EQHIP_E_INVALID, "out_pv / out_dv01 are required");
        hipStream_t st = ctx->stream;
        HIPCHK(hipSetDevice(ctx->device));
        const int n = pf->n_trades, nf = pf->flow_offset[n], np = n_pillars;
        // inputs, then outputs, packed at 256-byte offsets: one upload and one download
        std::lock_guard<std::mutex> lock(ctx->staging_lock);
        size_t off = 0;
        auto place = [&](size_t bytes) {
            const size_t o = off;
            off += (bytes + 255) & ~(size_t)255;
            return o;
        };
        const size_t o_p = place((size_t)np * 8), o_z = place((size_t)np * 8), o_off = place((size_t)(n + 1) * 4),
                     o_pay = place((size_t)nf * 8), o_acc = place((size_t)nf * 8), o_n = place((size_t)n * 8),
                     o_k = place((size_t)n * 8), o_s = place((size_t)n * 8);
        const size_t in_bytes = off;
        const size_t o_pv = place((size_t)n * 8), o_dv = place((size_t)n * 8),
                     o_bk = place(out_bucketed ? (size_t)n * np * 8 : 0);
        const size_t out_bytes = off - in_bytes;
        reserve(ctx->host_stage, off);
        reserve(ctx->dev_stage, off);
        char* h = ctx->host_stage.as<char>();
        char* d = ctx->dev_stage.as<char>();
        std::memcpy(h + o_p, pillars, (size_t)np * 8);
        std::memcpy(h + o_z, zero_rates, (size_t)np * 8);
        std::memcpy(h + o_off, pf->flow_offset, (size_t)(n + 1) * 4);
        std::memcpy(h + o_pay, pf->payment_times, (size_t)nf * 8);
        std::memcpy(h + o_acc, pf->accruals, (size_t)nf * 8);
        std::memcpy(h + o_n, pf->notional, (size_t)n * 8);
        std::memcpy(h + o_k, pf->fixed_rate, (size_t)n * 8);
        std::memcpy(h + o_s, pf->start, (size_t)n * 8);
        HIPCHK(hipMemcpyAsync(d, h, in_bytes, hipMemcpyHostToDevice, st));
        PvArgs g{};
        g.pillars = (const double*)(d + o_p);
        g.zeros = (const double*)(d + o_z);
        g.np = np;
        g.n_trades = n;
        g.flow_off = (const int*)(d + o_off);
        g.pay = (const double*)(d + o_pay);
        g.acc = (const double*)(d + o_acc);
        g.notional = (const double*)(d + o_n);
        g.fixed = (const double*)(d + o_k);
        g.start = (const double*)(d + o_s);
        g.bump = bump;
        g.valuation_time = valuation_time;
        g.pv = (double*)(d + o_pv);
        g.dv01 = (double*)(d + o_dv);
        g.bucketed = out_bucketed ? (double*)(d + o_bk) : nullptr;
        g.bucketed_on = out_bucketed ? 1 : 0;
        int total = n * (2 + np);
        hipLaunchKernelGGL(pv_dv01_kernel, dim3((total + 127) / 128), dim3(128), 0, st, g);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h + in_bytes, d + in_bytes, out_bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        std::memcpy(out_pv, h + o_pv, (size_t)n * 8);
        std::memcpy(out_dv01, h + o_dv, (size_t)n * 8);
        if (out_bucketed) std::memcpy(out_bucketed, h + o_bk, (size_t)n * np * 8);
    });
}

int eqhip_pv_sensitivities(eqhip_context* ctx, const double* pillars, const double* zero_rates, int32_t n_pillars,
                           const eqhip_portfolio* pf, const eqhip_model_hw1f* model, double* out_pillar_delta,
                           double* out_model_npv, double* out_model_greeks) {
    return guarded([&] {
        check_ctx(ctx);
        validate_portfolio(pf);
        if (n_pillars < 1 || !pillars || !zero_rates) raise(EQHIP_E_INVALID, "Curve: necesita al menos un pillar");
        for (int i = 1; i < n_pillars; ++i)
            if (!(pillars[i] > pillars[i - 1])) raise(EQHIP_E_INVALID, "Curve: pillars debe ser estrictamente creciente");
        if (!out_pillar_delta) raise(EQHIP_E_INVALID, "out_pillar_delta is required");
        if (model && (!out_model_npv || !out_model_greeks))
            raise(EQHIP_E_INVALID, "out_model_npv / out_model_greeks are required with a model");
        if (model && !(model->a > 0.0)) raise(EQHIP_E_INVALID, "HW1F needs a > 0");
        hipStream_t st = ctx->stream;
        const int n = pf->n_trades, nf = pf->flow_offset[n];
        DevBuf dp, dz, doff, dpay, dacc, dn, dk, ds, dd, dm, dg;
        upload(dp, std::vector<double>(pillars, pillars + n_pillars), st);
        upload(dz, std::vector<double>(zero_rates, zero_rates + n_pillars), st);
        upload(doff, std::vector<int>(pf->flow_offset, pf->flow_offset + n + 1), st);
        upload(dpay, std::vector<double>(pf->payment_times, pf->payment_times + nf), st);
        upload(dacc, std::vector<double>(pf->accruals, pf->accruals + nf), st);
        upload(dn, std::vector<double>(pf->notional, pf->notional + n), st);
        upload(dk, std::vector<double>(pf->fixed_rate, pf->fixed_rate + n), st);
        upload(ds, std::vector<double>(pf->start, pf->start + n), st);
        dd.alloc((size_t)n * n_pillars * 8);
        dm.alloc((size_t)n * 8);
        dg.alloc((size_t)n * 4 * 8);
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
        const eqhip_model_hw1f mm = model ? *model : eqhip_model_hw1f{1.0, 0.0, 0.0, 0.0};
        hipLaunchKernelGGL(pv_sens_kernel, dim3((n + 127) / 128), dim3(128), 0, st, g, mm.a, mm.b, mm.sigma, mm.r0,
                           model ? 1 : 0, dd.as<double>(), dm.as<double>(), dg.as<double>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out_pillar_delta, dd.p, (size_t)n * n_pillars * 8, hipMemcpyDeviceToHost, st));
        if (model) {
            HIPCHK(hipMemcpyAsync(out_model_npv, dm.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(out_model_greeks, dg.p, (size_t)n * 32, hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipStreamSynchronize(st));
    });
}

int eqhip_batch_create(eqhip_context* ctx, const eqhip_model_hw1f* model, int32_t n_jobs,
                       const eqhip_portfolio* portfolios, const uint64_t* n_paths, const double* dates,
                       int32_t n_dates, uint32_t n_steps, uint64_t seed, int32_t rng, int32_t disc,
                       int32_t want_trade_ee, eqhip_batch** out) {
    return guarded([&] {
        if (!out) raise(EQHIP_E_INVALID, "out is null");
        *out = nullptr;
        check_ctx(ctx);
        if (n_jobs < 1 || !portfolios || !n_paths) raise(EQHIP_E_INVALID, "a batch needs at least one job");
        if (rng != EQHIP_RNG_PHILOX) raise(EQHIP_E_INVALID, "batches draw from the Philox stream (rng = EQHIP_RNG_PHILOX)");
        auto B = std::make_unique<eqhip_batch>();
        B->ctx = ctx;
        B->device = ctx->device;
        B->model = model ? *model : eqhip_model_hw1f{};
        B->n_steps = n_steps;
        B->seed = seed;
        B->rng = rng;
        B->disc = disc;
        for (int j = 0; j < n_jobs; ++j) B->max_paths = std::max<uint64_t>(B->max_paths, n_paths[j]);
        for (int j = 0; j < n_jobs; ++j) {
            auto P = std::make_unique<eqhip_plan>();
            P->ctx = ctx;
            P->device = ctx->device;
            eqhip_mc mc{n_paths[j], n_steps, seed, rng, disc};
            build_plan(*P, model, &portfolios[j], dates, n_dates, &mc, want_trade_ee, true);
            B->det = P->det;
            B->dt = P->dt;
            B->jobs.push_back(std::move(P));
        }
        const uint32_t steps = B->det ? 0u : n_steps;
        std::vector<int> slot_of_step(steps + 1);
        for (uint32_t g = 0; g <= steps; ++g) slot_of_step[g] = (int)g;
        upload(B->slot_of_step, slot_of_step, ctx->stream);
        B->R.alloc((size_t)(steps + 1) * B->max_paths * 8);
        B->counter.alloc(sizeof(unsigned int));
        for (auto& e : B->ev) HIPCHK(hipEventCreate(&e));
        layout_batch(*B);
        *out = B.release();
    });
}

int eqhip_batch_execute(eqhip_batch* batch, const eqhip_credit* credit, eqhip_outputs* outs) {
    return guarded([&] {
        if (!batch) raise(EQHIP_E_INVALID, "batch is null");
        execute_batch(*batch, credit, outs);
    });
}

int eqhip_batch_timing(const eqhip_batch* batch, eqhip_timing* out, char* tune_log, size_t len) {
    return guarded([&] {
        if (!batch || !out) raise(EQHIP_E_INVALID, "null argument");
        *out = batch->timing;
        if (tune_log && len) {
            std::string log = std::string(batch->persistent ? "chosen: persistent" : "chosen: per-unit") +
                              ", units/CU=" + std::to_string(batch->target_units_per_cu);
            for (auto& l : batch->tune_log) log += "\n" + l;
            size_t n = std::min(len - 1, log.size());
            std::memcpy(tune_log, log.data(), n);
            tune_log[n] = '\0';
        }
    });
}

int eqhip_batch_destroy(eqhip_batch* batch) {
    return guarded([&] {
        if (!batch) return;
        (void)hipSetDevice(batch->device);
        delete batch;
    });
}

int eqhip_cva_adjoint(eqhip_plan* plan, const double* host_draws, int32_t set, const eqhip_credit* credit,
                      double* out_cva, double* out_grad, double* out_check) {
    return guarded([&] {
        if (!plan) raise(EQHIP_E_INVALID, "plan is null");
        cva_adjoint(*plan, host_draws, set, credit, out_cva, out_grad, out_check);
    });
}

}  // extern "C"
