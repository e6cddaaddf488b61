// Measures of the exposure path and their GPU arm (cpp/engine/src/measure.cpp:20-462,
// cpp/engine/src/engine.cpp:73-113). Every numeric result comes from libeqhip (include/eqhip.h).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <exception>
#include <map>
#include <mutex>
#include <thread>

#include "eqhip.h"
#include "engine.hpp"

namespace engine {

namespace {

[[noreturn]] void throw_eqhip(int rc) {
    char buf[2048];
    eqhip_last_error(buf, sizeof(buf));
    std::string msg = std::string("eqhip: ") + buf;
    if (rc == EQHIP_E_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

void check(int rc) {
    if (rc != EQHIP_OK) throw_eqhip(rc);
}

// One eqhip context (one HIP stream) per host thread, created on first use.
eqhip_context* gpu_context() {
    struct Holder {
        eqhip_context* ctx = nullptr;
        ~Holder() {
            if (ctx) eqhip_context_destroy(ctx);
        }
    };
    thread_local Holder holder;
    if (!holder.ctx) check(eqhip_context_create(0, &holder.ctx));
    return holder.ctx;
}

// Worker threads for run_on_gpu_streams. They live until the process ends (the pool is never
// destroyed, so no thread is joined at exit), and each keeps its gpu_context() and the context's
// pooled buffers between calls.
class StreamWorkers {
  public:
    explicit StreamWorkers(int n) {
        for (int i = 0; i < n; ++i) std::thread([this] { loop(); }).detach();
        n_workers_ = n;
    }

    void run(const std::function<void()>& drain) {
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = &drain;
            running_ = n_workers_;
            ++generation_;
        }
        wake_.notify_all();
        drain();
        std::unique_lock<std::mutex> g(m_);
        done_.wait(g, [&] { return running_ == 0; });
        job_ = nullptr;
    }

  private:
    void loop() {
        std::uint64_t seen = 0;
        for (;;) {
            const std::function<void()>* job;
            {
                std::unique_lock<std::mutex> g(m_);
                wake_.wait(g, [&] { return generation_ != seen; });
                seen = generation_;
                job = job_;
            }
            (*job)();
            std::lock_guard<std::mutex> g(m_);
            if (--running_ == 0) done_.notify_all();
        }
    }

    std::mutex m_;
    std::condition_variable wake_, done_;
    const std::function<void()>* job_ = nullptr;
    std::uint64_t generation_ = 0;
    int running_ = 0;
    int n_workers_ = 0;
};

// backend "cpu" is the reference's engine-core; this library is its GPU arm only (api.rs:151-159).
void require_gpu(const ExecutionContext& execution) {
    if (execution.backend() != "gpu")
        throw std::invalid_argument(
            "ExecutionContext backend '" + execution.backend() +
            "' is served by the reference engine (engine-core); this library implements backend 'gpu'");
}

int disc_of(const ExecutionContext& execution) {
    return execution.discretisation() == "exact" ? EQHIP_DISC_EXACT : EQHIP_DISC_EULER;
}

const HullWhite1FModel& require_hw1f(const IModel& model, const char* who) {
    const auto* hw1 = dynamic_cast<const HullWhite1FModel*>(&model);
    if (!hw1) throw std::invalid_argument(std::string(who) + ": modelo no soportado: " + model.type_name());
    return *hw1;
}

const IrSwapProduct& require_irs(const IProduct& product, const char* who) {
    const auto* irs = dynamic_cast<const IrSwapProduct*>(&product);
    if (!irs) throw std::invalid_argument(std::string(who) + ": producto no soportado: " + product.type_name());
    return *irs;
}

double par_rate_from_market(const MarketSnapshot& market, const IrSwapProduct& p) {
    // measure.cpp:124-134
    double numerator = market.discount_factor(p.start()) - market.discount_factor(p.payment_times().back());
    double denominator = 0.0;
    for (std::size_t i = 0; i < p.payment_times().size(); ++i)
        denominator += p.accruals()[i] * market.discount_factor(p.payment_times()[i]);
    return numerator / denominator;
}

double effective_fixed_rate(const MarketSnapshot& market, const IrSwapProduct& p) {
    return p.use_par_rate() ? par_rate_from_market(market, p) : p.fixed_rate();
}

// measure.cpp:36-65: the model's own exposure wrapper
ExposureProfile compute_exposure_profile(const IModel& model, const IrSwapProduct& irs, const PricingContext& pricing,
                                         const ExecutionContext& execution) {
    if (const auto* hw2 = dynamic_cast<const HullWhite2FModel*>(&model))
        return irs_hull_white_2f_exposure_profile(execution, hw2->a(), hw2->b(), hw2->sigma(), hw2->eta(), hw2->rho(),
                                                  hw2->r0(), irs.notional(), irs.fixed_rate(), irs.use_par_rate(),
                                                  irs.start(), irs.payment_times(), irs.accruals(), reset_dates(irs),
                                                  pricing.n_steps(), pricing.n_paths(), pricing.seed());
    const auto& hw1 = require_hw1f(model, "ExposureProfileMeasure");
    return irs_hull_white_exposure_profile(execution, hw1.a(), hw1.b(), hw1.sigma(), hw1.r0(), irs.notional(),
                                           irs.fixed_rate(), irs.use_par_rate(), irs.start(), irs.payment_times(),
                                           irs.accruals(), reset_dates(irs), pricing.n_steps(), pricing.n_paths(),
                                           pricing.seed());
}

}  // namespace

std::vector<double> reset_dates(const IrSwapProduct& product) {
    std::vector<double> dates;
    dates.reserve(product.payment_times().size());
    dates.push_back(product.start());
    for (std::size_t i = 0; i + 1 < product.payment_times().size(); ++i) dates.push_back(product.payment_times()[i]);
    return dates;
}

// ---- GPU arm of the FFI wrappers ---------------------------------------------------------------

ExposureProfile irs_hull_white_exposure_profile(const ExecutionContext& execution, double a, double b, double sigma,
                                                double r0, double notional, double fixed_rate, bool use_par_rate,
                                                double start, const std::vector<double>& payment_times,
                                                const std::vector<double>& accruals,
                                                const std::vector<double>& monitoring_times, std::size_t n_steps,
                                                std::uint64_t n_paths, std::uint64_t seed) {
    require_gpu(execution);
    std::size_t nd = monitoring_times.size();
    ExposureProfile out;
    out.times.resize(nd);
    out.ee.resize(nd);
    out.pfe_95.resize(nd);
    int32_t n_out = 0;
    check(eqhip_irs_hull_white_exposure_profile(
        gpu_context(), a, b, sigma, r0, notional, fixed_rate, use_par_rate ? 1 : 0, start, payment_times.data(),
        accruals.data(), (int32_t)payment_times.size(), monitoring_times.data(), (int32_t)nd, (uint32_t)n_steps,
        n_paths, seed, EQHIP_RNG_PHILOX, disc_of(execution), nullptr, 0, out.times.data(), out.ee.data(),
        out.pfe_95.data(), &n_out));
    out.times.resize(n_out);
    out.ee.resize(n_out);
    out.pfe_95.resize(n_out);
    return out;
}

std::vector<ExposureProfile> irs_hull_white_exposure_profile_batch(
    const ExecutionContext& execution, double a, double b, double sigma, double r0,
    const std::vector<double>& notionals, const std::vector<double>& fixed_rates, double start,
    const std::vector<double>& payment_times, const std::vector<double>& accruals,
    const std::vector<double>& monitoring_times, std::size_t n_steps, std::uint64_t n_paths, std::uint64_t seed) {
    require_gpu(execution);
    if (notionals.size() != fixed_rates.size())
        throw std::invalid_argument("el lote requiere un fixed_rate por notional");
    std::size_t nt = notionals.size(), nd = monitoring_times.size();
    std::vector<double> times(nd), ee(nt * nd), pfe(nt * nd);
    int32_t n_out = 0;
    check(eqhip_irs_hull_white_exposure_profile_batch(
        gpu_context(), a, b, sigma, r0, notionals.data(), fixed_rates.data(), (int32_t)nt, start,
        payment_times.data(), accruals.data(), (int32_t)payment_times.size(), monitoring_times.data(), (int32_t)nd,
        (uint32_t)n_steps, n_paths, seed, EQHIP_RNG_PHILOX, disc_of(execution), nullptr, 0, times.data(), ee.data(),
        pfe.data(), &n_out));
    std::vector<ExposureProfile> out(nt);
    for (std::size_t j = 0; j < nt; ++j) {
        out[j].times.assign(times.begin(), times.begin() + n_out);
        out[j].ee.assign(ee.begin() + j * nd, ee.begin() + j * nd + n_out);
        out[j].pfe_95.assign(pfe.begin() + j * nd, pfe.begin() + j * nd + n_out);
    }
    return out;
}

double unilateral_cva_from_exposure(const ExecutionContext& execution, double a, double b, double sigma, double r0,
                                    const std::vector<double>& times, const std::vector<double>& ee,
                                    double hazard_rate, double recovery_rate) {
    require_gpu(execution);
    if (times.size() != ee.size()) throw std::invalid_argument("times y ee deben tener el mismo largo");
    double cva = 0.0;
    check(eqhip_unilateral_cva_from_exposure(gpu_context(), a, b, sigma, r0, times.data(), ee.data(),
                                             (int32_t)times.size(), hazard_rate, recovery_rate, &cva));
    return cva;
}

std::vector<double> unilateral_cva_from_exposure_batch(const ExecutionContext& execution, double a, double b,
                                                       double sigma, double r0,
                                                       const std::vector<ExposureProfile>& profiles,
                                                       double hazard_rate, double recovery_rate) {
    require_gpu(execution);
    std::vector<double> out(profiles.size());
    if (profiles.empty()) return out;
    const std::vector<double>& times = profiles.front().times;
    std::vector<double> ee;
    ee.reserve(profiles.size() * times.size());
    for (const auto& p : profiles) {
        if (p.times != times || p.ee.size() != times.size())
            throw std::invalid_argument("unilateral_cva_from_exposure_batch: profiles must share their times");
        ee.insert(ee.end(), p.ee.begin(), p.ee.end());
    }
    check(eqhip_unilateral_cva_from_exposure_batch(gpu_context(), a, b, sigma, r0, times.data(), ee.data(),
                                                   (int32_t)profiles.size(), (int32_t)times.size(), hazard_rate,
                                                   recovery_rate, out.data()));
    return out;
}

ExposureProfile irs_hull_white_2f_exposure_profile(const ExecutionContext& execution, double a, double b,
                                                   double sigma, double eta, double rho, double r0, double notional,
                                                   double fixed_rate, bool use_par_rate, double start,
                                                   const std::vector<double>& payment_times,
                                                   const std::vector<double>& accruals,
                                                   const std::vector<double>& monitoring_times, std::size_t n_steps,
                                                   std::uint64_t n_paths, std::uint64_t seed) {
    require_gpu(execution);
    std::size_t nd = monitoring_times.size();
    ExposureProfile out;
    out.times.resize(nd);
    out.ee.resize(nd);
    out.pfe_95.resize(nd);
    int32_t n_out = 0;
    check(eqhip_irs_hull_white_2f_exposure_profile(
        gpu_context(), a, b, sigma, eta, rho, r0, notional, fixed_rate, use_par_rate ? 1 : 0, start,
        payment_times.data(), accruals.data(), (int32_t)payment_times.size(), monitoring_times.data(), (int32_t)nd,
        (uint32_t)n_steps, n_paths, seed, EQHIP_RNG_PHILOX, disc_of(execution), nullptr, 0, out.times.data(),
        out.ee.data(), out.pfe_95.data(), &n_out));
    out.times.resize(n_out);
    out.ee.resize(n_out);
    out.pfe_95.resize(n_out);
    return out;
}

std::vector<ExposureProfile> irs_hull_white_2f_exposure_profile_batch(
    const ExecutionContext& execution, double a, double b, double sigma, double eta, double rho, double r0,
    const std::vector<double>& notionals, const std::vector<double>& fixed_rates, double start,
    const std::vector<double>& payment_times, const std::vector<double>& accruals,
    const std::vector<double>& monitoring_times, std::size_t n_steps, std::uint64_t n_paths, std::uint64_t seed) {
    require_gpu(execution);
    if (notionals.size() != fixed_rates.size())
        throw std::invalid_argument("el lote requiere un fixed_rate por notional");
    std::size_t nt = notionals.size(), nd = monitoring_times.size();
    std::vector<double> times(nd), ee(nt * nd), pfe(nt * nd);
    int32_t n_out = 0;
    check(eqhip_irs_hull_white_2f_exposure_profile_batch(
        gpu_context(), a, b, sigma, eta, rho, r0, notionals.data(), fixed_rates.data(), (int32_t)nt, start,
        payment_times.data(), accruals.data(), (int32_t)payment_times.size(), monitoring_times.data(), (int32_t)nd,
        (uint32_t)n_steps, n_paths, seed, EQHIP_RNG_PHILOX, disc_of(execution), nullptr, 0, times.data(), ee.data(),
        pfe.data(), &n_out));
    std::vector<ExposureProfile> out(nt);
    for (std::size_t j = 0; j < nt; ++j) {
        out[j].times.assign(times.begin(), times.begin() + n_out);
        out[j].ee.assign(ee.begin() + j * nd, ee.begin() + j * nd + n_out);
        out[j].pfe_95.assign(pfe.begin() + j * nd, pfe.begin() + j * nd + n_out);
    }
    return out;
}

double unilateral_cva_from_exposure_2f(const ExecutionContext& execution, double a, double b, double sigma,
                                       double eta, double rho, double r0, const std::vector<double>& times,
                                       const std::vector<double>& ee, double hazard_rate, double recovery_rate) {
    require_gpu(execution);
    if (times.size() != ee.size()) throw std::invalid_argument("times y ee deben tener el mismo largo");
    double cva = 0.0;
    check(eqhip_unilateral_cva_from_exposure_2f(gpu_context(), a, b, sigma, eta, rho, r0, times.data(), ee.data(),
                                                (int32_t)times.size(), hazard_rate, recovery_rate, &cva));
    return cva;
}

std::vector<double> unilateral_cva_from_exposure_2f_batch(const ExecutionContext& execution, double a, double b,
                                                          double sigma, double eta, double rho, double r0,
                                                          const std::vector<ExposureProfile>& profiles,
                                                          double hazard_rate, double recovery_rate) {
    require_gpu(execution);
    std::vector<double> out(profiles.size());
    if (profiles.empty()) return out;
    const std::vector<double>& times = profiles.front().times;
    std::vector<double> ee;
    ee.reserve(profiles.size() * times.size());
    for (const auto& p : profiles) {
        if (p.times != times || p.ee.size() != times.size())
            throw std::invalid_argument("unilateral_cva_from_exposure_2f_batch: profiles must share their times");
        ee.insert(ee.end(), p.ee.begin(), p.ee.end());
    }
    check(eqhip_unilateral_cva_from_exposure_2f_batch(gpu_context(), a, b, sigma, eta, rho, r0, times.data(),
                                                      ee.data(), (int32_t)profiles.size(), (int32_t)times.size(),
                                                      hazard_rate, recovery_rate, out.data()));
    return out;
}

void run_on_gpu_streams(std::size_t n, const std::function<void(std::size_t)>& task,
                        const std::vector<std::size_t>& claim_order) {
    if (!claim_order.empty() && claim_order.size() != n)
        throw std::logic_error("run_on_gpu_streams: claim_order must list every task once");
    static std::mutex busy;
    std::unique_lock<std::mutex> lock(busy, std::try_to_lock);
    if (n < 2 || !lock.owns_lock()) {
        for (std::size_t i = 0; i < n; ++i) task(i);
        return;
    }
    static StreamWorkers* workers = new StreamWorkers(kGpuStreams - 1);
    std::vector<std::exception_ptr> errors(n);
    std::atomic<std::size_t> next{0};
    const std::function<void()> drain = [&] {
        for (std::size_t c; (c = next.fetch_add(1)) < n;) {
            const std::size_t i = claim_order.empty() ? c : claim_order[c];
            try {
                task(i);
            } catch (...) {
                errors[i] = std::current_exception();
            }
        }
    };
    workers->run(drain);
    for (const std::exception_ptr& e : errors)
        if (e) std::rethrow_exception(e);
}

std::vector<ExposureProfile> shared_exposure_profiles(const IModel& model,
                                                      const std::vector<const IrSwapProduct*>& products,
                                                      const PricingContext& pricing,
                                                      const ExecutionContext& execution) {
    require_gpu(execution);
    // one scenario set for every calendar: each group is valued in a plan of its own (the per-trade
    // PFE rows of one group at a time, not of the whole book at once) whose dates end at the latest
    // reset date of all, so every plan has the same grid. Under HW1F the paths are simulated once and
    // every plan reads them; under G2++ each plan simulates the same counter-based Philox paths again
    double t_last = -INFINITY;
    for (const IrSwapProduct* p : products)
        for (double t : reset_dates(*p)) t_last = std::max(t_last, t);
    std::map<std::string, std::vector<std::size_t>> groups;
    for (std::size_t j = 0; j < products.size(); ++j) {
        const IrSwapProduct& p = *products[j];
        std::string key;
        for (double t : p.payment_times()) key += std::to_string(t) + ",";
        key += "|" + std::to_string(p.start());
        for (double a : p.accruals()) key += "," + std::to_string(a);
        groups[key].push_back(j);
    }
    std::vector<ExposureProfile> profiles(products.size());
    // HW1F: the paths are simulated once, into a scenario set every group values on
    eqhip_scenarios* scen = nullptr;
    const auto* hw1_model = dynamic_cast<const HullWhite1FModel*>(&model);
    if (hw1_model && t_last > 0.0) {
        eqhip_model_hw1f m{hw1_model->a(), hw1_model->b(), hw1_model->sigma(), hw1_model->r0()};
        eqhip_mc mc{pricing.n_paths(), (uint32_t)pricing.n_steps(), pricing.seed(), EQHIP_RNG_PHILOX,
                    disc_of(execution)};
        check(eqhip_scenarios_create(gpu_context(), &m, t_last, &mc, &scen));
    }
    struct ScenariosGuard {
        eqhip_scenarios* s;
        ~ScenariosGuard() {
            if (s) eqhip_scenarios_destroy(s);
        }
    } scen_guard{scen};
    // the groups spread over the GPU streams, largest first; each writes only its own trades' profiles
    std::vector<const std::pair<const std::string, std::vector<std::size_t>>*> group_list;
    for (const auto& kv : groups) group_list.push_back(&kv);
    std::vector<std::size_t> largest_first(group_list.size());
    for (std::size_t g = 0; g < group_list.size(); ++g) largest_first[g] = g;
    std::stable_sort(largest_first.begin(), largest_first.end(), [&](std::size_t x, std::size_t y) {
        return group_list[x]->second.size() > group_list[y]->second.size();
    });
    run_on_gpu_streams(group_list.size(), [&](std::size_t g) {
        const auto& kv = *group_list[g];
        const IrSwapProduct& first = *products[kv.second.front()];
        const std::vector<double> own = reset_dates(first);
        std::vector<double> dates = own;
        const bool extra = t_last > own.back();
        if (extra) dates.push_back(t_last);
        std::vector<double> notionals, fixed_rates;
        for (std::size_t j : kv.second) {
            notionals.push_back(products[j]->notional());
            fixed_rates.push_back(products[j]->fixed_rate());
        }
        const std::size_t nt = notionals.size(), nd = dates.size();
        std::vector<double> times(nd), ee(nt * nd), pfe(nt * nd);
        int32_t n_out = 0;
        if (scen) {
            check(eqhip_irs_hull_white_exposure_profile_batch_on(
                gpu_context(), scen, notionals.data(), fixed_rates.data(), (int32_t)nt, first.start(),
                first.payment_times().data(), first.accruals().data(), (int32_t)first.payment_times().size(),
                dates.data(), (int32_t)nd, times.data(), ee.data(), pfe.data(), &n_out));
        } else if (const auto* hw2 = dynamic_cast<const HullWhite2FModel*>(&model)) {
            check(eqhip_irs_hull_white_2f_exposure_profile_batch(
                gpu_context(), hw2->a(), hw2->b(), hw2->sigma(), hw2->eta(), hw2->rho(), hw2->r0(), notionals.data(),
                fixed_rates.data(), (int32_t)nt, first.start(), first.payment_times().data(), first.accruals().data(),
                (int32_t)first.payment_times().size(), dates.data(), (int32_t)nd, (uint32_t)pricing.n_steps(),
                pricing.n_paths(), pricing.seed(), EQHIP_RNG_PHILOX, disc_of(execution), nullptr, 1, times.data(),
                ee.data(), pfe.data(), &n_out));
        } else {
            const auto& hw1 = require_hw1f(model, "ExposureProfileMeasure");
            check(eqhip_irs_hull_white_exposure_profile_batch(
                gpu_context(), hw1.a(), hw1.b(), hw1.sigma(), hw1.r0(), notionals.data(), fixed_rates.data(),
                (int32_t)nt, first.start(), first.payment_times().data(), first.accruals().data(),
                (int32_t)first.payment_times().size(), dates.data(), (int32_t)nd, (uint32_t)pricing.n_steps(),
                pricing.n_paths(), pricing.seed(), EQHIP_RNG_PHILOX, disc_of(execution), nullptr, 1, times.data(),
                ee.data(), pfe.data(), &n_out));
        }
        const std::size_t keep = std::min<std::size_t>((std::size_t)n_out, own.size());
        for (std::size_t i = 0; i < nt; ++i) {
            ExposureProfile& pr = profiles[kv.second[i]];
            pr.times.assign(times.begin(), times.begin() + keep);
            pr.ee.assign(ee.begin() + i * nd, ee.begin() + i * nd + keep);
            pr.pfe_95.assign(pfe.begin() + i * nd, pfe.begin() + i * nd + keep);
        }
    }, largest_first);
    return profiles;
}

// ---- batched helpers (measure.cpp:238-325) ----------------------------------------------------

std::vector<ExposureProfile> compute_exposure_profile_batch(const IModel& model,
                                                            const std::vector<const IrSwapProduct*>& products,
                                                            const PricingContext& pricing,
                                                            const ExecutionContext& execution) {
    const IrSwapProduct& first = *products.front();
    std::vector<double> notionals, fixed_rates;
    for (const auto* p : products) {
        notionals.push_back(p->notional());
        fixed_rates.push_back(p->fixed_rate());
    }
    if (const auto* hw2 = dynamic_cast<const HullWhite2FModel*>(&model))
        return irs_hull_white_2f_exposure_profile_batch(execution, hw2->a(), hw2->b(), hw2->sigma(), hw2->eta(),
                                                        hw2->rho(), hw2->r0(), notionals, fixed_rates, first.start(),
                                                        first.payment_times(), first.accruals(), reset_dates(first),
                                                        pricing.n_steps(), pricing.n_paths(), pricing.seed());
    const auto& hw1 = require_hw1f(model, "ExposureProfileMeasure");
    return irs_hull_white_exposure_profile_batch(execution, hw1.a(), hw1.b(), hw1.sigma(), hw1.r0(), notionals,
                                                 fixed_rates, first.start(), first.payment_times(), first.accruals(),
                                                 reset_dates(first), pricing.n_steps(), pricing.n_paths(),
                                                 pricing.seed());
}

std::vector<double> compute_cva_from_exposure_batch(const IModel& model, const ExecutionContext& execution,
                                                    const std::vector<ExposureProfile>& profiles,
                                                    double hazard_rate, double recovery_rate) {
    if (const auto* hw2 = dynamic_cast<const HullWhite2FModel*>(&model))
        return unilateral_cva_from_exposure_2f_batch(execution, hw2->a(), hw2->b(), hw2->sigma(), hw2->eta(),
                                                     hw2->rho(), hw2->r0(), profiles, hazard_rate, recovery_rate);
    const auto& hw1 = require_hw1f(model, "UnilateralCvaMeasure");
    return unilateral_cva_from_exposure_batch(execution, hw1.a(), hw1.b(), hw1.sigma(), hw1.r0(), profiles,
                                              hazard_rate, recovery_rate);
}

PvDv01 compute_pv_dv01_batch(const MarketSnapshot& market, const std::vector<const IrSwapProduct*>& products,
                             double bump, bool bucketed, const ExecutionContext& execution, double valuation_time) {
    require_gpu(execution);
    std::size_t n = products.size(), np = market.pillars().size();
    std::vector<int32_t> set_id(n), off(n + 1, 0);
    std::vector<double> notional(n), fixed(n), start(n), pay, acc;
    for (std::size_t j = 0; j < n; ++j) {
        const IrSwapProduct& p = *products[j];
        set_id[j] = (int32_t)j;
        notional[j] = p.notional();
        fixed[j] = effective_fixed_rate(market, p);  // fixed once under the base curve (measure.cpp:195-200)
        start[j] = p.start();
        pay.insert(pay.end(), p.payment_times().begin(), p.payment_times().end());
        acc.insert(acc.end(), p.accruals().begin(), p.accruals().end());
        off[j + 1] = (int32_t)pay.size();
    }
    eqhip_portfolio pf{(int32_t)n, (int32_t)n, set_id.data(), notional.data(), fixed.data(), start.data(),
                       off.data(), pay.data(), acc.data()};
    PvDv01 out;
    out.pv.resize(n);
    out.dv01.resize(n);
    std::vector<double> bk(bucketed ? n * np : 0);
    check(eqhip_pv_dv01_at(gpu_context(), market.pillars().data(), market.zero_rates().data(), (int32_t)np, &pf, bump,
                           valuation_time, out.pv.data(), out.dv01.data(), bucketed ? bk.data() : nullptr));
    if (bucketed)
        for (std::size_t j = 0; j < n; ++j) out.bucketed.emplace_back(bk.begin() + j * np, bk.begin() + (j + 1) * np);
    return out;
}

// ---- measures -----------------------------------------------------------------------------------

MeasureResult ExposureProfileMeasure::evaluate(const IModel& model, const IProduct& product, const MarketSnapshot&,
                                               const PricingContext& pricing,
                                               const ExecutionContext& execution) const {
    const auto& irs = require_irs(product, "ExposureProfileMeasure");
    ExposureProfile profile = compute_exposure_profile(model, irs, pricing, execution);
    MeasureResult r;
    r.times = std::move(profile.times);
    r.primary = std::move(profile.ee);
    r.secondary = std::move(profile.pfe_95);
    return r;
}

MeasureResult UnilateralCvaMeasure::evaluate(const IModel& model, const IProduct& product,
                                             const MarketSnapshot& market, const PricingContext& pricing,
                                             const ExecutionContext& execution) const {
    ExposureProfileMeasure exposure(Params{});
    MeasureResult r = exposure.evaluate(model, product, market, pricing, execution);
    if (const auto* hw2 = dynamic_cast<const HullWhite2FModel*>(&model)) {  // measure.cpp:79-85
        r.scalar = unilateral_cva_from_exposure_2f(execution, hw2->a(), hw2->b(), hw2->sigma(), hw2->eta(),
                                                   hw2->rho(), hw2->r0(), r.times, r.primary, market.hazard_rate(),
                                                   market.recovery_rate());
    } else {
        const auto& hw1 = require_hw1f(model, "UnilateralCvaMeasure");
        r.scalar = unilateral_cva_from_exposure(execution, hw1.a(), hw1.b(), hw1.sigma(), hw1.r0(), r.times,
                                                r.primary, market.hazard_rate(), market.recovery_rate());
    }
    r.has_scalar = true;
    return r;
}

MeasureResult PresentValueMeasure::evaluate(const IModel&, const IProduct& product, const MarketSnapshot& market,
                                            const PricingContext& pricing, const ExecutionContext& execution) const {
    const auto& irs = require_irs(product, "PresentValueMeasure");
    // valued at the pricing date on today's curve (measure.cpp:366-380)
    PvDv01 v = compute_pv_dv01_batch(market, {&irs}, 0.0001, false, execution, pricing.pricing_date());
    MeasureResult r;
    r.has_scalar = true;
    r.scalar = v.pv[0];
    return r;
}

MeasureResult Dv01Measure::evaluate(const IModel&, const IProduct& product, const MarketSnapshot& market,
                                    const PricingContext& pricing, const ExecutionContext& execution) const {
    const auto& irs = require_irs(product, "Dv01Measure");
    // the parallel DV01 at the pricing date, the bucketed ladder at t = 0 (measure.cpp:425-445)
    PvDv01 v = compute_pv_dv01_batch(market, {&irs}, bump_, bucketed_, execution,
                                     bucketed_ ? 0.0 : pricing.pricing_date());
    MeasureResult r;
    if (bucketed_) {
        r.times = market.pillars();
        r.primary = v.bucketed[0];
    } else {
        r.has_scalar = true;
        r.scalar = v.dv01[0];
    }
    return r;
}

// UnilateralCVA of one fixed-rate swap as a singleton netting set on the measure's grid (reset-date
// monitoring, the pricing context's steps / paths / seed), with its pathwise derivatives from the
// GPU adjoint (eqhip_cva_adjoint): grad = dCVA / d{a, b, sigma, r0, hazard_rate, recovery_rate}.
CvaAdjoint unilateral_cva_adjoint(const IModel& model, const IProduct& product, const MarketSnapshot& market,
                                  const PricingContext& pricing, const ExecutionContext& execution) {
    require_gpu(execution);
    const auto& hw1 = require_hw1f(model, "UnilateralCVA (pathwise)");
    const auto& irs = require_irs(product, "UnilateralCVA (pathwise)");
    if (irs.use_par_rate())
        throw std::invalid_argument(
            "compute_greek: metodo 'pathwise' de 'UnilateralCVA' requiere un fixed_rate explicito: con tasa par la "
            "tasa fija depende de los parametros del modelo y el adjunto la mantiene fija (use method='auto' o "
            "'bump_and_reval')");
    const std::vector<double> dates = reset_dates(irs);
    const eqhip_model_hw1f m{hw1.a(), hw1.b(), hw1.sigma(), hw1.r0()};
    const int32_t set_id = 0;
    const double notional = irs.notional(), k = irs.fixed_rate(), start = irs.start();
    const int32_t off[2] = {0, (int32_t)irs.payment_times().size()};
    const eqhip_portfolio pf{1, 1, &set_id, &notional, &k, &start, off, irs.payment_times().data(),
                             irs.accruals().data()};
    const eqhip_mc mc{pricing.n_paths(), (uint32_t)pricing.n_steps(), pricing.seed(), EQHIP_RNG_PHILOX,
                      disc_of(execution)};
    eqhip_plan* plan = nullptr;
    check(eqhip_plan_create(gpu_context(), &m, &pf, dates.data(), (int32_t)dates.size(), &mc, 0, &plan));
    struct Guard {
        eqhip_plan* p;
        ~Guard() { eqhip_plan_destroy(p); }
    } guard{plan};
    const eqhip_credit credit{market.hazard_rate(), market.recovery_rate()};
    CvaAdjoint out{};
    check(eqhip_cva_adjoint(plan, nullptr, 0, &credit, &out.cva, out.grad, nullptr));
    return out;
}

void register_builtins(Registries& registries) {
    registries.models.register_type<HullWhite1FModel>("HullWhite1F");
    registries.models.register_type<HullWhite2FModel>("HullWhite2F");
    registries.products.register_type<IrSwapProduct>("IRSwap");
    registries.measures.register_type<ExposureProfileMeasure>("ExposureProfile");
    registries.measures.register_type<UnilateralCvaMeasure>("UnilateralCVA");
    registries.measures.register_type<PresentValueMeasure>("PV");
    registries.measures.register_type<Dv01Measure>("DV01");
    // the Greek measure resolves its inner metric by name, so it keeps the registries (bootstrap.cpp:33-39)
    const Registries* regs = &registries;
    registries.measures.register_factory(
        "Greek", [regs](const Params& params) { return std::make_unique<GreekMeasure>(params, *regs); });
}

}  // namespace engine
