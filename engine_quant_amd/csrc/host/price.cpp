// price / price_batch / price_many (cpp/engine/src/price.cpp:301-578): the caller's list of measure
// names becomes a list of requests, each served by one evaluation of a registered measure (requests of
// the same type with equal parameters share it) and shown whole or as the EE / PFE95 part of an exposure
// profile. On the GPU one exposure simulation serves every exposure-derived request of a call:
// ExpectedExposure, PFE95, ExposureProfile and UnilateralCVA share it (the reference simulates again
// for UnilateralCVA, measure.cpp:348-365; with a seeded counter-based stream the shared run gives the
// same numbers). A batch is one schedule; price_many splits a portfolio into such batches.
#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>

#include "engine.hpp"

namespace engine {

namespace {

// ---- requests and evaluations -------------------------------------------------------------------
enum class View { Whole, Expected, Potential };

// Names that show a part of another measure's result (price.cpp:29-61: the legacy aliases of
// ExposureProfile's primary and secondary profiles).
struct PartName {
    const char* name;
    const char* type;
    View view;
};
constexpr PartName kPartNames[] = {{"ExpectedExposure", "ExposureProfile", View::Expected},
                                   {"PFE95", "ExposureProfile", View::Potential}};

struct Request {
    std::string shown_as;
    View view = View::Whole;
    std::size_t eval = 0;  // the evaluation that serves it
};

struct Evaluation {
    std::string type;
    Params params;
};

struct Plan {
    std::vector<Request> requests;
    std::vector<Evaluation> evals;
};

// Resolves every spec; `accepts(type)` says whether a registered type can be served here.
template <class Accepts>
Plan plan_requests(const std::vector<MeasureSpec>& specs, Accepts accepts, const std::string& unknown_prefix) {
    Plan plan;
    for (const MeasureSpec& spec : specs) {
        std::string type = spec.name;
        View view = View::Whole;
        for (const PartName& part : kPartNames)
            if (spec.name == part.name) {
                type = part.type;
                view = part.view;
            }
        if (!accepts(type)) throw std::invalid_argument(unknown_prefix + spec.name + "'");
        std::size_t e = 0;
        while (e < plan.evals.size() && !(plan.evals[e].type == type && plan.evals[e].params == spec.params)) ++e;
        if (e == plan.evals.size()) plan.evals.push_back(Evaluation{type, spec.params});
        plan.requests.push_back(Request{spec.alias.value_or(spec.name), view, e});
    }
    return plan;
}

MeasureResult shown(const MeasureResult& r, View view) {
    if (view == View::Whole) return r;
    MeasureResult part;
    part.times = r.times;
    part.primary = (view == View::Expected) ? r.primary : r.secondary;
    return part;
}

bool from_exposure_run(const std::string& type) { return type == "ExposureProfile" || type == "UnilateralCVA"; }

// ---- batches of one schedule --------------------------------------------------------------------
// Everything a calendar group shares: the product type and, for swaps, start, payment times, accruals.
struct CalendarKey {
    std::string type;
    double start = 0.0;
    const std::vector<double>* pay = nullptr;
    const std::vector<double>* acc = nullptr;
};

CalendarKey calendar_of(const IProduct& p) {
    CalendarKey k{p.type_name()};
    if (const auto* s = dynamic_cast<const IrSwapProduct*>(&p)) {
        k.start = s->start();
        k.pay = &s->payment_times();
        k.acc = &s->accruals();
    }
    return k;
}

// A strict weak order on calendars (exact comparison of every time and accrual).
bool calendar_less(const CalendarKey& x, const CalendarKey& y) {
    static const std::vector<double> none;
    const auto& xp = x.pay ? *x.pay : none;
    const auto& yp = y.pay ? *y.pay : none;
    const auto& xa = x.acc ? *x.acc : none;
    const auto& ya = y.acc ? *y.acc : none;
    return std::tie(x.type, x.start, xp, xa) < std::tie(y.type, y.start, yp, ya);
}

// The swaps of a batch, which must share one product type, one schedule and explicit fixed rates; the
// messages are the reference's (price.cpp:110-150).
std::vector<const IrSwapProduct*> batch_swaps(const std::vector<const IProduct*>& products) {
    if (products.empty()) throw std::invalid_argument("engine::price_batch: el lote no puede estar vacío");
    const std::string& type = products.front()->type_name();
    std::vector<const IrSwapProduct*> swaps;
    swaps.reserve(products.size());
    for (const IProduct* p : products) {
        if (p->type_name() != type)
            throw std::invalid_argument("engine::price_batch: todos los trades del lote deben ser del mismo tipo de "
                                        "producto ('" + type + "' vs '" + p->type_name() + "')");
        const auto* s = dynamic_cast<const IrSwapProduct*>(p);
        if (s == nullptr)
            throw std::invalid_argument("engine::price_batch: tipo de producto no soportado para lote: " + type);
        if (s->use_par_rate())
            throw std::invalid_argument("engine::price_batch: los trades del lote deben traer fixed_rate explícito "
                                        "(use_par_rate no soportado en lote)");
        swaps.push_back(s);
    }
    const CalendarKey first = calendar_of(*swaps.front());
    for (const IrSwapProduct* s : swaps) {
        const CalendarKey k = calendar_of(*s);
        if (calendar_less(k, first) || calendar_less(first, k))
            throw std::invalid_argument("engine::price_batch: todos los trades del lote deben compartir calendario "
                                        "(start/payment_times/accruals)");
    }
    return swaps;
}

// The portfolio's calendar groups, each listing its trades in input order; groups are ordered by their
// first trade (the order in which a serial loop over the portfolio meets them).
std::vector<std::vector<std::size_t>> calendar_groups(const std::vector<const IProduct*>& products) {
    std::vector<CalendarKey> keys;
    keys.reserve(products.size());
    for (const IProduct* p : products) keys.push_back(calendar_of(*p));
    std::vector<std::size_t> by_key(products.size());
    std::iota(by_key.begin(), by_key.end(), std::size_t{0});
    std::stable_sort(by_key.begin(), by_key.end(),
                     [&](std::size_t x, std::size_t y) { return calendar_less(keys[x], keys[y]); });
    std::vector<std::vector<std::size_t>> groups;
    for (std::size_t i = 0; i < by_key.size(); ++i) {
        const bool new_group = i == 0 || calendar_less(keys[by_key[i - 1]], keys[by_key[i]]);
        if (new_group) groups.emplace_back();
        groups.back().push_back(by_key[i]);
    }
    std::sort(groups.begin(), groups.end(), [](const auto& x, const auto& y) { return x.front() < y.front(); });
    return groups;
}

const char* const kBatchTypes[] = {"PV", "DV01", "UnilateralCVA", "ExposureProfile"};

// One batch of one schedule. `profiles` (nullable) are the batch's exposure profiles when the caller
// already has them (a shared scenario set).
PriceBatchResult value_batch(const Registries& registries, const std::vector<const IProduct*>& products,
                             const std::vector<MeasureSpec>& specs, const IModel& model, const MarketSnapshot& market,
                             const PricingContext& pricing, const ExecutionContext& execution,
                             const std::vector<ExposureProfile>* profiles) {
    if (products.empty()) throw std::invalid_argument("engine::price_batch: el lote no puede estar vacío");
    const Plan plan = plan_requests(
        specs,
        [&](const std::string& type) {
            return registries.measures.contains(type) ||
                   std::find(std::begin(kBatchTypes), std::end(kBatchTypes), type) != std::end(kBatchTypes);
        },
        "engine::price_batch: medida desconocida o no soportada en lote: '");
    const std::vector<const IrSwapProduct*> swaps = batch_swaps(products);
    const std::size_t n = swaps.size();

    std::optional<std::vector<ExposureProfile>> run;  // one simulation for the whole batch
    if (profiles) run = *profiles;
    auto exposure = [&]() -> const std::vector<ExposureProfile>& {
        if (!run) run = compute_exposure_profile_batch(model, swaps, pricing, execution);
        return *run;
    };
    // PV / DV01 ladders: PV and the parallel DV01 at the pricing date, the bucketed ladder at t = 0, so a
    // ladder is keyed by (bump, bucketed, valuation time); PV is the same in every ladder of its time
    std::map<std::tuple<double, bool, double>, PvDv01> ladders;
    auto ladder = [&](double bump, bool bucketed, double vt) -> const PvDv01& {
        const auto key = std::make_tuple(bump, bucketed, vt);
        auto it = ladders.find(key);
        if (it == ladders.end())
            it = ladders.emplace(key, compute_pv_dv01_batch(market, swaps, bump, bucketed, execution, vt)).first;
        return it->second;
    };
    const double pd = pricing.pricing_date();

    std::vector<std::vector<MeasureResult>> per_eval(plan.evals.size(), std::vector<MeasureResult>(n));
    for (std::size_t e = 0; e < plan.evals.size(); ++e) {
        const Evaluation& ev = plan.evals[e];
        std::vector<MeasureResult>& out = per_eval[e];
        if (ev.type == "ExposureProfile") {
            const auto& p = exposure();
            for (std::size_t j = 0; j < n; ++j) {
                out[j].times = p[j].times;
                out[j].primary = p[j].ee;
                out[j].secondary = p[j].pfe_95;
            }
        } else if (ev.type == "UnilateralCVA") {
            const std::vector<double> cva = compute_cva_from_exposure_batch(model, execution, exposure(),
                                                                            market.hazard_rate(), market.recovery_rate());
            for (std::size_t j = 0; j < n; ++j) {
                out[j].has_scalar = true;
                out[j].scalar = cva[j];
            }
        } else if (ev.type == "PV") {
            const PvDv01* any = nullptr;
            for (const auto& [key, l] : ladders)
                if (std::get<2>(key) == pd) any = &l;
            const PvDv01& l = any ? *any : ladder(0.0001, false, pd);
            for (std::size_t j = 0; j < n; ++j) {
                out[j].has_scalar = true;
                out[j].scalar = l.pv[j];
            }
        } else if (ev.type == "DV01") {
            const double bump = get_double(ev.params, "bump", 0.0001);
            const bool bucketed = get_bool(ev.params, "bucketed", false);
            const PvDv01& l = ladder(bump, bucketed, bucketed ? 0.0 : pd);
            for (std::size_t j = 0; j < n; ++j) {
                if (bucketed) {
                    out[j].times = market.pillars();
                    out[j].primary = l.bucketed[j];
                } else {
                    out[j].has_scalar = true;
                    out[j].scalar = l.dv01[j];
                }
            }
        } else {
            const auto measure = registries.measures.create(ev.type, ev.params);
            for (std::size_t j = 0; j < n; ++j) out[j] = measure->evaluate(model, *swaps[j], market, pricing, execution);
        }
    }
    PriceBatchResult result(n);
    for (std::size_t j = 0; j < n; ++j) {
        result[j].trade_index = j;
        for (const Request& r : plan.requests)
            result[j].measures.push_back(PriceResultEntry{r.shown_as, shown(per_eval[r.eval][j], r.view)});
    }
    return result;
}

// Prices each calendar group as one batch (the groups' `batch(k)` results land in their trades' rows).
template <class Batch>
PriceBatchResult over_groups(const std::vector<std::vector<std::size_t>>& groups, std::size_t n_trades,
                             const ExecutionContext& execution, const std::vector<double>& cost, Batch batch) {
    PriceBatchResult result(n_trades);
    auto price_group = [&](std::size_t k) {
        PriceBatchResult g = batch(k);
        for (std::size_t j = 0; j < groups[k].size(); ++j)
            result[groups[k][j]] = PriceBatchResultEntry{groups[k][j], std::move(g[j].measures)};
    };
    if (execution.backend() != "gpu") {
        for (std::size_t k = 0; k < groups.size(); ++k) price_group(k);
        return result;
    }
    // GPU: the groups run over several streams at once, the most trade-flows first; an error is the one
    // a serial loop over the groups would report
    std::vector<std::size_t> heavy_first(groups.size());
    std::iota(heavy_first.begin(), heavy_first.end(), std::size_t{0});
    if (!cost.empty())
        std::stable_sort(heavy_first.begin(), heavy_first.end(),
                         [&](std::size_t x, std::size_t y) { return cost[x] > cost[y]; });
    run_on_gpu_streams(groups.size(), price_group, heavy_first);
    return result;
}

}  // namespace

std::vector<std::string> price_measure_names(const Registries& registries) {
    std::vector<std::string> names = registries.measures.list();
    for (const PartName& part : kPartNames) names.push_back(part.name);
    return names;
}

PriceResult price(const Registries& registries, const IProduct& product, const std::vector<MeasureSpec>& specs,
                  const IModel& model, const MarketSnapshot& market, const PricingContext& pricing,
                  const ExecutionContext& execution) {
    const Plan plan = plan_requests(
        specs, [&](const std::string& type) { return registries.measures.contains(type); },
        "engine::price: medida desconocida: '");
    // on the GPU, one exposure run (with the CVA when any is asked) serves every parameterless
    // exposure-derived evaluation
    std::optional<MeasureResult> run;
    if (execution.backend() == "gpu") {
        bool exposure = false, cva = false;
        for (const Evaluation& ev : plan.evals) {
            exposure |= from_exposure_run(ev.type);
            cva |= ev.type == "UnilateralCVA";
        }
        if (exposure)
            run = registries.measures.create(cva ? "UnilateralCVA" : "ExposureProfile", Params{})
                      ->evaluate(model, product, market, pricing, execution);
    }
    std::vector<MeasureResult> results;
    results.reserve(plan.evals.size());
    for (const Evaluation& ev : plan.evals) {
        if (run && from_exposure_run(ev.type) && ev.params.empty()) {
            MeasureResult r = *run;
            if (ev.type == "ExposureProfile") {
                r.has_scalar = false;
                r.scalar = 0.0;
            }
            results.push_back(std::move(r));
        } else {
            results.push_back(registries.measures.create(ev.type, ev.params)
                                  ->evaluate(model, product, market, pricing, execution));
        }
    }
    PriceResult out;
    for (const Request& r : plan.requests) out.push_back(PriceResultEntry{r.shown_as, shown(results[r.eval], r.view)});
    return out;
}

PriceBatchResult price_batch(const Registries& registries, const std::vector<const IProduct*>& products,
                             const std::vector<MeasureSpec>& measures, const IModel& model,
                             const MarketSnapshot& market, const PricingContext& pricing,
                             const ExecutionContext& execution) {
    return value_batch(registries, products, measures, model, market, pricing, execution, nullptr);
}

PriceBatchResult price_many(const Registries& registries, const std::vector<const IProduct*>& products,
                            const std::vector<MeasureSpec>& measures, const IModel& model,
                            const MarketSnapshot& market, const PricingContext& pricing,
                            const ExecutionContext& execution) {
    if (products.empty()) throw std::invalid_argument("engine::price_many: la lista de trades no puede estar vacía");
    const std::vector<std::vector<std::size_t>> groups = calendar_groups(products);
    auto members = [&](std::size_t k) {
        std::vector<const IProduct*> g;
        for (std::size_t i : groups[k]) g.push_back(products[i]);
        return g;
    };

    if (execution.simulation() != "shared") {
        // each group simulates on its own grid (the reference's price_many)
        std::vector<double> cost(groups.size());
        for (std::size_t k = 0; k < groups.size(); ++k) {
            const auto* s = dynamic_cast<const IrSwapProduct*>(products[groups[k].front()]);
            cost[k] = (double)groups[k].size() * (s ? (double)s->payment_times().size() : 1.0);
        }
        return over_groups(groups, products.size(), execution, cost, [&](std::size_t k) {
            return value_batch(registries, members(k), measures, model, market, pricing, execution, nullptr);
        });
    }

    // one scenario set for every group: every trade's profile on the common grid and paths, then each
    // group's measures from its trades' profiles
    std::vector<const IrSwapProduct*> swaps;
    for (const IProduct* p : products) {
        const auto* s = dynamic_cast<const IrSwapProduct*>(p);
        if (!s)
            throw std::invalid_argument("engine::price_many: simulation \"shared\" serves IRSwap trades only, got '" +
                                        p->type_name() + "'");
        if (s->use_par_rate())
            throw std::invalid_argument(
                "engine::price_many: los trades deben traer fixed_rate explícito (use_par_rate no soportado en lote)");
        swaps.push_back(s);
    }
    bool needs_run = false;
    for (const MeasureSpec& spec : measures) {
        std::string type = spec.name;
        for (const PartName& part : kPartNames)
            if (spec.name == part.name) type = part.type;
        needs_run |= from_exposure_run(type);
    }
    const std::vector<ExposureProfile> all =
        needs_run ? shared_exposure_profiles(model, swaps, pricing, execution) : std::vector<ExposureProfile>{};
    return over_groups(groups, products.size(), execution, {}, [&](std::size_t k) {
        std::vector<ExposureProfile> mine;
        if (needs_run)
            for (std::size_t i : groups[k]) mine.push_back(all[i]);
        return value_batch(registries, members(k), measures, model, market, pricing, execution,
                           needs_run ? &mine : nullptr);
    });
}

}  // namespace engine
