// price / price_batch / price_many, cpp/engine/src/price.cpp:29-578 (aliases, per-call cache,
// homogeneous batches, calendar grouping). On the GPU path one exposure simulation serves every
// exposure-derived measure of a call: ExpectedExposure / PFE95 / ExposureProfile / UnilateralCVA
// share it (the reference re-simulates for UnilateralCVA, measure.cpp:348-365; with a seeded
// counter-based stream the shared result is identical).
#include <algorithm>
#include <map>
#include <unordered_map>
#include <unordered_set>

#include "engine.hpp"

namespace engine {

namespace {

enum class Field { Scalar, Primary, Secondary };

struct ResolvedMeasure {
    std::string registered_type;
    Field field;
};

const std::unordered_map<std::string, ResolvedMeasure>& legacy_field_aliases() {
    static const std::unordered_map<std::string, ResolvedMeasure> aliases = {
        {"ExpectedExposure", {"ExposureProfile", Field::Primary}},
        {"PFE95", {"ExposureProfile", Field::Secondary}},
    };
    return aliases;
}

MeasureResult extract_field(const MeasureResult& source, Field field) {
    switch (field) {
        case Field::Scalar:
            return source;
        case Field::Primary: {
            MeasureResult out;
            out.times = source.times;
            out.primary = source.primary;
            return out;
        }
        case Field::Secondary: {
            MeasureResult out;
            out.times = source.times;
            out.primary = source.secondary;
            return out;
        }
    }
    throw std::logic_error("engine::price: Field desconocido");
}

ResolvedMeasure resolve_measure_name(const std::string& name) {
    const auto& aliases = legacy_field_aliases();
    auto it = aliases.find(name);
    if (it != aliases.end()) return it->second;
    return ResolvedMeasure{name, Field::Scalar};
}

// Deterministic serialisation of a Params bag for the per-call cache key (price.cpp:77-108).
std::string params_cache_key(const Params& params) {
    std::vector<std::string> parts;
    for (const auto& kv : params) {
        std::string part = kv.first + "=";
        std::visit(
            [&](const auto& v) {
                using T = std::decay_t<decltype(v)>;
                if constexpr (std::is_same_v<T, double>) {
                    part += std::to_string(v);
                } else if constexpr (std::is_same_v<T, bool>) {
                    part += v ? "true" : "false";
                } else if constexpr (std::is_same_v<T, std::string>) {
                    part += v;
                } else {
                    for (double d : v) part += std::to_string(d) + ",";
                }
            },
            kv.second);
        parts.push_back(std::move(part));
    }
    std::sort(parts.begin(), parts.end());
    std::string key;
    for (const auto& p : parts) key += p + "|";
    return key;
}

const std::string& display_name(const MeasureSpec& spec) { return spec.alias ? *spec.alias : spec.name; }

bool exposure_derived(const std::string& type) { return type == "ExposureProfile" || type == "UnilateralCVA"; }

std::vector<const IrSwapProduct*> require_homogeneous_irs_batch(const std::vector<const IProduct*>& products) {
    if (products.empty()) throw std::invalid_argument("engine::price_batch: el lote no puede estar vacío");
    std::vector<const IrSwapProduct*> irs_products;
    const std::string& type_name = products.front()->type_name();
    for (const IProduct* p : products) {
        if (p->type_name() != type_name)
            throw std::invalid_argument("engine::price_batch: todos los trades del lote deben ser del mismo tipo de producto ('" +
                                        type_name + "' vs '" + p->type_name() + "')");
        const auto* irs = dynamic_cast<const IrSwapProduct*>(p);
        if (!irs) throw std::invalid_argument("engine::price_batch: tipo de producto no soportado para lote: " + type_name);
        if (irs->use_par_rate())
            throw std::invalid_argument(
                "engine::price_batch: los trades del lote deben traer fixed_rate explícito (use_par_rate no soportado en lote)");
        irs_products.push_back(irs);
    }
    const IrSwapProduct& first = *irs_products.front();
    for (const auto* irs : irs_products)
        if (irs->start() != first.start() || irs->payment_times() != first.payment_times() ||
            irs->accruals() != first.accruals())
            throw std::invalid_argument(
                "engine::price_batch: todos los trades del lote deben compartir calendario (start/payment_times/accruals)");
    return irs_products;
}

const std::unordered_set<std::string>& known_batch_registered_types() {
    static const std::unordered_set<std::string> known = {"PV", "DV01", "UnilateralCVA", "ExposureProfile"};
    return known;
}

std::string calendar_group_key(const IrSwapProduct& irs) {
    std::string key = std::to_string(irs.start()) + "|";
    for (double t : irs.payment_times()) key += std::to_string(t) + ",";
    key += "|";
    for (double t : irs.accruals()) key += std::to_string(t) + ",";
    return key;
}

}  // namespace

std::vector<std::string> price_measure_names(const Registries& registries) {
    std::vector<std::string> names = registries.measures.list();
    for (const auto& kv : legacy_field_aliases()) names.push_back(kv.first);
    return names;
}

PriceResult price(const Registries& registries, const IProduct& product, const std::vector<MeasureSpec>& measures,
                  const IModel& model, const MarketSnapshot& market, const PricingContext& pricing,
                  const ExecutionContext& execution) {
    std::vector<ResolvedMeasure> resolved;
    for (const MeasureSpec& spec : measures) {
        ResolvedMeasure r = resolve_measure_name(spec.name);
        if (!registries.measures.contains(r.registered_type))
            throw std::invalid_argument("engine::price: medida desconocida: '" + spec.name + "'");
        resolved.push_back(std::move(r));
    }
    std::unordered_map<std::string, MeasureResult> computed;
    // one exposure simulation shared by every exposure-derived measure of the call
    bool want_cva = false, want_exposure = false;
    for (const auto& r : resolved) {
        want_cva |= r.registered_type == "UnilateralCVA";
        want_exposure |= exposure_derived(r.registered_type);
    }
    std::optional<MeasureResult> shared;
    if (want_exposure && execution.backend() == "gpu") {
        auto m = registries.measures.create(want_cva ? "UnilateralCVA" : "ExposureProfile", Params{});
        shared = m->evaluate(model, product, market, pricing, execution);
    }
    for (std::size_t i = 0; i < measures.size(); ++i) {
        const std::string key = resolved[i].registered_type + "#" + params_cache_key(measures[i].params);
        if (computed.count(key)) continue;
        if (shared && exposure_derived(resolved[i].registered_type) && measures[i].params.empty()) {
            MeasureResult r = *shared;
            if (resolved[i].registered_type == "ExposureProfile") {
                r.has_scalar = false;
                r.scalar = 0.0;
            }
            computed.emplace(key, std::move(r));
            continue;
        }
        auto measure = registries.measures.create(resolved[i].registered_type, measures[i].params);
        computed.emplace(key, measure->evaluate(model, product, market, pricing, execution));
    }
    PriceResult result;
    for (std::size_t i = 0; i < measures.size(); ++i) {
        const std::string key = resolved[i].registered_type + "#" + params_cache_key(measures[i].params);
        result.push_back(PriceResultEntry{display_name(measures[i]), extract_field(computed.at(key), resolved[i].field)});
    }
    return result;
}

static PriceBatchResult price_batch_with_profiles(const Registries& registries,
                                                  const std::vector<const IProduct*>& products,
                                                  const std::vector<MeasureSpec>& measures, const IModel& model,
                                                  const MarketSnapshot& market, const PricingContext& pricing,
                                                  const ExecutionContext& execution,
                                                  const std::vector<ExposureProfile>* given_profiles);

PriceBatchResult price_batch(const Registries& registries, const std::vector<const IProduct*>& products,
                             const std::vector<MeasureSpec>& measures, const IModel& model,
                             const MarketSnapshot& market, const PricingContext& pricing,
                             const ExecutionContext& execution) {
    return price_batch_with_profiles(registries, products, measures, model, market, pricing, execution, nullptr);
}

// price_batch; profiles (nullable) are the batch's exposure profiles when the caller already has them.
static PriceBatchResult price_batch_with_profiles(const Registries& registries,
                                                  const std::vector<const IProduct*>& products,
                                                  const std::vector<MeasureSpec>& measures, const IModel& model,
                                                  const MarketSnapshot& market, const PricingContext& pricing,
                                                  const ExecutionContext& execution,
                                                  const std::vector<ExposureProfile>* given_profiles) {
    if (products.empty()) throw std::invalid_argument("engine::price_batch: el lote no puede estar vacío");
    std::vector<ResolvedMeasure> resolved;
    for (const MeasureSpec& spec : measures) {
        ResolvedMeasure r = resolve_measure_name(spec.name);
        if (!known_batch_registered_types().count(r.registered_type) && !registries.measures.contains(r.registered_type))
            throw std::invalid_argument("engine::price_batch: medida desconocida o no soportada en lote: '" + spec.name + "'");
        resolved.push_back(std::move(r));
    }
    std::vector<const IrSwapProduct*> irs = require_homogeneous_irs_batch(products);

    std::unordered_map<std::string, std::vector<MeasureResult>> computed;
    std::optional<std::vector<ExposureProfile>> profiles;  // one simulation for the whole batch
    if (given_profiles) profiles = *given_profiles;
    auto get_profiles = [&]() -> const std::vector<ExposureProfile>& {
        if (!profiles) profiles = compute_exposure_profile_batch(model, irs, pricing, execution);
        return *profiles;
    };
    // one PV / DV01 ladder per (bump, bucketed); PV is the same in every ladder (the unbumped curve)
    std::map<std::pair<double, bool>, PvDv01> ladders;
    auto get_ladder = [&](double bump, bool bucketed) -> const PvDv01& {
        auto it = ladders.find({bump, bucketed});
        if (it == ladders.end())
            it = ladders.emplace(std::make_pair(bump, bucketed),
                                 compute_pv_dv01_batch(market, irs, bump, bucketed, execution)).first;
        return it->second;
    };
    for (std::size_t i = 0; i < measures.size(); ++i) {
        const std::string& type = resolved[i].registered_type;
        const std::string key = type + "#" + params_cache_key(measures[i].params);
        if (computed.count(key)) continue;
        std::vector<MeasureResult> results;
        if (type == "ExposureProfile") {
            for (const auto& p : get_profiles()) {
                MeasureResult r;
                r.times = p.times;
                r.primary = p.ee;
                r.secondary = p.pfe_95;
                results.push_back(std::move(r));
            }
        } else if (type == "UnilateralCVA") {
            std::vector<double> cvas = compute_cva_from_exposure_batch(model, execution, get_profiles(),
                                                                       market.hazard_rate(), market.recovery_rate());
            for (double c : cvas) {
                MeasureResult r;
                r.has_scalar = true;
                r.scalar = c;
                results.push_back(r);
            }
        } else if (type == "PV" || type == "DV01") {
            double bump = get_double(measures[i].params, "bump", 0.0001);
            bool bucketed = type == "DV01" && get_bool(measures[i].params, "bucketed", false);
            const PvDv01& v = (type == "PV" && !ladders.empty()) ? ladders.begin()->second : get_ladder(bump, bucketed);
            for (std::size_t j = 0; j < irs.size(); ++j) {
                MeasureResult r;
                if (type == "PV") {
                    r.has_scalar = true;
                    r.scalar = v.pv[j];
                } else if (bucketed) {
                    r.times = market.pillars();
                    r.primary = v.bucketed[j];
                } else {
                    r.has_scalar = true;
                    r.scalar = v.dv01[j];
                }
                results.push_back(std::move(r));
            }
        } else {
            auto measure = registries.measures.create(type, measures[i].params);
            for (const IrSwapProduct* p : irs) results.push_back(measure->evaluate(model, *p, market, pricing, execution));
        }
        computed.emplace(key, std::move(results));
    }
    PriceBatchResult result;
    for (std::size_t j = 0; j < irs.size(); ++j) {
        PriceResult row;
        for (std::size_t i = 0; i < measures.size(); ++i) {
            const std::string key = resolved[i].registered_type + "#" + params_cache_key(measures[i].params);
            row.push_back(PriceResultEntry{display_name(measures[i]), extract_field(computed.at(key)[j], resolved[i].field)});
        }
        result.push_back(PriceBatchResultEntry{j, std::move(row)});
    }
    return result;
}

// price_many with one scenario set for every calendar group: every trade's profile on the common grid
// and paths, then per group the CVA on the group's own dates and the measures that need no simulation,
// in price_batch's order and aliases.
static PriceBatchResult price_many_shared(const Registries& registries, const std::vector<const IProduct*>& products,
                                   const std::vector<MeasureSpec>& measures, const IModel& model,
                                   const MarketSnapshot& market, const PricingContext& pricing,
                                   const ExecutionContext& execution) {
    std::vector<const IrSwapProduct*> irs;
    for (const IProduct* p : products) {
        const auto* s = dynamic_cast<const IrSwapProduct*>(p);
        if (!s)
            throw std::invalid_argument("engine::price_many: simulation \"shared\" serves IRSwap trades only, got '" +
                                        p->type_name() + "'");
        if (s->use_par_rate())
            throw std::invalid_argument(
                "engine::price_many: los trades deben traer fixed_rate explícito (use_par_rate no soportado en lote)");
        irs.push_back(s);
    }
    bool needs_profiles = false;
    for (const MeasureSpec& spec : measures) {
        const std::string t = resolve_measure_name(spec.name).registered_type;
        if (t == "ExposureProfile" || t == "UnilateralCVA") needs_profiles = true;
    }
    std::vector<ExposureProfile> all;
    if (needs_profiles) all = shared_exposure_profiles(model, irs, pricing, execution);
    // the trades' calendar groups (price_many's grouping), each priced by price_batch's measure logic
    // on profiles taken from the shared run
    std::unordered_map<std::string, std::vector<std::size_t>> groups;
    std::vector<std::string> group_order;
    for (std::size_t i = 0; i < irs.size(); ++i) {
        auto ins = groups.try_emplace(calendar_group_key(*irs[i]));
        if (ins.second) group_order.push_back(ins.first->first);
        ins.first->second.push_back(i);
    }
    PriceBatchResult result(products.size());
    run_on_gpu_streams(group_order.size(), [&](std::size_t k) {
        const std::vector<std::size_t>& idx = groups.at(group_order[k]);
        std::vector<const IProduct*> group;
        std::vector<ExposureProfile> group_profiles;
        for (std::size_t i : idx) {
            group.push_back(products[i]);
            if (needs_profiles) group_profiles.push_back(all[i]);
        }
        PriceBatchResult g = price_batch_with_profiles(registries, group, measures, model, market, pricing, execution,
                                                       needs_profiles ? &group_profiles : nullptr);
        for (std::size_t j = 0; j < idx.size(); ++j) result[idx[j]] = PriceBatchResultEntry{idx[j], std::move(g[j].measures)};
    });
    return result;
}

PriceBatchResult price_many(const Registries& registries, const std::vector<const IProduct*>& products,
                            const std::vector<MeasureSpec>& measures, const IModel& model,
                            const MarketSnapshot& market, const PricingContext& pricing,
                            const ExecutionContext& execution) {
    if (products.empty()) throw std::invalid_argument("engine::price_many: la lista de trades no puede estar vacía");
    if (execution.simulation() == "shared") return price_many_shared(registries, products, measures, model, market,
                                                                      pricing, execution);
    std::unordered_map<std::string, std::vector<std::size_t>> groups;
    std::vector<std::string> group_order;
    for (std::size_t i = 0; i < products.size(); ++i) {
        const IProduct& p = *products[i];
        std::string key = p.type_name();
        if (const auto* irs = dynamic_cast<const IrSwapProduct*>(&p)) key += "#" + calendar_group_key(*irs);
        auto ins = groups.try_emplace(key);
        if (ins.second) group_order.push_back(key);
        ins.first->second.push_back(i);
    }
    // Each calendar group is an independent price_batch. On the GPU backend the groups run over
    // several streams at once, the ones with the most trade-dates first; each writes only its own rows
    // of the result, and an error is reported as the serial loop would report it.
    PriceBatchResult result(products.size());
    auto price_group = [&](std::size_t k) {
        const std::vector<std::size_t>& idx = groups.at(group_order[k]);
        std::vector<const IProduct*> group;
        for (std::size_t i : idx) group.push_back(products[i]);
        PriceBatchResult g = price_batch(registries, group, measures, model, market, pricing, execution);
        for (std::size_t j = 0; j < idx.size(); ++j) result[idx[j]] = PriceBatchResultEntry{idx[j], std::move(g[j].measures)};
    };
    if (execution.backend() != "gpu") {
        for (std::size_t k = 0; k < group_order.size(); ++k) price_group(k);
        return result;
    }
    std::vector<double> cost(group_order.size());
    for (std::size_t k = 0; k < group_order.size(); ++k) {
        const std::vector<std::size_t>& idx = groups.at(group_order[k]);
        const auto* irs = dynamic_cast<const IrSwapProduct*>(products[idx.front()]);
        cost[k] = (double)idx.size() * (irs ? (double)irs->payment_times().size() : 1.0);
    }
    std::vector<std::size_t> largest_first(group_order.size());
    for (std::size_t k = 0; k < largest_first.size(); ++k) largest_first[k] = k;
    std::stable_sort(largest_first.begin(), largest_first.end(),
                     [&](std::size_t x, std::size_t y) { return cost[x] > cost[y]; });
    run_on_gpu_streams(group_order.size(), price_group, largest_first);
    return result;
}

}  // namespace engine
