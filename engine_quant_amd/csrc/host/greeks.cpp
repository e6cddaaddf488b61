// Greeks of the registered measures (cpp/engine/src/greeks.cpp, cpp/engine/include/engine/greeks.hpp):
// risk-factor parsing, the default bump policy, bump-and-reval stencils of first and second order and
// the cross (4-point) stencil, the automatic sweep, and the "Greek" measure. Every evaluation runs the
// measure on the GPU; the Philox stream is counter-based, so the bumped runs draw the same shocks
// (common random numbers) without any extra bookkeeping. UnilateralCVA of a fixed-rate swap under
// HullWhite1F also has a pathwise route (method "pathwise"), served by the GPU CVA adjoint.
#include <algorithm>
#include <cmath>

#include "engine.hpp"

namespace engine {
namespace greeks {

namespace {

// Default bump policy, greeks.cpp:160-171: relative with an absolute floor for model parameters,
// one basis point for curve and credit, one day for theta.
double default_model_parameter_bump(double base_value) { return std::max(1e-2 * std::abs(base_value), 1e-4); }
double default_curve_bump() { return 0.0001; }
double default_time_shift_bump() { return 1.0 / 365.0; }

bool is_second_order_capable_kind(RiskFactorKind kind) {
    return kind == RiskFactorKind::ModelParameter || kind == RiskFactorKind::CurveParallel ||
           kind == RiskFactorKind::CurvePillar || kind == RiskFactorKind::CreditParameter;
}

// Measures that honour PricingContext::pricing_date (greeks.cpp:1081-1089 names PV and PayoffPriceQ;
// PayoffPriceQ is not part of this path).
bool metric_supports_time_shift(const std::string& metric_name) { return metric_name == "PV"; }

// greeks.cpp:205-259 (model parameters of HullWhite1F are their to_params keys, no aliases)
double resolve_bump(const IModel& model, const MarketSnapshot& market, const RiskFactor& risk_factor,
                    std::optional<double> override_h) {
    if (risk_factor.kind == RiskFactorKind::ModelParameter) {
        Params params = model.to_params();
        auto it = params.find(risk_factor.name);
        if (it == params.end() || !std::holds_alternative<double>(it->second))
            throw std::invalid_argument("compute_greek: el modelo '" + model.type_name() +
                                        "' no tiene el parametro de riesgo '" + to_string(risk_factor) + "'");
        return override_h.value_or(default_model_parameter_bump(std::get<double>(it->second)));
    }
    if (risk_factor.kind == RiskFactorKind::CurvePillar) {
        const std::size_t pillar_index = risk_factor.pillar_index.value_or(0);
        if (pillar_index >= market.pillars().size())
            throw std::invalid_argument("compute_greek: 'curve.pillar:" + std::to_string(pillar_index) +
                                        "' fuera de rango (la curva tiene " +
                                        std::to_string(market.pillars().size()) + " pillars)");
    }
    if (risk_factor.kind == RiskFactorKind::CurveParallel || risk_factor.kind == RiskFactorKind::CurvePillar ||
        risk_factor.kind == RiskFactorKind::CreditParameter)
        return override_h.value_or(default_curve_bump());
    throw std::invalid_argument("compute_greek: RiskFactorKind de '" + to_string(risk_factor) +
                                "' no soportado para bump-and-reval");
}

// (model, market) with one risk factor moved by h (greeks.cpp:262-330); a second call on a bumped
// state moves a second factor on top of the first (the cross stencil).
struct BumpedState {
    std::shared_ptr<IModel> owned_model;
    const IModel* model;
    MarketSnapshot market;
    BumpedState(const IModel& base_model, MarketSnapshot base_market)
        : model(&base_model), market(std::move(base_market)) {}
};

BumpedState bump_state(const Registries& registries, const BumpedState& base, const RiskFactor& risk_factor,
                       double h) {
    BumpedState state = base;
    switch (risk_factor.kind) {
        case RiskFactorKind::ModelParameter: {
            Params params = base.model->to_params();
            auto it = params.find(risk_factor.name);
            if (it == params.end() || !std::holds_alternative<double>(it->second))
                throw std::invalid_argument("compute_greek: el modelo '" + base.model->type_name() +
                                            "' no tiene el parametro de riesgo '" + to_string(risk_factor) + "'");
            params[risk_factor.name] = std::get<double>(it->second) + h;
            state.owned_model = registries.models.create(base.model->type_name(), params);
            state.model = state.owned_model.get();
            break;
        }
        case RiskFactorKind::CurveParallel:
            state.market = bump_market_parallel(base.market, h);
            break;
        case RiskFactorKind::CurvePillar:
            state.market = bump_market_pillar(base.market, *risk_factor.pillar_index, h);
            break;
        case RiskFactorKind::CreditParameter:
            state.market = bump_market_credit(base.market, risk_factor.name, h);
            break;
        default:
            throw std::invalid_argument("compute_greek: RiskFactorKind de '" + to_string(risk_factor) +
                                        "' no soportado para bump-and-reval");
    }
    return state;
}

void require_same_shape(const MeasureResult& a, const MeasureResult& b, const std::string& metric_name) {
    if (a.has_scalar != b.has_scalar || a.times.size() != b.times.size() || a.primary.size() != b.primary.size() ||
        a.secondary.size() != b.secondary.size())
        throw std::invalid_argument("compute_greek: la medida '" + metric_name +
                                    "' devolvio formas incompatibles entre evaluaciones bumpeadas");
}

// The pathwise route of this engine: UnilateralCVA x HullWhite1F x {a, b, sigma, r0}, first order.
std::optional<GreekResult> try_pathwise(const GreekRequest& request, const IModel& model, const IProduct& product,
                                        const MarketSnapshot& market, const PricingContext& pricing,
                                        const ExecutionContext& execution) {
    if (request.metric_name != "UnilateralCVA" || model.type_name() != "HullWhite1F" ||
        product.type_name() != "IRSwap" || execution.backend() != "gpu")
        return std::nullopt;
    static const char* names[4] = {"a", "b", "sigma", "r0"};
    int idx = -1;
    for (int i = 0; i < 4; ++i)
        if (request.risk_factor.name == names[i]) idx = i;
    if (idx < 0) return std::nullopt;
    CvaAdjoint adj = unilateral_cva_adjoint(model, product, market, pricing, execution);
    GreekResult result;
    result.has_scalar = true;
    result.value = adj.grad[idx];
    result.order = request.order;
    result.risk_factor = request.risk_factor;
    result.method_used = GreekMethod::Pathwise;
    return result;
}

GreekResult difference(const MeasureResult& up, const MeasureResult& down, const MeasureResult* base, double h,
                       const GreekRequest& request) {
    const bool gamma = base != nullptr;
    const double den = gamma ? h * h : 2.0 * h;
    auto d = [&](double u, double dn, double b) { return gamma ? (u - 2.0 * b + dn) / den : (u - dn) / den; };
    GreekResult result;
    result.has_scalar = up.has_scalar;
    if (up.has_scalar) result.value = d(up.scalar, down.scalar, gamma ? base->scalar : 0.0);
    result.times = up.times;
    for (std::size_t i = 0; i < up.primary.size(); ++i)
        result.primary.push_back(d(up.primary[i], down.primary[i], gamma ? base->primary[i] : 0.0));
    for (std::size_t i = 0; i < up.secondary.size(); ++i)
        result.secondary.push_back(d(up.secondary[i], down.secondary[i], gamma ? base->secondary[i] : 0.0));
    result.order = request.order;
    result.risk_factor = request.risk_factor;
    result.method_used = GreekMethod::BumpAndReval;
    result.bump_used = h;
    return result;
}

}  // namespace

// greeks.cpp:954-1010
RiskFactor parse_risk_factor(const std::string& text) {
    auto dot = text.find('.');
    if (dot == std::string::npos)
        throw std::invalid_argument("RiskFactor: '" + text +
                                    "' no tiene el formato '<scope>.<nombre>' esperado "
                                    "(scopes validos: model, curve, credit, time)");
    const std::string scope = text.substr(0, dot), rest = text.substr(dot + 1);
    if (scope == "model") {
        if (rest.empty())
            throw std::invalid_argument("RiskFactor: 'model.' requiere un nombre de parametro, p.ej. 'model.spot'");
        return RiskFactor{RiskFactorKind::ModelParameter, "model", rest, std::nullopt};
    }
    if (scope == "curve") {
        if (rest == "parallel") return RiskFactor{RiskFactorKind::CurveParallel, "curve", "", std::nullopt};
        const std::string prefix = "pillar:";
        if (rest.rfind(prefix, 0) == 0) {
            const std::string idx_str = rest.substr(prefix.size());
            try {
                std::size_t consumed = 0;
                long idx = std::stol(idx_str, &consumed);
                if (idx < 0 || consumed != idx_str.size()) throw std::invalid_argument("");
                return RiskFactor{RiskFactorKind::CurvePillar, "curve", "", static_cast<std::size_t>(idx)};
            } catch (...) {
                throw std::invalid_argument(
                    "RiskFactor: 'curve.pillar:<i>' requiere un indice entero no negativo, recibido 'curve." + rest +
                    "'");
            }
        }
        throw std::invalid_argument("RiskFactor: 'curve." + rest +
                                    "' desconocido (validos: 'curve.parallel', 'curve.pillar:<i>')");
    }
    if (scope == "credit") {
        if (rest == "hazard_rate" || rest == "recovery_rate")
            return RiskFactor{RiskFactorKind::CreditParameter, "credit", rest, std::nullopt};
        throw std::invalid_argument("RiskFactor: 'credit." + rest +
                                    "' desconocido (validos: 'credit.hazard_rate', 'credit.recovery_rate')");
    }
    if (scope == "time") {
        if (rest == "theta") return RiskFactor{RiskFactorKind::TimeShift, "time", "", std::nullopt};
        throw std::invalid_argument("RiskFactor: 'time." + rest + "' desconocido (valido: 'time.theta')");
    }
    throw std::invalid_argument("RiskFactor: scope '" + scope +
                                "' desconocido (validos: 'model', 'curve', 'credit', 'time')");
}

std::string to_string(const RiskFactor& risk_factor) {
    switch (risk_factor.kind) {
        case RiskFactorKind::ModelParameter: return "model." + risk_factor.name;
        case RiskFactorKind::CurveParallel: return "curve.parallel";
        case RiskFactorKind::CurvePillar: return "curve.pillar:" + std::to_string(risk_factor.pillar_index.value_or(0));
        case RiskFactorKind::CreditParameter: return "credit." + risk_factor.name;
        case RiskFactorKind::TimeShift: return "time.theta";
    }
    throw std::logic_error("to_string(RiskFactor): RiskFactorKind desconocido");
}

// greeks.cpp:1028-1051
GreekMethod parse_greek_method(const std::string& text) {
    if (text == "auto") return GreekMethod::Auto;
    if (text == "bump_and_reval") return GreekMethod::BumpAndReval;
    if (text == "pathwise") return GreekMethod::Pathwise;
    if (text == "aad") return GreekMethod::AadReverse;
    if (text == "likelihood_ratio_hessian") return GreekMethod::LikelihoodRatioHessian;
    if (text == "aad_forward_over_forward") return GreekMethod::AadForwardOverForward;
    throw std::invalid_argument("GreekMethod: '" + text +
                                "' desconocido (validos: 'auto', 'bump_and_reval', 'pathwise', 'aad', "
                                "'likelihood_ratio_hessian', 'aad_forward_over_forward')");
}

std::string to_string(GreekMethod method) {
    switch (method) {
        case GreekMethod::Auto: return "auto";
        case GreekMethod::BumpAndReval: return "bump_and_reval";
        case GreekMethod::Pathwise: return "pathwise";
        case GreekMethod::AadReverse: return "aad";
        case GreekMethod::LikelihoodRatioHessian: return "likelihood_ratio_hessian";
        case GreekMethod::AadForwardOverForward: return "aad_forward_over_forward";
    }
    throw std::logic_error("to_string(GreekMethod): GreekMethod desconocido");
}

// greeks.cpp:1062-1343
GreekResult compute_greek(const Registries& registries, const GreekRequest& request, const IModel& model,
                          const IProduct& product, const MarketSnapshot& market, const PricingContext& pricing,
                          const ExecutionContext& execution) {
    if (request.order.order < 1 || request.order.order > 2)
        throw std::invalid_argument("compute_greek: order debe ser 1 o 2 (pedido: order=" +
                                    std::to_string(request.order.order) + ")");
    if (request.order.order == 2 && request.order.cross_factor.has_value())
        throw std::invalid_argument("compute_greek: order=2 con cross_factor no soportado (tercera derivada, fuera "
                                    "de alcance)");
    const bool is_time_factor = request.risk_factor.kind == RiskFactorKind::TimeShift;
    if (is_time_factor) {
        if (request.order.order != 1 || request.order.cross_factor.has_value())
            throw std::invalid_argument("compute_greek: 'time.theta' solo soporta order=1 sin cross_factor");
        if (!metric_supports_time_shift(request.metric_name))
            throw std::invalid_argument("compute_greek: 'time.theta' todavia no esta cableado para la metrica '" +
                                        request.metric_name +
                                        "' (pedirlo daria un Theta silenciosamente nulo, asi que se rechaza)");
    } else {
        if (!is_second_order_capable_kind(request.risk_factor.kind))
            throw std::invalid_argument("compute_greek: RiskFactorKind de '" + to_string(request.risk_factor) +
                                        "' no soportado todavia");
        if (request.order.cross_factor.has_value() &&
            !is_second_order_capable_kind(request.order.cross_factor->kind))
            throw std::invalid_argument("compute_greek: cross_factor '" + to_string(*request.order.cross_factor) +
                                        "' tiene un RiskFactorKind no soportado para derivadas cruzadas");
    }
    if (request.method == GreekMethod::Pathwise || request.method == GreekMethod::AadReverse) {
        if (request.risk_factor.kind != RiskFactorKind::ModelParameter)
            throw std::invalid_argument("compute_greek: metodo '" + to_string(request.method) +
                                        "' solo soporta RiskFactorKind::ModelParameter (pedido: '" +
                                        to_string(request.risk_factor) + "')");
    }
    if (request.method == GreekMethod::AadReverse || request.method == GreekMethod::LikelihoodRatioHessian ||
        request.method == GreekMethod::AadForwardOverForward)
        throw std::invalid_argument("compute_greek: metodo '" + to_string(request.method) +
                                    "' no disponible para la metrica '" + request.metric_name + "' del modelo '" +
                                    model.type_name() + "' (use 'auto', 'bump_and_reval' o 'pathwise')");

    auto metric = registries.measures.create(request.metric_name, request.metric_params);

    if (is_time_factor) {  // one-sided: Metric(t + dt) - Metric(t), not divided (greeks.cpp:1142-1171)
        const double h = request.bump_override.value_or(default_time_shift_bump());
        PricingContext shifted(Params{{"pricing_date", pricing.pricing_date() + h},
                                      {"n_paths", static_cast<double>(pricing.n_paths())},
                                      {"n_steps", static_cast<double>(pricing.n_steps())},
                                      {"seed", static_cast<double>(pricing.seed())}});
        MeasureResult up = metric->evaluate(model, product, market, shifted, execution);
        MeasureResult down = metric->evaluate(model, product, market, pricing, execution);
        require_same_shape(up, down, request.metric_name);
        GreekResult result;
        result.has_scalar = up.has_scalar;
        if (up.has_scalar) result.value = up.scalar - down.scalar;
        result.times = up.times;
        for (std::size_t i = 0; i < up.primary.size(); ++i) result.primary.push_back(up.primary[i] - down.primary[i]);
        for (std::size_t i = 0; i < up.secondary.size(); ++i)
            result.secondary.push_back(up.secondary[i] - down.secondary[i]);
        result.order = request.order;
        result.risk_factor = request.risk_factor;
        result.method_used = GreekMethod::BumpAndReval;
        result.bump_used = h;
        return result;
    }

    const double h1 = resolve_bump(model, market, request.risk_factor, request.bump_override);

    if (request.method == GreekMethod::Pathwise) {
        if (request.order.order == 1 && !request.order.cross_factor.has_value()) {
            std::optional<GreekResult> r = try_pathwise(request, model, product, market, pricing, execution);
            if (r.has_value()) return *r;
        }
        throw std::invalid_argument(
            "compute_greek: metodo 'pathwise' no disponible para la metrica '" + request.metric_name +
            "' del modelo '" + model.type_name() + "' con order=" + std::to_string(request.order.order) +
            (request.order.cross_factor.has_value() ? " y cross_factor" : "") +
            " (este motor lo ofrece para UnilateralCVA de HullWhite1F, model.{a,b,sigma,r0}, order=1, backend gpu)");
    }

    const BumpedState base_state(model, market);
    if (request.order.cross_factor.has_value()) {  // 4-point stencil (greeks.cpp:1236-1291)
        const RiskFactor& cross = *request.order.cross_factor;
        const double h2 = resolve_bump(model, market, cross, std::nullopt);
        auto eval2 = [&](double s1, double s2) {
            BumpedState a = bump_state(registries, base_state, request.risk_factor, s1 * h1);
            BumpedState b = bump_state(registries, a, cross, s2 * h2);
            return metric->evaluate(*b.model, product, b.market, pricing, execution);
        };
        MeasureResult uu = eval2(1, 1), ud = eval2(1, -1), du = eval2(-1, 1), dd = eval2(-1, -1);
        require_same_shape(uu, ud, request.metric_name);
        require_same_shape(uu, du, request.metric_name);
        require_same_shape(uu, dd, request.metric_name);
        const double den = 4.0 * h1 * h2;
        GreekResult result;
        result.has_scalar = uu.has_scalar;
        if (uu.has_scalar) result.value = (uu.scalar - ud.scalar - du.scalar + dd.scalar) / den;
        result.times = uu.times;
        for (std::size_t i = 0; i < uu.primary.size(); ++i)
            result.primary.push_back((uu.primary[i] - ud.primary[i] - du.primary[i] + dd.primary[i]) / den);
        for (std::size_t i = 0; i < uu.secondary.size(); ++i)
            result.secondary.push_back((uu.secondary[i] - ud.secondary[i] - du.secondary[i] + dd.secondary[i]) / den);
        result.order = request.order;
        result.risk_factor = request.risk_factor;
        result.method_used = GreekMethod::BumpAndReval;
        result.bump_used = h1;
        result.warnings.push_back("cross_factor '" + to_string(cross) + "' bumpeado en +-" + std::to_string(h2));
        return result;
    }

    const BumpedState s_up = bump_state(registries, base_state, request.risk_factor, h1);
    const BumpedState s_down = bump_state(registries, base_state, request.risk_factor, -h1);
    MeasureResult up = metric->evaluate(*s_up.model, product, s_up.market, pricing, execution);
    MeasureResult down = metric->evaluate(*s_down.model, product, s_down.market, pricing, execution);
    require_same_shape(up, down, request.metric_name);
    if (!up.has_scalar && up.times.empty() && up.primary.empty() && up.secondary.empty())
        throw std::invalid_argument("compute_greek: la medida '" + request.metric_name +
                                    "' no produce ningun resultado (ni escalar ni perfil)");
    if (request.order.order == 2) {
        MeasureResult base = metric->evaluate(model, product, market, pricing, execution);
        require_same_shape(up, base, request.metric_name);
        return difference(up, down, &base, h1, request);
    }
    return difference(up, down, nullptr, h1, request);
}

// greeks.cpp:1345-1430: every double model parameter (sorted), curve.parallel (+ pillars on request),
// both credit parameters and theta; a candidate that throws goes to `skipped` with its reason.
GreeksReport compute_all_greeks(const Registries& registries, const std::string& metric_name,
                                const Params& metric_params, const IModel& model, const IProduct& product,
                                const MarketSnapshot& market, const PricingContext& pricing,
                                const ExecutionContext& execution, bool include_curve_buckets,
                                bool include_second_order) {
    GreeksReport report;
    auto try_request = [&](const GreekRequest& request) {
        try {
            report.greeks.push_back(compute_greek(registries, request, model, product, market, pricing, execution));
        } catch (const std::exception& e) {
            report.skipped.push_back(to_string(request.risk_factor) + " (order=" +
                                     std::to_string(request.order.order) + "): " + e.what());
        }
    };
    auto candidate = [&](const RiskFactor& rf, int order) {
        GreekRequest request;
        request.metric_name = metric_name;
        request.metric_params = metric_params;
        request.risk_factor = rf;
        request.order = GreekOrder{order, std::nullopt};
        request.method = GreekMethod::Auto;
        try_request(request);
    };
    const Params model_params = model.to_params();
    std::vector<std::string> keys;
    for (const auto& kv : model_params)
        if (std::holds_alternative<double>(kv.second)) keys.push_back(kv.first);
    std::sort(keys.begin(), keys.end());
    for (const std::string& key : keys) {
        const RiskFactor rf{RiskFactorKind::ModelParameter, "model", key, std::nullopt};
        candidate(rf, 1);
        if (include_second_order) candidate(rf, 2);
    }
    candidate(RiskFactor{RiskFactorKind::CurveParallel, "curve", "", std::nullopt}, 1);
    if (include_curve_buckets)
        for (std::size_t i = 0; i < market.pillars().size(); ++i)
            candidate(RiskFactor{RiskFactorKind::CurvePillar, "curve", "", i}, 1);
    candidate(RiskFactor{RiskFactorKind::CreditParameter, "credit", "hazard_rate", std::nullopt}, 1);
    candidate(RiskFactor{RiskFactorKind::CreditParameter, "credit", "recovery_rate", std::nullopt}, 1);
    candidate(RiskFactor{RiskFactorKind::TimeShift, "time", "", std::nullopt}, 1);
    return report;
}

}  // namespace greeks

// greeks.cpp:1670-1714
GreekMeasure::GreekMeasure(const Params& params, const Registries& registries) : registries_(registries) {
    request_.metric_name = get_string(params, "metric");
    Params inner;
    const std::string prefix = "metric.";
    for (const auto& kv : params)
        if (kv.first.rfind(prefix, 0) == 0) inner[kv.first.substr(prefix.size())] = kv.second;
    request_.metric_params = std::move(inner);
    request_.risk_factor = greeks::parse_risk_factor(get_string(params, "risk_factor"));
    request_.order = greeks::GreekOrder{static_cast<int>(get_double(params, "order", 1.0)), std::nullopt};
    if (contains(params, "cross_factor"))
        request_.order.cross_factor = greeks::parse_risk_factor(get_string(params, "cross_factor"));
    request_.method = greeks::parse_greek_method(get_string(params, "method", "auto"));
    if (contains(params, "bump")) request_.bump_override = get_double(params, "bump");
}

MeasureResult GreekMeasure::evaluate(const IModel& model, const IProduct& product, const MarketSnapshot& market,
                                     const PricingContext& pricing, const ExecutionContext& execution) const {
    greeks::GreekResult out = greeks::compute_greek(registries_, request_, model, product, market, pricing, execution);
    MeasureResult result;
    result.has_scalar = out.has_scalar;
    result.scalar = out.value;
    result.times = out.times;
    result.primary = out.primary;
    result.secondary = out.secondary;
    result.bump_used = out.bump_used;
    return result;
}

}  // namespace engine
