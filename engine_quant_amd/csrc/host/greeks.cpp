// Greeks of the registered measures, by finite differences over GPU evaluations, plus a pathwise route
// for UnilateralCVA served by the GPU CVA adjoint.
//
// Design. A request is turned into a stencil: a list of evaluation points, each a weight and a signed
// multiple of the step along one or two factor axes, and a divisor. An axis is a direction in the space of
// (model, market, pricing date): a model parameter, the whole zero curve, one pillar, a credit parameter or
// the clock. Every point is one evaluation of the metric at the shifted inputs; the Greek is the weighted
// sum of the points' results over the divisor, element by element over the scalar and the two profiles.
// The Philox stream is counter-based, so every shifted evaluation draws the same shocks: common random
// numbers without bookkeeping.
//
// What callers rely on is kept as the reference states it (cpp/engine/src/greeks.cpp, greeks.hpp): the
// risk-factor strings, the default steps (greeks.cpp:160-171), the stencils (central first order,
// second order about the base, the four-point cross, one-sided theta) and the error texts.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <memory>

#include "engine.hpp"

namespace engine {
namespace greeks {

namespace {

// ---- the risk-factor language -----------------------------------------------------------------
// "<scope>.<rest>" with scopes model, curve, credit and time.

[[noreturn]] void bad_factor(const std::string& msg) { throw std::invalid_argument("RiskFactor: " + msg); }

RiskFactor model_factor(const std::string& rest) {
    if (rest.empty()) bad_factor("'model.' requiere un nombre de parametro, p.ej. 'model.spot'");
    return RiskFactor{RiskFactorKind::ModelParameter, "model", rest, std::nullopt};
}

RiskFactor curve_factor(const std::string& rest) {
    static const std::string pillar_tag = "pillar:";
    if (rest == "parallel") return RiskFactor{RiskFactorKind::CurveParallel, "curve", "", std::nullopt};
    if (rest.compare(0, pillar_tag.size(), pillar_tag) == 0) {
        const std::string digits = rest.substr(pillar_tag.size());
        const bool ok = !digits.empty() && digits.size() <= 9 &&
                        std::all_of(digits.begin(), digits.end(), [](unsigned char c) { return std::isdigit(c); });
        if (!ok)
            bad_factor("'curve.pillar:<i>' requiere un indice entero no negativo, recibido 'curve." + rest + "'");
        return RiskFactor{RiskFactorKind::CurvePillar, "curve", "", static_cast<std::size_t>(std::stoul(digits))};
    }
    bad_factor("'curve." + rest + "' desconocido (validos: 'curve.parallel', 'curve.pillar:<i>')");
}

RiskFactor credit_factor(const std::string& rest) {
    if (rest != "hazard_rate" && rest != "recovery_rate")
        bad_factor("'credit." + rest + "' desconocido (validos: 'credit.hazard_rate', 'credit.recovery_rate')");
    return RiskFactor{RiskFactorKind::CreditParameter, "credit", rest, std::nullopt};
}

RiskFactor time_factor(const std::string& rest) {
    if (rest != "theta") bad_factor("'time." + rest + "' desconocido (valido: 'time.theta')");
    return RiskFactor{RiskFactorKind::TimeShift, "time", "", std::nullopt};
}

struct ScopeRule {
    const char* scope;
    RiskFactor (*make)(const std::string&);
};
constexpr ScopeRule kScopes[] = {
    {"model", model_factor}, {"curve", curve_factor}, {"credit", credit_factor}, {"time", time_factor}};

// ---- axes -------------------------------------------------------------------------------------
// The inputs one evaluation sees. A shifted model is owned by the point that built it.
struct Inputs {
    std::shared_ptr<const IModel> owned;
    const IModel* model = nullptr;
    MarketSnapshot market;
    PricingContext pricing;
};

std::string factor_error(const RiskFactor& rf) {
    return "compute_greek: RiskFactorKind de '" + to_string(rf) + "' no soportado para bump-and-reval";
}

// A direction in input space: its default step at the base inputs, and how it moves a set of inputs.
class Axis {
public:
    virtual ~Axis() = default;
    virtual double default_step() const = 0;
    virtual void move(Inputs& in, double delta) const = 0;
};

class ModelParameterAxis final : public Axis {
public:
    ModelParameterAxis(const Registries& registries, const IModel& base, const RiskFactor& rf)
        : registries_(registries), key_(rf.name) {
        base_value_ = read(base, rf);
    }
    double default_step() const override { return std::max(1e-2 * std::abs(base_value_), 1e-4); }
    void move(Inputs& in, double delta) const override {
        Params p = in.model->to_params();
        p[key_] = std::get<double>(p.at(key_)) + delta;
        in.owned = registries_.models.create(in.model->type_name(), p);
        in.model = in.owned.get();
    }

private:
    static double read(const IModel& model, const RiskFactor& rf) {
        const Params p = model.to_params();
        const auto it = p.find(rf.name);
        if (it == p.end() || !std::holds_alternative<double>(it->second))
            throw std::invalid_argument("compute_greek: el modelo '" + model.type_name() +
                                        "' no tiene el parametro de riesgo '" + to_string(rf) + "'");
        return std::get<double>(it->second);
    }
    const Registries& registries_;
    std::string key_;
    double base_value_ = 0.0;
};

class CurveAxis final : public Axis {
public:
    explicit CurveAxis(std::optional<std::size_t> pillar) : pillar_(pillar) {}
    double default_step() const override { return 1e-4; }
    void move(Inputs& in, double delta) const override {
        in.market = pillar_ ? bump_market_pillar(in.market, *pillar_, delta) : bump_market_parallel(in.market, delta);
    }

private:
    std::optional<std::size_t> pillar_;
};

class CreditAxis final : public Axis {
public:
    explicit CreditAxis(std::string name) : name_(std::move(name)) {}
    double default_step() const override { return 1e-4; }
    void move(Inputs& in, double delta) const override { in.market = bump_market_credit(in.market, name_, delta); }

private:
    std::string name_;
};

// The clock: the pricing date moves, the curve observed today stays (one day by default).
class ClockAxis final : public Axis {
public:
    double default_step() const override { return 1.0 / 365.0; }
    void move(Inputs& in, double delta) const override {
        const PricingContext& p = in.pricing;
        in.pricing = PricingContext(Params{{"pricing_date", p.pricing_date() + delta},
                                           {"n_paths", static_cast<double>(p.n_paths())},
                                           {"n_steps", static_cast<double>(p.n_steps())},
                                           {"seed", static_cast<double>(p.seed())}});
    }
};

std::unique_ptr<Axis> axis_for(const Registries& registries, const RiskFactor& rf, const IModel& model,
                               const MarketSnapshot& market) {
    switch (rf.kind) {
        case RiskFactorKind::ModelParameter:
            return std::make_unique<ModelParameterAxis>(registries, model, rf);
        case RiskFactorKind::CurveParallel:
            return std::make_unique<CurveAxis>(std::nullopt);
        case RiskFactorKind::CurvePillar: {
            const std::size_t i = rf.pillar_index.value_or(0);
            if (i >= market.pillars().size())
                throw std::invalid_argument("compute_greek: 'curve.pillar:" + std::to_string(i) +
                                            "' fuera de rango (la curva tiene " +
                                            std::to_string(market.pillars().size()) + " pillars)");
            return std::make_unique<CurveAxis>(i);
        }
        case RiskFactorKind::CreditParameter:
            return std::make_unique<CreditAxis>(rf.name);
        case RiskFactorKind::TimeShift:
            return std::make_unique<ClockAxis>();
    }
    throw std::invalid_argument(factor_error(rf));
}

// ---- stencils ---------------------------------------------------------------------------------
struct Point {
    double weight;
    int m1, m2;  // signed multiples of the first and second axis step
};

struct Stencil {
    std::vector<Point> points;
    double divisor;
};

Stencil first_order(double h) { return {{{1.0, 1, 0}, {-1.0, -1, 0}}, 2.0 * h}; }
Stencil second_order(double h) { return {{{1.0, 1, 0}, {-2.0, 0, 0}, {1.0, -1, 0}}, h * h}; }
Stencil cross_order(double h1, double h2) {
    return {{{1.0, 1, 1}, {-1.0, 1, -1}, {-1.0, -1, 1}, {1.0, -1, -1}}, 4.0 * h1 * h2};
}
Stencil one_sided(double) { return {{{1.0, 1, 0}, {-1.0, 0, 0}}, 1.0}; }  // M(x + h) - M(x), not divided

bool same_shape(const MeasureResult& a, const MeasureResult& b) {
    return a.has_scalar == b.has_scalar && a.times.size() == b.times.size() &&
           a.primary.size() == b.primary.size() && a.secondary.size() == b.secondary.size();
}

// Evaluates the metric at every point of the stencil and combines the results.
GreekResult apply_stencil(const IMeasure& metric, const std::string& metric_name, const Stencil& stencil,
                          const Axis& a1, double h1, const Axis* a2, double h2, const IModel& model,
                          const IProduct& product, const MarketSnapshot& market, const PricingContext& pricing,
                          const ExecutionContext& execution) {
    std::vector<MeasureResult> at;
    at.reserve(stencil.points.size());
    for (const Point& pt : stencil.points) {
        Inputs in{nullptr, &model, market, pricing};
        if (pt.m1 != 0) a1.move(in, pt.m1 * h1);
        if (pt.m2 != 0) a2->move(in, pt.m2 * h2);
        at.push_back(metric.evaluate(*in.model, product, in.market, in.pricing, execution));
        if (!same_shape(at.front(), at.back()))
            throw std::invalid_argument("compute_greek: la medida '" + metric_name +
                                        "' devolvio formas incompatibles entre evaluaciones bumpeadas");
    }
    const MeasureResult& shape = at.front();
    if (!shape.has_scalar && shape.times.empty() && shape.primary.empty() && shape.secondary.empty())
        throw std::invalid_argument("compute_greek: la medida '" + metric_name +
                                    "' no produce ningun resultado (ni escalar ni perfil)");
    auto combine = [&](auto field) {
        double acc = 0.0;
        for (std::size_t p = 0; p < at.size(); ++p) acc += stencil.points[p].weight * field(at[p]);
        return acc / stencil.divisor;
    };
    GreekResult out;
    out.has_scalar = shape.has_scalar;
    if (shape.has_scalar) out.value = combine([](const MeasureResult& r) { return r.scalar; });
    out.times = shape.times;
    for (std::size_t i = 0; i < shape.primary.size(); ++i)
        out.primary.push_back(combine([i](const MeasureResult& r) { return r.primary[i]; }));
    for (std::size_t i = 0; i < shape.secondary.size(); ++i)
        out.secondary.push_back(combine([i](const MeasureResult& r) { return r.secondary[i]; }));
    out.method_used = GreekMethod::BumpAndReval;
    out.bump_used = h1;
    return out;
}

// ---- the pathwise route: UnilateralCVA of a swap under HullWhite1F in a, b, sigma, r0 -----------
std::optional<GreekResult> pathwise(const GreekRequest& request, const IModel& model, const IProduct& product,
                                    const MarketSnapshot& market, const PricingContext& pricing,
                                    const ExecutionContext& execution) {
    static const std::string adjoint_order[4] = {"a", "b", "sigma", "r0"};
    const bool served = request.metric_name == "UnilateralCVA" && model.type_name() == "HullWhite1F" &&
                        product.type_name() == "IRSwap" && execution.backend() == "gpu" &&
                        request.order.order == 1 && !request.order.cross_factor;
    if (!served) return std::nullopt;
    const auto at = std::find(std::begin(adjoint_order), std::end(adjoint_order), request.risk_factor.name);
    if (at == std::end(adjoint_order)) return std::nullopt;
    const CvaAdjoint adj = unilateral_cva_adjoint(model, product, market, pricing, execution);
    GreekResult out;
    out.has_scalar = true;
    out.value = adj.grad[at - std::begin(adjoint_order)];
    out.method_used = GreekMethod::Pathwise;
    return out;
}

// Checks of the request that do not need the inputs; the messages are the reference's.
void validate(const GreekRequest& request) {
    const int order = request.order.order;
    const bool cross = request.order.cross_factor.has_value();
    if (order != 1 && order != 2)
        throw std::invalid_argument("compute_greek: order debe ser 1 o 2 (pedido: order=" + std::to_string(order) + ")");
    if (order == 2 && cross)
        throw std::invalid_argument(
            "compute_greek: order=2 con cross_factor no soportado (tercera derivada, fuera de alcance)");
    if (request.risk_factor.kind == RiskFactorKind::TimeShift) {
        if (order != 1 || cross)
            throw std::invalid_argument("compute_greek: 'time.theta' solo soporta order=1 sin cross_factor");
        // the metrics that honour PricingContext::pricing_date (greeks.cpp:1081-1089: PV on this path)
        if (request.metric_name != "PV")
            throw std::invalid_argument("compute_greek: 'time.theta' todavia no esta cableado para la metrica '" +
                                        request.metric_name +
                                        "' (pedirlo daria un Theta silenciosamente nulo, asi que se rechaza)");
    }
    if (cross && request.order.cross_factor->kind == RiskFactorKind::TimeShift)
        throw std::invalid_argument("compute_greek: cross_factor '" + to_string(*request.order.cross_factor) +
                                    "' tiene un RiskFactorKind no soportado para derivadas cruzadas");
    if (request.method == GreekMethod::Pathwise && request.risk_factor.kind != RiskFactorKind::ModelParameter)
        throw std::invalid_argument("compute_greek: metodo 'pathwise' solo soporta RiskFactorKind::ModelParameter "
                                    "(pedido: '" + to_string(request.risk_factor) + "')");
}

}  // namespace

RiskFactor parse_risk_factor(const std::string& text) {
    const std::size_t dot = text.find('.');
    if (dot == std::string::npos)
        bad_factor("'" + text + "' no tiene el formato '<scope>.<nombre>' esperado "
                   "(scopes validos: model, curve, credit, time)");
    const std::string scope = text.substr(0, dot);
    for (const ScopeRule& rule : kScopes)
        if (scope == rule.scope) return rule.make(text.substr(dot + 1));
    bad_factor("scope '" + scope + "' desconocido (validos: 'model', 'curve', 'credit', 'time')");
}

std::string to_string(const RiskFactor& rf) {
    switch (rf.kind) {
        case RiskFactorKind::CurvePillar:
            return "curve.pillar:" + std::to_string(rf.pillar_index.value_or(0));
        case RiskFactorKind::CurveParallel:
            return "curve.parallel";
        case RiskFactorKind::TimeShift:
            return "time.theta";
        case RiskFactorKind::ModelParameter:
        case RiskFactorKind::CreditParameter:
            return rf.scope + "." + rf.name;
    }
    throw std::logic_error("to_string(RiskFactor): RiskFactorKind desconocido");
}

namespace {
struct MethodName {
    const char* text;
    GreekMethod method;
};
constexpr MethodName kMethods[] = {
    {"auto", GreekMethod::Auto}, {"bump_and_reval", GreekMethod::BumpAndReval}, {"pathwise", GreekMethod::Pathwise}};
// names of the reference's other methods (greeks.hpp): known, never served here
constexpr const char* kUnserved[] = {"aad", "likelihood_ratio_hessian", "aad_forward_over_forward"};
}  // namespace

GreekMethod parse_greek_method(const std::string& text) {
    for (const MethodName& m : kMethods)
        if (text == m.text) return m.method;
    for (const char* name : kUnserved)
        if (text == name)
            throw std::invalid_argument("compute_greek: metodo '" + text +
                                        "' no disponible en este motor (use 'auto', 'bump_and_reval' o 'pathwise')");
    throw std::invalid_argument("GreekMethod: '" + text + "' desconocido (validos: 'auto', 'bump_and_reval', 'pathwise')");
}

std::string to_string(GreekMethod method) {
    for (const MethodName& m : kMethods)
        if (m.method == method) return m.text;
    throw std::logic_error("to_string(GreekMethod): GreekMethod desconocido");
}

GreekResult compute_greek(const Registries& registries, const GreekRequest& request, const IModel& model,
                          const IProduct& product, const MarketSnapshot& market, const PricingContext& pricing,
                          const ExecutionContext& execution) {
    validate(request);
    const auto metric = registries.measures.create(request.metric_name, request.metric_params);
    const auto a1 = axis_for(registries, request.risk_factor, model, market);
    const double h1 = request.bump_override.value_or(a1->default_step());
    const bool clock = request.risk_factor.kind == RiskFactorKind::TimeShift;

    GreekResult out;
    if (request.method == GreekMethod::Pathwise) {
        std::optional<GreekResult> pw = pathwise(request, model, product, market, pricing, execution);
        if (!pw)
            throw std::invalid_argument(
                "compute_greek: metodo 'pathwise' no disponible para la metrica '" + request.metric_name +
                "' del modelo '" + model.type_name() + "' con order=" + std::to_string(request.order.order) +
                (request.order.cross_factor ? " y cross_factor" : "") +
                " (este motor lo ofrece para UnilateralCVA de HullWhite1F, model.{a,b,sigma,r0}, order=1, backend gpu)");
        out = std::move(*pw);
    } else if (request.order.cross_factor) {
        const RiskFactor& second = *request.order.cross_factor;
        const auto a2 = axis_for(registries, second, model, market);
        const double h2 = a2->default_step();
        out = apply_stencil(*metric, request.metric_name, cross_order(h1, h2), *a1, h1, a2.get(), h2, model, product,
                            market, pricing, execution);
        out.warnings.push_back("cross_factor '" + to_string(second) + "' bumpeado en +-" + std::to_string(h2));
    } else {
        const Stencil st = clock ? one_sided(h1) : (request.order.order == 2 ? second_order(h1) : first_order(h1));
        out = apply_stencil(*metric, request.metric_name, st, *a1, h1, nullptr, 0.0, model, product, market, pricing,
                            execution);
    }
    out.order = request.order;
    out.risk_factor = request.risk_factor;
    return out;
}

// The sweep: every double parameter of the model in key order, the parallel curve (and each pillar on
// request), both credit parameters and theta, each at first order (and second order on request for
// the model parameters); a factor whose Greek raises is listed in `skipped` with the reason.
GreeksReport compute_all_greeks(const Registries& registries, const std::string& metric_name,
                                const Params& metric_params, const IModel& model, const IProduct& product,
                                const MarketSnapshot& market, const PricingContext& pricing,
                                const ExecutionContext& execution, bool include_curve_buckets,
                                bool include_second_order) {
    std::vector<std::pair<RiskFactor, int>> plan;
    std::vector<std::string> keys;
    for (const auto& kv : model.to_params())
        if (std::holds_alternative<double>(kv.second)) keys.push_back(kv.first);
    std::sort(keys.begin(), keys.end());
    for (const std::string& key : keys) {
        const RiskFactor rf = model_factor(key);
        plan.emplace_back(rf, 1);
        if (include_second_order) plan.emplace_back(rf, 2);
    }
    plan.emplace_back(curve_factor("parallel"), 1);
    for (std::size_t i = 0; include_curve_buckets && i < market.pillars().size(); ++i)
        plan.emplace_back(curve_factor("pillar:" + std::to_string(i)), 1);
    plan.emplace_back(credit_factor("hazard_rate"), 1);
    plan.emplace_back(credit_factor("recovery_rate"), 1);
    plan.emplace_back(time_factor("theta"), 1);

    GreeksReport report;
    for (const auto& [rf, order] : plan) {
        GreekRequest request;
        request.metric_name = metric_name;
        request.metric_params = metric_params;
        request.risk_factor = rf;
        request.order.order = order;
        try {
            report.greeks.push_back(compute_greek(registries, request, model, product, market, pricing, execution));
        } catch (const std::exception& e) {
            report.skipped.push_back(to_string(rf) + " (order=" + std::to_string(order) + "): " + e.what());
        }
    }
    return report;
}

}  // namespace greeks

// The "Greek" measure (greeks.hpp:304-318): its parameters are metric, risk_factor, order, cross_factor,
// method and bump; the inner metric's own parameters arrive prefixed "metric.".
GreekMeasure::GreekMeasure(const Params& params, const Registries& registries) : registries_(registries) {
    static const std::string inner = "metric.";
    for (const auto& [key, value] : params)
        if (key.size() > inner.size() && key.compare(0, inner.size(), inner) == 0)
            request_.metric_params.emplace(key.substr(inner.size()), value);
    request_.metric_name = get_string(params, "metric");
    request_.risk_factor = greeks::parse_risk_factor(get_string(params, "risk_factor"));
    request_.order.order = static_cast<int>(get_double(params, "order", 1.0));
    if (contains(params, "cross_factor"))
        request_.order.cross_factor = greeks::parse_risk_factor(get_string(params, "cross_factor"));
    request_.method = greeks::parse_greek_method(get_string(params, "method", "auto"));
    if (contains(params, "bump")) request_.bump_override = get_double(params, "bump");
}

MeasureResult GreekMeasure::evaluate(const IModel& model, const IProduct& product, const MarketSnapshot& market,
                                     const PricingContext& pricing, const ExecutionContext& execution) const {
    const greeks::GreekResult g =
        greeks::compute_greek(registries_, request_, model, product, market, pricing, execution);
    MeasureResult r;
    r.has_scalar = g.has_scalar;
    r.scalar = g.value;
    r.times = g.times;
    r.primary = g.primary;
    r.secondary = g.secondary;
    r.bump_used = g.bump_used;
    return r;
}

}  // namespace engine
