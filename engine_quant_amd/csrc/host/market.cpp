// Curve / MarketSnapshot, cpp/engine/src/market.cpp:11-41,67-111.
#include <algorithm>
#include <cmath>

#include "engine.hpp"

namespace engine {

Curve::Curve(std::vector<double> pillars, std::vector<double> zero_rates)
    : pillars_(std::move(pillars)), zero_rates_(std::move(zero_rates)) {
    if (pillars_.size() != zero_rates_.size())
        throw std::invalid_argument("Curve: pillars y zero_rates deben tener el mismo tamano");
    if (pillars_.empty()) throw std::invalid_argument("Curve: necesita al menos un pillar");
    for (std::size_t i = 1; i < pillars_.size(); ++i)
        if (pillars_[i] <= pillars_[i - 1])
            throw std::invalid_argument("Curve: pillars debe ser estrictamente creciente");
}

// Linear in the zero rate between pillars, flat outside (market.cpp:26-37).
double Curve::zero_rate(double t) const {
    if (t <= pillars_.front()) return zero_rates_.front();
    if (t >= pillars_.back()) return zero_rates_.back();
    std::size_t i = (std::size_t)(std::lower_bound(pillars_.begin(), pillars_.end(), t) - pillars_.begin());
    if (i == 0) i = 1;
    double w = (t - pillars_[i - 1]) / (pillars_[i] - pillars_[i - 1]);
    return zero_rates_[i - 1] + w * (zero_rates_[i] - zero_rates_[i - 1]);
}

double Curve::discount_factor(double t) const { return std::exp(-zero_rate(t) * t); }

MarketSnapshot::MarketSnapshot(std::vector<double> pillars, std::vector<double> zero_rates, double hazard_rate,
                               double recovery_rate)
    : curve_(std::move(pillars), std::move(zero_rates)), hazard_rate_(hazard_rate), recovery_rate_(recovery_rate) {}

MarketSnapshot::MarketSnapshot(const Params& params)
    : MarketSnapshot(get_vector(params, "pillars"), get_vector(params, "zero_rates"),
                     get_double(params, "hazard_rate", 0.0), get_double(params, "recovery_rate", 0.0)) {}

MarketSnapshot bump_market_parallel(const MarketSnapshot& market, double bump) {
    std::vector<double> z = market.zero_rates();
    for (double& v : z) v += bump;
    return MarketSnapshot(market.pillars(), std::move(z), market.hazard_rate(), market.recovery_rate());
}

MarketSnapshot bump_market_pillar(const MarketSnapshot& market, std::size_t pillar_index, double bump) {
    if (pillar_index >= market.pillars().size())
        throw std::invalid_argument("bump_market_pillar: indice de pillar " + std::to_string(pillar_index) +
                                    " fuera de rango (la curva tiene " + std::to_string(market.pillars().size()) +
                                    " pillars)");
    std::vector<double> z = market.zero_rates();
    z[pillar_index] += bump;
    return MarketSnapshot(market.pillars(), std::move(z), market.hazard_rate(), market.recovery_rate());
}

// market.cpp:113-127
MarketSnapshot bump_market_credit(const MarketSnapshot& market, const std::string& name, double bump) {
    double hazard_rate = market.hazard_rate();
    double recovery_rate = market.recovery_rate();
    if (name == "hazard_rate") {
        hazard_rate += bump;
    } else if (name == "recovery_rate") {
        recovery_rate += bump;
    } else {
        throw std::invalid_argument("bump_market_credit: parametro de credito '" + name +
                                    "' desconocido (validos: 'hazard_rate', 'recovery_rate')");
    }
    return MarketSnapshot(market.pillars(), market.zero_rates(), hazard_rate, recovery_rate);
}

}  // namespace engine
