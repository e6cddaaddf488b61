// C++ host mirror of engine_quant's registry-driven pricing API for the exposure path
// (cpp/engine/include/engine/*.hpp): Params bag, Registry<IModel/IProduct/IMeasure>,
// MarketSnapshot, PricingContext, ExecutionContext, the PV / DV01 / ExposureProfile /
// UnilateralCVA measures and price / price_batch / price_many. Names, argument meaning and error
// behaviour follow the reference; the numerics run through the C ABI of include/eqhip.h on the
// GPU. Backend "cpu" names the reference's own engine-core and is not served by this library.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <variant>
#include <vector>

namespace engine {

// ---- params.hpp:15-39 -------------------------------------------------------------------------
using ParamValue = std::variant<double, std::vector<double>, bool, std::string>;
using Params = std::unordered_map<std::string, ParamValue>;

double get_double(const Params& params, const std::string& key);
double get_double(const Params& params, const std::string& key, double default_value);
bool get_bool(const Params& params, const std::string& key, bool default_value);
const std::vector<double>& get_vector(const Params& params, const std::string& key);
const std::string& get_string(const Params& params, const std::string& key);
std::string get_string(const Params& params, const std::string& key, std::string default_value);
bool contains(const Params& params, const std::string& key);

// ---- registry.hpp:17-52 -----------------------------------------------------------------------
template <typename Interface>
class Registry {
public:
    using Factory = std::function<std::unique_ptr<Interface>(const Params&)>;

    void register_factory(std::string name, Factory factory) { factories_[std::move(name)] = std::move(factory); }

    template <typename Concrete>
    void register_type(std::string name) {
        register_factory(std::move(name), [](const Params& p) { return std::make_unique<Concrete>(p); });
    }

    std::unique_ptr<Interface> create(const std::string& name, const Params& params = {}) const {
        auto it = factories_.find(name);
        if (it == factories_.end()) throw std::out_of_range("Registry: tipo no registrado: " + name);
        return it->second(params);
    }

    bool contains(const std::string& name) const { return factories_.count(name) > 0; }

    std::vector<std::string> list() const {
        std::vector<std::string> names;
        names.reserve(factories_.size());
        for (const auto& kv : factories_) names.push_back(kv.first);
        return names;
    }

private:
    std::unordered_map<std::string, Factory> factories_;
};

// ---- market.hpp / market.cpp:11-41,95-111 -----------------------------------------------------
class Curve {
public:
    Curve(std::vector<double> pillars, std::vector<double> zero_rates);
    const std::vector<double>& pillars() const { return pillars_; }
    const std::vector<double>& zero_rates() const { return zero_rates_; }
    double zero_rate(double t) const;
    double discount_factor(double t) const;

private:
    std::vector<double> pillars_;
    std::vector<double> zero_rates_;
};

class MarketSnapshot {
public:
    MarketSnapshot(std::vector<double> pillars, std::vector<double> zero_rates, double hazard_rate = 0.0,
                   double recovery_rate = 0.0);
    explicit MarketSnapshot(const Params& params);
    const Curve& discount_curve() const { return curve_; }
    const std::vector<double>& pillars() const { return curve_.pillars(); }
    const std::vector<double>& zero_rates() const { return curve_.zero_rates(); }
    double zero_rate(double t) const { return curve_.zero_rate(t); }
    double discount_factor(double t) const { return curve_.discount_factor(t); }
    double hazard_rate() const { return hazard_rate_; }
    double recovery_rate() const { return recovery_rate_; }

private:
    Curve curve_;
    double hazard_rate_;
    double recovery_rate_;
};

MarketSnapshot bump_market_parallel(const MarketSnapshot& market, double bump);
MarketSnapshot bump_market_pillar(const MarketSnapshot& market, std::size_t pillar_index, double bump);
MarketSnapshot bump_market_credit(const MarketSnapshot& market, const std::string& name, double bump);

// ---- model.hpp:41-58, product.hpp:49-69 -------------------------------------------------------
class IModel {
public:
    virtual ~IModel() = default;
    virtual std::string type_name() const = 0;
    virtual Params to_params() const = 0;
};

class HullWhite1FModel : public IModel {
public:
    explicit HullWhite1FModel(const Params& params);
    std::string type_name() const override { return "HullWhite1F"; }
    Params to_params() const override;
    double a() const { return a_; }
    double b() const { return b_; }
    double sigma() const { return sigma_; }
    double r0() const { return r0_; }

private:
    double a_, b_, sigma_, r0_;
};

// model.cpp:54-71: G2++ with r = x + y + r0, state (x, y) from (0, 0).
class HullWhite2FModel : public IModel {
public:
    explicit HullWhite2FModel(const Params& params);
    std::string type_name() const override { return "HullWhite2F"; }
    Params to_params() const override;
    double a() const { return a_; }
    double b() const { return b_; }
    double sigma() const { return sigma_; }
    double eta() const { return eta_; }
    double rho() const { return rho_; }
    double r0() const { return r0_; }

private:
    double a_, b_, sigma_, eta_, rho_, r0_;
};

class IProduct {
public:
    virtual ~IProduct() = default;
    virtual std::string type_name() const = 0;
};

class IrSwapProduct : public IProduct {
public:
    explicit IrSwapProduct(const Params& params);
    std::string type_name() const override { return "IRSwap"; }
    double notional() const { return notional_; }
    double fixed_rate() const { return fixed_rate_; }
    bool use_par_rate() const { return use_par_rate_; }
    double start() const { return start_; }
    const std::vector<double>& payment_times() const { return payment_times_; }
    const std::vector<double>& accruals() const { return accruals_; }

private:
    double notional_, fixed_rate_;
    bool use_par_rate_;
    double start_;
    std::vector<double> payment_times_, accruals_;
};

// ---- pricing_context.cpp:7-18, execution_context.cpp:20-39 ------------------------------------
class PricingContext {
public:
    explicit PricingContext(const Params& params);
    double pricing_date() const { return pricing_date_; }
    std::uint64_t n_paths() const { return n_paths_; }
    std::size_t n_steps() const { return n_steps_; }
    std::uint64_t seed() const { return seed_; }

private:
    double pricing_date_;
    std::uint64_t n_paths_;
    std::size_t n_steps_;
    std::uint64_t seed_;
};

// backend in {"cpu","gpu","auto"} ("auto" resolves at construction); precision "fp64" only.
// Execution keys added for the GPU path: rng in {"philox4x32-10"}, discretisation in
// {"euler","exact"}.
class ExecutionContext {
public:
    explicit ExecutionContext(const Params& params);
    const std::string& backend() const { return backend_; }
    const std::string& precision() const { return precision_; }
    const std::string& rng() const { return rng_; }
    const std::string& discretisation() const { return discretisation_; }
    // "per_group" (the reference's price_many: each calendar group simulated on its own grid) or
    // "shared" (every group on one scenario set: the grid of the latest date and the same paths)
    const std::string& simulation() const { return simulation_; }

private:
    std::string backend_, precision_, rng_, discretisation_, simulation_;
};

bool is_gpu_backend_available();

// ---- engine.hpp:42 and the FFI wrappers of engine.cpp:73-113 (GPU arm) ------------------------
struct ExposureProfile {
    std::vector<double> times;
    std::vector<double> ee;
    std::vector<double> pfe_95;
};

ExposureProfile irs_hull_white_exposure_profile(const ExecutionContext& execution, double a, double b,
                                                double sigma, double r0, double notional, double fixed_rate,
                                                bool use_par_rate, double start,
                                                const std::vector<double>& payment_times,
                                                const std::vector<double>& accruals,
                                                const std::vector<double>& monitoring_times, std::size_t n_steps,
                                                std::uint64_t n_paths, std::uint64_t seed);
// Runs task(0) .. task(n - 1) on the calling thread and up to kGpuStreams - 1 persistent worker threads,
// each with its own eqhip context and HIP stream, so that the GPU work and host set-up of independent
// tasks overlap. Tasks are claimed in the order claim_order lists them (index order when it is empty;
// largest first balances the streams). Once every task has ended, the exception of the lowest failed
// index is rethrown (the one a serial loop would have thrown). While another call is using the
// workers, or for n < 2, the tasks run serially on the calling thread.
constexpr int kGpuStreams = 4;
void run_on_gpu_streams(std::size_t n, const std::function<void(std::size_t)>& task,
                        const std::vector<std::size_t>& claim_order = {});
// Every trade's profile at its own reset dates, all on one scenario set: the grid of the latest reset
// date of all trades and the same Philox paths for every calendar group (the shared-simulation form of
// price_many).
std::vector<ExposureProfile> shared_exposure_profiles(const IModel& model,
                                                      const std::vector<const IrSwapProduct*>& products,
                                                      const PricingContext& pricing,
                                                      const ExecutionContext& execution);
std::vector<ExposureProfile> irs_hull_white_exposure_profile_batch(
    const ExecutionContext& execution, double a, double b, double sigma, double r0,
    const std::vector<double>& notionals, const std::vector<double>& fixed_rates, double start,
    const std::vector<double>& payment_times, const std::vector<double>& accruals,
    const std::vector<double>& monitoring_times, std::size_t n_steps, std::uint64_t n_paths, std::uint64_t seed);
double unilateral_cva_from_exposure(const ExecutionContext& execution, double a, double b, double sigma, double r0,
                                    const std::vector<double>& times, const std::vector<double>& ee,
                                    double hazard_rate, double recovery_rate);
std::vector<double> unilateral_cva_from_exposure_batch(const ExecutionContext& execution, double a, double b,
                                                       double sigma, double r0,
                                                       const std::vector<ExposureProfile>& profiles,
                                                       double hazard_rate, double recovery_rate);
// engine.cpp:226-266 and their batch forms (the HullWhite2F arm)
ExposureProfile irs_hull_white_2f_exposure_profile(const ExecutionContext& execution, double a, double b,
                                                   double sigma, double eta, double rho, double r0, double notional,
                                                   double fixed_rate, bool use_par_rate, double start,
                                                   const std::vector<double>& payment_times,
                                                   const std::vector<double>& accruals,
                                                   const std::vector<double>& monitoring_times, std::size_t n_steps,
                                                   std::uint64_t n_paths, std::uint64_t seed);
std::vector<ExposureProfile> irs_hull_white_2f_exposure_profile_batch(
    const ExecutionContext& execution, double a, double b, double sigma, double eta, double rho, double r0,
    const std::vector<double>& notionals, const std::vector<double>& fixed_rates, double start,
    const std::vector<double>& payment_times, const std::vector<double>& accruals,
    const std::vector<double>& monitoring_times, std::size_t n_steps, std::uint64_t n_paths, std::uint64_t seed);
double unilateral_cva_from_exposure_2f(const ExecutionContext& execution, double a, double b, double sigma,
                                       double eta, double rho, double r0, const std::vector<double>& times,
                                       const std::vector<double>& ee, double hazard_rate, double recovery_rate);
std::vector<double> unilateral_cva_from_exposure_2f_batch(const ExecutionContext& execution, double a, double b,
                                                          double sigma, double eta, double rho, double r0,
                                                          const std::vector<ExposureProfile>& profiles,
                                                          double hazard_rate, double recovery_rate);

// ---- measure.hpp:30-172 -----------------------------------------------------------------------
struct MeasureResult {
    std::vector<double> times;
    std::vector<double> primary;
    std::vector<double> secondary;
    bool has_scalar = false;
    double scalar = 0.0;
    std::optional<double> bump_used;
};

class IMeasure {
public:
    virtual ~IMeasure() = default;
    virtual std::string type_name() const = 0;
    virtual MeasureResult evaluate(const IModel& model, const IProduct& product, const MarketSnapshot& market,
                                   const PricingContext& pricing, const ExecutionContext& execution) const = 0;
};

class ExposureProfileMeasure : public IMeasure {
public:
    explicit ExposureProfileMeasure(const Params&) {}
    std::string type_name() const override { return "ExposureProfile"; }
    MeasureResult evaluate(const IModel& model, const IProduct& product, const MarketSnapshot& market,
                           const PricingContext& pricing, const ExecutionContext& execution) const override;
};

class UnilateralCvaMeasure : public IMeasure {
public:
    explicit UnilateralCvaMeasure(const Params&) {}
    std::string type_name() const override { return "UnilateralCVA"; }
    MeasureResult evaluate(const IModel& model, const IProduct& product, const MarketSnapshot& market,
                           const PricingContext& pricing, const ExecutionContext& execution) const override;
};

class PresentValueMeasure : public IMeasure {
public:
    explicit PresentValueMeasure(const Params&) {}
    std::string type_name() const override { return "PV"; }
    MeasureResult evaluate(const IModel& model, const IProduct& product, const MarketSnapshot& market,
                           const PricingContext& pricing, const ExecutionContext& execution) const override;
};

class Dv01Measure : public IMeasure {
public:
    explicit Dv01Measure(const Params& params)
        : bump_(get_double(params, "bump", 0.0001)), bucketed_(get_bool(params, "bucketed", false)) {}
    std::string type_name() const override { return "DV01"; }
    MeasureResult evaluate(const IModel& model, const IProduct& product, const MarketSnapshot& market,
                           const PricingContext& pricing, const ExecutionContext& execution) const override;
    double bump() const { return bump_; }
    bool bucketed() const { return bucketed_; }

private:
    double bump_;
    bool bucketed_;
};

// Reset dates of an IRS: start + every payment time but the last (measure.cpp:20-29).
std::vector<double> reset_dates(const IrSwapProduct& product);

// Batched GPU evaluation of the four vectorised measures over a homogeneous IRS batch
// (measure.cpp:238-325).
std::vector<ExposureProfile> compute_exposure_profile_batch(const IModel& model,
                                                            const std::vector<const IrSwapProduct*>& products,
                                                            const PricingContext& pricing,
                                                            const ExecutionContext& execution);
std::vector<double> compute_cva_from_exposure_batch(const IModel& model, const ExecutionContext& execution,
                                                    const std::vector<ExposureProfile>& profiles,
                                                    double hazard_rate, double recovery_rate);
// PV, DV01 (parallel) and bucketed DV01 of each product from the market curve, on the GPU.
struct PvDv01 {
    std::vector<double> pv, dv01;
    std::vector<std::vector<double>> bucketed;
};
PvDv01 compute_pv_dv01_batch(const MarketSnapshot& market, const std::vector<const IrSwapProduct*>& products,
                             double bump, bool bucketed, const ExecutionContext& execution,
                             double valuation_time = 0.0);

// ---- bootstrap.hpp ----------------------------------------------------------------------------
struct Registries {
    Registry<IModel> models;
    Registry<IProduct> products;
    Registry<IMeasure> measures;
};

void register_builtins(Registries& registries);

// ---- greeks.hpp:17-320 ------------------------------------------------------------------------
// Sensitivities of any registered measure by bump-and-reval on the GPU measures (the Philox stream
// is counter-based, so the bumped runs see the same shocks: common random numbers by construction),
// and, for UnilateralCVA of a fixed-rate swap under HullWhite1F, pathwise by the CVA adjoint.
namespace greeks {

enum class RiskFactorKind { ModelParameter, CurveParallel, CurvePillar, CreditParameter, TimeShift };

struct RiskFactor {
    RiskFactorKind kind = RiskFactorKind::ModelParameter;
    std::string scope;
    std::string name;
    std::optional<std::size_t> pillar_index;
};

RiskFactor parse_risk_factor(const std::string& text);
std::string to_string(const RiskFactor& risk_factor);

struct GreekOrder {
    int order = 1;
    std::optional<RiskFactor> cross_factor;
};

// The methods this engine serves. "auto" is bump-and-reval; the reference's other method names
// ("aad", "likelihood_ratio_hessian", "aad_forward_over_forward") are refused by parse_greek_method.
enum class GreekMethod { Auto, BumpAndReval, Pathwise };

GreekMethod parse_greek_method(const std::string& text);
std::string to_string(GreekMethod method);

struct GreekRequest {
    std::string metric_name;
    Params metric_params;
    RiskFactor risk_factor;
    GreekOrder order;
    GreekMethod method = GreekMethod::Auto;
    std::optional<double> bump_override;
};

struct GreekResult {
    bool has_scalar = false;
    double value = 0.0;
    std::vector<double> times;
    std::vector<double> primary;
    std::vector<double> secondary;
    GreekOrder order;
    RiskFactor risk_factor;
    GreekMethod method_used = GreekMethod::BumpAndReval;
    std::optional<double> bump_used;
    std::vector<std::string> warnings;
};

GreekResult compute_greek(const Registries& registries, const GreekRequest& request, const IModel& model,
                          const IProduct& product, const MarketSnapshot& market, const PricingContext& pricing,
                          const ExecutionContext& execution);

struct GreeksReport {
    std::vector<GreekResult> greeks;
    std::vector<std::string> skipped;
};

GreeksReport compute_all_greeks(const Registries& registries, const std::string& metric_name,
                                const Params& metric_params, const IModel& model, const IProduct& product,
                                const MarketSnapshot& market, const PricingContext& pricing,
                                const ExecutionContext& execution, bool include_curve_buckets = false,
                                bool include_second_order = false);

}  // namespace greeks

struct CvaAdjoint {
    double cva = 0.0;
    double grad[6] = {0, 0, 0, 0, 0, 0};  // d/d{a, b, sigma, r0, hazard_rate, recovery_rate}
};
CvaAdjoint unilateral_cva_adjoint(const IModel& model, const IProduct& product, const MarketSnapshot& market,
                                  const PricingContext& pricing, const ExecutionContext& execution);

// The "Greek" measure (greeks.hpp:304-318): a thin wrapper over compute_greek; the inner metric's
// own parameters arrive as "metric.<key>".
class GreekMeasure : public IMeasure {
public:
    GreekMeasure(const Params& params, const Registries& registries);
    std::string type_name() const override { return "Greek"; }
    MeasureResult evaluate(const IModel& model, const IProduct& product, const MarketSnapshot& market,
                           const PricingContext& pricing, const ExecutionContext& execution) const override;

private:
    greeks::GreekRequest request_;
    const Registries& registries_;
};

// ---- price.hpp --------------------------------------------------------------------------------
struct PriceResultEntry {
    std::string measure_name;
    MeasureResult result;
};
using PriceResult = std::vector<PriceResultEntry>;

struct MeasureSpec {
    std::string name;
    Params params;
    std::optional<std::string> alias = std::nullopt;
};

struct PriceBatchResultEntry {
    std::size_t trade_index = 0;
    PriceResult measures;
};
using PriceBatchResult = std::vector<PriceBatchResultEntry>;

std::vector<std::string> price_measure_names(const Registries& registries);

PriceResult price(const Registries& registries, const IProduct& product, const std::vector<MeasureSpec>& measures,
                  const IModel& model, const MarketSnapshot& market, const PricingContext& pricing,
                  const ExecutionContext& execution);

PriceBatchResult price_batch(const Registries& registries, const std::vector<const IProduct*>& products,
                             const std::vector<MeasureSpec>& measures, const IModel& model,
                             const MarketSnapshot& market, const PricingContext& pricing,
                             const ExecutionContext& execution);

PriceBatchResult price_many(const Registries& registries, const std::vector<const IProduct*>& products,
                            const std::vector<MeasureSpec>& measures, const IModel& model,
                            const MarketSnapshot& market, const PricingContext& pricing,
                            const ExecutionContext& execution);

}  // namespace engine
