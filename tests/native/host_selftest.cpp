// Sanitizer run of the C++ host (test infrastructure): the registry, Params accessors, contexts, the
// market and its bumps, the Greek grammar and every measure through price / price_batch / price_many /
// compute_all_greeks, built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitizers.py. Without a GPU the pricing calls must fail with the no-device error; with
// --gpu they run on the device (the kernels and the C ABI are not instrumented, the host code is).
// Exit status 0 when every expectation holds; the sanitizers abort on their first finding.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "host/engine.hpp"

using namespace engine;

static int fails = 0;

template <class F>
static void expect_error(const char* what, F f) {
    try {
        f();
    } catch (const std::exception&) {
        return;
    }
    std::fprintf(stderr, "expected an error: %s\n", what);
    ++fails;
}

static void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "check failed: %s\n", what);
        ++fails;
    }
}

static std::unique_ptr<IProduct> make_swap(const Registries& reg, double notional, double fixed, double start,
                                           int periods, double freq) {
    std::vector<double> pay, acc;
    for (int i = 1; i <= periods; ++i) {
        pay.push_back(start + i / freq);
        acc.push_back(1.0 / freq);
    }
    Params p{{"notional", notional}, {"payment_times", pay}, {"accruals", acc}, {"start", start}};
    if (!std::isnan(fixed)) p["fixed_rate"] = fixed;
    return reg.products.create("IRSwap", p);
}

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::string(argv[1]) == "--gpu";
    Registries reg;
    register_builtins(reg);

    // registry and models
    const Params hw{{"a", 0.1}, {"b", 0.03}, {"sigma", 0.01}, {"r0", 0.02}};
    auto m1 = reg.models.create("HullWhite1F", hw);
    auto m2 = reg.models.create(
        "HullWhite2F", Params{{"a", 0.1}, {"b", 0.2}, {"sigma", 0.01}, {"eta", 0.012}, {"rho", -0.7}, {"r0", 0.03}});
    expect(m1->type_name() == "HullWhite1F" && m2->to_params().size() == 6, "models");
    expect_error("unknown model", [&] { reg.models.create("Vasicek", hw); });
    expect_error("G2++ without eta", [&] { reg.models.create("HullWhite2F", hw); });
    expect_error("unknown measure type", [&] { reg.measures.create("Nope"); });

    // Params accessors
    Params p{{"x", 1.5}, {"v", std::vector<double>{1.0, 2.0}}, {"f", true}, {"s", std::string("text")}};
    expect(get_double(p, "x") == 1.5 && get_vector(p, "v").size() == 2 && get_bool(p, "f", false) &&
               get_string(p, "s") == "text" && get_double(p, "missing", 3.0) == 3.0 && contains(p, "x"),
           "params");
    expect_error("double from string", [&] { get_double(p, "s"); });
    expect_error("missing key", [&] { get_vector(p, "nope"); });

    // products
    auto spot = make_swap(reg, 1e6, 0.02, 0.0, 5, 1.0);
    auto fwd = make_swap(reg, -2e6, 0.021, 0.5, 6, 2.0);
    auto par = make_swap(reg, 1e6, NAN, 0.0, 3, 1.0);
    expect_error("accruals length", [&] {
        reg.products.create("IRSwap", Params{{"notional", 1.0}, {"payment_times", std::vector<double>{1.0, 2.0}},
                                              {"accruals", std::vector<double>{1.0}}});
    });
    // an empty schedule is a valid product (product.cpp:10-18) that no measure can price
    auto empty = reg.products.create("IRSwap", Params{{"notional", 1.0}, {"fixed_rate", 0.02},
                                                      {"payment_times", std::vector<double>{}},
                                                      {"accruals", std::vector<double>{}}});

    // Greek grammar
    for (const char* s : {"model.a", "model.sigma", "curve.parallel", "curve.pillar:0", "curve.pillar:12",
                          "credit.hazard_rate", "credit.recovery_rate", "time.theta", "a", "model.", "curve.pillar:",
                          "curve.pillar:-1", "curve.pillar:1x", "curve.pillar: 1", "fx.spot", "", ".", "curve.",
                          "credit.spread", "time.rho"}) {
        try {
            const greeks::RiskFactor rf = greeks::parse_risk_factor(s);
            expect(!greeks::to_string(rf).empty(), "risk factor text");
        } catch (const std::exception&) {
        }
    }
    for (const char* s : {"auto", "bump_and_reval", "pathwise", "aad", "finite", "", "likelihood_ratio_hessian"}) {
        try {
            (void)greeks::to_string(greeks::parse_greek_method(s));
        } catch (const std::exception&) {
        }
    }

    // market and contexts
    MarketSnapshot mk({0.5, 1.0, 2.0, 5.0, 10.0}, {0.015, 0.018, 0.02, 0.022, 0.025}, 0.02, 0.4);
    expect(std::fabs(bump_market_parallel(mk, 1e-4).zero_rate(3.0) - mk.zero_rate(3.0) - 1e-4) < 1e-15,
           "parallel bump");
    (void)bump_market_pillar(mk, 4, 1e-4);
    (void)bump_market_credit(mk, "hazard_rate", 1e-4);
    expect_error("pillar out of range", [&] { bump_market_pillar(mk, 5, 1e-4); });
    expect_error("unknown credit parameter", [&] { bump_market_credit(mk, "spread", 1e-4); });
    expect_error("decreasing pillars", [&] { MarketSnapshot({2.0, 1.0}, {0.02, 0.02}); });
    expect_error("empty curve", [&] { MarketSnapshot(std::vector<double>{}, std::vector<double>{}); });
    for (double t : {0.0, 0.25, 0.5, 3.0, 10.0, 40.0}) expect(mk.discount_factor(t) > 0.0, "discount factor");
    const PricingContext pr(Params{{"n_paths", 4096.0}, {"n_steps", 52.0}, {"seed", 7.0}});
    expect_error("no paths", [&] { PricingContext(Params{{"n_paths", 0.0}, {"n_steps", 10.0}, {"seed", 1.0}}); });
    expect_error("unknown backend", [&] { ExecutionContext(Params{{"backend", std::string("tpu")}}); });
    expect_error("fp32", [&] {
        ExecutionContext(Params{{"backend", std::string("gpu")}, {"precision", std::string("fp32")}});
    });
    const ExecutionContext ex(Params{{"backend", std::string("gpu")}, {"discretisation", std::string("exact")}});
    const ExecutionContext shared(Params{{"backend", std::string("gpu")}, {"simulation", std::string("shared")}});
    const ExecutionContext cpu(Params{{"backend", std::string("cpu")}});

    const std::vector<MeasureSpec> all = {
        {"PV", {}},
        {"DV01", {}},
        {"DV01", Params{{"bucketed", true}}, std::string("DV01 bucketed")},
        {"ExposureProfile", {}},
        {"ExpectedExposure", {}},
        {"PFE95", {}},
        {"UnilateralCVA", {}},
        {"Greek", Params{{"metric", std::string("UnilateralCVA")}, {"risk_factor", std::string("model.a")}}},
        {"Greek", Params{{"metric", std::string("PV")}, {"risk_factor", std::string("curve.pillar:2")}}, std::string("pillar")}};
    expect_error("unknown measure", [&] { price(reg, *spot, {{"NoSuchMeasure", {}}}, *m1, mk, pr, ex); });
    expect_error("empty schedule", [&] { price(reg, *empty, {{"PV", {}}, {"UnilateralCVA", {}}}, *m1, mk, pr, ex); });
    expect_error("cpu backend", [&] { price(reg, *spot, all, *m1, mk, pr, cpu); });
    expect_error("bad Greek", [&] {
        price(reg, *spot, {{"Greek", Params{{"metric", std::string("UnilateralCVA")}, {"risk_factor", std::string("a")}}}},
              *m1, mk, pr, ex);
    });
    if (!gpu) {
        expect_error("no device", [&] { price(reg, *spot, all, *m1, mk, pr, ex); });
    } else {
        const PriceResult r = price(reg, *spot, all, *m1, mk, pr, ex);
        expect(r.size() == all.size(), "every measure priced");
        const PriceResult r2 = price(reg, *fwd, {{"ExpectedExposure", {}}, {"UnilateralCVA", {}}}, *m2, mk, pr, ex);
        expect(r2.size() == 2, "G2++ measures");
        std::vector<std::unique_ptr<IProduct>> book;
        for (int i = 0; i < 24; ++i)
            book.push_back(make_swap(reg, (i % 2 ? -1.0 : 1.0) * 1e6 * (1 + i), 0.018 + 0.0005 * (i % 7),
                                     (i % 3) * 0.25, 2 + i % 9, (i % 2) ? 2.0 : 1.0));
        std::vector<const IProduct*> ptrs;
        for (auto& b : book) ptrs.push_back(b.get());
        const std::vector<MeasureSpec> expo = {{"PV", {}}, {"ExpectedExposure", {}}, {"PFE95", {}}, {"UnilateralCVA", {}}};
        expect(price_many(reg, ptrs, expo, *m1, mk, pr, ex).size() == ptrs.size(), "price_many per group");
        expect(price_many(reg, ptrs, expo, *m1, mk, pr, shared).size() == ptrs.size(), "price_many shared");
        std::vector<const IProduct*> same(6, spot.get());
        expect(price_batch(reg, same, expo, *m1, mk, pr, ex).size() == same.size(), "price_batch");
        const greeks::GreeksReport g =
            greeks::compute_all_greeks(reg, "UnilateralCVA", {}, *m1, *spot, mk, pr, ex, true, true);
        expect(!g.greeks.empty(), "all greeks");
        const CvaAdjoint adj = unilateral_cva_adjoint(*m1, *spot, mk, pr, ex);
        expect(adj.cva > 0.0, "adjoint");
        expect_error("par-rate swap by the adjoint", [&] { unilateral_cva_adjoint(*m1, *par, mk, pr, ex); });
    }
    if (fails) std::fprintf(stderr, "%d expectations failed\n", fails);
    else std::printf("host self-test passed (%s)\n", gpu ? "gpu" : "no device");
    return fails ? 1 : 0;
}
