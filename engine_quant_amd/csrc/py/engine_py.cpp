// pybind11 module exposing the C++ host with the shape of the reference's nanobind `engine`
// module (clients/python/src/engine_py_ext.cpp): Engine.create_model / create_product / price /
// price_batch / price_many, MarketSnapshot, PricingContext, ExecutionContext, MeasureResult.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <unordered_set>

#include "host/engine.hpp"

namespace py = pybind11;

namespace {

engine::Params dict_to_params(const py::dict& d) {
    engine::Params params;
    for (auto item : d) {
        std::string key = py::cast<std::string>(item.first);
        py::handle v = item.second;
        if (py::isinstance<py::bool_>(v)) {
            params[key] = py::cast<bool>(v);
        } else if (py::isinstance<py::float_>(v) || py::isinstance<py::int_>(v)) {
            params[key] = py::cast<double>(v);
        } else if (py::isinstance<py::str>(v)) {
            params[key] = py::cast<std::string>(v);
        } else {
            params[key] = py::cast<std::vector<double>>(v);
        }
    }
    return params;
}

// A measure is a bare name or a tuple (name, params[, alias]) (engine_py_ext.cpp to_measure_spec).
std::vector<engine::MeasureSpec> to_measure_specs(const py::list& measures) {
    std::vector<engine::MeasureSpec> specs;
    for (auto m : measures) {
        if (py::isinstance<py::str>(m)) {
            specs.push_back(engine::MeasureSpec{py::cast<std::string>(m), {}, std::nullopt});
            continue;
        }
        py::tuple t = py::cast<py::tuple>(m);
        if (t.size() < 1 || t.size() > 3)
            throw std::invalid_argument("measure spec must be name, (name, params) or (name, params, alias)");
        engine::MeasureSpec spec{py::cast<std::string>(t[0]), {}, std::nullopt};
        if (t.size() >= 2 && !t[1].is_none()) spec.params = dict_to_params(py::cast<py::dict>(t[1]));
        if (t.size() == 3 && !t[2].is_none()) spec.alias = py::cast<std::string>(t[2]);
        specs.push_back(std::move(spec));
    }
    return specs;
}

py::dict calc_result_to_dict(const engine::PriceResult& result, const char* caller) {
    py::dict out;
    std::unordered_set<std::string> seen;
    for (const auto& e : result) {
        if (!seen.insert(e.measure_name).second)
            throw std::invalid_argument(std::string(caller) + ": dos medidas resuelven al mismo nombre de salida '" +
                                        e.measure_name + "' -- dales un alias distinto");
        out[py::str(e.measure_name)] = py::cast(e.result);
    }
    return out;
}

class Engine {
public:
    Engine() { engine::register_builtins(registries_); }
    std::vector<std::string> list_models() const { return registries_.models.list(); }
    std::vector<std::string> list_products() const { return registries_.products.list(); }
    std::vector<std::string> list_measures() const { return engine::price_measure_names(registries_); }
    std::unique_ptr<engine::IModel> create_model(const std::string& name, const py::dict& params) const {
        return registries_.models.create(name, dict_to_params(params));
    }
    std::unique_ptr<engine::IProduct> create_product(const std::string& name, const py::dict& params) const {
        return registries_.products.create(name, dict_to_params(params));
    }
    py::dict price(const engine::IProduct& product, const py::list& measures, const engine::IModel& model,
                   const engine::MarketSnapshot& market, const engine::PricingContext& pricing,
                   const engine::ExecutionContext& execution) const {
        auto specs = to_measure_specs(measures);
        engine::PriceResult r;
        {
            py::gil_scoped_release release;
            r = engine::price(registries_, product, specs, model, market, pricing, execution);
        }
        return calc_result_to_dict(r, "Engine.price");
    }
    py::list batch(bool many, const std::vector<const engine::IProduct*>& products, const py::list& measures,
                   const engine::IModel& model, const engine::MarketSnapshot& market,
                   const engine::PricingContext& pricing, const engine::ExecutionContext& execution) const {
        auto specs = to_measure_specs(measures);
        engine::PriceBatchResult r;
        {
            py::gil_scoped_release release;
            r = many ? engine::price_many(registries_, products, specs, model, market, pricing, execution)
                     : engine::price_batch(registries_, products, specs, model, market, pricing, execution);
        }
        py::list out;
        for (const auto& e : r) {
            py::object br = py::module_::import("types").attr("SimpleNamespace")(
                py::arg("trade_index") = e.trade_index,
                py::arg("measures") = calc_result_to_dict(e.measures, many ? "Engine.price_many" : "Engine.price_batch"));
            out.append(br);
        }
        return out;
    }

    // engine_py_ext.cpp:298-312: the automatic Greek sweep of one metric
    engine::greeks::GreeksReport all_greeks(const engine::IProduct& product, const std::string& metric_name,
                                            const engine::IModel& model, const engine::MarketSnapshot& market,
                                            const engine::PricingContext& pricing,
                                            const engine::ExecutionContext& execution, const py::dict& metric_params,
                                            bool include_curve_buckets, bool include_second_order) const {
        engine::Params mp = dict_to_params(metric_params);
        py::gil_scoped_release release;
        return engine::greeks::compute_all_greeks(registries_, metric_name, mp, model, product, market, pricing,
                                                  execution, include_curve_buckets, include_second_order);
    }

private:
    engine::Registries registries_;
};

}  // namespace

PYBIND11_MODULE(_engine, m) {
    m.doc() = "C++ host of the MI355X exposure path (registry, measures, price) over libeqhip";
    m.def("is_gpu_backend_available", &engine::is_gpu_backend_available);

    py::class_<engine::IModel>(m, "Model")
        .def_property_readonly("type_name", &engine::IModel::type_name)
        .def("__repr__", [](const engine::IModel& s) { return "<Model '" + s.type_name() + "'>"; });
    py::class_<engine::IProduct>(m, "Product")
        .def_property_readonly("type_name", &engine::IProduct::type_name)
        .def("__repr__", [](const engine::IProduct& s) { return "<Product '" + s.type_name() + "'>"; });

    py::class_<engine::MeasureResult>(m, "MeasureResult")
        .def_readonly("times", &engine::MeasureResult::times)
        .def_readonly("primary", &engine::MeasureResult::primary)
        .def_readonly("secondary", &engine::MeasureResult::secondary)
        .def_readonly("has_scalar", &engine::MeasureResult::has_scalar)
        .def_readonly("scalar", &engine::MeasureResult::scalar)
        .def_readonly("bump_used", &engine::MeasureResult::bump_used)
        .def("__repr__", [](const engine::MeasureResult& s) {
            return "<MeasureResult times=" + std::to_string(s.times.size()) +
                   (s.has_scalar ? " scalar=" + std::to_string(s.scalar) : std::string()) + ">";
        });

    py::class_<engine::MarketSnapshot>(m, "MarketSnapshot")
        .def(py::init<std::vector<double>, std::vector<double>, double, double>(), py::arg("pillars"),
             py::arg("zero_rates"), py::arg("hazard_rate") = 0.0, py::arg("recovery_rate") = 0.0)
        .def_property_readonly("pillars", &engine::MarketSnapshot::pillars)
        .def_property_readonly("zero_rates", &engine::MarketSnapshot::zero_rates)
        .def_property_readonly("hazard_rate", &engine::MarketSnapshot::hazard_rate)
        .def_property_readonly("recovery_rate", &engine::MarketSnapshot::recovery_rate)
        .def("zero_rate", &engine::MarketSnapshot::zero_rate, py::arg("t"))
        .def("discount_factor", &engine::MarketSnapshot::discount_factor, py::arg("t"));

    py::class_<engine::PricingContext>(m, "PricingContext")
        .def(py::init([](const py::dict& d) { return engine::PricingContext(dict_to_params(d)); }), py::arg("params"))
        .def_property_readonly("pricing_date", &engine::PricingContext::pricing_date)
        .def_property_readonly("n_paths", &engine::PricingContext::n_paths)
        .def_property_readonly("n_steps", &engine::PricingContext::n_steps)
        .def_property_readonly("seed", &engine::PricingContext::seed);

    py::class_<engine::ExecutionContext>(m, "ExecutionContext")
        .def(py::init([](const py::dict& d) { return engine::ExecutionContext(dict_to_params(d)); }), py::arg("params"))
        .def_property_readonly("backend", &engine::ExecutionContext::backend)
        .def_property_readonly("precision", &engine::ExecutionContext::precision)
        .def_property_readonly("rng", &engine::ExecutionContext::rng)
        .def_property_readonly("discretisation", &engine::ExecutionContext::discretisation);

    py::class_<engine::greeks::GreekResult>(m, "GreekResult")
        .def_readonly("has_scalar", &engine::greeks::GreekResult::has_scalar)
        .def_readonly("value", &engine::greeks::GreekResult::value)
        .def_readonly("times", &engine::greeks::GreekResult::times)
        .def_readonly("primary", &engine::greeks::GreekResult::primary)
        .def_readonly("secondary", &engine::greeks::GreekResult::secondary)
        .def_readonly("bump_used", &engine::greeks::GreekResult::bump_used)
        .def_readonly("warnings", &engine::greeks::GreekResult::warnings)
        .def_property_readonly("order", [](const engine::greeks::GreekResult& g) { return g.order.order; })
        .def_property_readonly("risk_factor",
                               [](const engine::greeks::GreekResult& g) { return engine::greeks::to_string(g.risk_factor); })
        .def_property_readonly("method_used",
                               [](const engine::greeks::GreekResult& g) { return engine::greeks::to_string(g.method_used); })
        .def("__repr__", [](const engine::greeks::GreekResult& g) {
            return "<GreekResult " + engine::greeks::to_string(g.risk_factor) + " order=" + std::to_string(g.order.order) +
                   " method=" + engine::greeks::to_string(g.method_used) + ">";
        });
    py::class_<engine::greeks::GreeksReport>(m, "GreeksReport")
        .def_readonly("greeks", &engine::greeks::GreeksReport::greeks)
        .def_readonly("skipped", &engine::greeks::GreeksReport::skipped)
        .def("__repr__", [](const engine::greeks::GreeksReport& r) {
            return "<GreeksReport greeks=" + std::to_string(r.greeks.size()) + " skipped=" + std::to_string(r.skipped.size()) +
                   ">";
        });

    py::class_<Engine>(m, "Engine")
        .def(py::init<>())
        .def("list_models", &Engine::list_models)
        .def("list_products", &Engine::list_products)
        .def("list_measures", &Engine::list_measures)
        .def("create_model", &Engine::create_model, py::arg("name"), py::arg("params"))
        .def("create_product", &Engine::create_product, py::arg("name"), py::arg("params"))
        .def("price", &Engine::price, py::arg("product"), py::arg("measures"), py::arg("model"), py::arg("market"),
             py::arg("pricing"), py::arg("execution"))
        .def("price_batch",
             [](const Engine& e, const std::vector<const engine::IProduct*>& products, const py::list& measures,
                const engine::IModel& model, const engine::MarketSnapshot& market,
                const engine::PricingContext& pricing, const engine::ExecutionContext& execution) {
                 return e.batch(false, products, measures, model, market, pricing, execution);
             },
             py::arg("products"), py::arg("measures"), py::arg("model"), py::arg("market"), py::arg("pricing"),
             py::arg("execution"))
        .def("price_many",
             [](const Engine& e, const std::vector<const engine::IProduct*>& products, const py::list& measures,
                const engine::IModel& model, const engine::MarketSnapshot& market,
                const engine::PricingContext& pricing, const engine::ExecutionContext& execution) {
                 return e.batch(true, products, measures, model, market, pricing, execution);
             },
             py::arg("products"), py::arg("measures"), py::arg("model"), py::arg("market"), py::arg("pricing"),
             py::arg("execution"))
        .def("all_greeks", &Engine::all_greeks, py::arg("product"), py::arg("metric_name"), py::arg("model"),
             py::arg("market"), py::arg("pricing"), py::arg("execution"), py::arg("metric_params") = py::dict(),
             py::arg("include_curve_buckets") = false, py::arg("include_second_order") = false);

    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const std::out_of_range& e) {
            PyErr_SetString(PyExc_KeyError, e.what());
        }
    });
}
