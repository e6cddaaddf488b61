// PricingContext (cpp/engine/src/pricing_context.cpp:7-18) and ExecutionContext
// (execution_context.cpp:20-39), with the GPU path's rng / discretisation keys.
#include <algorithm>
#include <cctype>

#include "eqhip.h"
#include "engine.hpp"

namespace engine {

namespace {
std::string to_lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}
}  // namespace

bool is_gpu_backend_available() { return eqhip_is_gpu_backend_available() != 0; }

PricingContext::PricingContext(const Params& params)
    : pricing_date_(get_double(params, "pricing_date", 0.0)),
      n_paths_(static_cast<std::uint64_t>(get_double(params, "n_paths"))),
      n_steps_(static_cast<std::size_t>(get_double(params, "n_steps"))),
      seed_(static_cast<std::uint64_t>(get_double(params, "seed"))) {
    if (n_paths_ == 0) throw std::invalid_argument("PricingContext: n_paths debe ser mayor que 0");
    if (n_steps_ == 0) throw std::invalid_argument("PricingContext: n_steps debe ser mayor que 0");
}

ExecutionContext::ExecutionContext(const Params& params) {
    std::string requested = to_lower(get_string(params, "backend"));
    if (requested == "auto") {
        backend_ = is_gpu_backend_available() ? "gpu" : "cpu";
    } else if (requested == "cpu" || requested == "gpu") {
        backend_ = requested;
    } else {
        throw std::invalid_argument("ExecutionContext: backend desconocido: '" + requested +
                                    "' (valores validos: \"cpu\", \"gpu\", \"auto\")");
    }
    std::string precision = to_lower(get_string(params, "precision", "fp64"));
    if (precision != "fp64")
        throw std::invalid_argument("ExecutionContext: precision no soportada: '" + precision +
                                    "' (este build solo soporta \"FP64\")");
    precision_ = "FP64";
    rng_ = to_lower(get_string(params, "rng", "philox4x32-10"));
    if (rng_ != "philox4x32-10")
        throw std::invalid_argument("ExecutionContext: rng desconocido: '" + rng_ + "' (valido: \"philox4x32-10\")");
    discretisation_ = to_lower(get_string(params, "discretisation", "euler"));
    if (discretisation_ != "euler" && discretisation_ != "exact")
        throw std::invalid_argument("ExecutionContext: discretisation desconocida: '" + discretisation_ +
                                    "' (valores validos: \"euler\", \"exact\")");
    simulation_ = to_lower(get_string(params, "simulation", "per_group"));
    if (simulation_ != "per_group" && simulation_ != "shared")
        throw std::invalid_argument("ExecutionContext: simulation desconocida: '" + simulation_ +
                                    "' (valores validos: \"per_group\", \"shared\")");
}

}  // namespace engine
