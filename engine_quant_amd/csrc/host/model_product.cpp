// HullWhite1FModel, HullWhite2FModel (cpp/engine/src/model.cpp:39-71) and IrSwapProduct (product.cpp:7-25).
#include "engine.hpp"

namespace engine {

HullWhite1FModel::HullWhite1FModel(const Params& params)
    : a_(get_double(params, "a")), b_(get_double(params, "b")), sigma_(get_double(params, "sigma")),
      r0_(get_double(params, "r0")) {}

Params HullWhite1FModel::to_params() const { return Params{{"a", a_}, {"b", b_}, {"sigma", sigma_}, {"r0", r0_}}; }

HullWhite2FModel::HullWhite2FModel(const Params& params)
    : a_(get_double(params, "a")), b_(get_double(params, "b")), sigma_(get_double(params, "sigma")),
      eta_(get_double(params, "eta")), rho_(get_double(params, "rho")), r0_(get_double(params, "r0")) {}

Params HullWhite2FModel::to_params() const {
    return Params{{"a", a_}, {"b", b_}, {"sigma", sigma_}, {"eta", eta_}, {"rho", rho_}, {"r0", r0_}};
}

// fixed_rate absent => par swap (use_par_rate), product.cpp:10.
IrSwapProduct::IrSwapProduct(const Params& params)
    : notional_(get_double(params, "notional")),
      fixed_rate_(get_double(params, "fixed_rate", 0.0)),
      use_par_rate_(!contains(params, "fixed_rate")),
      start_(get_double(params, "start", 0.0)),
      payment_times_(get_vector(params, "payment_times")),
      accruals_(get_vector(params, "accruals")) {
    if (payment_times_.size() != accruals_.size())
        throw std::invalid_argument("IrSwapProduct: payment_times y accruals deben tener el mismo largo");
}

}  // namespace engine
