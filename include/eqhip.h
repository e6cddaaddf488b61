/*
 * eqhip — C ABI of the MI355X (gfx950) Monte Carlo exposure library.
 *
 * This is the drop-in boundary under engine_quant's registry-driven pricing API: it replaces
 * the cxx bridge functions the C++ measures call (rust/crates/engine-ffi/src/lib.rs:326-354,
 * 412, 511-541) in their GPU arm (rust/crates/engine-core/src/api.rs:151-159, 229-236,
 * 289-296, 342-349). Plain pointers and sizes only; every entry point returns 0 on success and
 * a non-zero EQHIP_E_* code on failure, with a message readable through eqhip_last_error()
 * (thread-local, like g_last_error in cpp/engine/src/abi.cpp:43-45). No exception crosses it.
 *
 * All arithmetic is FP64. Outputs are caller-allocated host buffers unless stated otherwise.
 * A context owns one HIP stream on one device; use one context per host thread.
 */
#ifndef EQHIP_H
#define EQHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EQHIP_OK 0
#define EQHIP_E_INVALID 1   /* invalid argument (the reference's std::invalid_argument / assert!) */
#define EQHIP_E_NO_DEVICE 2 /* backend "gpu" requested with no usable device: never a CPU run */
#define EQHIP_E_HIP 3       /* HIP runtime error */
#define EQHIP_E_CAPACITY 4  /* problem exceeds a documented capacity limit of the library */

/* Random-number source of the short-rate shocks. */
#define EQHIP_RNG_HOST_DRAWS 0 /* caller supplies Z[step][path] (shared-draw parity mode) */
#define EQHIP_RNG_PHILOX 1     /* in-register Philox4x32-10 + Box-Muller, key = seed */

/* Discretisation of dr = a(b - r)dt + sigma dW. */
#define EQHIP_DISC_EULER 0 /* Euler-Maruyama, as kernel.rs:41-50 (reference semantics) */
#define EQHIP_DISC_EXACT 1 /* exact Gaussian transition of the Vasicek/HW1F short rate */

/* Flags of eqhip_plan_create's want_trade_ee argument (a plain 0 / 1 keeps its old meaning). */
#define EQHIP_PLAN_TRADE_EE 1  /* per-trade EE / CVA outputs */
#define EQHIP_PLAN_KEEP_CUBE 2 /* keep the signed netting-set values [date][set][path] resident
                                  after each execute (eqhip_plan_exposure_cube, eqhip_plan_xva) */
#define EQHIP_PLAN_AUTOTUNE 4  /* the first execute times the valuation's variants (paths per lane
                                  1 / 2 / 4, trades per work item 2 / 4 / 8) that fit the plan on its
                                  own paths and keeps the fastest; results do not depend on it */
#define EQHIP_PLAN_TRADE_PFE 8 /* per-trade PFE95 output (HullWhite1F): the valuation also writes the
                                  legs of every live schedule class (16 B per class, path and date, in
                                  groups of dates sized to a quarter of the free device memory) and one
                                  exact select per live (trade, date) forms the trade's values from them */

typedef struct eqhip_context eqhip_context;
typedef struct eqhip_plan eqhip_plan;

/* HullWhite1F model parameters, cpp/engine/src/model.cpp:39-43. */
typedef struct {
    double a, b, sigma, r0;
} eqhip_model_hw1f;

/* HullWhite2F (G2++) model parameters, cpp/engine/src/model.cpp:54-60: r = x + y + r0 with
 * dx = -a x dt + sigma dW1, dy = -b y dt + eta dW2, d<W1,W2> = rho dt, x0 = y0 = 0
 * (rust/crates/engine-core/src/models/hull_white_2f.rs). */
typedef struct {
    double a, b, sigma, eta, rho, r0;
} eqhip_model_hw2f;

/* Portfolio of vanilla IRS in structure-of-arrays form. Trade j belongs to netting set
 * set_id[j]; its schedule is payment_times/accruals[flow_offset[j] .. flow_offset[j+1]) with
 * floating-leg start start[j]. Payer swaps have notional > 0, receivers notional < 0 (the
 * reference has no direction flag, cpp/engine/src/product.cpp:7-18). */
typedef struct {
    int32_t n_trades;
    int32_t n_sets;
    const int32_t* set_id;
    const double* notional;
    const double* fixed_rate;
    const double* start;
    const int32_t* flow_offset;
    const double* payment_times;
    const double* accruals;
} eqhip_portfolio;

/* Monte Carlo settings (PricingContext n_paths/n_steps/seed, cpp/engine/src/pricing_context.cpp
 * :7-18, plus the execution keys rng and discretisation). */
typedef struct {
    uint64_t n_paths;
    uint32_t n_steps;
    uint64_t seed;
    int32_t rng;
    int32_t disc;
} eqhip_mc;

/* Flat hazard / recovery (MarketSnapshot credit data, cpp/engine/include/engine/market.hpp). */
typedef struct {
    double hazard_rate, recovery_rate;
} eqhip_credit;

/* Results of one exposure run. set_ee / set_pfe are required; every other pointer may be NULL.
 * Layouts are row-major: [n_sets][n_dates], [n_trades][n_dates], [n_dates][n_trades][n_paths]. */
typedef struct {
    double* set_ee;
    double* set_pfe;
    double* set_cva;
    double* trade_ee;
    double* trade_cva;
    double* path_mtm; /* per-path, per-trade mark-to-market before max(.,0) (small runs only) */
    double* set_mtm;  /* per-path netting-set value before max(.,0) (small runs only) */
    double* trade_pfe; /* per-trade PFE95 [n_trades][n_dates] (ee_pfe_per_trade, exposure.rs:151-159);
                          needs a plan created with EQHIP_PLAN_TRADE_PFE */
} eqhip_outputs;

/* Per-kernel device times (milliseconds, HIP events on the context's stream) of the last
 * eqhip_plan_execute on that plan. */
typedef struct {
    double simulate_ms;
    double value_ms;
    double reduce_ms; /* EE sums + PFE radix select + per-trade EE finalisation */
    double cva_ms;
    double total_ms;
    uint64_t n_value_blocks;
    uint32_t paths_per_tile;
    uint32_t tiles_per_block;
    double xva_rows_ms;   /* eqhip_plan_xva: CSA adjustment, positive parts and ENE sums of the cube rows */
    double xva_select_ms; /* eqhip_plan_xva: EE sums and the quantile of every row */
    uint32_t trades_per_item; /* trades of one class and sign per work item of the per-trade step */
    uint32_t n_parts;         /* value plans the valuation runs in (more than one for a portfolio whose
                                 schedules do not fit one block: groups of classes, one after another) */
    double trade_pfe_ms;      /* per-trade PFE95 selects (the class legs they read are written by the
                                 valuation, whose time value_ms includes that) */
    uint32_t trade_pfe_groups; /* date groups of the class-leg buffer (its size: a quarter of the free
                                  device memory at plan creation, or EQHIP_TRADE_PFE_BYTES when smaller) */
} eqhip_timing;

/* ---- library / context -------------------------------------------------------------------- */

/* Replaces engine::ffi::is_gpu_backend_available (engine-ffi/src/lib.rs:412, api.rs:362-364):
 * here it is a runtime probe (a gfx950 device is visible), not a compile-time flag. */
int eqhip_is_gpu_backend_available(void);

/* Copies the calling thread's last error message into buf (always NUL-terminated). Returns the
 * full message length. */
size_t eqhip_last_error(char* buf, size_t len);

int eqhip_context_create(int device, eqhip_context** out);
int eqhip_context_destroy(eqhip_context* ctx);

/* No counterpart in the reference: a self-check for callers and tests. y[i] = the valuation kernels'
 * FP64 exp of x[i] (the e^x of every per-path zero-coupon bond A exp(-B r)), computed on the device;
 * within 0.81 ulp of e^x for |x| <= 708, inf / 0 beyond, NaN propagated. The kernels compile the exp in
 * two instruction forms with the same arithmetic; y[i] is NaN wherever the two differ in any bit. */
int eqhip_bond_exp(eqhip_context* ctx, const double* x, double* y, int32_t n);

/* ---- reference-shaped entry points -------------------------------------------------------- */

/* Replaces engine::ffi::irs_hull_white_exposure_profile (engine-ffi/src/lib.rs:326-342;
 * api.rs:126-173): EE / PFE95 of one payer IRS at its monitoring dates. use_par_rate != 0
 * sets the fixed rate to the model par rate at r0 (api.rs:53-57). host_draws ([n_steps][n_paths])
 * is read only when rng == EQHIP_RNG_HOST_DRAWS. out_times/out_ee/out_pfe hold n_dates values;
 * *n_out receives the number written (1 when every monitoring date is <= 0, exposure.rs:98-108).
 * Monitoring dates must be reset dates of the swap (exposure.rs:85-90) unless allow_any_date. */
int eqhip_irs_hull_white_exposure_profile(
    eqhip_context* ctx, double a, double b, double sigma, double r0, double notional,
    double fixed_rate, int use_par_rate, double start, const double* payment_times,
    const double* accruals, int32_t n_flows, const double* monitoring_times, int32_t n_dates,
    uint32_t n_steps, uint64_t n_paths, uint64_t seed, int32_t rng, int32_t disc,
    const double* host_draws, int32_t allow_any_date, double* out_times, double* out_ee,
    double* out_pfe, int32_t* n_out);

/* Replaces engine::ffi::irs_hull_white_exposure_profile_batch (engine-ffi/src/lib.rs:511-527;
 * api.rs:204-247): homogeneous batch (shared schedule), one simulation, one profile per trade.
 * out_ee / out_pfe are [n_trades][n_dates]. */
int eqhip_irs_hull_white_exposure_profile_batch(
    eqhip_context* ctx, double a, double b, double sigma, double r0, const double* notionals,
    const double* fixed_rates, int32_t n_trades, double start, const double* payment_times,
    const double* accruals, int32_t n_flows, const double* monitoring_times, int32_t n_dates,
    uint32_t n_steps, uint64_t n_paths, uint64_t seed, int32_t rng, int32_t disc,
    const double* host_draws, int32_t allow_any_date, double* out_times, double* out_ee,
    double* out_pfe, int32_t* n_out);

/* A scenario set: the HW1F short-rate paths of one model on the grid of t_max / n_steps (every grid
 * step kept on the device), simulated once from the Philox stream of mc->seed. Replaces the
 * per-calendar-group re-simulation of price_many (cpp/engine/src/price.cpp:534-578; measure.cpp:348-365)
 * when every group is valued on one scenario set. The set is read-only once created: any context of
 * the same device may value on it, from any thread, at the same time. */
typedef struct eqhip_scenarios eqhip_scenarios;
int eqhip_scenarios_create(eqhip_context* ctx, const eqhip_model_hw1f* model, double t_max, const eqhip_mc* mc,
                           eqhip_scenarios** out);
int eqhip_scenarios_destroy(eqhip_scenarios* scenarios);

/* eqhip_irs_hull_white_exposure_profile_batch on a scenario set instead of a simulation of its own:
 * the same homogeneous batch and outputs, with the set's model, grid, paths and seed. The dates may be
 * any dates in [0, t_max] (as with allow_any_date) and the last of them must be the set's t_max (so
 * the batch's grid is the set's). Equals the batch call with the same arguments and allow_any_date,
 * bit for bit. */
int eqhip_irs_hull_white_exposure_profile_batch_on(
    eqhip_context* ctx, const eqhip_scenarios* scenarios, const double* notionals, const double* fixed_rates,
    int32_t n_trades, double start, const double* payment_times, const double* accruals, int32_t n_flows,
    const double* monitoring_times, int32_t n_dates, double* out_times, double* out_ee, double* out_pfe,
    int32_t* n_out);

/* Replaces engine::ffi::unilateral_cva_from_exposure (engine-ffi/src/lib.rs:344-354;
 * api.rs:273-304; exposure.rs:242-275). */
int eqhip_unilateral_cva_from_exposure(eqhip_context* ctx, double a, double b, double sigma,
                                       double r0, const double* times, const double* ee,
                                       int32_t n_dates, double hazard_rate, double recovery_rate,
                                       double* out_cva);

/* Replaces engine::ffi::unilateral_cva_from_exposure_batch (engine-ffi/src/lib.rs:532-541;
 * api.rs:327-357): ee is [n_profiles][n_dates] over shared times. */
int eqhip_unilateral_cva_from_exposure_batch(eqhip_context* ctx, double a, double b,
                                             double sigma, double r0, const double* times,
                                             const double* ee, int32_t n_profiles,
                                             int32_t n_dates, double hazard_rate,
                                             double recovery_rate, double* out_cva);

/* Allocation-free form of eqhip_unilateral_cva_from_exposure_batch. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc), whose EE profiles are already on the device. d_times [n_dates],
 * d_ee [n_profiles][n_dates] and d_out_cva [n_profiles] are device memory of the context's GPU. The
 * per-date weights go to a fixed device array of the library (at most 16384 dates: more is refused with
 * EQHIP_E_CAPACITY); the call makes no allocation and no free. stream: the HIP stream (hipStream_t) to
 * enqueue on, and then the call returns without synchronising; NULL uses the context's stream and
 * waits for the result. Calls that overlap on one GPU share that array and race: order them. */
int eqhip_unilateral_cva_from_exposure_batch_noalloc(eqhip_context* ctx, double a, double b, double sigma,
                                                     double r0, const double* d_times, const double* d_ee,
                                                     int32_t n_profiles, int32_t n_dates, double hazard_rate,
                                                     double recovery_rate, double* d_out_cva, void* stream);

/* Replaces engine::ffi::irs_hull_white_2f_exposure_profile (engine-ffi/src/lib.rs:547-565;
 * api.rs:587-665; exposure.rs:305-370): the G2++ profile of one payer IRS. use_par_rate != 0 sets
 * the fixed rate to the par rate at the initial state (0, 0). host_draws is [2][n_steps][n_paths]
 * (every x shock, then every independent y shock, as exposure.rs:339-344 draws them) and is read
 * only when rng == EQHIP_RNG_HOST_DRAWS. A date on grid index 0 averages n_paths equal values, as
 * the reference's path[0] holds n_paths zero states. Other arguments as in
 * eqhip_irs_hull_white_exposure_profile. */
int eqhip_irs_hull_white_2f_exposure_profile(
    eqhip_context* ctx, double a, double b, double sigma, double eta, double rho, double r0,
    double notional, double fixed_rate, int use_par_rate, double start, const double* payment_times,
    const double* accruals, int32_t n_flows, const double* monitoring_times, int32_t n_dates,
    uint32_t n_steps, uint64_t n_paths, uint64_t seed, int32_t rng, int32_t disc,
    const double* host_draws, int32_t allow_any_date, double* out_times, double* out_ee,
    double* out_pfe, int32_t* n_out);

/* Replaces engine::ffi::irs_hull_white_2f_exposure_profile_batch (engine-ffi/src/lib.rs:677-694;
 * exposure.rs:378-447): homogeneous batch, one G2++ simulation, out_ee / out_pfe
 * [n_trades][n_dates]. */
int eqhip_irs_hull_white_2f_exposure_profile_batch(
    eqhip_context* ctx, double a, double b, double sigma, double eta, double rho, double r0,
    const double* notionals, const double* fixed_rates, int32_t n_trades, double start,
    const double* payment_times, const double* accruals, int32_t n_flows,
    const double* monitoring_times, int32_t n_dates, uint32_t n_steps, uint64_t n_paths,
    uint64_t seed, int32_t rng, int32_t disc, const double* host_draws, int32_t allow_any_date,
    double* out_times, double* out_ee, double* out_pfe, int32_t* n_out);

/* Replaces engine::ffi::unilateral_cva_from_exposure_2f (engine-ffi/src/lib.rs:567-579;
 * exposure.rs:452-466): the discount is the G2++ bond at the initial state (0, 0). */
int eqhip_unilateral_cva_from_exposure_2f(eqhip_context* ctx, double a, double b, double sigma,
                                          double eta, double rho, double r0, const double* times,
                                          const double* ee, int32_t n_dates, double hazard_rate,
                                          double recovery_rate, double* out_cva);

/* Replaces engine::ffi::unilateral_cva_from_exposure_2f_batch (engine-ffi/src/lib.rs:696-708;
 * exposure.rs:468-480): ee is [n_profiles][n_dates] over shared times. */
int eqhip_unilateral_cva_from_exposure_2f_batch(eqhip_context* ctx, double a, double b,
                                                double sigma, double eta, double rho, double r0,
                                                const double* times, const double* ee,
                                                int32_t n_profiles, int32_t n_dates,
                                                double hazard_rate, double recovery_rate,
                                                double* out_cva);

/* PV and DV01 of every trade from the market zero curve, cpp/engine/src/measure.cpp:124-222
 * (npv_from_market, compute_dv01, dv01_bucketed_from_market) with market.cpp:26-41,95-111.
 * fixed_rate in the portfolio is the effective rate (par under the market for par swaps).
 * out_bucketed ([n_trades][n_pillars]) may be NULL. */
int eqhip_pv_dv01(eqhip_context* ctx, const double* pillars, const double* zero_rates,
                  int32_t n_pillars, const eqhip_portfolio* portfolio, double bump, double* out_pv,
                  double* out_dv01, double* out_bucketed);

/* eqhip_pv_dv01 with PV and the parallel DV01 valued at valuation_time v on the curve observed today:
 * every discount factor is divided by DF(v), the valuation-time argument of npv_from_market /
 * compute_dv01 (cpp/engine/src/measure.cpp:146-200, reached from PresentValueMeasure and Dv01Measure
 * through PricingContext::pricing_date, measure.cpp:366-380,425-445). v = 0 is eqhip_pv_dv01 bit for
 * bit. A trade with v > start is refused with EQHIP_E_INVALID and the reference's message (its first
 * floating coupon would already be fixed). The bucketed ladder stays at t = 0, as the reference's
 * dv01_bucketed_from_market does. */
int eqhip_pv_dv01_at(eqhip_context* ctx, const double* pillars, const double* zero_rates,
                     int32_t n_pillars, const eqhip_portfolio* portfolio, double bump, double valuation_time,
                     double* out_pv, double* out_dv01, double* out_bucketed);

/* Analytic first-order sensitivities of every trade's PV:
 * out_pillar_delta [n_trades][n_pillars] = dPV/dz_p, the derivative of npv_from_market
 * (cpp/engine/src/measure.cpp:160-184) with respect to each pillar zero rate, which the one-sided
 * bucketed DV01 of measure.cpp:211-222 approximates as bump * dPV/dz_p; with a model (nullable),
 * out_model_npv [n_trades] is the HW1F model NPV at t = 0 and out_model_greeks [n_trades][4] its
 * derivatives in (a, b, sigma, r0), the quantities the reference obtains by reverse-mode AD
 * (rust/crates/engine-core/src/api.rs:394-415, 519-552). */
int eqhip_pv_sensitivities(eqhip_context* ctx, const double* pillars, const double* zero_rates,
                           int32_t n_pillars, const eqhip_portfolio* portfolio,
                           const eqhip_model_hw1f* model, double* out_pillar_delta,
                           double* out_model_npv, double* out_model_greeks);

/* ---- portfolio / netting-set exposure ----------------------------------------------------- */

/* Builds the device-resident plan of one exposure job: schedule classes, cash-flow date table,
 * ZCB factor tables, work buffers. Monitoring dates are arbitrary (between-reset valuation
 * carries the last floating fixing on each path). */
int eqhip_plan_create(eqhip_context* ctx, const eqhip_model_hw1f* model,
                      const eqhip_portfolio* portfolio, const double* dates, int32_t n_dates,
                      const eqhip_mc* mc, int32_t want_trade_ee, eqhip_plan** out);

/* The same plan under G2++ (the portfolio form of exposure.rs:305-370 and measure.cpp:53-64,
 * 257-268): every trade's value is the reference's swap valuation with the two-factor bond. */
int eqhip_plan_create_2f(eqhip_context* ctx, const eqhip_model_hw2f* model,
                         const eqhip_portfolio* portfolio, const double* dates, int32_t n_dates,
                         const eqhip_mc* mc, int32_t want_trade_ee, eqhip_plan** out);

/* Runs the job on resident data and copies the results into out. host_draws ([n_steps][n_paths],
 * [2][n_steps][n_paths] for a G2++ plan) is read only in EQHIP_RNG_HOST_DRAWS mode. */
int eqhip_plan_execute(eqhip_plan* plan, const double* host_draws, const eqhip_credit* credit,
                       eqhip_outputs* out);

int eqhip_plan_timing(const eqhip_plan* plan, eqhip_timing* out);

/* Sets the plan's valuation variant: paths_per_lane (1, 2, 4: a tile of 64 x that many paths) and
 * trades_per_item (2, 4, 8). Refused with EQHIP_E_CAPACITY when it does not fit the plan (LDS, work
 * items; four paths per lane take two trades per item; two-factor plans one path per lane). Results
 * do not depend on the variant: netting-set values bit for bit, per-trade EE to rounding.
 * paths_per_lane = 0 selects path mode (trades_per_item ignored): one thread per path with the
 * schedule families and netting sets in registers, for plans without per-trade outputs, with at most
 * 4 families and 8 netting sets, which take it by default. It needs no LDS, so it also serves such
 * plans too diverse for a block's LDS. Its netting-set sums run in cash-flow-date order, so its values
 * agree with the tiled variants' to rounding (bit for bit between its own runs, batched or not). No
 * counterpart in the reference (it has no GPU valuation kernel); it fixes what EQHIP_PLAN_AUTOTUNE
 * would otherwise choose. */
int eqhip_plan_set_variant(eqhip_plan* plan, int32_t paths_per_lane, int32_t trades_per_item);
int eqhip_plan_destroy(eqhip_plan* plan);

/* One-shot convenience: plan_create + plan_execute + plan_destroy. */
int eqhip_exposure_run(eqhip_context* ctx, const eqhip_model_hw1f* model,
                       const eqhip_portfolio* portfolio, const double* dates, int32_t n_dates,
                       const eqhip_mc* mc, const double* host_draws, const eqhip_credit* credit,
                       eqhip_outputs* out);

/* ---- exposure cube and the XVA reducer -------------------------------------------------------- */

/* The first collateral agreement of an XVA request (quant-domain/src/xva.rs:55-75); the reducer
 * adjusts values by its variation margin (quant-engine/src/xva.rs:48-64). */
typedef struct {
    double threshold, minimum_transfer_amount, independent_amount;
    int32_t variation_margin;
} eqhip_csa;

/* DefaultSpec (quant-domain/src/xva.rs:97-121): survival_probabilities, when given, override the
 * flat hazard rate at the dates they cover. */
typedef struct {
    double hazard_rate, recovery_rate;
    const double* survival_probabilities;
    int32_t n_survival;
} eqhip_default_spec;

/* The numeric part of an XvaRequest (quant-domain/src/xva.rs:234-262). initial_margin_total is the
 * sum over the request's CSAs of max(initial_margin, 0); csa is the first CSA or NULL. */
typedef struct {
    eqhip_default_spec counterparty_default, own_default;
    double funding_spread, borrowing_spread, capital_factor;
    double initial_margin_total;
    const eqhip_csa* csa;
    double confidence_level;
} eqhip_xva_params;

/* Replaces the reductions of quant-engine xva::calculate (quant-engine/src/xva.rs:97-246) over an
 * ExposureCubeInput (quant-domain/src/xva.rs:176-220): q_paths (and p_paths, nullable, the P cube for
 * DVA) are row-major [n_times][n_paths]; discount_factors (nullable) [n_times]. Writes EE and the
 * confidence quantile of the positive exposure per time (out_ee, out_pfe [n_times]) and
 * out_measures[5] = {cva, dva, fva, mva, kva}. The Q/P tags, source hashes and the artifact record
 * of the request stay with the caller. */
int eqhip_xva_calculate(eqhip_context* ctx, const eqhip_xva_params* params, const double* times,
                        int32_t n_times, uint64_t n_paths, const double* q_paths,
                        const double* discount_factors, const double* p_paths, double* out_ee,
                        double* out_pfe, double* out_measures);

/* The exposure cube of netting set `set` from the plan's last execute, in the reducer's layout: the
 * signed netting-set values [n_dates][n_paths] and the model discount factors P(0, t_k) (nullable).
 * This is the Monte Carlo cube that replaces the flat one-path PV cube of
 * quant-engine/src/lib.rs:751-779. Needs a plan created with EQHIP_PLAN_KEEP_CUBE. */
int eqhip_plan_exposure_cube(eqhip_plan* plan, int32_t set, double* out_paths, double* out_discount);

/* eqhip_xva_calculate on the plan's resident cube of `set` (Q measure, DVA from the same cube,
 * discount factors P(0, t_k) of the model): the cube never leaves the device. */
int eqhip_plan_xva(eqhip_plan* plan, int32_t set, const eqhip_xva_params* params, double* out_ee,
                   double* out_pfe, double* out_measures);

/* ---- sensitivities ---------------------------------------------------------------------------- */

/* Pathwise adjoint of netting set `set`'s unilateral CVA (exposure.rs:242-275 over the plan's EE):
 * runs the plan's exposure job, then a reverse sweep per path that stores the short rate only every
 * 12 grid steps (plus the fixing dates) and re-simulates each segment backwards.
 * out_cva: the CVA; out_grad[6]: dCVA / d{a, b, sigma, r0, hazard_rate, recovery_rate};
 * out_check (nullable): sum_k w_k sum_p max(V,0) recomputed by the adjoint pass (equals out_cva up
 * to rounding). Curve-pillar sensitivities of this CVA are identically zero: it discounts with the
 * model bond P(0,t|r0), not the market curve (exposure.rs:264-275). Requires one grid step per
 * monitoring date. The reference computes CVA Greeks by bump-and-revalue with common random
 * numbers (cpp/engine/src/greeks.cpp:160-169). HullWhite1F plans only (EQHIP_E_INVALID on a G2++
 * plan). */
int eqhip_cva_adjoint(eqhip_plan* plan, const double* host_draws, int32_t set, const eqhip_credit* credit,
                      double* out_cva, double* out_grad, double* out_check);

/* ---- batched netting sets ------------------------------------------------------------------ */

/* Many exposure jobs (each its own portfolio and path count; shared model, monitoring dates, grid,
 * seed and discretisation, Philox draws) as one simulation, one valuation launch over every job's
 * work units (persistent or one block per unit, chosen by an autotuner on first execute) and one
 * select launch. Job j uses the first n_paths[j] paths of the shared stream, so its set-level
 * EE / PFE95 / CVA are bitwise those of an unbatched plan with the same settings.
 * No counterpart in the reference, which prices one calendar group at a time
 * (cpp/engine/src/price.cpp:534-578). */
typedef struct eqhip_batch eqhip_batch;
int eqhip_batch_create(eqhip_context* ctx, const eqhip_model_hw1f* model, int32_t n_jobs,
                       const eqhip_portfolio* portfolios, const uint64_t* n_paths, const double* dates,
                       int32_t n_dates, uint32_t n_steps, uint64_t seed, int32_t rng, int32_t disc,
                       int32_t want_trade_ee, eqhip_batch** out);
/* outs[n_jobs]; per-path outputs are not produced by batches. */
int eqhip_batch_execute(eqhip_batch* batch, const eqhip_credit* credit, eqhip_outputs* outs);
/* Timing of the last execute; tune_log (nullable) receives the autotuner's choice and trials. */
int eqhip_batch_timing(const eqhip_batch* batch, eqhip_timing* out, char* tune_log, size_t len);
int eqhip_batch_destroy(eqhip_batch* batch);

/* ---- Bermudan payer swaptions and the learned continuation value ---------------------------
 * No reference counterpart: the reference's only regression method is the quadratic
 * Longstaff-Schwartz fit of rust/crates/engine-core/src/payoff/lsm.rs:73-107 (GBM payoffs). These
 * entry points hold a book of Bermudan payer swaptions under the same HW1F model, fit a degree-3
 * LSM exercise policy, and estimate exposures either by the per-date polynomial regression (the
 * fallback and judge) or by a per-trade continuation network 3 -> 64 -> 64 -> 64 -> 1 (tanh) on
 * (short rate, log numeraire, time), trained and evaluated with BF16 matrix-core products and FP32
 * accumulation. Cash settlement: exercising at year T pays
 * N [(1 - P(T, M)) - K sum_{m=1..M-T} P(T, T+m)], M = first_exercise + n_exercise, and ends the
 * trade, so exposure counts only paths not exercised before t. */
typedef struct eqhip_berm eqhip_berm;

typedef struct {
    int32_t n_trades;
    const double* notional;
    const double* strike;
    const int32_t* first_exercise; /* E >= 1 (years) */
    const int32_t* n_exercise;     /* exercise years E .. E + n_exercise - 1, annual */
} eqhip_berm_book;

typedef struct {
    uint64_t n_train, n_infer;      /* paths of the training set and of the inference set */
    uint32_t steps_per_year;        /* grid; monitoring dates are k / steps_per_year, k = 1..n_steps */
    uint32_t horizon_years;         /* n_steps = steps_per_year * horizon_years, >= every maturity */
    uint64_t seed_train, seed_infer;
    int32_t rng;                    /* EQHIP_RNG_PHILOX, or EQHIP_RNG_HOST_DRAWS (z_train, z_infer) */
} eqhip_berm_mc;

typedef struct {
    uint32_t steps;                 /* Adam steps */
    uint32_t batch;                 /* samples per network per step, multiple of 64 */
    float lr, lr_final;             /* cosine decay from lr to lr_final */
    float beta1, beta2, eps;
    uint64_t seed;                  /* initial weights and sample draws */
    /* inputs u0 = (r - r_shift) r_scale, u1 = lb lb_scale + lb_shift, u2 = t t_scale + t_shift;
     * the network's output is the continuation value per unit notional times y_scale */
    double r_shift, r_scale, lb_scale, lb_shift, t_scale, t_shift, y_scale;
} eqhip_berm_mlp_cfg;

typedef struct {
    double simulate_ms, policy_ms, apply_ms, judge_ms, train_ms, infer_ms;
    double lattice_ms;        /* eqhip_berm_exposure_mlp_lattice, wall time of the call */
    double lattice_net_ms;    /* its network evaluation at the nodes (HIP events) */
    double lattice_interp_ms; /* its per-path interpolation and exposure sums (HIP events) */
} eqhip_berm_timing;

/* Simulates both path sets (exact scheme, trapezoid log-numeraire) and keeps them resident.
 * z_train [n_steps][n_train] and z_infer [n_steps][n_infer] are read only in host-draw mode. */
int eqhip_berm_create(eqhip_context* ctx, const eqhip_model_hw1f* model, const eqhip_berm_book* book,
                      const eqhip_berm_mc* mc, const double* z_train, const double* z_infer,
                      eqhip_berm** out);
/* LSM exercise policy fitted backwards on the training set, then applied to the inference set.
 * Nullable outputs: ex_coef [n_trades][max n_exercise][4], price_train / price_infer [n_trades]. */
int eqhip_berm_fit_policy(eqhip_berm* berm, double* ex_coef, double* price_train, double* price_infer);
/* Exercise years and discounted pathwise cashflows of set 0 (training) or 1 (inference):
 * tau [n_trades][n_paths] (-1 = never exercised), y0 [n_trades][n_paths]; either may be NULL. */
int eqhip_berm_pathwise(eqhip_berm* berm, int32_t set, int32_t* tau, double* y0);
/* Polynomial judge on set 0 or 1: per (trade, date) degree-3 regression of the pathwise value on
 * the standardised short rate over the alive paths; ee [n_trades][n_steps], coef (nullable)
 * [n_trades][n_steps][4]. The coefficients are kept for eqhip_berm_mse. */
int eqhip_berm_exposure_lsm(eqhip_berm* berm, int32_t set, double* ee, double* coef);
/* Initialises (seeded) and trains the networks on the training set with Adam. loss_curve
 * (nullable) [steps]: mean over trades of each step's minibatch loss. */
int eqhip_berm_train(eqhip_berm* berm, const eqhip_berm_mlp_cfg* cfg, float* loss_curve);
/* Network exposure on set 0 or 1: ee [n_trades][n_steps]. */
int eqhip_berm_exposure_mlp(eqhip_berm* berm, int32_t set, double* ee);
/* The same exposure with each network evaluated on a per-date lattice: nr x nl nodes spanning the
 * (r, log-numeraire) box of the set's paths at that date (BF16 MFMA, two 16-sample tiles per wave at
 * a time), each alive path's value the bilinear interpolation of its four surrounding nodes.
 * nr, nl powers of two >= 2 with 32 <= nr nl <= 1024. No counterpart in the reference (its LSM
 * has no network); it serves the network's exposure at a cost per path independent of its size. */
int eqhip_berm_exposure_mlp_lattice(eqhip_berm* berm, int32_t set, int32_t nr, int32_t nl, double* ee);
/* Out-of-sample accuracy on set 1 against the pathwise values: per-trade mean squared error per
 * unit notional of the network and of the polynomials last fitted by eqhip_berm_exposure_lsm. */
int eqhip_berm_mse(eqhip_berm* berm, double* mse_mlp, double* mse_poly);
/* Network parameters, float32 [n_trades][8641] (W1 [64][4] incl. bias, W2, b2, W3, b3, w4, b4). */
int eqhip_berm_get_params(eqhip_berm* berm, float* params);
int eqhip_berm_set_params(eqhip_berm* berm, const eqhip_berm_mlp_cfg* cfg, const float* params);
/* Loss and gradient of every network on explicit samples (n per network, multiple of 64):
 * u [n_trades][n][3], target / mask [n_trades][n]; loss [n_trades], grad [n_trades][8641]. */
int eqhip_berm_mlp_grad(eqhip_berm* berm, const float* u, const float* target, const float* mask, uint32_t n,
                        float* loss, float* grad);
/* Network outputs on explicit inputs: u [n_trades][n][3] -> y [n_trades][n]. */
int eqhip_berm_mlp_forward(eqhip_berm* berm, const float* u, uint32_t n, float* y);
/* One Adam step of every network from the gradient of eqhip_berm_mlp_grad's samples (for tests);
 * step counts from 1. */
int eqhip_berm_adam_step(eqhip_berm* berm, const float* u, const float* target, const float* mask, uint32_t n,
                         uint32_t step, float lr);
int eqhip_berm_get_timing(const eqhip_berm* berm, eqhip_berm_timing* out);
int eqhip_berm_destroy(eqhip_berm* berm);

#ifdef __cplusplus
}
#endif

#endif
