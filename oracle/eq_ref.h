/*
 * eq_ref — CPU FP64 restatement of engine_quant's Monte Carlo exposure path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the oracle the HIP path is checked against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker or the timed CPU baseline. The product (engine_quant_amd, libeqhip) never
 * links, loads or calls it.
 *
 * Every function cites the reference file:line whose arithmetic it restates; the expression
 * order of the reference is kept so that the restatement is a faithful CPU twin.
 *
 * Parity pins (tests/test_oracle.py): reference golden values (PV/DV01 exact, EE/PFE/CVA of the
 * seed-7 case within 3 Monte Carlo standard errors), closed forms (ZCB vs numerical integral,
 * Euler conditional mean, MC ZCB), Philox4x32-10 known-answer vectors, brute-force order
 * statistics, the martingale property of the between-reset floating leg, and the
 * single-trade-netting-set reduction.
 */
#ifndef EQ_REF_H
#define EQ_REF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Hull-White 1F (Vasicek form), models/hull_white.rs:50-80 ---- */
double eqr_hw_b(double a, double tau);
double eqr_hw_a(double a, double b, double sigma, double tau, double bf);
double eqr_zcb(double a, double b, double sigma, double r, double t, double maturity);

/* ---- Philox4x32-10 (Salmon et al., SC'11) + Box-Muller normals ---- */
void eqr_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Normals Z[step][path] for steps [0,n_steps), paths [0,n_paths): counter = (step/2, path_lo,
 * path_hi, 0), key = (seed_lo, seed_hi). */
void eqr_philox_normals(uint64_t seed, uint64_t n_paths, uint32_t n_steps, double* z);
/* The two uniforms of one Philox draw, as used by Box-Muller: u1 in (0,1], u2 in [0,1). */
void eqr_philox_uniforms(uint64_t seed, uint64_t path, uint32_t pair, double* u1, double* u2);

/* ---- short-rate simulation, kernel.rs:41-50, hull_white.rs:87-97 ----
 * scheme 0 = Euler-Maruyama (reference), 1 = exact discretisation.
 * z: [n_steps][n_paths] standard normals; r_out: [n_steps+1][n_paths]. */
void eqr_simulate(double a, double b, double sigma, double r0, double dt, uint32_t n_steps,
                  uint64_t n_paths, int scheme, const double* z, double* r_out);

/* ---- vanilla payer IRS, products/irs.rs ---- */
int eqr_is_reset_date(double start, const double* pay, int m, double t);
double eqr_par_rate(double a, double b, double sigma, double r, double start,
                    const double* pay, const double* acc, int m);
/* NPV at a reset date t (reference semantics: remaining_from(t) then npv, irs.rs:46-68,163-187). */
double eqr_irs_npv_reset(double a, double b, double sigma, double r, double t, double notional,
                         double fixed_rate, double start, const double* pay, const double* acc, int m);
/* Extended valuation at any date t (between resets the floating leg carries the fixing taken
 * at the start of the period in progress, from r_fix = short rate at that date). */
double eqr_irs_value(double a, double b, double sigma, double r, double t, double r_fix,
                     double notional, double fixed_rate, double start, const double* pay,
                     const double* acc, int m);
/* Start date of the period in progress at t, or -1 when no fixing is needed at t. */
double eqr_fixing_date(double t, double start, const double* pay, int m);

/* ---- exposure aggregation, exposure.rs:42-49 ---- */
void eqr_ee_pfe(double* exposures, uint64_t n, double* ee, double* pfe);

/* ---- CVA, exposure.rs:242-275 ---- */
double eqr_unilateral_cva(double a, double b, double sigma, double r0, const double* times,
                          const double* ee, int n, double hazard_rate, double recovery_rate);

/* ---- reference exposure profile of one swap, exposure.rs:73-138 + api.rs:42-66 ----
 * use_par_rate: K = par rate under the model at r0 (api.rs:53-57).
 * z: [n_steps][n_paths]. Returns the number of dates written (1 for the t_max<=0 branch). */
int eqr_exposure_profile(double a, double b, double sigma, double r0, double notional,
                         double fixed_rate, int use_par_rate, double start, const double* pay,
                         const double* acc, int m, const double* times, int n_dates,
                         uint32_t n_steps, uint64_t n_paths, const double* z, double* ee,
                         double* pfe);

/* ---- market curve PV / DV01, market.cpp:26-41,95-111, measure.cpp:124-222 ---- */
double eqr_curve_zero_rate(const double* pillars, const double* zeros, int np, double t);
double eqr_curve_df(const double* pillars, const double* zeros, int np, double t);
double eqr_par_rate_from_market(const double* pillars, const double* zeros, int np, double start,
                                const double* pay, const double* acc, int m);
double eqr_npv_from_market(const double* pillars, const double* zeros, int np, double notional,
                           double fixed_rate, double start, const double* pay, const double* acc,
                           int m);
/* The same at valuation time v (every DF divided by DF(v)); NaN when v > start (refused by the
 * reference, measure.cpp:166-174). */
double eqr_npv_from_market_at(const double* pillars, const double* zeros, int np, double notional,
                              double fixed_rate, double start, const double* pay, const double* acc, int m,
                              double valuation_time);
double eqr_dv01_at(const double* pillars, const double* zeros, int np, double notional, double fixed_rate,
                   double start, const double* pay, const double* acc, int m, double bump, double valuation_time);
/* DV01 = NPV(z+bump) - NPV(z) at the same effective fixed rate; bucketed: one pillar at a time. */
double eqr_dv01(const double* pillars, const double* zeros, int np, double notional,
                double fixed_rate, double start, const double* pay, const double* acc, int m,
                double bump);
void eqr_dv01_bucketed(const double* pillars, const double* zeros, int np, double notional,
                       double fixed_rate, double start, const double* pay, const double* acc,
                       int m, double bump, double* out);

/* ---- portfolio / netting-set exposure (extended semantics, per-trade restatement) ----
 * trades j: set_id[j], notional[j] (signed: receiver < 0), fixed_rate[j], start[j], schedule
 * flows pay/acc[flow_off[j] .. flow_off[j+1]).
 * dates: monitoring times; grid dt = max(dates)/n_steps; r at date = r[round(t/dt)].
 * Outputs (nullable unless noted):
 *   set_ee, set_pfe [n_sets][n_dates] (required); set_cva [n_sets];
 *   trade_ee, trade_pfe [n_trades][n_dates]; trade_cva [n_trades];
 *   path_mtm [n_dates][n_trades][n_paths]; set_mtm [n_dates][n_sets][n_paths].
 * n_threads: OpenMP threads over paths (<=0: library default). */
int eqr_portfolio_exposure(double a, double b, double sigma, double r0, int n_trades, int n_sets,
                           const int32_t* set_id, const double* notional,
                           const double* fixed_rate, const double* start,
                           const int32_t* flow_off, const double* pay, const double* acc,
                           const double* dates, int n_dates, uint32_t n_steps, uint64_t n_paths,
                           int scheme, const double* z, double hazard_rate, double recovery_rate,
                           double* set_ee, double* set_pfe, double* set_cva, double* trade_ee,
                           double* trade_pfe, double* trade_cva, double* path_mtm,
                           double* set_mtm, int n_threads);

/* ---- the same run as a fast CPU engine organises it (eq_ref_fused.c) ----
 * One simulation per path, every zero-coupon bond of a distinct live cash-flow time once per
 * (path, date), OpenMP over paths, PFE95 by selection. z = NULL draws the Philox stream of `seed`
 * per path (the normals eqr_philox_normals would produce). share_classes = 1: trades with identical
 * schedules share one floating leg and one annuity per (path, date). Outputs: set_ee, set_pfe
 * [n_sets][n_dates] (required); set_cva [n_sets], trade_ee [n_trades][n_dates], trade_cva, set_mtm
 * [n_dates][n_sets][n_paths] (nullable; a date on grid index 0 holds path 0's value and zeros). */
int eqr_portfolio_exposure_fused(double a, double b, double sigma, double r0, int n_trades, int n_sets,
                                 const int32_t* set_id, const double* notional, const double* fixed_rate,
                                 const double* start, const int32_t* flow_off, const double* pay,
                                 const double* acc, const double* dates, int n_dates, uint32_t n_steps,
                                 uint64_t n_paths, int scheme, const double* z, uint64_t seed,
                                 double hazard_rate, double recovery_rate, int share_classes, double* set_ee,
                                 double* set_pfe, double* set_cva, double* trade_ee, double* trade_cva,
                                 double* set_mtm, int n_threads);

/* ---- Hull-White 2F (G2++), models/hull_white_2f.rs; state (x, y) from (0, 0) ---- */
double eqr_hw2_var(double z, double vol, double tau);
double eqr_hw2_cross(double a, double b, double sigma, double eta, double rho, double tau);
double eqr_zcb2(double a, double b, double sigma, double eta, double rho, double r0, double x, double y,
                double t, double maturity);
/* z[2][n_steps][n_paths] from Philox counter stream 1 (x shocks, then independent y shocks). */
void eqr_philox_normals2(uint64_t seed, uint64_t n_paths, uint32_t n_steps, double* z);
/* scheme 0 Euler (reference), 1 exact; x_out, y_out [n_steps+1][n_paths]. */
void eqr_simulate2(double a, double b, double sigma, double eta, double rho, double dt, uint32_t n_steps,
                   uint64_t n_paths, int scheme, const double* z, double* x_out, double* y_out);
double eqr_par_rate2(double a, double b, double sigma, double eta, double rho, double r0, double start,
                     const double* pay, const double* acc, int m);
double eqr_irs_npv_reset2(double a, double b, double sigma, double eta, double rho, double r0, double x,
                          double y, double t, double notional, double fixed_rate, double start,
                          const double* pay, const double* acc, int m);
double eqr_irs_value2(double a, double b, double sigma, double eta, double rho, double r0, double x, double y,
                      double t, double xf, double yf, double notional, double fixed_rate, double start,
                      const double* pay, const double* acc, int m);
double eqr_unilateral_cva2(double a, double b, double sigma, double eta, double rho, double r0,
                           const double* times, const double* ee, int n, double hazard_rate,
                           double recovery_rate);
int eqr_exposure_profile2(double a, double b, double sigma, double eta, double rho, double r0,
                          double notional, double fixed_rate, int use_par_rate, double start,
                          const double* pay, const double* acc, int m, const double* times, int n_dates,
                          uint32_t n_steps, uint64_t n_paths, const double* z, double* ee, double* pfe);
/* eqr_portfolio_exposure under G2++; z is [2][n_steps][n_paths]. */
int eqr_portfolio_exposure2(double a, double b, double sigma, double eta, double rho, double r0, int n_trades,
                            int n_sets, const int32_t* set_id, const double* notional, const double* fixed_rate,
                            const double* start, const int32_t* flow_off, const double* pay, const double* acc,
                            const double* dates, int n_dates, uint32_t n_steps, uint64_t n_paths, int scheme,
                            const double* z, double hazard_rate, double recovery_rate, double* set_ee,
                            double* set_pfe, double* set_cva, double* trade_ee, double* trade_pfe,
                            double* trade_cva, double* path_mtm, double* set_mtm, int n_threads);

/* ---- XVA reducer over an exposure cube, quant-engine/src/xva.rs:14-246 ---- */
double eqr_collateral_adjust(double value, int has_csa, int variation_margin, double threshold, double mta,
                             double ia);
void eqr_xva_calculate(const double* times, int n_times, uint64_t n_paths, const double* q, const double* df,
                       const double* p, double cp_hazard, double cp_recovery, const double* cp_sp, int cp_nsp,
                       double own_hazard, double own_recovery, const double* own_sp, int own_nsp,
                       double funding_spread, double borrowing_spread, double capital_factor,
                       double initial_margin_total, int has_csa, int variation_margin, double threshold,
                       double mta, double ia, double confidence, double* out_ee, double* out_pfe,
                       double* out_measures);

/* ---- Bermudan payer swaptions, Longstaff-Schwartz with a degree-3 basis ---- */
void eqr_berm_simulate(double a, double b, double sigma, double r0, double dt, uint32_t n_steps,
                       uint64_t n_paths, const double* z, double* r_out, double* lb_out);
double eqr_berm_x(double a, double b, double sigma, double r0, double t, double r);
int eqr_berm_policy(double a, double b, double sigma, double r0, int n_trades, const double* notional,
                    const double* strike, const int32_t* first_ex, const int32_t* n_ex, int max_ex,
                    uint32_t steps_per_year, uint64_t n_paths, const double* r, const double* lb,
                    double* ex_coef, int32_t* tau_year, double* y0, double* price);
int eqr_berm_apply(double a, double b, double sigma, double r0, int n_trades, const double* notional,
                   const double* strike, const int32_t* first_ex, const int32_t* n_ex, int max_ex,
                   uint32_t steps_per_year, uint64_t n_paths, const double* r, const double* lb,
                   const double* ex_coef, int32_t* tau_year, double* y0);
uint64_t eqr_berm_judge_fit(double a, double b, double sigma, double r0, uint64_t n_paths,
                            uint64_t step, double t, const double* r, const double* lb,
                            const int32_t* tau_year, const double* y0, double* coef);
double eqr_berm_ee(double a, double b, double sigma, double r0, uint64_t n_paths, uint64_t step,
                   double t, const double* r, const int32_t* tau_year, const double* coef);

/* Maximum OpenMP threads the oracle would use (for the CPU-baseline report). */
int eqr_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif
