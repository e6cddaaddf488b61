// Kernel argument blocks and launch constants shared by the kernels and the host plan.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace eqhip {

constexpr int kValueThreads = 1024;  // 16 wave64s per valuation block
constexpr int kValueWaves = kValueThreads / 64;
constexpr int kPathsPerTile = 64;    // paths of a tile per path-per-lane: a tile is 64 * PPL paths
constexpr int kTradeSlots = kValueThreads;  // (item, path slice) work slots per block, one per thread
// Valuation variants: PPL paths per lane (1, 2, 4) and TPI trades of one class and sign per work
// item (2, 4, 8); four paths per lane take two trades per item. Two-factor plans use PPL = 1.
__host__ __device__ constexpr int kVariant(int ppl, int tpi) { return ppl * 16 + tpi; }
// Path mode: plans without per-trade outputs, with at most kPathFamilies schedule families and
// kPathSets netting sets, are valued by value_path_kernel: path_ppt(NS) paths per thread, the family
// sums and set values in registers, no LDS and no barriers. A tile is then kPathThreads x
// path_ppt(NS) paths, NS the set registers of the launch.
constexpr int kPathVariant = 1;
constexpr int kPathFamilies = 4;
constexpr int kPathSets = 8;
constexpr int kPathThreads = 256;
__host__ __device__ constexpr int path_set_regs(int ns) { return ns <= 1 ? 1 : (ns <= 2 ? 2 : (ns <= 4 ? 4 : 8)); }
__host__ __device__ constexpr int path_ppt(int ns) { return path_set_regs(ns) > 4 ? 2 : 4; }
constexpr int kDefaultPpl = 1, kDefaultTpi = 4;
constexpr int kMaxSplitSets = 8;     // netting sets whose class sums are split across the waves
constexpr int kSelectThreads = 1024;  // EE / PFE select block
constexpr int kPvLdsPillars = 256;    // curves up to this many pillars are staged in LDS by pv_dv01_kernel
constexpr int kSelectCand = 4096;     // candidates ranked in LDS by the select kernel
constexpr int kSelectBracket = 2048;  // keys of the first pass's bracket kept in LDS by the select kernel
constexpr int kSelectWin = 8;         // binades of the select kernel's first-pass mantissa histogram
constexpr double kDateTolHost = 1e-9; // date matching tolerance, products/irs.rs:151-154,173

struct SimArgs {
    double a, b, sigma, r0;
    double dt, sqrt_dt;          // Euler
    double decay, mean_c, sd;    // exact transition
    uint64_t n_paths;
    uint32_t n_steps;
    uint64_t seed;
    const int* slot_of_step;     // [n_steps + 1], -1 when r is not stored at that step
    const double* Z;             // [n_steps][n_paths] host draws
    double* R;                   // [n_slots][n_paths]
};

// G2++ (models/hull_white_2f.rs:1-147): x and y start at 0, r = x + y + r0. Euler (the reference's
// scheme): x' = (x + (-a x) dt) + sigma (zx sqrt(dt)), y' likewise with b, eta and
// zy = zy' sqrt(1 - rho^2) + zx rho. Exact: each factor's own transition, the two Gaussian
// increments correlated by their exact covariance.
struct Sim2Args {
    double a, b, sigma, eta, rho;
    double dt, sqrt_dt, sqrt_1mrho2;                  // Euler
    double ex, ey, sdx, sdy, c_xy, sqrt_1mc2;         // exact
    uint64_t n_paths;
    uint32_t n_steps;
    uint64_t seed;
    const int* slot_of_step;
    const double* Z;  // host draws [2][n_steps][n_paths]: every x shock, then every y shock
    double* X;        // [n_slots][n_paths]
    double* Y;        // [n_slots][n_paths]
};

// One cash-flow date of a valuation date in path mode: the bond P = A exp(-B r) of date d, then for
// every family f its flow at d (annuity_f += P tau[f]; tau[f] = 0 when f has no flow at d, an exact
// no-op), then the netting-set entries of the classes ending at d, family by family: n[f] entries
// from e0 on, V_s += alpha (st_f - P) - beta annuity_f.
struct PathBond {
    int d;
    int e0;
    int n[kPathFamilies];
    int pad[2];
    double tau[kPathFamilies];
};

// One exposure plan as the valuation kernel sees it. A launch takes an array of plans and a list
// of work units (plan, date, path chunk), so one launch can serve many netting-set plans.
struct ValuePlan {
    const double* R;          // [n_slots][r_stride]
    uint64_t r_stride;        // row stride of R (>= n_paths when R is shared by a batch)
    uint64_t n_paths;
    uint64_t n_tiles;         // ceil(n_paths / kPathsPerTile)
    int n_dates, D, C, NF, S, Fmax;
    int n_chunks, tiles_per_block;
    int set_mode;             // 0: S <= kMaxSplitSets, class sums split across waves; 1: one wave per set
    int variant;              // kVariant(PPL, TPI) of the valuation unit
    int x_mode;               // 0: X = max(V, 0) and set_mtm = V; a part of a plan valued in parts: 1 stores V
                              //    in X, 2 adds V to X, 3 adds, stores max(total, 0) and writes the total
    int two_factor;           // 1: G2++ (R holds x, R2 holds y; P = exp((c - B1 x) - B2 y) with AB = (c, B1),
                              //    AB2 = B2; fixings likewise with fix_AB, fix_B2)
    const double* R2;         // [n_slots][r_stride] second factor (two-factor plans)
    const double* AB2;        // [n_dates][D]
    const double* fix_B2;
    const double* dates;      // [n_dates]
    const double2* AB;        // [n_dates][D] (A, B) of P(t_k, T_d)
    const int* date_slot;     // [n_dates]
    const int* date_dfirst;   // [n_dates] first live cash-flow date
    const int* date_det;      // [n_dates] 1 when the date sits at grid index 0
    const int4* fdesc;        // [n_dates][NF] mode (0 dead, 1 reset, 2 forward, 3 mid), first live flow,
                              //   start/next date, fixing index
    const int* fam_off;       // [NF + 1] family flow ranges
    const int* flow_d;        // family flows: cash-flow date index
    const double* flow_tau;   //               accrual
    const int* flow_cls;      //               class whose last flow this is, or -1
    // per date: the live flows of every family in family order, split into kValueWaves segments
    // (one per wave), and the classes already matured
    const int* lflow_off;     // [n_dates + 1]
    const int4* lflow;        // (cash-flow date, class or -1, family, 0)
    const double* ltau;       // accrual of the flow
    const int4* lseg;         // [n_dates][kValueWaves] (begin, end, first of the earlier segments whose
                              //   totals carry into this one, number of leading flows that take the carry)
    const int* zcls_off;      // [n_dates + 1]
    const int* zcls;          // matured classes
    const int* fix_off;       // [n_dates + 1]
    const int* fix_slot;      // R slot of the fixing date
    const double2* fix_AB;    // (A, B) of the fixing period
    const int* set_off;       // [S + 1] CSR of (class, alpha, beta), each set's classes latest maturity first
    const int* set_nlive;     // [n_dates][S] classes of the set still alive at the date (a prefix)
    const int* set_cls;
    const double* set_alpha;  // sum of notionals of the set's trades in the class
    const double* set_beta;   // sum of notional * fixed rate
    double* X;                // [S][n_dates][n_paths] max(V_set, 0)
    double* set_mtm;          // optional [n_dates][S][n_paths]
    int want_trade, n_items, n_trades;
    const int4* items;        // (class, first sorted trade, count, sign), latest maturity last
    const int* item_first;    // [n_dates] first item alive at the date (the live items are a suffix)
    const double* cls_mat;    // [C] maturity of the class
    const double* tr_K;       // sorted trades: fixed rate, |notional|, notional, notional * rate
    const double* tr_absN;
    const double* tr_N;
    const double* tr_NK;
    const int* tr_orig;       // sorted -> caller trade index
    double* partial;          // [n_dates][n_chunks][n_trades] per-trade EE partial sums
    double* path_mtm;         // optional [n_dates][n_trades][n_paths]
    // per-trade PFE: the legs (F, A) of the classes alive at each date of a date group,
    // [cube_row[k]][cube_C][n_paths] (a part's classes start at cube_cls0), or null
    double2* ccube;
    const int* cube_row;      // [n_dates] the date's row in its date group
    const int* cls_live0;     // [n_dates] first class alive at the date (classes are numbered by maturity)
    int cube_C, cube_cls0;
    // path mode (variant kPathVariant): each date's live cash-flow dates in date order (PathBond),
    // and the netting-set entries of the classes ending at them (set, alpha, beta)
    const int* pe_off;        // [n_dates + 1]
    const PathBond* pe;
    const int* pc_set;
    const double2* pc_ab;     // (alpha, beta)
};

// Dynamic LDS of one value_kernel block for a tile of TW = 64 PPL paths: sP [D][TW], sF [Fmax][TW],
// sC [TW][C] (double2), set partials [S_split][waves][TW] and segment totals [waves][TW] (the per-date
// descriptor tables are read from global memory through the scalar cache).
__host__ __device__ inline size_t value_lds_bytes(int D, int Fmax, int C, int S_split, int ppl = 1) {
    const size_t tw = (size_t)kPathsPerTile * ppl;
    return 8 * (tw * (D + Fmax) + 2 * tw * C + tw * kValueWaves * S_split + tw * kValueWaves);
}

// LDS that holds the per-slot trade sums of a unit for their reduction (floor of the allocation)
__host__ __device__ inline size_t value_trade_scratch_bytes(int tpi) { return 8 * (size_t)kTradeSlots * tpi; }

struct PvArgs {
    const double* pillars;
    const double* zeros;
    int np;
    int n_trades;
    const int* flow_off;
    const double* pay;
    const double* acc;
    const double* notional;
    const double* fixed;
    const double* start;
    double bump;
    double valuation_time;  // PV and parallel DV01 at this time (every DF divided by DF(v)); bucketed at 0
    double* pv;
    double* dv01;
    double* bucketed;
    int bucketed_on;
};

// Pathwise adjoint of CVA for one netting set (reverse sweep with short-rate checkpoints).
constexpr int kAdjCheckpoint = 12;   // grid steps between stored short rates
constexpr int kAdjThreads = 256;

// A, B of P(t,T) = A exp(-B r) and their derivatives in a, b, sigma (B does not depend on b, sigma).
struct ZcbSens {
    double A, B, B_a, lnA_a, lnA_b, lnA_s;
};

struct AdjArgs {
    double a, b, sigma, r0;
    double dt, sqrt_dt, decay, mean_c, sd, dsd_da, dsd_ds;
    int disc, rng;
    uint64_t seed;
    const double* Z;              // host draws [n_steps][n_paths] (rng == host draws)
    uint64_t n_paths;
    uint32_t n_steps;
    const double* CP;             // [n_steps / K + 1][n_paths] r at steps 0, K, 2K, ...
    const double* RF;             // [n_fix_slots][n_paths] r at the fixing steps
    double* G;                    // [n_fix_slots][n_paths] pending adjoints of the fixing rates
    int n_fix_slots;
    const int* date_of_step;      // [n_steps + 1] monitoring date index or -1
    const int* fixslot_of_step;   // [n_steps + 1] fixing slot or -1
    int D;
    const int* date_dfirst;       // [n_dates]
    const double* C0;             // [n_dates] constant part of V_s (reset floating legs)
    const double* nu;             // [n_dates][D] V_s = C0 + sum_d nu_d P_d + mid-period terms
    const ZcbSens* zs;            // [n_dates][D]
    const int* mid_off;           // [n_dates + 1]
    const int4* mid;              // (next cash-flow date, fixing slot, fixing sens index, -)
    const double* mid_alpha;      // sum of notionals of the family's trades in the set
    const ZcbSens* fix_zs;        // fixing-period sensitivities
    const double* wk;             // [n_dates] (1 - R)(S_{k-1} - S_k) P(0,t_k) / n_eff(k)
    const int* date_det;          // [n_dates]
    double* partial;              // [blocks][5] d/da, d/db, d/dsigma, d/dr0, sum of w V+
};

template <int DISC, int RNG>
__global__ void simulate_kernel(SimArgs g);
__global__ void zcb_sens_kernel(const double* tau, int n, double a, double b, double sigma, ZcbSens* out);
template <int DISC, int RNG>
__global__ void adjoint_kernel(AdjArgs g);
__global__ void ab_kernel(const double* tau, int n, double a, double b, double sigma, double2* out);
template <int DISC, int RNG>
__global__ void simulate2_kernel(Sim2Args g);
__global__ void ab2_kernel(const double* tau, int n, double a, double b, double sigma, double eta, double rho,
                           double r0, double2* c_b1, double* b2);
__global__ void cva_weights2_kernel(const double* times, int n_dates, double a, double b, double sigma, double eta,
                                    double rho, double r0, double hazard, double2* w);
constexpr int kSimPaths = 64;  // paths per block of the Philox simulation kernels (4 threads per path)
constexpr int kSimBatch = 32;  // steps per batch of shocks in LDS (one factor; 16 for two factors)
template <int DISC> __global__ void simulate_philox_kernel(SimArgs g);
template <int DISC> __global__ void simulate2_philox_kernel(Sim2Args g);
__global__ void value_kernel(const ValuePlan* plans, const int4* units);
template <int TPI>
__global__ void value2_kernel(const ValuePlan* plans, const int4* units);
template <int TPI>
__global__ void value_p1_kernel(const ValuePlan* plans, const int4* units);    // one path per lane
template <int NFAC, int NFM, int NS>
__global__ void value_path_kernel(const ValuePlan* plans, const int4* units);  // path mode
template <int TPI>
__global__ void value_cube_kernel(const ValuePlan* plans, const int4* units);  // one path per lane
__global__ void value_kernel_persistent(const ValuePlan* plans, const int4* units, int n_units,
                                        unsigned int* counter);
__global__ void reset_counter_kernel(unsigned int* counter);
// One EE / PFE95 row: n values at x, rank q, outputs ee / pfe (device pointers).
struct SelectRow {
    const double* x;
    uint64_t n;
    uint64_t q;
    double* ee;
    double* pfe;
};
__global__ void select_kernel(const SelectRow* rows);
// One per-trade PFE95 row: the positive parts max(s F - (s K) A, 0) of a trade of sign s = sign(N) over
// the n paths of its class's legs c ([n] (F, A) pairs), rank q; pfe = |N| times that order statistic.
struct TradeSelectRow {
    const double2* c;
    uint64_t n;
    uint64_t q;
    double s, sK, absN;
    double* pfe;
};
__global__ void trade_select_kernel(const TradeSelectRow* rows);

// Homogeneous batch: one schedule, each trade its own netting set (the reference's
// npv_batch_over_paths, irs.rs:87-114), at dates that are reset dates, precede the start or follow
// maturity. Per (path, date) one floating leg P(t,S) - P(t,T_n) and one annuity sum_i P(t,T_i) tau_i
// over the remaining flows; per trade N * float - (N K) * annuity, positive part into X.
constexpr int kFlatThreads = 256;
struct FlatArgs {
    const double* R;       // short rates [slot][path], row stride r_stride
    uint64_t r_stride;
    uint64_t n_paths;
    int n_dates;
    int m;                 // flows of the schedule
    int n_trades;
    double r0;
    const int* date_slot;  // R row of each date
    const int* date_det;   // 1: the date sits on grid step 0 (r = r0)
    const int* first;      // first remaining flow of each date; m when matured
    const double2* AB;     // [date][m + 1]: (A, B) of P(t, T_i) for i < m, of P(t, S) at i = m
    const double* acc;     // [m] accruals
    const double* N;       // [n_trades] signed notionals
    const double* NK;      // [n_trades] notional * fixed rate
    double* X;             // [trade][date][path] positive parts
};
__global__ void value_flat_kernel(FlatArgs g);
__global__ void bond_exp_kernel(const double* x, double* y, int n);

// One time row of an exposure cube for the XVA reducer (quant-engine/src/xva.rs:150-183): the row's
// values after the first CSA's variation margin, their positive parts (written out for the quantile)
// and the fixed-order sums of their negative parts, for the Q row and, when given, the P row.
struct XvaRowArgs {
    const double* q;     // row k at q + k * q_stride, n values
    uint64_t q_stride;
    const double* p;     // the P cube's rows (nullable), same shape
    uint64_t p_stride;
    uint64_t n;
    double* pos;         // [rows][n] positive parts of the adjusted Q rows
    double* ene_q;       // [rows] sum of negative parts of the adjusted Q row
    double* ene_p;       // [rows] the same for the P row (when p)
    int csa_vm;          // 1: a CSA with variation margin adjusts the values
    double threshold, mta, ia;  // max(x, 0) of the CSA's threshold, minimum transfer, independent amount
};
__global__ void xva_rows_kernel(XvaRowArgs g);
__global__ void trade_ee_kernel(const double* partial, int n_trades, int n_dates, int n_chunks,
                                const uint64_t* n_eff, double* trade_ee);
__global__ void cva_weights_kernel(const double* times, int n_dates, double a, double b, double sigma,
                                   double r0, double hazard, double2* w);
__global__ void cva_kernel(const double* ee, int n_profiles, int n_dates, const double2* w, double recovery,
                           double* cva);
__global__ void pv_dv01_kernel(PvArgs g);
__global__ void pv_sens_kernel(PvArgs g, double a, double b, double sigma, double r0, int want_model,
                               double* pillar_delta, double* model_npv, double* model_greeks);

}  // namespace eqhip
