/*
 * eq_ref_fused — the oracle's portfolio exposure run organised as a fast CPU engine would be.
 *
 * TEST INFRASTRUCTURE ONLY: the timed CPU baseline beside the GPU path, and a checker fast enough
 * for full-size portfolios. The product never links, loads or calls this file's library.
 *
 * Same semantics and expression order as eqr_portfolio_exposure (eq_ref.c, which restates
 * exposure.rs:42-294, irs.rs:46-187 and hull_white.rs:50-97), reorganised:
 *   - one simulation of the short rate per path for the whole portfolio;
 *   - per (path, date) every zero-coupon bond P(t_k, T_u) of a distinct live cash-flow time T_u is
 *     evaluated once, A(tau) exp(-(B(tau) r)) with A, B tabulated per (date, time), and shared by every
 *     trade: the factor eqr_zcb would recompute for each trade, bit for bit;
 *   - per (path, date) every fixing factor P(T_fix, T_next | r(T_fix)) of a period in progress once;
 *   - OpenMP over paths; PFE95 by selection (quickselect) instead of a full sort, the same order
 *     statistic sorted[ceil(0.95 n) - 1];
 *   - share_classes = 1: trades with identical (start, payment times, accruals) share one floating
 *     leg and one annuity per (path, date), each trade then costs N F - (N K) A. The fixed leg rounds
 *     as (N K) * sum(P tau) instead of sum((N K) * (P tau)), so values agree with the per-trade
 *     restatement to rounding, not bit for bit.
 * With share_classes = 0 the per-path netting-set values, set EE and set PFE equal
 * eqr_portfolio_exposure's bit for bit; per-trade EE is summed per thread and then across threads,
 * so it equals the path-order sum to rounding.
 */
#include "eq_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FUSED_TOL 1e-9 /* date-matching tolerance of products/irs.rs:151-154,173 */

static int cmp_dbl(const void* x, const void* y) {
    double a = *(const double*)x, b = *(const double*)y;
    return (a > b) - (a < b);
}

/* k-th smallest of x[0..n) (x is reordered): Hoare partition around a median of three. */
static double select_kth(double* x, uint64_t n, uint64_t k) {
    uint64_t lo = 0, hi = n - 1;
    while (hi > lo) {
        if (hi - lo < 16) {
            qsort(x + lo, (size_t)(hi - lo + 1), sizeof(double), cmp_dbl);
            return x[k];
        }
        uint64_t mid = lo + (hi - lo) / 2;
        double a = x[lo], b = x[mid], c = x[hi];
        double pivot = (a < b) ? ((b < c) ? b : ((a < c) ? c : a)) : ((a < c) ? a : ((b < c) ? c : b));
        uint64_t i = lo, j = hi;
        for (;;) {
            while (x[i] < pivot) ++i;
            while (x[j] > pivot) --j;
            if (i >= j) break;
            double t = x[i];
            x[i] = x[j];
            x[j] = t;
            ++i;
            --j;
        }
        /* [lo, j] <= pivot <= [j + 1, hi] */
        if (k <= j)
            hi = j;
        else
            lo = j + 1;
    }
    return x[k];
}

/* Per (date, trade or class) valuation descriptor. */
typedef struct {
    int mode;  /* 0 matured, 1 floating leg from P(t,t), 2 forward start P(t,start), 3 period in progress */
    int first; /* first flow with T_i > t + tol */
    int ustart; /* unique-time index of start (mode 2) */
    int fix;   /* index into the date's fixing list (mode 3) */
} leg_desc;

typedef struct {
    double prev, next; /* fixing date and the period's payment time */
    uint64_t grid;     /* grid index of the fixing date */
    double A, B;       /* P(prev, next | r) = A exp(-(B r)) */
} fixing;

static int find_time(const double* ut, int nu, double t) {
    int lo = 0, hi = nu - 1;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (ut[mid] < t)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

/* The valuation branch of value_m (eq_ref.c) for a schedule at date t. */
static leg_desc describe(double t, double start, const double* pay, int m, const double* ut, int nu,
                         fixing** fx, int* n_fx, int* cap_fx, double dt, int deterministic) {
    leg_desc d = {0, 0, -1, -1};
    if (t >= pay[m - 1] - FUSED_TOL) return d;
    int first = 0;
    while (first < m && !(pay[first] > t + FUSED_TOL)) ++first;
    d.first = first;
    if (fabs(t - start) < FUSED_TOL) {
        d.mode = 1;
    } else if (t < start) {
        d.mode = 2;
        d.ustart = find_time(ut, nu, start);
    } else {
        double prev = (first == 0) ? start : pay[first - 1];
        if (fabs(t - prev) < FUSED_TOL) {
            d.mode = 1;
        } else {
            d.mode = 3;
            double next = pay[first];
            int f = 0;
            for (; f < *n_fx; ++f)
                if ((*fx)[f].prev == prev && (*fx)[f].next == next) break;
            if (f == *n_fx) {
                if (*n_fx == *cap_fx) {
                    *cap_fx = *cap_fx ? 2 * *cap_fx : 16;
                    *fx = (fixing*)realloc(*fx, sizeof(fixing) * (size_t)*cap_fx);
                }
                fixing* g = &(*fx)[f];
                g->prev = prev;
                g->next = next;
                g->grid = deterministic ? 0 : (uint64_t)round(prev / dt);
                g->A = g->B = 0.0; /* tabulated once the date's list is complete */
                (*n_fx)++;
            }
            d.fix = f;
        }
    }
    return d;
}

int eqr_portfolio_exposure_fused(double a, double b, double sigma, double r0, int n_trades, int n_sets,
                                 const int32_t* set_id, const double* notional, const double* fixed_rate,
                                 const double* start, const int32_t* flow_off, const double* pay,
                                 const double* acc, const double* dates, int n_dates, uint32_t n_steps,
                                 uint64_t n_paths, int scheme, const double* z, uint64_t seed,
                                 double hazard_rate, double recovery_rate, int share_classes, double* set_ee,
                                 double* set_pfe, double* set_cva, double* trade_ee, double* trade_cva,
                                 double* set_mtm, int n_threads) {
    if (n_trades <= 0 || n_dates <= 0 || n_paths == 0) return 1;
    double t_max = -INFINITY;
    for (int k = 0; k < n_dates; ++k) t_max = fmax(t_max, dates[k]);
    int deterministic = !(t_max > 0.0);
    double dt = deterministic ? 1.0 : t_max / (double)n_steps;
    uint32_t steps = deterministic ? 0 : n_steps;
    uint64_t* idx = (uint64_t*)malloc(sizeof(uint64_t) * n_dates);
    for (int k = 0; k < n_dates; ++k) idx[k] = deterministic ? 0 : (uint64_t)round(dates[k] / dt);

    /* schedule classes (each trade its own class unless shared) */
    int* cls = (int*)malloc(sizeof(int) * n_trades);
    int* rep = (int*)malloc(sizeof(int) * n_trades); /* a trade of each class */
    int n_cls = 0;
    for (int j = 0; j < n_trades; ++j) {
        int c = -1;
        if (share_classes) {
            int mj = flow_off[j + 1] - flow_off[j];
            for (int q = 0; q < n_cls && c < 0; ++q) {
                int i = rep[q];
                int mi = flow_off[i + 1] - flow_off[i];
                if (mi != mj || start[i] != start[j]) continue;
                if (memcmp(pay + flow_off[i], pay + flow_off[j], sizeof(double) * mj) == 0 &&
                    memcmp(acc + flow_off[i], acc + flow_off[j], sizeof(double) * mj) == 0)
                    c = q;
            }
        }
        if (c < 0) {
            c = n_cls++;
            rep[c] = j;
        }
        cls[j] = c;
    }

    /* distinct cash-flow times (payment times and starts), ascending */
    int n_all = flow_off[n_trades] + n_trades;
    double* ut = (double*)malloc(sizeof(double) * n_all);
    memcpy(ut, pay, sizeof(double) * flow_off[n_trades]);
    memcpy(ut + flow_off[n_trades], start, sizeof(double) * n_trades);
    qsort(ut, (size_t)n_all, sizeof(double), cmp_dbl);
    int nu = 0;
    for (int i = 0; i < n_all; ++i)
        if (nu == 0 || ut[i] != ut[nu - 1]) ut[nu++] = ut[i];
    int* uflow = (int*)malloc(sizeof(int) * (flow_off[n_trades] > 0 ? flow_off[n_trades] : 1));
    for (int f = 0; f < flow_off[n_trades]; ++f) uflow[f] = find_time(ut, nu, pay[f]);

    /* per date: first live time, (A, B) of P(t_k, T_u), class descriptors, fixings */
    int* ufirst = (int*)malloc(sizeof(int) * n_dates);
    double* tabA = (double*)malloc(sizeof(double) * (size_t)n_dates * nu);
    double* tabB = (double*)malloc(sizeof(double) * (size_t)n_dates * nu);
    leg_desc* desc = (leg_desc*)malloc(sizeof(leg_desc) * (size_t)n_dates * n_cls);
    int* fix_off = (int*)malloc(sizeof(int) * (n_dates + 1));
    fixing* fixes = NULL;
    int n_fix_all = 0, cap_all = 0;
    for (int k = 0; k < n_dates; ++k) {
        double t = dates[k];
        int u0 = 0; /* every time at or after t: the live flows and forward starts are among them */
        while (u0 < nu && ut[u0] < t) ++u0;
        ufirst[k] = u0;
        for (int u = 0; u < nu; ++u) {
            double tau = ut[u] - t; /* the maturity - t of eqr_zcb */
            double bf = eqr_hw_b(a, tau);
            tabB[(size_t)k * nu + u] = bf;
            tabA[(size_t)k * nu + u] = eqr_hw_a(a, b, sigma, tau, bf);
        }
        fixing* fx = NULL;
        int n_fx = 0, cap_fx = 0;
        for (int c = 0; c < n_cls; ++c) {
            int j = rep[c];
            desc[(size_t)k * n_cls + c] = describe(t, start[j], pay + flow_off[j], flow_off[j + 1] - flow_off[j],
                                                   ut, nu, &fx, &n_fx, &cap_fx, dt, deterministic);
        }
        fix_off[k] = n_fix_all;
        for (int f = 0; f < n_fx; ++f) {
            double tau = fx[f].next - fx[f].prev;
            fx[f].B = eqr_hw_b(a, tau);
            fx[f].A = eqr_hw_a(a, b, sigma, tau, fx[f].B);
            if (n_fix_all == cap_all) {
                cap_all = cap_all ? 2 * cap_all : 64;
                fixes = (fixing*)realloc(fixes, sizeof(fixing) * (size_t)cap_all);
            }
            fixes[n_fix_all++] = fx[f];
        }
        free(fx);
    }
    fix_off[n_dates] = n_fix_all;

    double* xs = (double*)malloc(sizeof(double) * (size_t)n_sets * n_dates * n_paths);
    int want_trade = (trade_ee != NULL) || (trade_cva != NULL);
    int nth = 1;
#ifdef _OPENMP
    nth = (n_threads > 0) ? n_threads : omp_get_max_threads();
#endif
    double* tpart = want_trade ? (double*)calloc((size_t)nth * n_trades * n_dates, sizeof(double)) : NULL;
    const int max_fx_date = 1 + n_fix_all;

#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* rp = (double*)malloc(sizeof(double) * (steps + 1));
        double* zl = (double*)malloc(sizeof(double) * (steps + 2));
        double* P = (double*)malloc(sizeof(double) * nu);
        double* PF = (double*)malloc(sizeof(double) * max_fx_date);
        double* cF = (double*)malloc(sizeof(double) * n_cls); /* class floating-leg factor */
        double* cA = (double*)malloc(sizeof(double) * n_cls); /* class annuity (shared mode) */
        double* vset = (double*)malloc(sizeof(double) * n_sets);
        double* tp = want_trade ? tpart + (size_t)tid * n_trades * n_dates : NULL;
        const double two_pi = 6.283185307179586476925286766559;
        const double two_m53 = 1.0 / 9007199254740992.0;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t pp = 0; pp < (int64_t)n_paths; ++pp) {
            uint64_t p = (uint64_t)pp;
            /* the path's shocks: the caller's Z[step][path], or its Philox stream drawn here */
            const double* zp = z;
            uint64_t zstride = n_paths, zcol = p;
            if (!z && steps > 0) {
                for (uint32_t s = 0; s < steps; s += 2) {
                    uint32_t ctr[4] = {s / 2u, (uint32_t)p, (uint32_t)(p >> 32), 0u};
                    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
                    uint32_t o[4];
                    eqr_philox4x32_10(ctr, key, o);
                    uint64_t wa = ((uint64_t)o[1] << 32) | o[0];
                    uint64_t wb = ((uint64_t)o[3] << 32) | o[2];
                    double u1 = (double)((wa >> 11) + 1u) * two_m53, u2 = (double)(wb >> 11) * two_m53;
                    double rad = sqrt(-2.0 * log(u1)), th = two_pi * u2;
                    zl[s] = rad * cos(th);
                    zl[s + 1] = rad * sin(th);
                }
                zp = zl;
                zstride = 1;
                zcol = 0;
            }
            /* simulate_one of eq_ref.c */
            double r = r0;
            rp[0] = r;
            if (scheme == 0) {
                double sq = sqrt(dt);
                for (uint32_t s = 0; s < steps; ++s) {
                    double drift = a * (b - r);
                    double dw = zp[(uint64_t)s * zstride + zcol] * sq;
                    r = (r + drift * dt) + sigma * dw;
                    rp[s + 1] = r;
                }
            } else {
                double e = exp(-(a * dt));
                double mean_c = b * (1.0 - e);
                double sd = sigma * sqrt((1.0 - exp(-(2.0 * a * dt))) / (2.0 * a));
                for (uint32_t s = 0; s < steps; ++s) {
                    r = (r * e + mean_c) + sd * zp[(uint64_t)s * zstride + zcol];
                    rp[s + 1] = r;
                }
            }
            for (int k = 0; k < n_dates; ++k) {
                const int row1 = (idx[k] == 0); /* grid index 0: path 0 alone (exposure.rs:98-123) */
                if (row1 && p != 0) {
                    for (int s = 0; s < n_sets; ++s) {
                        xs[((size_t)s * n_dates + k) * n_paths + p] = 0.0;
                        if (set_mtm) set_mtm[((size_t)k * n_sets + s) * n_paths + p] = 0.0;
                    }
                    continue;
                }
                double rk = rp[idx[k]];
                const double* A = tabA + (size_t)k * nu;
                const double* B = tabB + (size_t)k * nu;
                for (int u = ufirst[k]; u < nu; ++u) P[u] = A[u] * exp(-(B[u] * rk));
                /* P(t, t) as eqr_zcb evaluates it: tau = 0 */
                double bt = eqr_hw_b(a, 0.0);
                double p_tt = eqr_hw_a(a, b, sigma, 0.0, bt) * exp(-(bt * rk));
                const fixing* fk = fixes + fix_off[k];
                for (int f = 0; f < fix_off[k + 1] - fix_off[k]; ++f)
                    PF[f] = fk[f].A * exp(-(fk[f].B * rp[fk[f].grid]));
                const leg_desc* dk = desc + (size_t)k * n_cls;
                for (int c = 0; c < n_cls; ++c) {
                    const leg_desc d = dk[c];
                    if (d.mode == 0) continue;
                    int j = rep[c];
                    int f0 = flow_off[j], m = flow_off[j + 1] - f0;
                    double p_end = P[uflow[f0 + m - 1]];
                    double fl;
                    if (d.mode == 1)
                        fl = p_tt - p_end;
                    else if (d.mode == 2)
                        fl = P[d.ustart] - p_end;
                    else
                        fl = P[uflow[f0 + d.first]] / PF[d.fix] - p_end;
                    cF[c] = fl;
                    if (share_classes) {
                        double ann = 0.0;
                        for (int i = d.first; i < m; ++i) {
                            double term = P[uflow[f0 + i]] * acc[f0 + i];
                            ann = (i == d.first) ? term : ann + term;
                        }
                        cA[c] = ann;
                    }
                }
                for (int s = 0; s < n_sets; ++s) vset[s] = 0.0;
                for (int j = 0; j < n_trades; ++j) {
                    const int c = cls[j];
                    const leg_desc d = dk[c];
                    double v;
                    if (d.mode == 0) {
                        v = 0.0;
                    } else {
                        double nk = notional[j] * fixed_rate[j];
                        double floating_leg = notional[j] * cF[c];
                        double fixed_leg;
                        if (share_classes) {
                            fixed_leg = nk * cA[c];
                        } else {
                            int f0 = flow_off[j], m = flow_off[j + 1] - f0;
                            fixed_leg = 0.0;
                            for (int i = d.first; i < m; ++i) {
                                double leg = nk * (P[uflow[f0 + i]] * acc[f0 + i]);
                                fixed_leg = (i == d.first) ? leg : fixed_leg + leg;
                            }
                        }
                        v = floating_leg - fixed_leg;
                    }
                    vset[set_id[j]] += v;
                    if (tp) tp[(size_t)j * n_dates + k] += fmax(v, 0.0);
                }
                for (int s = 0; s < n_sets; ++s) {
                    xs[((size_t)s * n_dates + k) * n_paths + p] = fmax(vset[s], 0.0);
                    if (set_mtm) set_mtm[((size_t)k * n_sets + s) * n_paths + p] = vset[s];
                }
            }
        }
        free(vset);
        free(cA);
        free(cF);
        free(PF);
        free(P);
        free(zl);
        free(rp);
    }

    /* set rows: EE in path order (as eqr_ee_pfe), PFE95 by selection of the same order statistic */
#ifdef _OPENMP
#pragma omp parallel num_threads(nth)
#endif
    {
        double* scratch = (double*)malloc(sizeof(double) * n_paths);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t row = 0; row < (int64_t)n_sets * n_dates; ++row) {
            int k = (int)(row % n_dates);
            uint64_t n = (idx[k] == 0) ? 1 : n_paths;
            const double* x = xs + (size_t)row * n_paths;
            double sum = 0.0;
            for (uint64_t i = 0; i < n; ++i) sum += x[i];
            set_ee[row] = sum / (double)n;
            memcpy(scratch, x, sizeof(double) * n);
            uint64_t q = (uint64_t)ceil(0.95 * (double)n);
            q = (q == 0) ? 0 : q - 1;
            if (q > n - 1) q = n - 1;
            set_pfe[row] = select_kth(scratch, n, q);
        }
        free(scratch);
    }
    if (set_cva)
        for (int s = 0; s < n_sets; ++s)
            set_cva[s] = eqr_unilateral_cva(a, b, sigma, r0, dates, set_ee + (size_t)s * n_dates, n_dates,
                                            hazard_rate, recovery_rate);
    if (want_trade) {
        double* tee = (double*)malloc(sizeof(double) * (size_t)n_trades * n_dates);
        for (size_t e = 0; e < (size_t)n_trades * n_dates; ++e) {
            double sum = 0.0;
            for (int th = 0; th < nth; ++th) sum += tpart[(size_t)th * n_trades * n_dates + e];
            int k = (int)(e % n_dates);
            tee[e] = sum / (double)((idx[k] == 0) ? 1 : n_paths);
        }
        if (trade_ee) memcpy(trade_ee, tee, sizeof(double) * (size_t)n_trades * n_dates);
        if (trade_cva)
            for (int j = 0; j < n_trades; ++j)
                trade_cva[j] = eqr_unilateral_cva(a, b, sigma, r0, dates, tee + (size_t)j * n_dates, n_dates,
                                                  hazard_rate, recovery_rate);
        free(tee);
    }
    free(tpart);
    free(xs);
    free(fixes);
    free(fix_off);
    free(desc);
    free(tabB);
    free(tabA);
    free(ufirst);
    free(uflow);
    free(ut);
    free(rep);
    free(cls);
    free(idx);
    return 0;
}
