/*
 * eq_ref — CPU FP64 restatement of engine_quant's Monte Carlo exposure path.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle the HIP path is checked against, and the CPU baseline
 * timed beside it. The product never links, loads or calls this file's library.
 *
 * Compiled with -ffp-contract=off so that every product and sum is rounded exactly as the
 * reference's Rust/Burn and C++ code rounds it (neither contracts a*b+c into an FMA).
 * File:line citations are relative to the reference repository root.
 */
#include "eq_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EQR_TOL 1e-9 /* date-matching tolerance of products/irs.rs:151-154,173 */

/* ------------------------------------------------------------------------------------------ */
/* Hull-White 1F (Vasicek form)                                                               */
/* ------------------------------------------------------------------------------------------ */

/* B(tau) = (1 - exp(-a tau)) / a, built as Burn does it: (-(exp(a*(-tau)))) + 1, then / a.
 * models/hull_white.rs:51-55 */
double eqr_hw_b(double a, double tau) {
    double neg_a_tau = a * (-tau);
    double one_minus_exp = (-exp(neg_a_tau)) + 1.0;
    return one_minus_exp / a;
}

/* A(tau) = exp(((B - tau)(a^2 b - sigma^2/2)) / a^2 - (sigma^2 B B) / (4a)).
 * models/hull_white.rs:59-69 */
double eqr_hw_a(double a, double b, double sigma, double tau, double bf) {
    double a2 = a * a;
    double sigma2 = sigma * sigma;
    double term1 = ((bf + (-tau)) * (a2 * b - sigma2 * 0.5)) / a2;
    double term2 = ((sigma2 * bf) * bf) / (a * 4.0);
    return exp(term1 - term2);
}

/* P(t,T) = A(T-t) * exp(-(B(T-t) * r)). models/hull_white.rs:75-80 */
double eqr_zcb(double a, double b, double sigma, double r, double t, double maturity) {
    double tau = maturity - t;
    double bf = eqr_hw_b(a, tau);
    double af = eqr_hw_a(a, b, sigma, tau, bf);
    return af * exp(-(bf * r));
}

/* ------------------------------------------------------------------------------------------ */
/* Hull-White 2F (G2++): r = x + y + r0, dx = -a x dt + sigma dW1, dy = -b y dt + eta dW2,     */
/* d<W1,W2> = rho dt, x0 = y0 = 0 (models/hull_white_2f.rs)                                    */
/* ------------------------------------------------------------------------------------------ */

/* vol^2/z^2 * [tau + (2/z)e^{-z tau} - (1/(2z))e^{-2 z tau} - 3/(2z)], hull_white_2f.rs:66-75 */
double eqr_hw2_var(double z, double vol, double tau) {
    double inv_z = pow(z, -1.0);
    double e1 = exp(z * (-tau));
    double e2 = exp(z * (-2.0 * tau));
    double bracket = (((inv_z * 2.0) * e1 - (inv_z * 0.5) * e2) - inv_z * 1.5) + tau;
    return ((vol * vol) / (z * z)) * bracket;
}

/* 2 rho sigma eta/(a b) * [tau + (e^{-a tau}-1)/a + (e^{-b tau}-1)/b - (e^{-(a+b)tau}-1)/(a+b)],
 * hull_white_2f.rs:79-91 */
double eqr_hw2_cross(double a, double b, double sigma, double eta, double rho, double tau) {
    double apb = a + b;
    double ta = (exp(a * (-tau)) + (-1.0)) / a;
    double tb = (exp(b * (-tau)) + (-1.0)) / b;
    double tab = (exp(apb * (-tau)) + (-1.0)) / apb;
    double bracket = ((ta + tau) + tb) - tab;
    double coef = ((sigma * eta) * (2.0 * rho)) / (a * b);
    return coef * bracket;
}

/* P(t,T | x, y) = exp(((r0 (-tau) + V/2) - B(a,tau) x) - B(b,tau) y), hull_white_2f.rs:96-106 */
double eqr_zcb2(double a, double b, double sigma, double eta, double rho, double r0, double x, double y,
                double t, double maturity) {
    double tau = maturity - t;
    double v = (eqr_hw2_var(a, sigma, tau) + eqr_hw2_var(b, eta, tau)) + eqr_hw2_cross(a, b, sigma, eta, rho, tau);
    double e = ((r0 * (-tau) + v * 0.5) - eqr_hw_b(a, tau) * x) - eqr_hw_b(b, tau) * y;
    return exp(e);
}

/* The swap valuations and the portfolio run below are written once over this descriptor:
 * nf = 1 is HW1F (b the long-run mean, state x = r), nf = 2 is G2++ (state (x, y)). */
typedef struct {
    int nf;
    double a, b, sigma, eta, rho, r0;
} eqr_model;

static double zcb_m(const eqr_model* M, double x, double y, double t, double maturity) {
    if (M->nf == 2) return eqr_zcb2(M->a, M->b, M->sigma, M->eta, M->rho, M->r0, x, y, t, maturity);
    return eqr_zcb(M->a, M->b, M->sigma, x, t, maturity);
}

/* ------------------------------------------------------------------------------------------ */
/* Philox4x32-10 + Box-Muller                                                                 */
/* (no reference function: the reference's exposure path draws from burn-ndarray's global RNG,  */
/*  exposure.rs:113-116; this counter-based stream replaces it in both the oracle and the GPU) */
/* ------------------------------------------------------------------------------------------ */

void eqr_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Counter (pair, path lo, path hi, stream): stream 0 feeds HW1F (two consecutive steps per pair),
 * stream 1 feeds G2++ (the x and y shocks of one step per pair). */
static void philox_uniforms_s(uint64_t seed, uint64_t path, uint32_t pair, uint32_t stream, double* u1,
                              double* u2) {
    uint32_t ctr[4] = {pair, (uint32_t)path, (uint32_t)(path >> 32), stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    eqr_philox4x32_10(ctr, key, out);
    uint64_t wa = ((uint64_t)out[1] << 32) | out[0];
    uint64_t wb = ((uint64_t)out[3] << 32) | out[2];
    const double two_m53 = 1.0 / 9007199254740992.0;
    *u1 = (double)((wa >> 11) + 1u) * two_m53; /* (0, 1] */
    *u2 = (double)(wb >> 11) * two_m53;        /* [0, 1) */
}

void eqr_philox_uniforms(uint64_t seed, uint64_t path, uint32_t pair, double* u1, double* u2) {
    philox_uniforms_s(seed, path, pair, 0u, u1, u2);
}

void eqr_philox_normals(uint64_t seed, uint64_t n_paths, uint32_t n_steps, double* z) {
    const double two_pi = 6.283185307179586476925286766559;
    for (uint64_t p = 0; p < n_paths; ++p) {
        for (uint32_t s = 0; s < n_steps; s += 2) {
            double u1, u2;
            eqr_philox_uniforms(seed, p, s / 2u, &u1, &u2);
            double rad = sqrt(-2.0 * log(u1));
            double th = two_pi * u2;
            z[(uint64_t)s * n_paths + p] = rad * cos(th);
            if (s + 1 < n_steps) z[(uint64_t)(s + 1) * n_paths + p] = rad * sin(th);
        }
    }
}

/* z[2][n_steps][n_paths]: every x shock (cos branch of step s's pair), then every independent
 * y shock (sin branch); the layout of the reference's two shock lists, exposure.rs:339-344. */
void eqr_philox_normals2(uint64_t seed, uint64_t n_paths, uint32_t n_steps, double* z) {
    const double two_pi = 6.283185307179586476925286766559;
    for (uint64_t p = 0; p < n_paths; ++p) {
        for (uint32_t s = 0; s < n_steps; ++s) {
            double u1, u2;
            philox_uniforms_s(seed, p, s, 1u, &u1, &u2);
            double rad = sqrt(-2.0 * log(u1));
            double th = two_pi * u2;
            z[(uint64_t)s * n_paths + p] = rad * cos(th);
            z[((uint64_t)n_steps + s) * n_paths + p] = rad * sin(th);
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Simulation                                                                                 */
/* ------------------------------------------------------------------------------------------ */

/* Euler: r' = (r + (a*(b - r))*dt) + sigma*(z*sqrt(dt)), kernel.rs:41-50, hull_white.rs:91-95.
 * Exact: r' = (r*e + b*(1-e)) + sd*z with e = exp(-a dt), sd = sigma*sqrt((1-exp(-2a dt))/(2a)). */
static void simulate_one(double a, double b, double sigma, double r0, double dt,
                         uint32_t n_steps, uint64_t n_paths, uint64_t p, int scheme,
                         const double* z, double* r_path /* [n_steps+1], stride 1 */) {
    double r = r0;
    r_path[0] = r;
    if (scheme == 0) {
        double sq = sqrt(dt);
        for (uint32_t s = 0; s < n_steps; ++s) {
            double drift = a * (b - r);
            double dw = z[(uint64_t)s * n_paths + p] * sq;
            r = (r + drift * dt) + sigma * dw;
            r_path[s + 1] = r;
        }
    } else {
        double e = exp(-(a * dt));
        double mean_c = b * (1.0 - e);
        double sd = sigma * sqrt((1.0 - exp(-(2.0 * a * dt))) / (2.0 * a));
        for (uint32_t s = 0; s < n_steps; ++s) {
            r = (r * e + mean_c) + sd * z[(uint64_t)s * n_paths + p];
            r_path[s + 1] = r;
        }
    }
}

void eqr_simulate(double a, double b, double sigma, double r0, double dt, uint32_t n_steps,
                  uint64_t n_paths, int scheme, const double* z, double* r_out) {
    double* tmp = (double*)malloc(sizeof(double) * (n_steps + 1));
    for (uint64_t p = 0; p < n_paths; ++p) {
        simulate_one(a, b, sigma, r0, dt, n_steps, n_paths, p, scheme, z, tmp);
        for (uint32_t s = 0; s <= n_steps; ++s) r_out[(uint64_t)s * n_paths + p] = tmp[s];
    }
    free(tmp);
}

/* ------------------------------------------------------------------------------------------ */
/* IRS                                                                                        */
/* ------------------------------------------------------------------------------------------ */

/* products/irs.rs:150-155 */
int eqr_is_reset_date(double start, const double* pay, int m, double t) {
    if (fabs(t - start) < EQR_TOL) return 1;
    for (int i = 0; i + 1 < m; ++i)
        if (fabs(pay[i] - t) < EQR_TOL) return 1;
    return 0;
}

/* products/irs.rs:126-145: (P(s,s) - P(s,T_n)) / sum_i P(s,T_i)*tau_i, reduce in order. */
static double par_rate_m(const eqr_model* M, double x, double y, double start, const double* pay,
                         const double* acc, int m) {
    double p_start = zcb_m(M, x, y, start, start);
    double p_end = zcb_m(M, x, y, start, pay[m - 1]);
    double numerator = p_start - p_end;
    double den = 0.0;
    for (int i = 0; i < m; ++i) {
        double leg = zcb_m(M, x, y, start, pay[i]) * acc[i];
        den = (i == 0) ? leg : den + leg;
    }
    return numerator / den;
}

double eqr_par_rate(double a, double b, double sigma, double r, double start,
                    const double* pay, const double* acc, int m) {
    eqr_model M = {1, a, b, sigma, 0.0, 0.0, 0.0};
    return par_rate_m(&M, r, 0.0, start, pay, acc, m);
}

/* irs.rs:126-145 at the initial G2++ state (0, 0) (api.rs build_irs_swap_2f). */
double eqr_par_rate2(double a, double b, double sigma, double eta, double rho, double r0, double start,
                     const double* pay, const double* acc, int m) {
    eqr_model M = {2, a, b, sigma, eta, rho, r0};
    return par_rate_m(&M, 0.0, 0.0, start, pay, acc, m);
}

/* G2++ paths of one path p, x_path/y_path [n_steps+1].
 * Euler (hull_white_2f.rs:116-140 with kernel.rs:41-50): zc = zy*sqrt(1-rho^2) + zx*rho,
 *   x' = (x + (-a x) dt) + sigma (zx sqrt(dt)), y' = (y + (-b y) dt) + eta (zc sqrt(dt)).
 * Exact (no reference function): the Gaussian transition of the OU pair,
 *   x' = x e^{-a dt} + sx zx, y' = y e^{-b dt} + sy (c zx + sqrt(1-c^2) zy),
 *   sx^2 = sigma^2 (1-e^{-2a dt})/(2a), sy^2 = eta^2 (1-e^{-2b dt})/(2b),
 *   c = rho sigma eta (1-e^{-(a+b)dt})/(a+b) / (sx sy). */
static void simulate2_one(const eqr_model* M, double dt, uint32_t n_steps, uint64_t n_paths, uint64_t p,
                          int scheme, const double* z, double* xp, double* yp) {
    double x = 0.0, y = 0.0;
    xp[0] = x;
    yp[0] = y;
    const double a = M->a, b = M->b, sigma = M->sigma, eta = M->eta, rho = M->rho;
    if (scheme == 0) {
        double sq = sqrt(dt), s1 = sqrt(1.0 - rho * rho);
        for (uint32_t s = 0; s < n_steps; ++s) {
            double zx = z[(uint64_t)s * n_paths + p];
            double zi = z[((uint64_t)n_steps + s) * n_paths + p];
            double zc = zi * s1 + zx * rho;
            double drift_x = (-a) * x;
            double drift_y = (-b) * y;
            x = (x + drift_x * dt) + sigma * (zx * sq);
            y = (y + drift_y * dt) + eta * (zc * sq);
            xp[s + 1] = x;
            yp[s + 1] = y;
        }
    } else {
        double ex = exp(-(a * dt)), ey = exp(-(b * dt));
        double sx = sigma * sqrt((1.0 - exp(-(2.0 * a * dt))) / (2.0 * a));
        double sy = eta * sqrt((1.0 - exp(-(2.0 * b * dt))) / (2.0 * b));
        double cov = rho * sigma * eta * (1.0 - exp(-((a + b) * dt))) / (a + b);
        double c = (sx > 0.0 && sy > 0.0) ? cov / (sx * sy) : 0.0;
        if (c > 1.0) c = 1.0;
        if (c < -1.0) c = -1.0;
        double s1 = sqrt(1.0 - c * c);
        for (uint32_t s = 0; s < n_steps; ++s) {
            double zx = z[(uint64_t)s * n_paths + p];
            double zi = z[((uint64_t)n_steps + s) * n_paths + p];
            double zc = c * zx + s1 * zi;
            x = x * ex + sx * zx;
            y = y * ey + sy * zc;
            xp[s + 1] = x;
            yp[s + 1] = y;
        }
    }
}

void eqr_simulate2(double a, double b, double sigma, double eta, double rho, double dt, uint32_t n_steps,
                   uint64_t n_paths, int scheme, const double* z, double* x_out, double* y_out) {
    eqr_model M = {2, a, b, sigma, eta, rho, 0.0};
    double* xp = (double*)malloc(sizeof(double) * (n_steps + 1));
    double* yp = (double*)malloc(sizeof(double) * (n_steps + 1));
    for (uint64_t p = 0; p < n_paths; ++p) {
        simulate2_one(&M, dt, n_steps, n_paths, p, scheme, z, xp, yp);
        for (uint32_t s = 0; s <= n_steps; ++s) {
            x_out[(uint64_t)s * n_paths + p] = xp[s];
            y_out[(uint64_t)s * n_paths + p] = yp[s];
        }
    }
    free(xp);
    free(yp);
}

/* remaining_from(t) (irs.rs:163-187) followed by npv (irs.rs:46-68):
 * N*(P(t,t) - P(t,T_n)) - sum_{period start >= t-1e-9} (N*K)*(P(t,T_i)*tau_i). */
static double npv_reset_m(const eqr_model* M, double x, double y, double t, double notional,
                          double fixed_rate, double start, const double* pay, const double* acc, int m) {
    /* remaining_from keeps the periods whose start >= t - 1e-9 and sets start := t */
    int first = -1;
    double period_start = start;
    for (int i = 0; i < m; ++i) {
        if (period_start >= t - EQR_TOL) {
            first = i;
            break;
        }
        period_start = pay[i];
    }
    if (first < 0) return 0.0; /* no period left: the reference would reduce an empty leg */
    double p_start = zcb_m(M, x, y, t, t);
    double p_end = zcb_m(M, x, y, t, pay[m - 1]);
    double floating_leg = notional * (p_start - p_end);
    double fixed_leg = 0.0;
    for (int i = first; i < m; ++i) {
        double p_i = zcb_m(M, x, y, t, pay[i]);
        double leg = (notional * fixed_rate) * (p_i * acc[i]);
        fixed_leg = (i == first) ? leg : fixed_leg + leg;
    }
    return floating_leg - fixed_leg;
}

double eqr_irs_npv_reset(double a, double b, double sigma, double r, double t, double notional,
                         double fixed_rate, double start, const double* pay, const double* acc,
                         int m) {
    eqr_model M = {1, a, b, sigma, 0.0, 0.0, 0.0};
    return npv_reset_m(&M, r, 0.0, t, notional, fixed_rate, start, pay, acc, m);
}

double eqr_irs_npv_reset2(double a, double b, double sigma, double eta, double rho, double r0, double x,
                          double y, double t, double notional, double fixed_rate, double start,
                          const double* pay, const double* acc, int m) {
    eqr_model M = {2, a, b, sigma, eta, rho, r0};
    return npv_reset_m(&M, x, y, t, notional, fixed_rate, start, pay, acc, m);
}

double eqr_fixing_date(double t, double start, const double* pay, int m) {
    if (t <= start + EQR_TOL) return -1.0;       /* not started: no fixing yet */
    if (t >= pay[m - 1] - EQR_TOL) return -1.0;  /* matured */
    double prev = start;
    for (int i = 0; i < m; ++i) {
        if (pay[i] > t + EQR_TOL) {
            if (fabs(t - prev) < EQR_TOL) return -1.0; /* reset date: fixing happens now */
            return prev;
        }
        prev = pay[i];
    }
    return -1.0;
}

/* Extended valuation (no reference function beyond reset dates; at reset dates it is exactly
 * eqr_irs_npv_reset). Cash flows alive at t are those with T_i > t + 1e-9.
 *   t < start            : float = P(t,start) - P(t,T_n)      (forward-starting swap)
 *   t at a reset date    : float = 1 - P(t,T_n)               (irs.rs:46-68 with start := t)
 *   T_{k-1} < t < T_k    : float = P(t,T_k)/P(T_{k-1},T_k | r_fix) - P(t,T_n)
 *   t >= T_n             : value 0
 *   value = N*float - sum_{T_i > t} (N*K)*(P(t,T_i)*tau_i). */
static double value_m(const eqr_model* M, double x, double y, double t, double xf, double yf,
                      double notional, double fixed_rate, double start, const double* pay, const double* acc,
                      int m) {
    if (t >= pay[m - 1] - EQR_TOL) return 0.0;
    int first = 0;
    while (first < m && !(pay[first] > t + EQR_TOL)) ++first;
    double p_end = zcb_m(M, x, y, t, pay[m - 1]);
    double fl;
    if (fabs(t - start) < EQR_TOL) {
        double p_start = zcb_m(M, x, y, t, t);
        fl = p_start - p_end;
    } else if (t < start) {
        double p_start = zcb_m(M, x, y, t, start);
        fl = p_start - p_end;
    } else {
        double prev = (first == 0) ? start : pay[first - 1];
        if (fabs(t - prev) < EQR_TOL) {
            double p_start = zcb_m(M, x, y, t, t);
            fl = p_start - p_end;
        } else {
            double p_next = zcb_m(M, x, y, t, pay[first]);
            double p_fix = zcb_m(M, xf, yf, prev, pay[first]);
            fl = p_next / p_fix - p_end;
        }
    }
    double floating_leg = notional * fl;
    double fixed_leg = 0.0;
    for (int i = first; i < m; ++i) {
        double p_i = zcb_m(M, x, y, t, pay[i]);
        double leg = (notional * fixed_rate) * (p_i * acc[i]);
        fixed_leg = (i == first) ? leg : fixed_leg + leg;
    }
    return floating_leg - fixed_leg;
}

double eqr_irs_value(double a, double b, double sigma, double r, double t, double r_fix,
                     double notional, double fixed_rate, double start, const double* pay,
                     const double* acc, int m) {
    eqr_model M = {1, a, b, sigma, 0.0, 0.0, 0.0};
    return value_m(&M, r, 0.0, t, r_fix, 0.0, notional, fixed_rate, start, pay, acc, m);
}

/* The same under G2++: state (x, y) at t, (xf, yf) at the fixing date of the period in progress. */
double eqr_irs_value2(double a, double b, double sigma, double eta, double rho, double r0, double x, double y,
                      double t, double xf, double yf, double notional, double fixed_rate, double start,
                      const double* pay, const double* acc, int m) {
    eqr_model M = {2, a, b, sigma, eta, rho, r0};
    return value_m(&M, x, y, t, xf, yf, notional, fixed_rate, start, pay, acc, m);
}

/* ------------------------------------------------------------------------------------------ */
/* EE / PFE95, exposure.rs:42-49                                                              */
/* ------------------------------------------------------------------------------------------ */

static int cmp_double(const void* x, const void* y) {
    double a = *(const double*)x, b = *(const double*)y;
    return (a > b) - (a < b);
}

/* mean = sum in input order / n; PFE = sorted[min(ceil(0.95 n) - 1, n - 1)] (saturating). */
void eqr_ee_pfe(double* exposures, uint64_t n, double* ee, double* pfe) {
    double sum = 0.0;
    for (uint64_t i = 0; i < n; ++i) sum += exposures[i];
    *ee = sum / (double)n;
    qsort(exposures, (size_t)n, sizeof(double), cmp_double);
    uint64_t q = (uint64_t)ceil(0.95 * (double)n);
    q = (q == 0) ? 0 : q - 1;
    if (q > n - 1) q = n - 1;
    *pfe = exposures[q];
}

/* ------------------------------------------------------------------------------------------ */
/* CVA, exposure.rs:242-275                                                                   */
/* ------------------------------------------------------------------------------------------ */

/* The discount is the model's bond from the initial state: r0 under HW1F, (0, 0) under G2++
 * (exposure.rs:452-466). */
static double cva_m(const eqr_model* M, const double* times, const double* ee, int n, double hazard_rate,
                    double recovery_rate) {
    double cva = 0.0;
    double prev_survival = 1.0;
    for (int i = 0; i < n; ++i) {
        double t = times[i];
        double survival = exp(-hazard_rate * t);
        double default_prob = prev_survival - survival;
        double disc = (M->nf == 2) ? zcb_m(M, 0.0, 0.0, 0.0, t) : zcb_m(M, M->r0, 0.0, 0.0, t);
        cva += (1.0 - recovery_rate) * ee[i] * default_prob * disc;
        prev_survival = survival;
    }
    return cva;
}

double eqr_unilateral_cva(double a, double b, double sigma, double r0, const double* times,
                          const double* ee, int n, double hazard_rate, double recovery_rate) {
    eqr_model M = {1, a, b, sigma, 0.0, 0.0, r0};
    return cva_m(&M, times, ee, n, hazard_rate, recovery_rate);
}

double eqr_unilateral_cva2(double a, double b, double sigma, double eta, double rho, double r0,
                           const double* times, const double* ee, int n, double hazard_rate,
                           double recovery_rate) {
    eqr_model M = {2, a, b, sigma, eta, rho, r0};
    return cva_m(&M, times, ee, n, hazard_rate, recovery_rate);
}

/* ------------------------------------------------------------------------------------------ */
/* Reference exposure profile of one swap, exposure.rs:73-138, api.rs:42-66,97-118            */
/* ------------------------------------------------------------------------------------------ */

int eqr_exposure_profile(double a, double b, double sigma, double r0, double notional,
                         double fixed_rate, int use_par_rate, double start, const double* pay,
                         const double* acc, int m, const double* times, int n_dates,
                         uint32_t n_steps, uint64_t n_paths, const double* z, double* ee,
                         double* pfe) {
    double k = use_par_rate ? eqr_par_rate(a, b, sigma, r0, start, pay, acc, m) : fixed_rate;
    double t_max = -INFINITY;
    for (int i = 0; i < n_dates; ++i) t_max = fmax(t_max, times[i]);
    if (t_max <= 0.0) {
        double v = eqr_irs_npv_reset(a, b, sigma, r0, times[0], notional, k, start, pay, acc, m);
        v = fmax(v, 0.0);
        ee[0] = v;
        pfe[0] = v;
        return 1;
    }
    double dt = t_max / (double)n_steps;
    double* rp = (double*)malloc(sizeof(double) * (n_steps + 1));
    double* r_all = (double*)malloc(sizeof(double) * (n_steps + 1) * n_paths);
    for (uint64_t p = 0; p < n_paths; ++p) {
        simulate_one(a, b, sigma, r0, dt, n_steps, n_paths, p, 0, z, rp);
        for (uint32_t s = 0; s <= n_steps; ++s) r_all[(uint64_t)s * n_paths + p] = rp[s];
    }
    double* x = (double*)malloc(sizeof(double) * n_paths);
    for (int d = 0; d < n_dates; ++d) {
        double t = times[d];
        uint64_t idx = (uint64_t)round(t / dt);
        uint64_t n = (idx == 0) ? 1 : n_paths; /* path[0] is r0 with shape [1] */
        for (uint64_t p = 0; p < n; ++p) {
            double v = eqr_irs_npv_reset(a, b, sigma, r_all[idx * n_paths + p], t, notional, k,
                                         start, pay, acc, m);
            x[p] = fmax(v, 0.0);
        }
        eqr_ee_pfe(x, n, &ee[d], &pfe[d]);
    }
    free(x);
    free(r_all);
    free(rp);
    return n_dates;
}

/* G2++ profile of one swap, exposure.rs:305-370 with api.rs:587-611 (build_irs_swap_2f): Euler
 * paths from (0, 0); z is [2][n_steps][n_paths]. path[0] holds n_paths zero states, so a date on
 * grid index 0 averages n_paths equal values (unlike HW1F, whose path[0] is a single r0). */
int eqr_exposure_profile2(double a, double b, double sigma, double eta, double rho, double r0,
                          double notional, double fixed_rate, int use_par_rate, double start,
                          const double* pay, const double* acc, int m, const double* times, int n_dates,
                          uint32_t n_steps, uint64_t n_paths, const double* z, double* ee, double* pfe) {
    eqr_model M = {2, a, b, sigma, eta, rho, r0};
    double k = use_par_rate ? par_rate_m(&M, 0.0, 0.0, start, pay, acc, m) : fixed_rate;
    double t_max = -INFINITY;
    for (int i = 0; i < n_dates; ++i) t_max = fmax(t_max, times[i]);
    if (t_max <= 0.0) {
        double v = npv_reset_m(&M, 0.0, 0.0, times[0], notional, k, start, pay, acc, m);
        v = fmax(v, 0.0);
        ee[0] = v;
        pfe[0] = v;
        return 1;
    }
    double dt = t_max / (double)n_steps;
    double* xp = (double*)malloc(sizeof(double) * (n_steps + 1));
    double* yp = (double*)malloc(sizeof(double) * (n_steps + 1));
    double* x_all = (double*)malloc(sizeof(double) * (n_steps + 1) * n_paths);
    double* y_all = (double*)malloc(sizeof(double) * (n_steps + 1) * n_paths);
    for (uint64_t p = 0; p < n_paths; ++p) {
        simulate2_one(&M, dt, n_steps, n_paths, p, 0, z, xp, yp);
        for (uint32_t s = 0; s <= n_steps; ++s) {
            x_all[(uint64_t)s * n_paths + p] = xp[s];
            y_all[(uint64_t)s * n_paths + p] = yp[s];
        }
    }
    double* x = (double*)malloc(sizeof(double) * n_paths);
    for (int d = 0; d < n_dates; ++d) {
        double t = times[d];
        uint64_t idx = (uint64_t)round(t / dt);
        for (uint64_t p = 0; p < n_paths; ++p) {
            double v = npv_reset_m(&M, x_all[idx * n_paths + p], y_all[idx * n_paths + p], t, notional, k, start,
                                   pay, acc, m);
            x[p] = fmax(v, 0.0);
        }
        eqr_ee_pfe(x, n_paths, &ee[d], &pfe[d]);
    }
    free(x);
    free(y_all);
    free(x_all);
    free(yp);
    free(xp);
    return n_dates;
}

/* ------------------------------------------------------------------------------------------ */
/* Market curve, market.cpp:26-41,95-111, measure.cpp:124-222                                 */
/* ------------------------------------------------------------------------------------------ */

double eqr_curve_zero_rate(const double* pillars, const double* zeros, int np, double t) {
    if (t <= pillars[0]) return zeros[0];
    if (t >= pillars[np - 1]) return zeros[np - 1];
    int i = 0; /* lower_bound: first pillar >= t */
    while (i < np && pillars[i] < t) ++i;
    if (i == 0) i = 1;
    double t0 = pillars[i - 1], t1 = pillars[i];
    double z0 = zeros[i - 1], z1 = zeros[i];
    double w = (t - t0) / (t1 - t0);
    return z0 + w * (z1 - z0);
}

double eqr_curve_df(const double* pillars, const double* zeros, int np, double t) {
    return exp(-eqr_curve_zero_rate(pillars, zeros, np, t) * t);
}

double eqr_par_rate_from_market(const double* pillars, const double* zeros, int np, double start,
                                const double* pay, const double* acc, int m) {
    double numerator = eqr_curve_df(pillars, zeros, np, start) -
                       eqr_curve_df(pillars, zeros, np, pay[m - 1]);
    double denominator = 0.0;
    for (int i = 0; i < m; ++i) denominator += acc[i] * eqr_curve_df(pillars, zeros, np, pay[i]);
    return numerator / denominator;
}

/* Value at valuation_time v of the swap's cash flows on the curve observed today: every discount
 * factor DF(t) is divided by DF(v) (measure.cpp:146-184); v = 0 divides by DF(0) = 1, the t = 0 form
 * bit for bit. A swap already started (v > start) has fixed its first floating coupon, which the
 * two-bond floating leg cannot represent: the reference refuses it (measure.cpp:166-174), the caller
 * checks it here (the function returns NaN). */
double eqr_npv_from_market_at(const double* pillars, const double* zeros, int np, double notional,
                              double fixed_rate, double start, const double* pay, const double* acc, int m,
                              double valuation_time) {
    if (valuation_time > start) return NAN;
    double dv = eqr_curve_df(pillars, zeros, np, valuation_time);
    double p_start = eqr_curve_df(pillars, zeros, np, start) / dv;
    double p_end = eqr_curve_df(pillars, zeros, np, pay[m - 1]) / dv;
    double floating_leg = notional * (p_start - p_end);
    double fixed_leg = 0.0;
    for (int i = 0; i < m; ++i)
        fixed_leg += notional * fixed_rate * acc[i] * (eqr_curve_df(pillars, zeros, np, pay[i]) / dv);
    return floating_leg - fixed_leg;
}

double eqr_npv_from_market(const double* pillars, const double* zeros, int np, double notional,
                           double fixed_rate, double start, const double* pay, const double* acc,
                           int m) {
    return eqr_npv_from_market_at(pillars, zeros, np, notional, fixed_rate, start, pay, acc, m, 0.0);
}

/* DV01 at valuation_time: NPV(z + bump) - NPV(z), the same fixed rate in both (measure.cpp:195-200). */
double eqr_dv01_at(const double* pillars, const double* zeros, int np, double notional, double fixed_rate,
                   double start, const double* pay, const double* acc, int m, double bump, double valuation_time) {
    double* bumped = (double*)malloc(sizeof(double) * np);
    for (int i = 0; i < np; ++i) bumped[i] = zeros[i] + bump;
    double base = eqr_npv_from_market_at(pillars, zeros, np, notional, fixed_rate, start, pay, acc, m, valuation_time);
    double up = eqr_npv_from_market_at(pillars, bumped, np, notional, fixed_rate, start, pay, acc, m, valuation_time);
    free(bumped);
    return up - base;
}

double eqr_dv01(const double* pillars, const double* zeros, int np, double notional,
                double fixed_rate, double start, const double* pay, const double* acc, int m,
                double bump) {
    return eqr_dv01_at(pillars, zeros, np, notional, fixed_rate, start, pay, acc, m, bump, 0.0);
}

void eqr_dv01_bucketed(const double* pillars, const double* zeros, int np, double notional,
                       double fixed_rate, double start, const double* pay, const double* acc,
                       int m, double bump, double* out) {
    double* bumped = (double*)malloc(sizeof(double) * np);
    double base = eqr_npv_from_market(pillars, zeros, np, notional, fixed_rate, start, pay, acc, m);
    for (int j = 0; j < np; ++j) {
        memcpy(bumped, zeros, sizeof(double) * np);
        bumped[j] += bump;
        out[j] = eqr_npv_from_market(pillars, bumped, np, notional, fixed_rate, start, pay, acc, m) -
                 base;
    }
    free(bumped);
}

/* ------------------------------------------------------------------------------------------ */
/* Portfolio / netting-set exposure                                                           */
/* ------------------------------------------------------------------------------------------ */

int eqr_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* z is [n_steps][n_paths] under HW1F and [2][n_steps][n_paths] under G2++. A date on grid index 0
 * has one row under HW1F (path[0] is r0 alone) and n_paths rows under G2++ (path[0] is n_paths
 * zero states), one row when every date is <= 0 (exposure.rs:98-108, 325-337). */
static int portfolio_m(const eqr_model* M, int n_trades, int n_sets, const int32_t* set_id,
                       const double* notional, const double* fixed_rate, const double* start,
                       const int32_t* flow_off, const double* pay, const double* acc, const double* dates,
                       int n_dates, uint32_t n_steps, uint64_t n_paths, int scheme, const double* z,
                       double hazard_rate, double recovery_rate, double* set_ee, double* set_pfe,
                       double* set_cva, double* trade_ee, double* trade_pfe, double* trade_cva,
                       double* path_mtm, double* set_mtm, int n_threads) {
    double t_max = -INFINITY;
    for (int k = 0; k < n_dates; ++k) t_max = fmax(t_max, dates[k]);
    int deterministic = !(t_max > 0.0);
    double dt = deterministic ? 1.0 : t_max / (double)n_steps;
    uint64_t* idx = (uint64_t*)malloc(sizeof(uint64_t) * (n_dates > 0 ? n_dates : 1));
    for (int k = 0; k < n_dates; ++k) idx[k] = deterministic ? 0 : (uint64_t)round(dates[k] / dt);

    /* exposures: set rows [n_sets][n_dates][n_paths], trade rows likewise when requested */
    int want_trade = (trade_ee != NULL) || (trade_pfe != NULL) || (trade_cva != NULL);
    double* xs = (double*)malloc(sizeof(double) * (size_t)n_sets * n_dates * n_paths);
    double* xt = want_trade ? (double*)malloc(sizeof(double) * (size_t)n_trades * n_dates * n_paths)
                            : NULL;
    uint32_t steps = deterministic ? 0 : n_steps;

#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
#endif
    {
        double* rp = (double*)malloc(sizeof(double) * (steps + 1));
        double* yp = (double*)malloc(sizeof(double) * (steps + 1));
        double* vset = (double*)malloc(sizeof(double) * (n_sets > 0 ? n_sets : 1));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t pp = 0; pp < (int64_t)n_paths; ++pp) {
            uint64_t p = (uint64_t)pp;
            if (M->nf == 2) {
                simulate2_one(M, dt, steps, n_paths, p, scheme, z, rp, yp);
            } else {
                simulate_one(M->a, M->b, M->sigma, M->r0, dt, steps, n_paths, p, scheme, z, rp);
                for (uint32_t s = 0; s <= steps; ++s) yp[s] = 0.0;
            }
            for (int k = 0; k < n_dates; ++k) {
                double t = dates[k];
                double r = rp[idx[k]], y = yp[idx[k]];
                for (int s = 0; s < n_sets; ++s) vset[s] = 0.0;
                for (int j = 0; j < n_trades; ++j) {
                    const double* pj = pay + flow_off[j];
                    const double* aj = acc + flow_off[j];
                    int m = flow_off[j + 1] - flow_off[j];
                    double t_fix = eqr_fixing_date(t, start[j], pj, m);
                    double r_fix = 0.0, y_fix = 0.0;
                    if (t_fix >= 0.0) {
                        uint64_t gf = deterministic ? 0 : (uint64_t)round(t_fix / dt);
                        r_fix = rp[gf];
                        y_fix = yp[gf];
                    }
                    double v = value_m(M, r, y, t, r_fix, y_fix, notional[j], fixed_rate[j], start[j], pj, aj, m);
                    vset[set_id[j]] += v;
                    if (xt) xt[((size_t)j * n_dates + k) * n_paths + p] = fmax(v, 0.0);
                    if (path_mtm) path_mtm[((size_t)k * n_trades + j) * n_paths + p] = v;
                }
                for (int s = 0; s < n_sets; ++s) {
                    xs[((size_t)s * n_dates + k) * n_paths + p] = fmax(vset[s], 0.0);
                    if (set_mtm) set_mtm[((size_t)k * n_sets + s) * n_paths + p] = vset[s];
                }
            }
        }
        free(vset);
        free(yp);
        free(rp);
    }

    double* scratch = (double*)malloc(sizeof(double) * n_paths);
    for (int s = 0; s < n_sets; ++s) {
        for (int k = 0; k < n_dates; ++k) {
            uint64_t n = (idx[k] == 0 && (M->nf == 1 || deterministic)) ? 1 : n_paths;
            memcpy(scratch, xs + ((size_t)s * n_dates + k) * n_paths, sizeof(double) * n);
            eqr_ee_pfe(scratch, n, &set_ee[(size_t)s * n_dates + k], &set_pfe[(size_t)s * n_dates + k]);
        }
        if (set_cva)
            set_cva[s] = cva_m(M, dates, set_ee + (size_t)s * n_dates, n_dates, hazard_rate, recovery_rate);
    }
    if (want_trade) {
        double* tee = (double*)malloc(sizeof(double) * (size_t)n_trades * n_dates);
        for (int j = 0; j < n_trades; ++j) {
            for (int k = 0; k < n_dates; ++k) {
                uint64_t n = (idx[k] == 0 && (M->nf == 1 || deterministic)) ? 1 : n_paths;
                double e, q;
                memcpy(scratch, xt + ((size_t)j * n_dates + k) * n_paths, sizeof(double) * n);
                eqr_ee_pfe(scratch, n, &e, &q);
                tee[(size_t)j * n_dates + k] = e;
                if (trade_ee) trade_ee[(size_t)j * n_dates + k] = e;
                if (trade_pfe) trade_pfe[(size_t)j * n_dates + k] = q;
            }
            if (trade_cva)
                trade_cva[j] = cva_m(M, dates, tee + (size_t)j * n_dates, n_dates, hazard_rate, recovery_rate);
        }
        free(tee);
    }
    free(scratch);
    free(xt);
    free(xs);
    free(idx);
    return 0;
}

int eqr_portfolio_exposure(double a, double b, double sigma, double r0, int n_trades, int n_sets,
                           const int32_t* set_id, const double* notional,
                           const double* fixed_rate, const double* start,
                           const int32_t* flow_off, const double* pay, const double* acc,
                           const double* dates, int n_dates, uint32_t n_steps, uint64_t n_paths,
                           int scheme, const double* z, double hazard_rate, double recovery_rate,
                           double* set_ee, double* set_pfe, double* set_cva, double* trade_ee,
                           double* trade_pfe, double* trade_cva, double* path_mtm,
                           double* set_mtm, int n_threads) {
    eqr_model M = {1, a, b, sigma, 0.0, 0.0, r0};
    return portfolio_m(&M, n_trades, n_sets, set_id, notional, fixed_rate, start, flow_off, pay, acc, dates,
                       n_dates, n_steps, n_paths, scheme, z, hazard_rate, recovery_rate, set_ee, set_pfe,
                       set_cva, trade_ee, trade_pfe, trade_cva, path_mtm, set_mtm, n_threads);
}

int eqr_portfolio_exposure2(double a, double b, double sigma, double eta, double rho, double r0, int n_trades,
                            int n_sets, const int32_t* set_id, const double* notional, const double* fixed_rate,
                            const double* start, const int32_t* flow_off, const double* pay, const double* acc,
                            const double* dates, int n_dates, uint32_t n_steps, uint64_t n_paths, int scheme,
                            const double* z, double hazard_rate, double recovery_rate, double* set_ee,
                            double* set_pfe, double* set_cva, double* trade_ee, double* trade_pfe,
                            double* trade_cva, double* path_mtm, double* set_mtm, int n_threads) {
    eqr_model M = {2, a, b, sigma, eta, rho, r0};
    return portfolio_m(&M, n_trades, n_sets, set_id, notional, fixed_rate, start, flow_off, pay, acc, dates,
                       n_dates, n_steps, n_paths, scheme, z, hazard_rate, recovery_rate, set_ee, set_pfe,
                       set_cva, trade_ee, trade_pfe, trade_cva, path_mtm, set_mtm, n_threads);
}

/* ------------------------------------------------------------------------------------------ */
/* XVA reducer over an exposure cube, quant-engine/src/xva.rs:14-246                           */
/* ------------------------------------------------------------------------------------------ */

/* xva.rs:48-64 */
double eqr_collateral_adjust(double value, int has_csa, int variation_margin, double threshold, double mta,
                             double ia) {
    if (!has_csa || !variation_margin) return value;
    double th = fmax(threshold, 0.0);
    double collateral = fmax(fabs(value) - th, 0.0);
    if (collateral < fmax(mta, 0.0)) return value;
    if (value >= 0.0) return value - collateral - fmax(ia, 0.0);
    return value + collateral + fmax(ia, 0.0);
}

/* xva.rs:31-46 */
static double xva_default_increment(double hazard, const double* sp, int n_sp, int i, const double* times) {
    if (i < n_sp) {
        double previous = (i == 0) ? 1.0 : sp[i - 1];
        return fmax(previous - sp[i], 0.0);
    }
    double h = fmax(hazard, 0.0);
    double start = (i == 0) ? 0.0 : times[i - 1];
    return exp(-(h * start)) - exp(-(h * fmax(times[i], start)));
}

static double clamp01(double r) { return fmin(fmax(r, 0.0), 1.0); }

/* xva.rs:97-183: per time row, the CSA-adjusted values' positive and negative parts, their means
 * (sums in path order), the quantile sorted[round((n-1) p)] of the positive parts (xva.rs:66-75),
 * and the running CVA / DVA / FVA / MVA / KVA sums in row order. q, p: [n_times][n_paths], p nullable;
 * df nullable (then exp(-0 t) = 1); out_measures[5] = cva, dva, fva, mva, kva. */
void eqr_xva_calculate(const double* times, int n_times, uint64_t n_paths, const double* q, const double* df,
                       const double* p, double cp_hazard, double cp_recovery, const double* cp_sp, int cp_nsp,
                       double own_hazard, double own_recovery, const double* own_sp, int own_nsp,
                       double funding_spread, double borrowing_spread, double capital_factor,
                       double initial_margin_total, int has_csa, int variation_margin, double threshold,
                       double mta, double ia, double confidence, double* out_ee, double* out_pfe,
                       double* out_measures) {
    double* pos = (double*)malloc(sizeof(double) * n_paths);
    double cva = 0.0, dva = 0.0, fva = 0.0, mva = 0.0, kva = 0.0;
    double pc = fmin(fmax(confidence, 0.0), 1.0);
    uint64_t index = (uint64_t)round((double)(n_paths - 1) * pc);
    for (int i = 0; i < n_times; ++i) {
        const double* row = q + (size_t)i * n_paths;
        double sp = 0.0, sn = 0.0;
        for (uint64_t j = 0; j < n_paths; ++j) {
            double a = eqr_collateral_adjust(row[j], has_csa, variation_margin, threshold, mta, ia);
            pos[j] = fmax(a, 0.0);
            sp += pos[j];
            sn += fmax(-a, 0.0);
        }
        double expected_positive = sp / (double)n_paths;
        double expected_negative = sn / (double)n_paths;
        qsort(pos, (size_t)n_paths, sizeof(double), cmp_double);
        double discount = df ? df[i] : exp(-0.0 * times[i]);
        out_ee[i] = expected_positive;
        out_pfe[i] = pos[index];
        double yf = (i == 0) ? fmax(times[i], 0.0) : fmax(times[i] - times[i - 1], 0.0);
        cva += discount * expected_positive * xva_default_increment(cp_hazard, cp_sp, cp_nsp, i, times) *
               (1.0 - clamp01(cp_recovery));
        fva += discount * expected_positive * funding_spread * yf;
        mva += discount * initial_margin_total * borrowing_spread * yf;
        kva += discount * expected_positive * fmax(capital_factor, 0.0) * yf;
        if (p) {
            const double* prow = p + (size_t)i * n_paths;
            double snp = 0.0;
            for (uint64_t j = 0; j < n_paths; ++j)
                snp += fmax(-eqr_collateral_adjust(prow[j], has_csa, variation_margin, threshold, mta, ia), 0.0);
            double expected_negative_p = snp / (double)n_paths;
            dva += discount * expected_negative_p * xva_default_increment(own_hazard, own_sp, own_nsp, i, times) *
                   (1.0 - clamp01(own_recovery));
        } else {
            dva += discount * expected_negative * xva_default_increment(own_hazard, own_sp, own_nsp, i, times) *
                   (1.0 - clamp01(own_recovery));
        }
    }
    out_measures[0] = cva;
    out_measures[1] = dva;
    out_measures[2] = fva;
    out_measures[3] = mva;
    out_measures[4] = kva;
    free(pos);
}

/* ------------------------------------------------------------------------------------------ */
/* Bermudan payer swaptions, Longstaff-Schwartz (no reference function: the reference's LSM is  */
/* the quadratic GBM one of payoff/lsm.rs:73-107; this restates the same regression method for   */
/* HW1F with a degree-3 basis in the standardized short rate)                                   */
/* ------------------------------------------------------------------------------------------ */

/* r and the trapezoid log-numeraire log B(t) = int_0^t r ds (hull_white.rs:114-120) on the exact
 * grid: out arrays [n_steps+1][n_paths]. */
void eqr_berm_simulate(double a, double b, double sigma, double r0, double dt, uint32_t n_steps,
                       uint64_t n_paths, const double* z, double* r_out, double* lb_out) {
    eqr_simulate(a, b, sigma, r0, dt, n_steps, n_paths, 1, z, r_out);
    for (uint64_t p = 0; p < n_paths; ++p) {
        double acc = 0.0;
        lb_out[p] = 0.0;
        for (uint32_t s = 0; s < n_steps; ++s) {
            acc += (r_out[(uint64_t)s * n_paths + p] + r_out[(uint64_t)(s + 1) * n_paths + p]) * (0.5 * dt);
            lb_out[(uint64_t)(s + 1) * n_paths + p] = acc;
        }
    }
}

/* standardized short rate x = (r - E[r_t]) / sd(r_t) */
double eqr_berm_x(double a, double b, double sigma, double r0, double t, double r) {
    double m = b + (r0 - b) * exp(-a * t);
    double s = sigma * sqrt((1.0 - exp(-2.0 * a * t)) / (2.0 * a));
    return (r - m) / s;
}

/* payer swap entered at an exercise date: N[(1 - P(T,T+L)) - K sum_{m=1..L} P(T,T+m)] */
static double berm_exercise_value(double a, double b, double sigma, double r, int L, double n, double k) {
    double ann = 0.0, pl = 0.0;
    for (int m = 1; m <= L; ++m) {
        pl = eqr_zcb(a, b, sigma, r, 0.0, (double)m);
        ann += pl;
    }
    return n * ((1.0 - pl) - k * ann);
}

/* solve the 4x4 normal equations (Gaussian elimination with partial pivoting); returns 0 if singular */
static int solve4(double M[4][4], double v[4], double c[4]) {
    double A[4][5];
    for (int i = 0; i < 4; ++i) {
        for (int j = 0; j < 4; ++j) A[i][j] = M[i][j];
        A[i][4] = v[i];
    }
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int i = col + 1; i < 4; ++i)
            if (fabs(A[i][col]) > fabs(A[piv][col])) piv = i;
        if (fabs(A[piv][col]) < 1e-300) return 0;
        for (int j = 0; j < 5; ++j) {
            double t = A[col][j];
            A[col][j] = A[piv][j];
            A[piv][j] = t;
        }
        for (int i = col + 1; i < 4; ++i) {
            double f = A[i][col] / A[col][col];
            for (int j = col; j < 5; ++j) A[i][j] -= f * A[col][j];
        }
    }
    for (int i = 3; i >= 0; --i) {
        double s = A[i][4];
        for (int j = i + 1; j < 4; ++j) s -= A[i][j] * c[j];
        c[i] = s / A[i][i];
    }
    return 1;
}

/* Exercise policy on training paths: backward over the trade's exercise years E..E+L-1 (annual);
 * at each, regress the discounted continuation cashflow on [1,x,x^2,x^3] over in-the-money paths,
 * exercise where U > 0 and U > C. r, lb: [n_steps+1][n_paths] with steps_per_year steps per year.
 * ex_coef: [n_trades][max_ex][4]; tau_year: [n_trades][n_paths] (-1 = never); y0: discounted to 0. */
int eqr_berm_policy(double a, double b, double sigma, double r0, int n_trades, const double* notional,
                    const double* strike, const int32_t* first_ex, const int32_t* n_ex, int max_ex,
                    uint32_t steps_per_year, uint64_t n_paths, const double* r, const double* lb,
                    double* ex_coef, int32_t* tau_year, double* y0, double* price) {
    for (int j = 0; j < n_trades; ++j) {
        int32_t* tau = tau_year + (size_t)j * n_paths;
        double* y = y0 + (size_t)j * n_paths;
        for (uint64_t p = 0; p < n_paths; ++p) { tau[p] = -1; y[p] = 0.0; }
        int E = first_ex[j], L = n_ex[j];
        for (int e = L - 1; e >= 0; --e) {
            int T = E + e;                    /* exercise year */
            int rem = L - e;                  /* years of the underlying swap left */
            uint64_t step = (uint64_t)T * steps_per_year;
            double M[4][4] = {{0}}, v[4] = {0}, c[4] = {0, 0, 0, 0};
            uint64_t n_itm = 0;
            for (uint64_t p = 0; p < n_paths; ++p) {
                double rr = r[step * n_paths + p];
                double u = berm_exercise_value(a, b, sigma, rr, rem, notional[j], strike[j]);
                if (!(u > 0.0)) continue;
                double x = eqr_berm_x(a, b, sigma, r0, (double)T, rr);
                double phi[4] = {1.0, x, x * x, x * x * x};
                double yy = y[p] * exp(lb[step * n_paths + p]);
                for (int i = 0; i < 4; ++i) {
                    for (int k = 0; k < 4; ++k) M[i][k] += phi[i] * phi[k];
                    v[i] += phi[i] * yy;
                }
                ++n_itm;
            }
            int fit = (e < L - 1) && n_itm >= 8 && solve4(M, v, c);
            if (!fit) c[0] = c[1] = c[2] = c[3] = 0.0;
            for (int i = 0; i < 4; ++i) ex_coef[((size_t)j * max_ex + e) * 4 + i] = c[i];
            for (uint64_t p = 0; p < n_paths; ++p) {
                double rr = r[step * n_paths + p];
                double u = berm_exercise_value(a, b, sigma, rr, rem, notional[j], strike[j]);
                if (!(u > 0.0)) continue;
                double x = eqr_berm_x(a, b, sigma, r0, (double)T, rr);
                double cont = c[0] + x * (c[1] + x * (c[2] + x * c[3]));
                if (u > cont) {
                    tau[p] = T;
                    y[p] = u * exp(-lb[step * n_paths + p]);
                }
            }
        }
        double s = 0.0;
        for (uint64_t p = 0; p < n_paths; ++p) s += y[p];
        price[j] = s / (double)n_paths;
    }
    return 0;
}

/* Forward application of a fitted policy to fresh paths (same outputs as eqr_berm_policy). */
int eqr_berm_apply(double a, double b, double sigma, double r0, int n_trades, const double* notional,
                   const double* strike, const int32_t* first_ex, const int32_t* n_ex, int max_ex,
                   uint32_t steps_per_year, uint64_t n_paths, const double* r, const double* lb,
                   const double* ex_coef, int32_t* tau_year, double* y0) {
    for (int j = 0; j < n_trades; ++j) {
        for (uint64_t p = 0; p < n_paths; ++p) {
            int32_t t_ex = -1;
            double yy = 0.0;
            for (int e = 0; e < n_ex[j]; ++e) {
                int T = first_ex[j] + e;
                uint64_t step = (uint64_t)T * steps_per_year;
                double rr = r[step * n_paths + p];
                double u = berm_exercise_value(a, b, sigma, rr, n_ex[j] - e, notional[j], strike[j]);
                if (!(u > 0.0)) continue;
                double x = eqr_berm_x(a, b, sigma, r0, (double)T, rr);
                const double* c = ex_coef + ((size_t)j * max_ex + e) * 4;
                double cont = c[0] + x * (c[1] + x * (c[2] + x * c[3]));
                if (u > cont) {
                    t_ex = T;
                    yy = u * exp(-lb[step * n_paths + p]);
                    break;
                }
            }
            tau_year[(size_t)j * n_paths + p] = t_ex;
            y0[(size_t)j * n_paths + p] = yy;
        }
    }
    return 0;
}

/* Continuation-value regression at monitoring date t (grid step `step`) over the paths alive at t
 * (not exercised before t): target y0 * B(t) on [1,x,x^2,x^3]. coef[4]; returns alive count. */
uint64_t eqr_berm_judge_fit(double a, double b, double sigma, double r0, uint64_t n_paths,
                            uint64_t step, double t, const double* r, const double* lb,
                            const int32_t* tau_year, const double* y0, double* coef) {
    double M[4][4] = {{0}}, v[4] = {0}, c[4] = {0, 0, 0, 0};
    uint64_t n = 0;
    for (uint64_t p = 0; p < n_paths; ++p) {
        if (tau_year[p] >= 0 && (double)tau_year[p] < t - 1e-9) continue;  /* exercised before t */
        double rr = r[step * n_paths + p];
        double x = eqr_berm_x(a, b, sigma, r0, t, rr);
        double phi[4] = {1.0, x, x * x, x * x * x};
        double yy = y0[p] * exp(lb[step * n_paths + p]);
        for (int i = 0; i < 4; ++i) {
            for (int k = 0; k < 4; ++k) M[i][k] += phi[i] * phi[k];
            v[i] += phi[i] * yy;
        }
        ++n;
    }
    if (!(n >= 8 && solve4(M, v, c))) c[0] = c[1] = c[2] = c[3] = 0.0;
    for (int i = 0; i < 4; ++i) coef[i] = c[i];
    return n;
}

/* EE at t from a continuation polynomial: (1/n) sum_{alive} max(C(x), 0). */
double eqr_berm_ee(double a, double b, double sigma, double r0, uint64_t n_paths, uint64_t step,
                   double t, const double* r, const int32_t* tau_year, const double* coef) {
    double s = 0.0;
    for (uint64_t p = 0; p < n_paths; ++p) {
        if (tau_year[p] >= 0 && (double)tau_year[p] < t - 1e-9) continue;
        double x = eqr_berm_x(a, b, sigma, r0, t, r[step * n_paths + p]);
        double c = coef[0] + x * (coef[1] + x * (coef[2] + x * coef[3]));
        s += (c > 0.0) ? c : 0.0;
    }
    return s / (double)n_paths;
}
