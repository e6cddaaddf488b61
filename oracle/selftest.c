/* Sanitizer run of the oracle (test infrastructure): every entry point of eq_ref.c / eq_ref_fused.c on
 * small mixed inputs (forward starts, dates between resets, grid index 0, several netting sets, both
 * schemes, G2++, the XVA reducer, the Bermudan LSM), built with AddressSanitizer and UndefinedBehavior-
 * Sanitizer by `make -C oracle sanitize`. A few cross-checks keep the calls honest: the fused run equals
 * the per-trade run, PFE95 is a value of its row. Exit status 0 when every check holds; the sanitizers
 * abort on the first finding. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "eq_ref.h"

static int fails = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            ++fails;                                                      \
        }                                                                 \
    } while (0)

static double* dalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

int main(void) {
    const double a = 0.1, b = 0.03, sigma = 0.01, r0 = 0.02;
    /* a book of 9 trades in 3 sets: spot, forward-starting and running (stub) schedules */
    enum { NT = 9, NS = 3 };
    int32_t set_id[NT], flow_off[NT + 1];
    double notional[NT], fixed[NT], start[NT], pay[200], acc[200];
    int nf = 0;
    flow_off[0] = 0;
    for (int j = 0; j < NT; ++j) {
        set_id[j] = j % NS;
        notional[j] = (j % 2 ? -1.0 : 1.0) * 1e6 * (1.0 + 0.1 * j);
        fixed[j] = 0.02 + 0.001 * (j % 4);
        const double f = (j % 3 == 0) ? 1.0 : (j % 3 == 1 ? 2.0 : 4.0);
        const int m = 2 + j % 5;
        start[j] = (j % 4 == 1) ? 0.75 : 0.0;
        const double first = (j % 4 == 2) ? 1.0 / 12.0 : 1.0 / f;
        for (int i = 0; i < m; ++i) {
            pay[nf] = start[j] + first + i / f;
            acc[nf] = (i == 0) ? first : 1.0 / f;
            ++nf;
        }
        flow_off[j + 1] = nf;
    }
    const int ND = 13;
    double dates[13];
    dates[0] = 0.0;
    for (int k = 1; k < ND; ++k) dates[k] = k / 6.0 + 1.0 / 24.0;
    const uint32_t n_steps = 48;
    const uint64_t n_paths = 257;
    double* z = dalloc((size_t)n_steps * n_paths);
    eqr_philox_normals(3, n_paths, n_steps, z);
    for (int scheme = 0; scheme < 2; ++scheme) {
        double *see = dalloc(NS * ND), *spfe = dalloc(NS * ND), *scva = dalloc(NS), *tee = dalloc(NT * ND),
               *tpfe = dalloc(NT * ND), *tcva = dalloc(NT), *pm = dalloc((size_t)ND * NT * n_paths),
               *sm = dalloc((size_t)ND * NS * n_paths);
        CHECK(eqr_portfolio_exposure(a, b, sigma, r0, NT, NS, set_id, notional, fixed, start, flow_off, pay, acc,
                                     dates, ND, n_steps, n_paths, scheme, z, 0.02, 0.4, see, spfe, scva, tee, tpfe,
                                     tcva, pm, sm, 2) == 0);
        for (int share = 0; share < 2; ++share) {
            double *fee = dalloc(NS * ND), *fpfe = dalloc(NS * ND), *fcva = dalloc(NS), *ftee = dalloc(NT * ND),
                   *ftcva = dalloc(NT), *fsm = dalloc((size_t)ND * NS * n_paths);
            CHECK(eqr_portfolio_exposure_fused(a, b, sigma, r0, NT, NS, set_id, notional, fixed, start, flow_off,
                                               pay, acc, dates, ND, n_steps, n_paths, scheme, share ? NULL : z, 3,
                                               0.02, 0.4, share, fee, fpfe, fcva, ftee, ftcva, fsm, 2) == 0);
            for (int i = 0; i < NS * ND; ++i)
                CHECK(fabs(fee[i] - see[i]) <= 1e-9 * (1.0 + fabs(see[i])));
            free(fee), free(fpfe), free(fcva), free(ftee), free(ftcva), free(fsm);
        }
        for (int k = 1; k < ND; ++k) { /* PFE95 is a value of its row */
            int found = 0;
            for (uint64_t p = 0; p < n_paths && !found; ++p) {
                const double v = sm[((size_t)k * NS) * n_paths + p];
                found = (v > 0.0 ? v : 0.0) == spfe[k];
            }
            CHECK(found);
        }
        free(see), free(spfe), free(scva), free(tee), free(tpfe), free(tcva), free(pm), free(sm);
    }
    /* one swap through the reference-shaped profile, and the t_max <= 0 branch */
    {
        const double p1[5] = {1, 2, 3, 4, 5}, a1[5] = {1, 1, 1, 1, 1}, t5[5] = {0, 1, 2, 3, 4}, t0[1] = {0.0};
        double ee[5], pfe[5];
        double* z2 = dalloc(52 * 300);
        eqr_philox_normals(7, 300, 52, z2);
        CHECK(eqr_exposure_profile(a, b, sigma, r0, 1e6, 0.0, 1, 0.0, p1, a1, 5, t5, 5, 52, 300, z2, ee, pfe) == 5);
        CHECK(eqr_exposure_profile(a, b, sigma, r0, 1e6, 0.02, 0, 0.0, p1, a1, 5, t0, 1, 52, 300, z2, ee, pfe) == 1);
        const double cva = eqr_unilateral_cva(a, b, sigma, r0, t5, ee, 5, 0.02, 0.4);
        CHECK(cva >= 0.0);
        const double pill[3] = {0.5, 2.0, 10.0}, zr[3] = {0.015, 0.02, 0.025};
        double bk[3];
        eqr_dv01_bucketed(pill, zr, 3, 1e6, 0.02, 0.0, p1, a1, 5, 1e-4, bk);
        CHECK(fabs(eqr_npv_from_market_at(pill, zr, 3, 1e6, 0.02, 0.0, p1, a1, 5, 0.0) -
                   eqr_npv_from_market(pill, zr, 3, 1e6, 0.02, 0.0, p1, a1, 5)) < 1e-9);
        CHECK(isnan(eqr_npv_from_market_at(pill, zr, 3, 1e6, 0.02, 0.5, p1, a1, 5, 1.0)));
        double e1[7] = {0, 0, 3, 1, 2, 5, 4}, ee1, pfe1;
        eqr_ee_pfe(e1, 7, &ee1, &pfe1);
        CHECK(pfe1 == 5.0);
        free(z2);
    }
    /* G2++ */
    {
        const double eta = 0.012, rho = -0.7;
        double* z2 = dalloc(2 * (size_t)n_steps * n_paths);
        eqr_philox_normals2(5, n_paths, n_steps, z2);
        double *see = dalloc(NS * ND), *spfe = dalloc(NS * ND), *scva = dalloc(NS), *tee = dalloc(NT * ND),
               *tpfe = dalloc(NT * ND), *tcva = dalloc(NT);
        for (int scheme = 0; scheme < 2; ++scheme)
            CHECK(eqr_portfolio_exposure2(a, 0.2, sigma, eta, rho, r0, NT, NS, set_id, notional, fixed, start,
                                          flow_off, pay, acc, dates, ND, n_steps, n_paths, scheme, z2, 0.02, 0.4, see,
                                          spfe, scva, tee, tpfe, tcva, NULL, NULL, 2) == 0);
        free(z2), free(see), free(spfe), free(scva), free(tee), free(tpfe), free(tcva);
    }
    /* XVA reducer */
    {
        enum { NTM = 4, NP = 33 };
        double times[NTM] = {0.25, 0.5, 1.0, 2.0}, q[NTM * NP], pq[NTM * NP], df[NTM] = {0.99, 0.98, 0.96, 0.93};
        for (int i = 0; i < NTM * NP; ++i) {
            q[i] = sin(0.37 * i) * 100.0;
            pq[i] = cos(0.11 * i) * 80.0;
        }
        const double sp[2] = {0.99, 0.97};
        double ee[NTM], pfe[NTM], m[5];
        eqr_xva_calculate(times, NTM, NP, q, df, pq, 0.02, 0.4, sp, 2, 0.01, 0.4, NULL, 0, 0.01, 0.02, 0.08, 1e4, 1,
                          1, 10.0, 1.0, 5.0, 0.95, ee, pfe, m);
        eqr_xva_calculate(times, NTM, NP, q, NULL, NULL, 0.02, 0.4, NULL, 0, 0.0, 0.4, NULL, 0, 0.0, 0.0, 0.0, 0.0, 0,
                          0, 0.0, 0.0, 0.0, 1.0, ee, pfe, m);
        CHECK(eqr_collateral_adjust(-5.0, 1, 1, 10.0, 1.0, 5.0) <= 0.0);
    }
    /* Bermudan LSM */
    {
        enum { NB = 3 };
        const uint32_t spy = 4, steps = 4 * 8;
        const uint64_t np = 513;
        double* zb = dalloc((size_t)steps * np);
        eqr_philox_normals(9, np, steps, zb);
        double *r = dalloc((size_t)(steps + 1) * np), *lb = dalloc((size_t)(steps + 1) * np);
        eqr_berm_simulate(a, b, sigma, r0, 1.0 / spy, steps, np, zb, r, lb);
        const double nb[NB] = {1e6, 2e6, 5e5}, kb[NB] = {0.028, 0.03, 0.025};
        const int32_t fe[NB] = {1, 2, 3}, ne[NB] = {4, 3, 5};
        const int max_ex = 5;
        double* coef = dalloc(NB * max_ex * 4);
        int32_t* tau = (int32_t*)calloc(NB * np, sizeof(int32_t));
        double *y0 = dalloc(NB * np), price[NB];
        CHECK(eqr_berm_policy(a, b, sigma, r0, NB, nb, kb, fe, ne, max_ex, spy, np, r, lb, coef, tau, y0, price) == 0);
        CHECK(eqr_berm_apply(a, b, sigma, r0, NB, nb, kb, fe, ne, max_ex, spy, np, r, lb, coef, tau, y0) == 0);
        for (int j = 0; j < NB; ++j)
            for (uint64_t k = 1; k <= (uint64_t)(fe[j] + ne[j] - 1) * spy; ++k) {
                double c4[4];
                eqr_berm_judge_fit(a, b, sigma, r0, np, k, (double)k / spy, r, lb, tau + j * np, y0 + j * np, c4);
                CHECK(eqr_berm_ee(a, b, sigma, r0, np, k, (double)k / spy, r, tau + j * np, c4) >= 0.0);
            }
        free(zb), free(r), free(lb), free(coef), free(tau), free(y0);
    }
    free(z);
    if (fails) fprintf(stderr, "%d checks failed\n", fails);
    else printf("oracle self-test passed\n");
    return fails ? 1 : 0;
}
