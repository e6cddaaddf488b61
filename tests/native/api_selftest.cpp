// Sanitizer run of the C ABI's host side (test infrastructure): the plan builder (schedule classes,
// families, flow lists, wave segments, set CSR, work items), plans in parts, the per-trade PFE rows,
// batches, the adjoint's linear form and the XVA reducer, with the host code of libeqhip compiled under
// AddressSanitizer and UndefinedBehaviorSanitizer (hipcc -Xarch_host -fsanitize=...; the device code is
// not instrumented). Run on a GPU: every call must return 0 and the sanitizers abort on their first
// finding. Results are checked elsewhere (tests/test_gpu_*.py); here only their sanity.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "eqhip.h"

static int fails = 0;
#define OK(call)                                                                          \
    do {                                                                                  \
        const int rc_ = (call);                                                           \
        if (rc_ != 0) {                                                                   \
            char msg_[512];                                                               \
            eqhip_last_error(msg_, sizeof msg_);                                          \
            std::fprintf(stderr, "line %d: %s -> %d (%s)\n", __LINE__, #call, rc_, msg_); \
            ++fails;                                                                      \
        }                                                                                 \
    } while (0)

struct Book {
    std::vector<int32_t> set_id, off{0};
    std::vector<double> notional, fixed, start, pay, acc;
    eqhip_portfolio view(int n_sets) const {
        return eqhip_portfolio{(int32_t)notional.size(), n_sets, set_id.data(), notional.data(), fixed.data(),
                               start.data(), off.data(), pay.data(), acc.data()};
    }
};

// forward starts, running trades with a short first stub, spot starts, on the monthly grid
static Book diverse(int n, int n_sets, unsigned seed, int max_years) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    Book b;
    for (int j = 0; j < n; ++j) {
        static const int freqs[3] = {1, 2, 4};
        const int f = freqs[g() % 3];
        const int m = 1 + (int)(g() % (unsigned)(max_years * f));
        const double kind = u(g);
        double s0 = 0.0, first = 1.0 / f;
        if (kind < 0.4) s0 = (1 + (int)(g() % 36)) / 12.0;
        else if (kind < 0.8) first = (1 + (int)(g() % (unsigned)std::max(1, 12 / f - 1))) / 12.0;
        double prev = s0;
        for (int i = 0; i < m; ++i) {
            const double t = std::round((s0 + first + (double)i / f) * 12.0) / 12.0;
            b.pay.push_back(t);
            b.acc.push_back(t - prev);
            prev = t;
        }
        b.off.push_back((int32_t)b.pay.size());
        b.start.push_back(s0);
        b.set_id.push_back((int32_t)(g() % (unsigned)n_sets));
        b.notional.push_back((u(g) < 0.5 ? -1.0 : 1.0) * 1e6 * (0.5 + u(g)));
        b.fixed.push_back(0.015 + 0.01 * u(g));
    }
    return b;
}

int main() {
    eqhip_context* ctx = nullptr;
    OK(eqhip_context_create(0, &ctx));
    if (!ctx) return 1;
    const eqhip_model_hw1f m1{0.1, 0.03, 0.01, 0.02};
    const eqhip_credit cr{0.02, 0.4};
    std::vector<double> monthly;
    for (int k = 1; k <= 120; ++k) monthly.push_back(k / 12.0);

    // a book too diverse for one block: plans in parts, per-trade EE and PFE, the cube, XVA, the adjoint
    {
        const Book b = diverse(720, 2, 11, 10);
        const eqhip_portfolio pf = b.view(2);
        const eqhip_mc mc{600, 120, 5, EQHIP_RNG_PHILOX, EQHIP_DISC_EXACT};
        eqhip_plan* plan = nullptr;
        OK(eqhip_plan_create(ctx, &m1, &pf, monthly.data(), 120, &mc,
                             EQHIP_PLAN_TRADE_EE | EQHIP_PLAN_TRADE_PFE | EQHIP_PLAN_KEEP_CUBE, &plan));
        if (plan) {
            std::vector<double> see(2 * 120), spfe(2 * 120), scva(2), tee(720 * 120), tcva(720), tpfe(720 * 120);
            eqhip_outputs o{see.data(), spfe.data(), scva.data(), tee.data(), tcva.data(), nullptr, nullptr, tpfe.data()};
            OK(eqhip_plan_execute(plan, nullptr, &cr, &o));
            eqhip_timing t{};
            OK(eqhip_plan_timing(plan, &t));
            if (t.n_parts < 2) ++fails;
            for (double v : tpfe) if (!(v >= 0.0)) ++fails;
            std::vector<double> cube(120 * 600), df(120);
            OK(eqhip_plan_exposure_cube(plan, 1, cube.data(), df.data()));
            const eqhip_csa csa{1e4, 1e3, 0.0, 1};
            const eqhip_xva_params xp{{0.02, 0.4, nullptr, 0}, {0.01, 0.4, nullptr, 0}, 0.01, 0.02, 0.08, 0.0, &csa, 0.95};
            std::vector<double> ee(120), pf95(120), meas(5);
            OK(eqhip_plan_xva(plan, 0, &xp, ee.data(), pf95.data(), meas.data()));
            OK(eqhip_plan_destroy(plan));
        }
        // the same book with the adjoint (one grid step per date)
        eqhip_plan* p2 = nullptr;
        OK(eqhip_plan_create(ctx, &m1, &pf, monthly.data(), 120, &mc, 0, &p2));
        if (p2) {
            double cva = 0.0, grad[6], check = 0.0;
            OK(eqhip_cva_adjoint(p2, nullptr, 0, &cr, &cva, grad, &check));
            if (!(std::fabs(cva - check) <= 1e-9 * std::fabs(cva))) ++fails;
            OK(eqhip_plan_set_variant(p2, 1, 2) == EQHIP_E_INVALID ? 0 : 1);  // a plan in parts keeps its variant
            OK(eqhip_plan_destroy(p2));
        }
    }
    // one class of 8193 trades (its work items in chunks), host draws, Euler, a date on grid index 0
    {
        Book b;
        for (int j = 0; j < 8193; ++j) {
            b.set_id.push_back(0);
            b.notional.push_back(1e5 + j);
            b.fixed.push_back(0.02);
            b.start.push_back(0.0);
            b.pay.insert(b.pay.end(), {1.0, 2.0});
            b.acc.insert(b.acc.end(), {1.0, 1.0});
            b.off.push_back((int32_t)b.pay.size());
        }
        const eqhip_portfolio pf = b.view(1);
        const double dates[3] = {0.0, 0.5, 1.5};
        const eqhip_mc mc{256, 6, 0, EQHIP_RNG_HOST_DRAWS, EQHIP_DISC_EULER};
        std::vector<double> z(6 * 256);
        std::mt19937_64 g(3);
        std::normal_distribution<double> nd;
        for (double& v : z) v = nd(g);
        eqhip_plan* plan = nullptr;
        OK(eqhip_plan_create(ctx, &m1, &pf, dates, 3, &mc, EQHIP_PLAN_TRADE_EE | EQHIP_PLAN_TRADE_PFE, &plan));
        if (plan) {
            std::vector<double> see(3), spfe(3), tee(8193 * 3), tpfe(8193 * 3);
            eqhip_outputs o{};
            o.set_ee = see.data();
            o.set_pfe = spfe.data();
            o.trade_ee = tee.data();
            o.trade_pfe = tpfe.data();
            OK(eqhip_plan_execute(plan, z.data(), nullptr, &o));
            OK(eqhip_plan_destroy(plan));
        }
    }
    // every valuation variant of a small book, autotuned and fixed
    {
        const Book b = diverse(40, 3, 5, 4);
        const eqhip_portfolio pf = b.view(3);
        const eqhip_mc mc{3000, 48, 9, EQHIP_RNG_PHILOX, EQHIP_DISC_EXACT};
        for (int ppl : {1, 2, 4})
            for (int tpi : {2, 4, 8}) {
                eqhip_plan* plan = nullptr;
                OK(eqhip_plan_create(ctx, &m1, &pf, monthly.data(), 48, &mc, EQHIP_PLAN_TRADE_EE, &plan));
                if (!plan) continue;
                const int rc = eqhip_plan_set_variant(plan, ppl, tpi);
                if (rc != 0 && rc != EQHIP_E_CAPACITY) ++fails;
                std::vector<double> see(3 * 48), spfe(3 * 48), tee(40 * 48);
                eqhip_outputs o{};
                o.set_ee = see.data();
                o.set_pfe = spfe.data();
                o.trade_ee = tee.data();
                OK(eqhip_plan_execute(plan, nullptr, nullptr, &o));
                OK(eqhip_plan_destroy(plan));
            }
        eqhip_plan* plan = nullptr;
        OK(eqhip_plan_create(ctx, &m1, &pf, monthly.data(), 48, &mc, EQHIP_PLAN_TRADE_EE | EQHIP_PLAN_AUTOTUNE, &plan));
        if (plan) {
            std::vector<double> see(3 * 48), spfe(3 * 48);
            eqhip_outputs o{};
            o.set_ee = see.data();
            o.set_pfe = spfe.data();
            OK(eqhip_plan_execute(plan, nullptr, nullptr, &o));
            OK(eqhip_plan_destroy(plan));
        }
    }
    // a batch of uneven jobs, autotuned
    {
        std::vector<Book> books;
        std::vector<eqhip_portfolio> pfs;
        std::vector<uint64_t> paths;
        for (int i = 0; i < 5; ++i) {  // a batch job must fit one block: regular schedules of a few tenors
            Book b;
            for (int j = 0; j < 10 + 37 * i; ++j) {
                const int years = 1 + (j * 7 + i) % 10, f = (j % 2) ? 2 : 1;
                for (int k = 1; k <= years * f; ++k) {
                    b.pay.push_back((double)k / f);
                    b.acc.push_back(1.0 / f);
                }
                b.off.push_back((int32_t)b.pay.size());
                b.set_id.push_back(0);
                b.notional.push_back((j % 3 ? 1.0 : -1.0) * 1e6);
                b.fixed.push_back(0.02 + 0.0003 * (j % 5));
                b.start.push_back(0.0);
            }
            books.push_back(b);
        }
        for (int i = 0; i < 5; ++i) {
            pfs.push_back(books[i].view(1));
            paths.push_back(1000 + 1731 * i);
        }
        std::vector<double> q;
        for (int k = 1; k <= 24; ++k) q.push_back(k / 4.0);
        eqhip_batch* bt = nullptr;
        OK(eqhip_batch_create(ctx, &m1, 5, pfs.data(), paths.data(), q.data(), 24, 24, 7, EQHIP_RNG_PHILOX,
                              EQHIP_DISC_EXACT, 1, &bt));
        if (bt) {
            std::vector<std::vector<double>> bufs(5 * 4);
            std::vector<eqhip_outputs> outs(5);
            for (int i = 0; i < 5; ++i) {
                bufs[4 * i].resize(24);
                bufs[4 * i + 1].resize(24);
                bufs[4 * i + 2].resize(1);
                bufs[4 * i + 3].resize(24 * (size_t)pfs[i].n_trades);
                outs[i] = eqhip_outputs{bufs[4 * i].data(), bufs[4 * i + 1].data(), bufs[4 * i + 2].data(),
                                        bufs[4 * i + 3].data(), nullptr, nullptr, nullptr, nullptr};
            }
            OK(eqhip_batch_execute(bt, &cr, outs.data()));
            OK(eqhip_batch_execute(bt, &cr, outs.data()));
            OK(eqhip_batch_destroy(bt));
        }
    }
    // G2++ plan
    {
        const Book b = diverse(30, 2, 8, 5);
        const eqhip_portfolio pf = b.view(2);
        const eqhip_model_hw2f m2{0.1, 0.2, 0.01, 0.012, -0.7, 0.03};
        const eqhip_mc mc{2000, 60, 3, EQHIP_RNG_PHILOX, EQHIP_DISC_EXACT};
        eqhip_plan* plan = nullptr;
        OK(eqhip_plan_create_2f(ctx, &m2, &pf, monthly.data(), 60, &mc, EQHIP_PLAN_TRADE_EE, &plan));
        if (plan) {
            std::vector<double> see(2 * 60), spfe(2 * 60), scva(2), tee(30 * 60), tcva(30);
            eqhip_outputs o{see.data(), spfe.data(), scva.data(), tee.data(), tcva.data(), nullptr, nullptr, nullptr};
            OK(eqhip_plan_execute(plan, nullptr, &cr, &o));
            OK(eqhip_plan_destroy(plan));
        }
    }
    OK(eqhip_context_destroy(ctx));
    if (fails) std::fprintf(stderr, "%d failures\n", fails);
    else std::printf("C ABI self-test passed\n");
    std::fflush(stdout);
    std::fflush(stderr);
    // end without the HIP runtime's exit-time teardown, which the sanitizer's allocator does not follow
    // (its check fires inside the runtime's own finalisation, after this program's work)
    std::_Exit(fails ? 1 : 0);
}
