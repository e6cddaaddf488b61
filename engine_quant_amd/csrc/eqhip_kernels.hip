// Exposure kernels for gfx950 (MI355X), FP64 throughout.
//
//   simulate_kernel    one thread per path; the short rate lives in a register, shocks come from
//                      a coalesced [step][path] buffer or from in-register Philox4x32-10; r is
//                      stored only at the grid steps the valuation needs ([slot][path]).
//   ab_kernel          A(tau), B(tau) of the affine ZCB for every (date, cash-flow date) pair.
//   value_kernel<PT>   one block per (monitoring date, path chunk); per tile of PT paths it
//                      stages P(t,T_d) for every live cash-flow date in LDS, builds each schedule
//                      class's floating-leg and annuity factors in LDS, writes the netting-set
//                      exposure max(V,0) for every set, and accumulates per-trade EE partial sums
//                      in registers (trades sorted by class, several trades per thread).
//   select_kernel      one block per (set, date) row: fixed-order EE sum + exact radix select of
//                      the PFE95 order statistic on the bit patterns of the non-negative doubles.
//   trade_ee_kernel    fixed-order sum of per-trade EE partials.
//   cva_kernel         unilateral CVA of each profile, exposure.rs:242-275.
//   pv_dv01_kernel     market-curve PV / DV01 / bucketed DV01 per trade, measure.cpp:124-222.
#include "eqhip_device.hpp"
#include "eqhip_kernels.hpp"

namespace eqhip {

// ---------------------------------------------------------------------------------------------
// Simulation. Euler: r' = (r + (a*(b-r))*dt) + sigma*(z*sqrt(dt))   (kernel.rs:41-50,
// hull_white.rs:91-95). Exact: r' = (r*e + b(1-e)) + sd*z. Contraction is off so that each step
// rounds like the reference's separate tensor ops.
// ---------------------------------------------------------------------------------------------
template <int DISC, int RNG>
__global__ void __launch_bounds__(256) simulate_kernel(SimArgs g) {
#pragma clang fp contract(off)
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n_paths) return;
    double r = g.r0;
    double z_next = 0.0;
    for (uint32_t s = 0;; ++s) {
        const int slot = g.slot_of_step[s];
        if (slot >= 0) g.R[(uint64_t)slot * g.n_paths + p] = r;
        if (s == g.n_steps) break;
        double z;
        if (RNG == 0) {
            z = g.Z[(uint64_t)s * g.n_paths + p];
        } else {
            if ((s & 1u) == 0u) {
                philox_normal_pair(g.seed, p, s >> 1, z, z_next);
            } else {
                z = z_next;
            }
        }
        if (DISC == 0) {
            double drift = g.a * (g.b - r);
            double dw = z * g.sqrt_dt;
            r = (r + drift * g.dt) + g.sigma * dw;
        } else {
            r = (r * g.decay + g.mean_c) + g.sd * z;
        }
    }
}

template __global__ void simulate_kernel<0, 0>(SimArgs);
template __global__ void simulate_kernel<0, 1>(SimArgs);
template __global__ void simulate_kernel<1, 0>(SimArgs);
template __global__ void simulate_kernel<1, 1>(SimArgs);

// A(tau), B(tau) for a flat list of year fractions (tau <= 0 entries are unused: A=B=0).
__global__ void ab_kernel(const double* tau, int n, double a, double b, double sigma,
                          double2* out) {
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double t = tau[i];
    double2 v;
    if (t < 0.0) {
        v.x = 0.0;
        v.y = 0.0;
    } else {
        double bf = hw_b(a, t);
        v.x = hw_a(a, b, sigma, t, bf);
        v.y = bf;
    }
    out[i] = v;
}

// ---------------------------------------------------------------------------------------------
// Valuation.
// ---------------------------------------------------------------------------------------------
template <int PT>
__global__ void __launch_bounds__(kValueThreads) value_kernel(ValueArgs g) {
    extern __shared__ double smem[];
    double* sR = smem;                    // [PT]
    double* sP = sR + PT;                 // [D][PT]     P(t_k, T_d) for live d
    double* sF = sP + (size_t)g.D * PT;   // [Fmax][PT]  fixing bonds P(T_{k-1}, T_k | r_fix)
    double2* sC = (double2*)(sF + (size_t)g.Fmax * PT);  // [C][PT]  (float, annuity)

    const int tid = threadIdx.x;
    const int k = blockIdx.x / g.n_chunks;
    const int chunk = blockIdx.x - k * g.n_chunks;
    const int dfirst = g.date_dfirst[k];
    const int n_live = g.D - dfirst;
    const int fbeg = g.fix_off[k];
    const int nfix = g.fix_off[k + 1] - fbeg;
    const int det = g.date_det[k];
    const uint64_t slot_base = (uint64_t)g.date_slot[k] * g.n_paths;
    const double2* ab = g.AB + (size_t)k * g.D;
    const int4* cdesc = g.cdesc + (size_t)k * g.C;
    const int2* cdate = g.cdate + (size_t)k * g.C;

    // per-trade accumulators: thread owns items tid + i*blockDim (i < kItemsPerThread), each up
    // to kTradesPerItem trades of one class
    double acc[kItemsPerThread][kTradesPerItem];
    double tN[kItemsPerThread][kTradesPerItem];
    double tNK[kItemsPerThread][kTradesPerItem];
    int icls[kItemsPerThread], icnt[kItemsPerThread];
#pragma unroll
    for (int i = 0; i < kItemsPerThread; ++i) {
        int it = tid + i * kValueThreads;
        icnt[i] = 0;
        icls[i] = 0;
        if (g.want_trade && it < g.n_items) {
            icls[i] = g.item_cls[it];
            icnt[i] = g.item_count[it];
            int f = g.item_first[it];
#pragma unroll
            for (int q = 0; q < kTradesPerItem; ++q) {
                tN[i][q] = (q < icnt[i]) ? g.tr_N[f + q] : 0.0;
                tNK[i][q] = (q < icnt[i]) ? g.tr_NK[f + q] : 0.0;
            }
        }
#pragma unroll
        for (int q = 0; q < kTradesPerItem; ++q) acc[i][q] = 0.0;
    }

    const uint64_t tile0 = (uint64_t)chunk * g.tiles_per_block;
    uint64_t tile_end = tile0 + g.tiles_per_block;
    if (tile_end > g.n_tiles) tile_end = g.n_tiles;

    for (uint64_t tile = tile0; tile < tile_end; ++tile) {
        const uint64_t p0 = tile * PT;
        __syncthreads();  // previous tile's readers are done with the LDS
        if (tid < PT) {
            uint64_t gp = p0 + tid;
            sR[tid] = (gp < g.n_paths) ? g.R[slot_base + gp] : 0.0;
        }
        // fixing bonds of the periods in progress
        for (int idx = tid; idx < nfix * PT; idx += kValueThreads) {
            int f = idx / PT, p = idx - f * PT;
            uint64_t gp = p0 + p;
            double rf = (gp < g.n_paths) ? g.R[(uint64_t)g.fix_slot[fbeg + f] * g.n_paths + gp] : 0.0;
            double2 fab = g.fix_AB[fbeg + f];
            sF[f * PT + p] = fab.x * exp(-(fab.y * rf));
        }
        __syncthreads();
        // P(t_k, T_d) for every live cash-flow date
        for (int idx = tid; idx < n_live * PT; idx += kValueThreads) {
            int dl = idx / PT, p = idx - dl * PT;
            int d = dfirst + dl;
            double2 v = ab[d];
            sP[d * PT + p] = v.x * exp(-(v.y * sR[p]));
        }
        __syncthreads();
        // schedule classes: floating-leg factor and fixed-leg annuity
        for (int idx = tid; idx < g.C * PT; idx += kValueThreads) {
            int c = idx / PT, p = idx - c * PT;
            int4 cd = cdesc[c];  // x mode, y first flow, z end flow, w fixing index
            double fl = 0.0, ann = 0.0;
            if (cd.x != 0) {
                for (int f = cd.y; f < cd.z; ++f) ann += sP[g.flow_d[f] * PT + p] * g.flow_tau[f];
                int2 dd = cdate[c];  // x start/next date, y maturity date
                double st;
                if (cd.x == 1) st = 1.0;
                else if (cd.x == 2) st = sP[dd.x * PT + p];
                else st = sP[dd.x * PT + p] / sF[cd.w * PT + p];
                fl = st - sP[dd.y * PT + p];
            }
            sC[c * PT + p] = make_double2(fl, ann);
        }
        __syncthreads();
        // netting sets: V_s = sum_c alpha_sc * float_c - beta_sc * annuity_c
        for (int idx = tid; idx < g.S * PT; idx += kValueThreads) {
            int s = idx / PT, p = idx - s * PT;
            uint64_t gp = p0 + p;
            double v = 0.0;
            for (int e = g.set_off[s]; e < g.set_off[s + 1]; ++e) {
                double2 c = sC[g.set_cls[e] * PT + p];
                v += g.set_alpha[e] * c.x - g.set_beta[e] * c.y;
            }
            if (gp < g.n_paths) {
                g.X[((uint64_t)s * g.n_dates + k) * g.n_paths + gp] = (v > 0.0) ? v : 0.0;
                if (g.set_mtm) g.set_mtm[((uint64_t)k * g.S + s) * g.n_paths + gp] = v;
            }
        }
        // per-trade EE partial sums (no barrier needed: reads sC only)
        if (g.want_trade) {
            uint64_t np_tile = g.n_paths - p0;
            const int pvalid = (np_tile < (uint64_t)PT) ? (int)np_tile : PT;
            int pend = pvalid;
            if (det) pend = (p0 == 0) ? 1 : 0;  // date at grid index 0: EE over path 0 only
#pragma unroll
            for (int i = 0; i < kItemsPerThread; ++i) {
                if (icnt[i] == 0) continue;
                const double2* row = sC + icls[i] * PT;
                for (int p = 0; p < pend; ++p) {
                    double2 c = row[p];
#pragma unroll
                    for (int q = 0; q < kTradesPerItem; ++q) {
                        double v = tN[i][q] * c.x - tNK[i][q] * c.y;
                        acc[i][q] += (v > 0.0) ? v : 0.0;
                    }
                }
                if (g.path_mtm) {
                    int f = g.item_first[tid + i * kValueThreads];
                    for (int q = 0; q < icnt[i]; ++q) {
                        int orig = g.tr_orig[f + q];
                        for (int p = 0; p < pvalid; ++p) {
                            double2 c = row[p];
                            g.path_mtm[((uint64_t)k * g.n_trades + orig) * g.n_paths + p0 + p] =
                                tN[i][q] * c.x - tNK[i][q] * c.y;
                        }
                    }
                }
            }
        }
    }
    if (g.want_trade) {
#pragma unroll
        for (int i = 0; i < kItemsPerThread; ++i) {
            if (icnt[i] == 0) continue;
            int f = g.item_first[tid + i * kValueThreads];
#pragma unroll
            for (int q = 0; q < kTradesPerItem; ++q) {
                if (q < icnt[i]) {
                    int orig = g.tr_orig[f + q];
                    g.partial[((uint64_t)orig * g.n_dates + k) * g.n_chunks + chunk] = acc[i][q];
                }
            }
        }
    }
}

template __global__ void value_kernel<16>(ValueArgs);
template __global__ void value_kernel<32>(ValueArgs);
template __global__ void value_kernel<64>(ValueArgs);

// ---------------------------------------------------------------------------------------------
// EE + PFE95 of one (set, date) row. EE = (fixed-order sum) / n. PFE = the element of rank
// q = min(ceil(0.95 n) - 1, n - 1) of the ascending order (exposure.rs:42-49), found exactly by
// an MSB-first radix select over the IEEE-754 bit patterns (monotone for x >= +0.0).
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kSelectThreads) select_kernel(const double* X, int n_dates,
                                                                uint64_t n_paths,
                                                                const uint64_t* n_eff,
                                                                const uint64_t* q_rank, double* ee,
                                                                double* pfe) {
    __shared__ unsigned int hist[2048];
    __shared__ unsigned int part[kSelectThreads];
    __shared__ double red[kSelectThreads];
    __shared__ unsigned long long s_prefix;
    __shared__ unsigned long long s_rank;

    const int tid = threadIdx.x;
    const int row = blockIdx.x;
    const int k = row % n_dates;
    const uint64_t n = n_eff[k];
    const unsigned long long* x = (const unsigned long long*)(X + (uint64_t)row * n_paths);

    double sum = 0.0;
    for (uint64_t i = tid; i < n; i += kSelectThreads) sum += __longlong_as_double(x[i]);
    red[tid] = sum;
    __syncthreads();
    for (int w = kSelectThreads / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] = red[tid] + red[tid + w];
        __syncthreads();
    }
    if (tid == 0) {
        ee[row] = red[0] / (double)n;
        s_prefix = 0ull;
        s_rank = q_rank[k];
    }
    __syncthreads();

    const int shifts[6] = {53, 42, 31, 20, 9, 0};
    const int widths[6] = {11, 11, 11, 11, 11, 9};
    for (int pass = 0; pass < 6; ++pass) {
        const int sh = shifts[pass];
        const int wd = widths[pass];
        const int hi = sh + wd;  // bits >= hi are fixed by the prefix
        const unsigned long long prefix = s_prefix;
        const unsigned int nb = 1u << wd;
        for (unsigned int b = tid; b < nb; b += kSelectThreads) hist[b] = 0u;
        __syncthreads();
        for (uint64_t i = tid; i < n; i += kSelectThreads) {
            unsigned long long key = x[i];
            bool match = (hi >= 64) || ((key >> hi) == (prefix >> hi));
            if (match) atomicAdd(&hist[(key >> sh) & (nb - 1u)], 1u);
        }
        __syncthreads();
        const unsigned int per = nb / kSelectThreads;  // 8 bins per thread (2 in the last pass)
        unsigned int local = 0;
        for (unsigned int b = 0; b < per; ++b) local += hist[tid * per + b];
        part[tid] = local;
        __syncthreads();
        if (tid == 0) {
            unsigned long long rank = s_rank;
            unsigned long long cum = 0;
            int t = 0;
            for (; t < kSelectThreads; ++t) {
                if (cum + part[t] > rank) break;
                cum += part[t];
            }
            unsigned int b = t * per;
            for (;; ++b) {
                if (cum + hist[b] > rank) break;
                cum += hist[b];
            }
            s_prefix = prefix | ((unsigned long long)b << sh);
            s_rank = rank - cum;
        }
        __syncthreads();
    }
    if (tid == 0) pfe[row] = __longlong_as_double(s_prefix);
}

// Per-trade EE: partial sums over chunks in chunk order, divided by the effective path count.
__global__ void trade_ee_kernel(const double* partial, int n_trades, int n_dates, int n_chunks,
                                const uint64_t* n_eff, double* trade_ee) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_trades * n_dates) return;
    int k = i % n_dates;
    const double* p = partial + (uint64_t)i * n_chunks;
    double s = 0.0;
    for (int c = 0; c < n_chunks; ++c) s += p[c];
    trade_ee[i] = s / (double)n_eff[k];
}

// CVA = sum_k (1-R) EE_k (S_{k-1} - S_k) P(0,t_k | r0), S = exp(-h t), exposure.rs:242-275.
__global__ void cva_kernel(const double* ee, int n_profiles, int n_dates, const double* times,
                           double a, double b, double sigma, double r0, double hazard,
                           double recovery, double* cva) {
#pragma clang fp contract(off)
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_profiles) return;
    double acc = 0.0, prev_survival = 1.0;
    for (int k = 0; k < n_dates; ++k) {
        double t = times[k];
        double survival = exp(-hazard * t);
        double default_prob = prev_survival - survival;
        double tau = t - 0.0;
        double bf = hw_b(a, tau);
        double disc = hw_a(a, b, sigma, tau, bf) * exp(-(bf * r0));
        acc += (1.0 - recovery) * ee[(uint64_t)j * n_dates + k] * default_prob * disc;
        prev_survival = survival;
    }
    cva[j] = acc;
}

// ---------------------------------------------------------------------------------------------
// Market-curve PV / DV01 (measure.cpp:160-222, market.cpp:26-41). Thread (trade, variant):
// variant 0 = PV, 1 = parallel bump, 2+i = pillar i bumped.
// ---------------------------------------------------------------------------------------------
__device__ double curve_df(const double* pillars, const double* zeros, int np, int bump_pillar,
                           double bump, double t) {
#pragma clang fp contract(off)
    auto zr = [&](int i) {
        double z = zeros[i];
        if (bump_pillar == -2 || bump_pillar == i) z = z + bump;
        return z;
    };
    double z;
    if (t <= pillars[0]) {
        z = zr(0);
    } else if (t >= pillars[np - 1]) {
        z = zr(np - 1);
    } else {
        int i = 0;
        while (i < np && pillars[i] < t) ++i;
        if (i == 0) i = 1;
        double t0 = pillars[i - 1], t1 = pillars[i];
        double z0 = zr(i - 1), z1 = zr(i);
        double w = (t - t0) / (t1 - t0);
        z = z0 + w * (z1 - z0);
    }
    return exp(-z * t);
}

__device__ double npv_from_curve(const double* pillars, const double* zeros, int np,
                                 int bump_pillar, double bump, double notional, double fixed_rate,
                                 double start, const double* pay, const double* acc, int m) {
#pragma clang fp contract(off)
    double dv = curve_df(pillars, zeros, np, bump_pillar, bump, 0.0);
    double p_start = curve_df(pillars, zeros, np, bump_pillar, bump, start) / dv;
    double p_end = curve_df(pillars, zeros, np, bump_pillar, bump, pay[m - 1]) / dv;
    double floating_leg = notional * (p_start - p_end);
    double fixed_leg = 0.0;
    for (int i = 0; i < m; ++i)
        fixed_leg += notional * fixed_rate * acc[i] *
                     (curve_df(pillars, zeros, np, bump_pillar, bump, pay[i]) / dv);
    return floating_leg - fixed_leg;
}

__global__ void pv_dv01_kernel(PvArgs g) {
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int nv = 2 + g.np;
    if (i >= g.n_trades * nv) return;
    int j = i / nv, v = i - j * nv;
    int f0 = g.flow_off[j], m = g.flow_off[j + 1] - f0;
    const double* pay = g.pay + f0;
    const double* acc = g.acc + f0;
    double base = npv_from_curve(g.pillars, g.zeros, g.np, -1, 0.0, g.notional[j], g.fixed[j],
                                 g.start[j], pay, acc, m);
    if (v == 0) {
        g.pv[j] = base;
        return;
    }
    int bp = (v == 1) ? -2 : v - 2;
    double bumped = npv_from_curve(g.pillars, g.zeros, g.np, bp, g.bump, g.notional[j], g.fixed[j],
                                   g.start[j], pay, acc, m);
    if (v == 1) g.dv01[j] = bumped - base;
    else if (g.bucketed) g.bucketed[(uint64_t)j * g.np + (v - 2)] = bumped - base;
}

}  // namespace eqhip
