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
// Valuation. One block = one work unit (plan, date k, chunk of path tiles); one lane = one path.
// Per tile of 64 paths:
//   A  P(t_k, T_d) = A exp(-B r) for every live cash-flow date d      -> LDS sP[d][lane]
//   B  one wave per schedule family: running annuity over the family's live flows; at the last
//      flow of each class, (float leg, annuity) of that class          -> LDS sC[c][lane]
//   C  netting sets: V_s = sum_c alpha_sc float_c - beta_sc annuity_c   -> X[s][k][path]
//   D  trades: x = s (float - K annuity), acc += max(x, 0) over the paths of the chunk
// Schedule families hold classes whose schedules are prefixes of one another, so one running sum
// yields every class's annuity, bit-identical to summing that class's flows in order.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kValueThreads) value_kernel(const ValuePlan* plans, const int4* units) {
    extern __shared__ double smem[];
    const int4 u = units[blockIdx.x];
    const ValuePlan& g = plans[u.x];
    const int k = u.y;
    const int chunk = u.z;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    double* sR = smem;                                     // [64]
    double* sP = sR + kPathsPerTile;                       // [D][64]
    double* sF = sP + (size_t)g.D * kPathsPerTile;         // [Fmax][64]
    double2* sC = (double2*)(sF + (size_t)g.Fmax * kPathsPerTile);  // [C][64]
    double* sPart = (double*)(sC + (size_t)g.C * kPathsPerTile);    // [S][waves][64] (split mode)

    const double t = g.dates[k];
    const int dfirst = g.date_dfirst[k];
    const int n_live = g.D - dfirst;
    const int fbeg = g.fix_off[k];
    const int nfix = g.fix_off[k + 1] - fbeg;
    const int det = g.date_det[k];
    const uint64_t n_paths = g.n_paths;
    const double* Rk = g.R + (uint64_t)g.date_slot[k] * n_paths;
    const double2* ab = g.AB + (size_t)k * g.D;
    const int4* fdesc = g.fdesc + (size_t)k * g.NF;

    // trade work items owned by this thread
    double acc[kItemsPerThread][kTradesPerItem];
    double tK[kItemsPerThread][kTradesPerItem];
    int icls[kItemsPerThread], icnt[kItemsPerThread], ifirst[kItemsPerThread];
    double isgn[kItemsPerThread];
#pragma unroll
    for (int i = 0; i < kItemsPerThread; ++i) {
        const int it = tid + i * kValueThreads;
        icnt[i] = 0;
        icls[i] = 0;
        ifirst[i] = 0;
        isgn[i] = 1.0;
        if (g.want_trade && it < g.n_items) {
            int4 item = g.items[it];
            icls[i] = item.x;
            ifirst[i] = item.y;
            isgn[i] = (double)item.w;
            // the class is alive at t when its last payment is still to come
            icnt[i] = (g.cls_mat[item.x] > t + kDateTol) ? item.z : 0;
            if (g.path_mtm) icnt[i] = item.z;
        }
#pragma unroll
        for (int q = 0; q < kTradesPerItem; ++q) {
            acc[i][q] = 0.0;
            tK[i][q] = (q < icnt[i]) ? g.tr_K[ifirst[i] + q] : 0.0;
        }
    }

    const uint64_t tile0 = (uint64_t)chunk * g.tiles_per_block;
    uint64_t tile_end = tile0 + g.tiles_per_block;
    if (tile_end > g.n_tiles) tile_end = g.n_tiles;

    for (uint64_t tile = tile0; tile < tile_end; ++tile) {
        const uint64_t p0 = tile * kPathsPerTile;
        const uint64_t gp = p0 + lane;
        const bool valid = gp < n_paths;
        __syncthreads();
        if (tid < kPathsPerTile) sR[lane] = valid ? Rk[gp] : 0.0;
        for (int f = wave; f < nfix; f += kValueWaves) {
            double rf = valid ? g.R[(uint64_t)g.fix_slot[fbeg + f] * n_paths + gp] : 0.0;
            double2 fab = g.fix_AB[fbeg + f];
            sF[f * kPathsPerTile + lane] = fab.x * exp(-(fab.y * rf));
        }
        __syncthreads();
        // A: zero-coupon bonds of the live cash-flow dates
        {
            const double r = sR[lane];
            for (int dl = wave; dl < n_live; dl += kValueWaves) {
                const int d = dfirst + dl;
                const double2 v = ab[d];
                sP[d * kPathsPerTile + lane] = v.x * exp(-(v.y * r));
            }
        }
        __syncthreads();
        // B: schedule families
        for (int F = wave; F < g.NF; F += kValueWaves) {
            const int4 fd = fdesc[F];
            const int f0 = g.fam_off[F];
            const int f1 = g.fam_off[F + 1];
            const int first = (fd.x == 0) ? f1 : f0 + fd.y;
            for (int i = f0; i < first; ++i) {  // classes already matured
                const int c = g.flow_cls[i];
                if (c >= 0) sC[c * kPathsPerTile + lane] = make_double2(0.0, 0.0);
            }
            if (fd.x == 0) continue;
            double st;
            if (fd.x == 1) st = 1.0;
            else if (fd.x == 2) st = sP[fd.z * kPathsPerTile + lane];
            else st = sP[fd.z * kPathsPerTile + lane] / sF[fd.w * kPathsPerTile + lane];
            double ann = 0.0;
            for (int i = first; i < f1; ++i) {
                const double P = sP[g.flow_d[i] * kPathsPerTile + lane];
                ann += P * g.flow_tau[i];
                const int c = g.flow_cls[i];
                if (c >= 0) sC[c * kPathsPerTile + lane] = make_double2(st - P, ann);
            }
        }
        __syncthreads();
        // C: netting sets
        if (g.set_mode == 0) {
            for (int s = 0; s < g.S; ++s) {
                const int beg = g.set_off[s], end = g.set_off[s + 1];
                const int per = (end - beg + kValueWaves - 1) / kValueWaves;
                const int cb = beg + wave * per;
                const int ce = min(end, cb + per);
                double v = 0.0;
                for (int e = cb; e < ce; ++e) {
                    const double2 c2 = sC[g.set_cls[e] * kPathsPerTile + lane];
                    v += g.set_alpha[e] * c2.x - g.set_beta[e] * c2.y;
                }
                sPart[(s * kValueWaves + wave) * kPathsPerTile + lane] = v;
            }
        } else {
            for (int s = wave; s < g.S; s += kValueWaves) {
                double v = 0.0;
                for (int e = g.set_off[s]; e < g.set_off[s + 1]; ++e) {
                    const double2 c2 = sC[g.set_cls[e] * kPathsPerTile + lane];
                    v += g.set_alpha[e] * c2.x - g.set_beta[e] * c2.y;
                }
                if (valid) {
                    g.X[((uint64_t)s * g.n_dates + k) * n_paths + gp] = (v > 0.0) ? v : 0.0;
                    if (g.set_mtm) g.set_mtm[((uint64_t)k * g.S + s) * n_paths + gp] = v;
                }
            }
        }
        // D: per-trade exposure sums (reads sC only)
        if (g.want_trade) {
            const uint64_t left = n_paths - p0;
            const int pvalid = (left < (uint64_t)kPathsPerTile) ? (int)left : kPathsPerTile;
            const int pend = det ? ((p0 == 0) ? 1 : 0) : pvalid;  // grid index 0: path 0 only
#pragma unroll
            for (int i = 0; i < kItemsPerThread; ++i) {
                if (icnt[i] == 0) continue;
                const double2* row = sC + icls[i] * kPathsPerTile;
                const double sg = isgn[i];
                for (int p = 0; p < pend; ++p) {
                    const double2 c2 = row[p];
                    const double sfl = sg * c2.x;
                    const double sann = sg * c2.y;
#pragma unroll
                    for (int q = 0; q < kTradesPerItem; ++q) {
                        const double x = fma(-tK[i][q], sann, sfl);
                        acc[i][q] += (x > 0.0) ? x : 0.0;
                    }
                }
                if (g.path_mtm) {
                    for (int q = 0; q < icnt[i]; ++q) {
                        const int j = ifirst[i] + q;
                        const double N = g.tr_N[j], NK = g.tr_NK[j];
                        double* dst = g.path_mtm + ((uint64_t)k * g.n_trades + g.tr_orig[j]) * n_paths + p0;
                        for (int p = 0; p < pvalid; ++p) dst[p] = N * row[p].x - NK * row[p].y;
                    }
                }
            }
        }
        if (g.set_mode == 0) {
            __syncthreads();
            if (tid < g.S * kPathsPerTile) {
                const int s = tid >> 6;
                double v = 0.0;
                for (int w = 0; w < kValueWaves; ++w) v += sPart[(s * kValueWaves + w) * kPathsPerTile + lane];
                if (valid) {
                    g.X[((uint64_t)s * g.n_dates + k) * n_paths + gp] = (v > 0.0) ? v : 0.0;
                    if (g.set_mtm) g.set_mtm[((uint64_t)k * g.S + s) * n_paths + gp] = v;
                }
            }
        }
    }
    if (g.want_trade) {
#pragma unroll
        for (int i = 0; i < kItemsPerThread; ++i) {
            const int it = tid + i * kValueThreads;
            if (!(it < g.n_items)) continue;
            const int4 item = g.items[it];
#pragma unroll
            for (int q = 0; q < kTradesPerItem; ++q) {
                if (q < item.z) {
                    const int j = item.y + q;
                    // a matured class contributes exact zeros
                    const double e = (icnt[i] == 0) ? 0.0 : g.tr_absN[j] * acc[i][q];
                    g.partial[((uint64_t)g.tr_orig[j] * g.n_dates + k) * g.n_chunks + chunk] = e;
                }
            }
        }
    }
}

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
