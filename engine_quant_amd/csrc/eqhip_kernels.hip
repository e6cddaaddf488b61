// Exposure kernels for gfx950 (MI355X), FP64 throughout.
//
//   simulate_kernel    one thread per path; the short rate lives in a register, shocks come from
//                      a coalesced [step][path] buffer or from in-register Philox4x32-10; r is
//                      stored only at the grid steps the valuation needs ([slot][path]).
//   simulate2_kernel   the same for G2++: x and y in registers, stored as two [slot][path] buffers.
//   ab_kernel          A(tau), B(tau) of the affine ZCB for every (date, cash-flow date) pair;
//   ab2_kernel         (c, B(a,tau)) and B(b,tau) of the G2++ bond likewise.
//   value_kernel       one 1024-thread block per work unit (plan, monitoring date, chunk of 64-path
//   value2_kernel      tiles); per tile it stages P(t,T_d) for every live cash-flow date and the
//                      fixing factors in LDS, builds each schedule class's floating-leg and annuity
//                      factors, writes the netting-set values (and their positive parts) for every
//                      set, and accumulates per-trade exposure sums over path slices. value2_kernel
//                      is the G2++ instance (two-factor bonds in stage A). A plan at one path per
//                      lane launches value_p1_kernel<TPI> (value2_kernel<TPI>), compiled for its
//                      trades-per-item variant alone.
//   value_path_kernel  path mode (plans without per-trade outputs, <= 4 schedule families and <= 8
//                      netting sets): each thread carries 4 paths (2 past 4 sets) through the date's
//                      live cash-flow dates in date order, family annuities and set values in
//                      registers, no LDS and no barriers.
//   value_flat_kernel  homogeneous batches (one schedule, one trade per set): one thread per path
//                      and date, one floating leg and annuity shared by the trades.
//   select_kernel      one block per row: fixed-order sum + exact radix select of an order
//                      statistic on the bit patterns of the non-negative doubles.
//   xva_rows_kernel    one block per exposure-cube row: CSA adjustment, positive parts and ENE.
//   trade_ee_kernel    fixed-order sum of per-trade EE partials.
//   cva_kernel         unilateral CVA of each profile (one wave per profile), exposure.rs:242-275.
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

// Philox paths with four threads per path drawing the shocks: the 4 waves of a block share 64 paths.
// Per batch of kSimBatch steps each wave draws a quarter of the batch's Philox pairs for every path into
// LDS, then wave 0 runs the paths' recursions over the batch. With four times the waves of one thread
// per path, the Box-Muller transcendentals (the bulk of the work) have the occupancy to hide their
// latency; the shocks and every step's rounding are those of simulate_kernel<DISC, 1>.
template <int DISC>
__global__ void __launch_bounds__(256) simulate_philox_kernel(SimArgs g) {
#pragma clang fp contract(off)
    __shared__ double sz[kSimBatch][kSimPaths];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t p = (uint64_t)blockIdx.x * kSimPaths + lane;
    const bool valid = p < g.n_paths;
    double r = g.r0;
    for (uint32_t s0 = 0; s0 < g.n_steps; s0 += kSimBatch) {
        for (int q = wave; q < kSimBatch / 2; q += 4) {  // pair (s0 + 2q) / 2: steps s0 + 2q, s0 + 2q + 1
            double z0 = 0.0, z1 = 0.0;
            if (valid) philox_normal_pair(g.seed, p, (s0 >> 1) + (uint32_t)q, z0, z1);
            sz[2 * q][lane] = z0;
            sz[2 * q + 1][lane] = z1;
        }
        __syncthreads();
        if (wave == 0 && valid) {
            const uint32_t cnt = min((uint32_t)kSimBatch, g.n_steps - s0);
            for (uint32_t u = 0; u < cnt; ++u) {
                const int slot = g.slot_of_step[s0 + u];
                if (slot >= 0) g.R[(uint64_t)slot * g.n_paths + p] = r;
                const double z = sz[u][lane];
                if (DISC == 0) {
                    double drift = g.a * (g.b - r);
                    double dw = z * g.sqrt_dt;
                    r = (r + drift * g.dt) + g.sigma * dw;
                } else {
                    r = (r * g.decay + g.mean_c) + g.sd * z;
                }
            }
        }
        __syncthreads();  // the next batch overwrites sz
    }
    if (wave == 0 && valid) {
        const int slot = g.slot_of_step[g.n_steps];
        if (slot >= 0) g.R[(uint64_t)slot * g.n_paths + p] = r;
    }
}

template __global__ void simulate_philox_kernel<0>(SimArgs);
template __global__ void simulate_philox_kernel<1>(SimArgs);

template __global__ void simulate_kernel<0, 0>(SimArgs);
template __global__ void simulate_kernel<0, 1>(SimArgs);
template __global__ void simulate_kernel<1, 0>(SimArgs);
template __global__ void simulate_kernel<1, 1>(SimArgs);

// G2++ simulation (see Sim2Args). One Philox pair per step gives both factors' shocks.
template <int DISC, int RNG>
__global__ void __launch_bounds__(256) simulate2_kernel(Sim2Args g) {
#pragma clang fp contract(off)
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n_paths) return;
    const uint64_t n = g.n_paths;
    double x = 0.0, y = 0.0;
    for (uint32_t s = 0;; ++s) {
        const int slot = g.slot_of_step[s];
        if (slot >= 0) {
            g.X[(uint64_t)slot * n + p] = x;
            g.Y[(uint64_t)slot * n + p] = y;
        }
        if (s == g.n_steps) break;
        double zx, zy;
        if (RNG == 0) {
            zx = g.Z[(uint64_t)s * n + p];
            zy = g.Z[((uint64_t)g.n_steps + s) * n + p];
        } else {
            philox_normal_pair(g.seed, p, s, zx, zy, 1u);
        }
        if (DISC == 0) {
            const double zc = zy * g.sqrt_1mrho2 + zx * g.rho;
            const double drift_x = (-g.a) * x;
            const double drift_y = (-g.b) * y;
            x = (x + drift_x * g.dt) + g.sigma * (zx * g.sqrt_dt);
            y = (y + drift_y * g.dt) + g.eta * (zc * g.sqrt_dt);
        } else {
            const double zc = g.c_xy * zx + g.sqrt_1mc2 * zy;
            x = x * g.ex + g.sdx * zx;
            y = y * g.ey + g.sdy * zc;
        }
    }
}

// simulate_philox_kernel for G2++: one Philox pair per step gives both factors' shocks.
template <int DISC>
__global__ void __launch_bounds__(256) simulate2_philox_kernel(Sim2Args g) {
#pragma clang fp contract(off)
    constexpr int kB = kSimBatch / 2;
    __shared__ double szx[kB][kSimPaths], szy[kB][kSimPaths];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t p = (uint64_t)blockIdx.x * kSimPaths + lane;
    const uint64_t n = g.n_paths;
    const bool valid = p < n;
    double x = 0.0, y = 0.0;
    for (uint32_t s0 = 0; s0 < g.n_steps; s0 += kB) {
        for (int q = wave; q < kB; q += 4) {
            double zx = 0.0, zy = 0.0;
            if (valid) philox_normal_pair(g.seed, p, s0 + (uint32_t)q, zx, zy, 1u);
            szx[q][lane] = zx;
            szy[q][lane] = zy;
        }
        __syncthreads();
        if (wave == 0 && valid) {
            const uint32_t cnt = min((uint32_t)kB, g.n_steps - s0);
            for (uint32_t u = 0; u < cnt; ++u) {
                const int slot = g.slot_of_step[s0 + u];
                if (slot >= 0) {
                    g.X[(uint64_t)slot * n + p] = x;
                    g.Y[(uint64_t)slot * n + p] = y;
                }
                const double zx = szx[u][lane], zy = szy[u][lane];
                if (DISC == 0) {
                    const double zc = zy * g.sqrt_1mrho2 + zx * g.rho;
                    const double drift_x = (-g.a) * x;
                    const double drift_y = (-g.b) * y;
                    x = (x + drift_x * g.dt) + g.sigma * (zx * g.sqrt_dt);
                    y = (y + drift_y * g.dt) + g.eta * (zc * g.sqrt_dt);
                } else {
                    const double zc = g.c_xy * zx + g.sqrt_1mc2 * zy;
                    x = x * g.ex + g.sdx * zx;
                    y = y * g.ey + g.sdy * zc;
                }
            }
        }
        __syncthreads();
    }
    if (wave == 0 && valid) {
        const int slot = g.slot_of_step[g.n_steps];
        if (slot >= 0) {
            g.X[(uint64_t)slot * n + p] = x;
            g.Y[(uint64_t)slot * n + p] = y;
        }
    }
}

template __global__ void simulate2_philox_kernel<0>(Sim2Args);
template __global__ void simulate2_philox_kernel<1>(Sim2Args);

template __global__ void simulate2_kernel<0, 0>(Sim2Args);
template __global__ void simulate2_kernel<0, 1>(Sim2Args);
template __global__ void simulate2_kernel<1, 0>(Sim2Args);
template __global__ void simulate2_kernel<1, 1>(Sim2Args);

// G2++ bond factors for a flat list of year fractions: (c, B(a,tau)) and B(b,tau), with
// P(t, t+tau) = exp((c - B(a,tau) x) - B(b,tau) y) (tau < 0 entries are unused: zeros).
__global__ void ab2_kernel(const double* tau, int n, double a, double b, double sigma, double eta, double rho,
                           double r0, double2* c_b1, double* b2) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double t = tau[i];
    if (t < 0.0) {
        c_b1[i] = make_double2(0.0, 0.0);
        b2[i] = 0.0;
        return;
    }
    c_b1[i] = make_double2(hw2_c(a, b, sigma, eta, rho, r0, t), hw_b(a, t));
    b2[i] = hw_b(b, t);
}

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

// A, B and their parameter derivatives (models/hull_white.rs:50-80 differentiated):
//   B   = (1 - e^{-a tau}) / a,            B_a   = (tau e^{-a tau} - B) / a
//   lnA = (B - tau)(b - s^2/(2a^2)) - s^2 B^2 / (4a)
//   lnA_b = B - tau,   lnA_s = -(B - tau) s / a^2 - s B^2 / (2a)
//   lnA_a = B_a (b - s^2/(2a^2)) + (B - tau) s^2 / a^3 - s^2 (2 a B B_a - B^2) / (4 a^2)
__device__ __forceinline__ ZcbSens zcb_sens(double a, double b, double s, double tau) {
    ZcbSens z;
    const double e = exp(-(a * tau));
    const double B = hw_b(a, tau);
    z.B = B;
    z.A = hw_a(a, b, s, tau, B);
    z.B_a = (tau * e - B) / a;
    const double s2 = s * s, a2 = a * a;
    z.lnA_b = B - tau;
    z.lnA_s = -(B - tau) * s / a2 - s * B * B / (2.0 * a);
    z.lnA_a = z.B_a * (b - s2 / (2.0 * a2)) + (B - tau) * s2 / (a2 * a) - s2 * (2.0 * a * B * z.B_a - B * B) / (4.0 * a2);
    return z;
}

__global__ void zcb_sens_kernel(const double* tau, int n, double a, double b, double sigma, ZcbSens* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double t = tau[i];
    out[i] = (t < 0.0) ? ZcbSens{0.0, 0.0, 0.0, 0.0, 0.0, 0.0} : zcb_sens(a, b, sigma, t);
}

// Reverse sweep of one path. The CVA estimator is sum_k w_k sum_p max(V_s(p,k), 0); its pathwise
// derivative flows through r at every monitoring date (and at the fixing dates of periods in
// progress) back to r0 and, at every step, into a, b, sigma. Short rates are recomputed per
// segment of kAdjCheckpoint steps from the stored checkpoint, shocks from the counter-based stream.
template <int DISC, int RNG>
__device__ void adjoint_path(const AdjArgs& g, uint64_t p, double acc[5]) {
    constexpr int K = kAdjCheckpoint;
    const uint64_t n = g.n_paths;
    for (int f = 0; f < g.n_fix_slots; ++f) g.G[(uint64_t)f * n + p] = 0.0;
    double rbar = 0.0, abar = 0.0, bbar = 0.0, sbar = 0.0, vsum = 0.0;
    const int n_seg = (int)((g.n_steps + K - 1) / K);

    auto date_adjoint = [&](int k, double r) {
        double w = g.wk[k];
        if (g.date_det[k] && p != 0) w = 0.0;
        if (w == 0.0) return;
        double V = g.C0[k], gr = 0.0, ga = 0.0, gb = 0.0, gs = 0.0;
        const ZcbSens* zk = g.zs + (size_t)k * g.D;
        const double* nk = g.nu + (size_t)k * g.D;
        for (int d = g.date_dfirst[k]; d < g.D; ++d) {
            const double nu = nk[d];
            if (nu == 0.0) continue;
            const ZcbSens z = zk[d];
            const double P = z.A * bond_exp(-(z.B * r));
            const double nP = nu * P;
            V += nP;
            gr -= nP * z.B;
            ga += nP * (z.lnA_a - r * z.B_a);
            gb += nP * z.lnA_b;
            gs += nP * z.lnA_s;
        }
        // mid-period legs N P(t, T_next) / P(t_fix, T_next): every entry of the date in the value, then,
        // where the exposure is positive, their adjoints into the fixing rows (recomputed, so a date
        // may hold any number of entries)
        const int m0 = g.mid_off[k], m1 = g.mid_off[k + 1];
        for (int m = m0; m < m1; ++m) {
            const int4 md = g.mid[m];
            const ZcbSens zn = zk[md.x];
            const ZcbSens zf = g.fix_zs[md.z];
            const double rf = g.RF[(uint64_t)md.y * n + p];
            const double Pn = zn.A * bond_exp(-(zn.B * r));
            const double Pf = zf.A * bond_exp(-(zf.B * rf));
            const double term = g.mid_alpha[m] * (Pn / Pf);
            V += term;
            gr -= term * zn.B;
            ga += term * ((zn.lnA_a - r * zn.B_a) - (zf.lnA_a - rf * zf.B_a));
            gb += term * (zn.lnA_b - zf.lnA_b);
            gs += term * (zn.lnA_s - zf.lnA_s);
        }
        if (V > 0.0) {
            vsum += w * V;
            rbar += w * gr;
            abar += w * ga;
            bbar += w * gb;
            sbar += w * gs;
            for (int m = m0; m < m1; ++m) {
                const int4 md = g.mid[m];
                const ZcbSens zn = zk[md.x];
                const ZcbSens zf = g.fix_zs[md.z];
                const double rf = g.RF[(uint64_t)md.y * n + p];
                const double term = g.mid_alpha[m] * ((zn.A * bond_exp(-(zn.B * r))) / (zf.A * bond_exp(-(zf.B * rf))));
                g.G[(uint64_t)md.w * n + p] += w * (term * zf.B);
            }
        }
    };
    auto step_inputs = [&](int j) {  // adjoints arriving at grid step j
        const int k = g.date_of_step[j];
        return k;
    };

    for (int seg = n_seg - 1; seg >= 0; --seg) {
        const int j0 = seg * K;
        const int len = min(K, (int)g.n_steps - j0);
        double rl[K + 1], zl[K];
        rl[0] = g.CP[(uint64_t)seg * n + p];
        double z_next = 0.0;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            if (i < len) {
                const int j = j0 + i;
                double z;
                if (RNG == 0) {
                    z = g.Z[(uint64_t)j * n + p];
                } else if ((j & 1) == 0) {
                    philox_normal_pair(g.seed, p, (uint32_t)(j >> 1), z, z_next);
                } else {
                    z = z_next;
                }
                zl[i] = z;
                const double r = rl[i];
                if (DISC == 0) rl[i + 1] = (r + (g.a * (g.b - r)) * g.dt) + g.sigma * (z * g.sqrt_dt);
                else rl[i + 1] = (r * g.decay + g.mean_c) + g.sd * z;
            }
        }
#pragma unroll
        for (int i = K; i >= 1; --i) {
            if (i <= len) {
                const int j = j0 + i;
                const int k = step_inputs(j);
                if (k >= 0) date_adjoint(k, rl[i]);
                const int fs = g.fixslot_of_step[j];
                if (fs >= 0) rbar += g.G[(uint64_t)fs * n + p];
                // back through r_j = phi(r_{j-1}, z_{j-1})
                const double r = rl[i - 1], z = zl[i - 1];
                if (DISC == 0) {
                    abar += rbar * ((g.b - r) * g.dt);
                    bbar += rbar * (g.a * g.dt);
                    sbar += rbar * (z * g.sqrt_dt);
                    rbar *= (1.0 - g.a * g.dt);
                } else {
                    abar += rbar * (g.dt * g.decay * (g.b - r) + z * g.dsd_da);
                    bbar += rbar * (1.0 - g.decay);
                    sbar += rbar * (z * g.dsd_ds);
                    rbar *= g.decay;
                }
            }
        }
    }
    {  // grid step 0: r = r0
        const int k = step_inputs(0);
        if (k >= 0) date_adjoint(k, g.r0);
        const int fs = g.fixslot_of_step[0];
        if (fs >= 0) rbar += g.G[(uint64_t)fs * n + p];
    }
    acc[0] += abar;
    acc[1] += bbar;
    acc[2] += sbar;
    acc[3] += rbar;
    acc[4] += vsum;
}

// Compiled for four waves per SIMD (128 VGPRs): the reverse sweep waits on its checkpoint and bond
// loads, and four waves hide them better than the two its unconstrained 190 VGPRs allow, the spills
// to scratch included (5.35 -> 5.05 ms per adjoint step; three or five waves were slower).
template <int DISC, int RNG>
__global__ void __launch_bounds__(kAdjThreads) __attribute__((amdgpu_waves_per_eu(4))) adjoint_kernel(AdjArgs g) {
    __shared__ double red[5][kAdjThreads];
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (p < g.n_paths) adjoint_path<DISC, RNG>(g, p, acc);
    for (int c = 0; c < 5; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int w = kAdjThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 5; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 5; ++c) g.partial[(uint64_t)blockIdx.x * 5 + c] = red[c][0];
}

template __global__ void adjoint_kernel<0, 0>(AdjArgs);
template __global__ void adjoint_kernel<0, 1>(AdjArgs);
template __global__ void adjoint_kernel<1, 0>(AdjArgs);
template __global__ void adjoint_kernel<1, 1>(AdjArgs);

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
// Values that are the same in every lane (LDS broadcasts of the date's flow list) moved to scalar
// registers, so branches on them are scalar and addresses built from them are uniform.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int4 uni(int4 v) { return make_int4(uni(v.x), uni(v.y), uni(v.z), uni(v.w)); }
__device__ __forceinline__ double uni(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
    return __builtin_bit_cast(double, lo | (hi << 32));
}

// Global memory touched inside the tile loop is accessed through pointers in the global address
// space. Generic (flat) accesses also count against the LDS wait counter, so every LDS wait after a
// flat load or store would wait for that memory access too.
template <class T> using gptr = const __attribute__((address_space(1))) T*;
template <class T> using gmut = __attribute__((address_space(1))) T*;
template <class T> __device__ __forceinline__ gptr<T> glb(const T* p) { return (gptr<T>)p; }
template <class T> __device__ __forceinline__ gmut<T> glb(T* p) { return (gmut<T>)p; }
// The per-date descriptor tables (flows, accruals, family descriptors, ZCB coefficients, set
// entries) are read-only and indexed by wave-uniform values: read through the constant address
// space they are scalar loads into SGPRs, served by the scalar cache after a unit's first tile.
template <class T> using cptr = const __attribute__((address_space(4))) T*;
template <class T> __device__ __forceinline__ cptr<T> cst(const T* p) { return (cptr<T>)p; }
// A wave-uniform pointer made opaque to the compiler at this point: loads through it after this point
// are issued after it, so that none is hoisted out of an enclosing loop.
template <class T> __device__ __forceinline__ void opaque(cptr<T>& p) {
    uint64_t b = (uint64_t)p;
    b = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b) |
        ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32)) << 32);
    asm volatile("" : "+s"(b));
    p = (cptr<T>)b;
}

// Positive parts in one instruction. min(max(a b + c, 0), 1) compiles to one fused multiply-add with
// the clamp modifier, which bounds the result to [0, 1] (NaN to 0). For operands scaled by 2^-512 it
// is 2^-512 max(x, 0), x the unscaled a b + c: scaling by a power of two is exact and commutes with
// the rounding of the fused multiply-add and of the sums that follow, for |x| in [2^-510, 2^512) per
// unit notional (a smaller x loses bits to subnormal range, a larger one saturates). The sums are
// scaled back by 2^512.
constexpr double kPosScale = 0x1p-512, kPosUnscale = 0x1p512;
__device__ __forceinline__ double fma_pos(double a, double b, double c) {
    return __builtin_fmin(__builtin_fmax(__builtin_fma(a, b, c), 0.0), 1.0);
}

// NFAC = 2: G2++ plans; the state is (x, y) in (R, R2) and the bond factors are
// P = exp((c - B(a,tau) x) - B(b,tau) y), with (c, B(a,tau)) in AB / fix_AB and B(b,tau) in AB2 / fix_B2.
// PPL: paths per lane, so that a tile is TW = 64 PPL paths and every phase between two barriers
// carries PPL independent chains per lane (the latency of each phase is paid once per TW paths).
// TPI: trades per work item of the per-trade step (one LDS read of a class's legs per TPI trades).
// The class legs and set values are formed with explicit fused multiply-adds in a fixed order, so
// every (PPL, TPI) variant writes the same netting-set values bit for bit.
template <int NFAC, int PPL, int TPI, bool CUBE = false>
__device__ __forceinline__ void value_unit(const ValuePlan& plan, const int k, const int chunk, double* smem) {
    constexpr int TW = kPathsPerTile * PPL;
    constexpr int FU = (PPL >= 4) ? 1 : 4 / PPL;  // flows (classes) per step of the unrolled loops
    // the descriptor read through the constant address space: no store of the kernel can alias it,
    // so a field is a scalar load wherever it is used instead of a register held for the whole unit
    cptr<ValuePlan> pg = cst(&plan);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    double* sP = smem;                                     // [D][TW]
    double* sF = sP + (size_t)pg->D * TW;                    // [Fmax][TW] reciprocal fixing factors
    double2* sC = (double2*)(sF + (size_t)pg->Fmax * TW);    // [TW][C] path-major
    const int CS = pg->C;
    double* sPart = (double*)(sC + (size_t)pg->C * TW);      // [S][waves][TW] (split mode)
    double* sSeg = sPart + (pg->set_mode == 0 ? (size_t)pg->S * kValueWaves * TW : 0);  // [waves][TW]

    const int dfirst = pg->date_dfirst[k];
    const int n_live = pg->D - dfirst;
    const int fbeg = pg->fix_off[k];
    const int nfix = pg->fix_off[k + 1] - fbeg;
    const int det = pg->date_det[k];
    const uint64_t n_paths = pg->n_paths;
    const gptr<double> Rk = glb(pg->R + (uint64_t)pg->date_slot[k] * pg->r_stride);
    const cptr<double> abd = cst((const double*)(pg->AB + (size_t)k * pg->D));  // (A, B) pairs
    const cptr<int> fdc = cst((const int*)(pg->fdesc + (size_t)k * pg->NF));     // int4 per family

    // Per-trade work. The items alive at t_k are a suffix of the maturity-sorted list (every item
    // when per-path values are requested); each gets ns equal slices of a tile's paths, ns as large
    // as the block's kTradeSlots work slots allow, so that the date's work is spread over every wave.
    // Slot tid = (slice tid / n_li, live item tid % n_li): the lanes of a wave read one path row of sC
    // at consecutive items, whose class ids are consecutive (classes are numbered by maturity), so
    // their reads fall in consecutive banks or share an address.
    const int item0 = !pg->want_trade ? pg->n_items : (pg->path_mtm ? 0 : pg->item_first[k]);
    const int n_li = pg->n_items - item0;
    const int ns = (n_li > 0) ? max(1, min(TW, kTradeSlots / n_li)) : 1;
    const int n_slots = n_li * ns;
    double acc[TPI], tK[TPI];
    int icls = 0, icnt = 0, ifirst = 0;
    double isgn = 1.0;
    int islice_b, islice_e;
    {
        const int sl = (n_li > 0) ? tid / n_li : 0;
        const int li = tid - sl * n_li;
        islice_b = (sl * TW) / ns;
        islice_e = ((sl + 1) * TW) / ns;
        if (tid < n_slots) {
            const int4 item = pg->items[item0 + li];
            icls = item.x;
            ifirst = item.y;
            isgn = (double)item.w;
            icnt = item.z;
        }
#pragma unroll
        for (int q = 0; q < TPI; ++q) {
            acc[q] = 0.0;
            tK[q] = (q < icnt) ? -(isgn * pg->tr_K[ifirst + q]) * kPosScale : 0.0;  // -sign * K, scaled
        }
    }

    // the date's live flows (cash-flow date, class or -1, family) and accruals, this wave's segment
    const int lb = pg->lflow_off[k];
    const cptr<int> lf = cst((const int*)(pg->lflow + lb));  // int4 per flow
    const cptr<double> lt = cst(pg->ltau + lb);
    const int zb = pg->zcls_off[k], n_zc = pg->zcls_off[k + 1] - zb;
    const int4 segd = pg->lseg[k * kValueWaves + wave];
    const int seg_b = uni(segd.x), seg_e = uni(segd.y), carry_lo = uni(segd.z), n_carry = uni(segd.w);
    __syncthreads();  // the previous unit of a persistent block may still read this LDS
    // classes already matured at t_k: zeroed once, no live flow writes them during this unit
    for (int i = wave; i < n_zc; i += kValueWaves)
#pragma unroll
        for (int pp = 0; pp < PPL; ++pp) sC[(lane + 64 * pp) * CS + cst(pg->zcls)[zb + i]] = make_double2(0.0, 0.0);

    // the first netting set's entries of this wave in the set step (the same on every tile)
    int cb_set0 = 0, ce_set0 = 0;
    if (pg->set_mode == 0 && pg->S > 0) {
        const int beg = cst(pg->set_off)[0];
        const int end = beg + cst(pg->set_nlive)[(size_t)k * pg->S];
        const int per = (end - beg + kValueWaves - 1) / kValueWaves;
        cb_set0 = beg + wave * per;
        ce_set0 = min(end, cb_set0 + per);
    }
    const uint64_t tile0 = (uint64_t)chunk * pg->tiles_per_block;
    uint64_t tile_end = tile0 + pg->tiles_per_block;
    if (tile_end > pg->n_tiles) tile_end = pg->n_tiles;

    // The short rates of the next tile are loaded one tile ahead (every wave: r at t_k for its own
    // lanes; wave f < nfix: r at fixing f), so that no tile starts by waiting on HBM.
    const gptr<double> Rfix = (wave < nfix) ? glb(pg->R + (uint64_t)pg->fix_slot[fbeg + wave] * pg->r_stride) : nullptr;
    const double2 fab_w = (wave < nfix) ? pg->fix_AB[fbeg + wave] : make_double2(1.0, 0.0);
    const double inv_a_w = 1.0 / fab_w.x;  // once per unit
    // two factors: the second state at t_k and at fixing `wave`, and that fixing's B(b, tau)
    const gptr<double> Yk = (NFAC == 2) ? glb(pg->R2 + (uint64_t)pg->date_slot[k] * pg->r_stride) : nullptr;
    const gptr<double> Yfix =
        (NFAC == 2 && wave < nfix) ? glb(pg->R2 + (uint64_t)pg->fix_slot[fbeg + wave] * pg->r_stride) : nullptr;
    const double fb2_w = (NFAC == 2 && wave < nfix) ? pg->fix_B2[fbeg + wave] : 0.0;
    double r_next[PPL], rf_next[PPL], y_next[PPL], yf_next[PPL];
    auto prefetch = [&](uint64_t tile, bool live) {  // r (and the fixing rates) of `tile` into the _next registers
#pragma unroll
        for (int pp = 0; pp < PPL; ++pp) {
            const uint64_t gq = tile * TW + lane + 64 * pp;
            const bool vq = live && gq < n_paths;
            r_next[pp] = vq ? Rk[gq] : 0.0;
            rf_next[pp] = (Rfix && vq) ? Rfix[gq] : 0.0;
            if (NFAC == 2) {
                y_next[pp] = vq ? Yk[gq] : 0.0;
                yf_next[pp] = (Yfix && vq) ? Yfix[gq] : 0.0;
            } else {
                y_next[pp] = yf_next[pp] = 0.0;
            }
        }
    };
    prefetch(tile0, tile0 < tile_end);
    // A: one list over the waves, the fixings of periods in progress (entry f < nfix, taken by wave f;
    // sF holds 1 / P(T_fix, T_next | r_fix) = exp(B r_fix) / A, so that a family start multiplies) and
    // then the zero-coupon bonds of the live cash-flow dates, for the tile starting at path pa, from
    // the rates in the _next registers. A wave takes its bonds two at a time (two independent exponential
    // chains).
    auto stage_a = [&](uint64_t pa) {
        if (NFAC == 2) {  // sF = 1 / P(T_fix, T_next | x_f, y_f), sP = P(t_k, T_d | x, y)
            const cptr<double> b2d = cst(pg->AB2 + (size_t)k * pg->D);
            int e = wave;
            for (; e < nfix; e += kValueWaves) {
                {
                    double c = fab_w.x, b1 = fab_w.y, b2 = fb2_w;
                    uint64_t row = 0;
                    if (e != wave) {  // more fixings than waves
                        row = (uint64_t)glb(pg->fix_slot)[fbeg + e] * pg->r_stride;
                        const gptr<double> fp = glb((const double*)(pg->fix_AB + fbeg + e));
                        c = fp[0];
                        b1 = fp[1];
                        b2 = glb(pg->fix_B2)[fbeg + e];
                    }
#pragma unroll
                    for (int pp = 0; pp < PPL; ++pp) {
                        const uint64_t gpa = pa + lane + 64 * pp;
                        double xf = rf_next[pp], yf = yf_next[pp];
                        if (e != wave) {
                            xf = (gpa < n_paths) ? glb(pg->R)[row + gpa] : 0.0;
                            yf = (gpa < n_paths) ? glb(pg->R2)[row + gpa] : 0.0;
                        }
                        sF[e * TW + lane + 64 * pp] = bond_exp_v(-((c - b1 * xf) - b2 * yf));
                    }
                }
            }
            // the bonds two at a time: two independent exponential chains per wave
            for (; e + kValueWaves < nfix + n_live; e += 2 * kValueWaves) {
                const int d0 = dfirst + (e - nfix), d1 = d0 + kValueWaves;
#pragma unroll
                for (int pp = 0; pp < PPL; ++pp) {
                    const double y0 = bond_exp_v((abd[2 * d0] - abd[2 * d0 + 1] * r_next[pp]) - b2d[d0] * y_next[pp]);
                    const double y1 = bond_exp_v((abd[2 * d1] - abd[2 * d1 + 1] * r_next[pp]) - b2d[d1] * y_next[pp]);
                    sP[d0 * TW + lane + 64 * pp] = y0;
                    sP[d1 * TW + lane + 64 * pp] = y1;
                }
            }
            if (e < nfix + n_live) {
                const int d = dfirst + (e - nfix);
#pragma unroll
                for (int pp = 0; pp < PPL; ++pp)
                    sP[d * TW + lane + 64 * pp] = bond_exp_v((abd[2 * d] - abd[2 * d + 1] * r_next[pp]) - b2d[d] * y_next[pp]);
            }
            return;
        }
        int e = wave;
        for (; e < nfix; e += kValueWaves) {
            {
                double inv_a = inv_a_w, b = fab_w.y;
                uint64_t row = 0;
                if (e != wave) {  // more fixings than waves
                    row = (uint64_t)glb(pg->fix_slot)[fbeg + e] * pg->r_stride;
                    const gptr<double> fp = glb((const double*)(pg->fix_AB + fbeg + e));
                    inv_a = 1.0 / fp[0];
                    b = fp[1];
                }
#pragma unroll
                for (int pp = 0; pp < PPL; ++pp) {
                    const uint64_t gpa = pa + lane + 64 * pp;
                    double rf = rf_next[pp];
                    if (e != wave) rf = (gpa < n_paths) ? glb(pg->R)[row + gpa] : 0.0;
                    sF[e * TW + lane + 64 * pp] = inv_a * bond_exp_v(b * rf);
                }
            }
        }
        // the bonds two at a time: two independent exponential chains per wave
        for (; e + kValueWaves < nfix + n_live; e += 2 * kValueWaves) {
            const int d0 = dfirst + (e - nfix), d1 = d0 + kValueWaves;
#pragma unroll
            for (int pp = 0; pp < PPL; ++pp) {
                const double x0 = -(abd[2 * d0 + 1] * r_next[pp]), x1 = -(abd[2 * d1 + 1] * r_next[pp]);
                const double y0 = bond_exp_v(x0), y1 = bond_exp_v(x1);
                sP[d0 * TW + lane + 64 * pp] = abd[2 * d0] * y0;
                sP[d1 * TW + lane + 64 * pp] = abd[2 * d1] * y1;
            }
        }
        if (e < nfix + n_live) {
            const int d = dfirst + (e - nfix);
#pragma unroll
            for (int pp = 0; pp < PPL; ++pp) sP[d * TW + lane + 64 * pp] = abd[2 * d] * bond_exp_v(-(abd[2 * d + 1] * r_next[pp]));
        }
    };
    // a netting-set value of this unit: stored as its positive part (and the signed value when asked),
    // or for a part of a plan valued in parts, added to the earlier parts' values at the same place
    const int x_mode = pg->x_mode;
    auto store_set = [&](int s, uint64_t gp, double v) {
        const uint64_t xi = ((uint64_t)s * pg->n_dates + k) * n_paths + gp;
        if (x_mode >= 2) v = glb(pg->X)[xi] + v;
        glb(pg->X)[xi] = (x_mode == 1 || x_mode == 2) ? v : ((v > 0.0) ? v : 0.0);
        if (pg->set_mtm && (x_mode == 0 || x_mode == 3)) glb(pg->set_mtm)[((uint64_t)k * pg->S + s) * n_paths + gp] = v;
    };
    // Split mode: the netting-set values of a tile are the fixed-order sums of the waves' partials in
    // sPart, taken by waves 0 .. S-1 during the next tile's pass 2 (wave 0 starts the date's first
    // family, so it carries nothing there; sPart is rewritten only by the next tile's C) and after
    // the last tile.
    auto reduce_sets = [&](uint64_t t) {
        if (pg->set_mode == 0 && wave < pg->S) {
            const int s = wave;
#pragma unroll
            for (int pp = 0; pp < PPL; ++pp) {
                double v = 0.0;
                for (int w = 0; w < kValueWaves; ++w) v += sPart[(s * kValueWaves + w) * TW + lane + 64 * pp];
                const uint64_t gp = t * TW + lane + 64 * pp;
                if (gp < n_paths) store_set(s, gp, v);
            }
        }
    };
    if (tile0 < tile_end) stage_a(tile0 * TW);
    __syncthreads();
    // Per tile: B pass 1 | B pass 2 and the previous tile's set reduction | A of the next tile, C and D,
    // with a barrier after each. A of the next tile can run beside C and D: it writes only sP and sF,
    // which this tile last read in pass 1; it goes first, so its exponentials' FMA chains overlap the
    // waits of C's descriptor and LDS reads.
    for (uint64_t tile = tile0; tile < tile_end; ++tile) {
        // The unit holds more values across the tile loop than the scalar registers do; with the
        // plan's address made opaque here, the fields a tile uses are scalar loads (scalar-cache hits)
        // where the tile uses them instead of registers held, or spilled, across the loop.
        opaque(pg);
        const uint64_t p0 = tile * TW;
        // next tile's short rates: in flight during passes 1 and 2, used by the next tile's A
        prefetch(tile + 1, tile + 1 < tile_end);
        // B: schedule families. Every class's (floating leg, annuity) is a running sum over its
        // family's live flows; the date's flows are split into one contiguous segment per wave, each
        // summed from zero, and a second pass adds the totals of the earlier segments of the same family.
        // the annuities of the segment's leading FU flows' classes, kept for pass 2 (class -1: none)
        int lead_c[FU];
        double lead_a[FU][PPL];
#pragma unroll
        for (int u = 0; u < FU; ++u) lead_c[u] = -1;
        {
            double run[PPL], st[PPL];
#pragma unroll
            for (int pp = 0; pp < PPL; ++pp) run[pp] = st[pp] = 0.0;
            int fam = -1;
            auto family_start = [&](int f) {  // P(t, start) or the carried fixing of the period in progress
                const int mode = fdc[4 * f], dz = fdc[4 * f + 2], fw = fdc[4 * f + 3];
#pragma unroll
                for (int pp = 0; pp < PPL; ++pp) {
                    const int lp = lane + 64 * pp;
                    st[pp] = (mode == 1) ? 1.0
                                         : (mode == 2 ? sP[dz * TW + lp] : sP[dz * TW + lp] * sF[fw * TW + lp]);
                    run[pp] = 0.0;
                }
            };
            int i = seg_b;
            for (; i + FU <= seg_e; i += FU) {  // all LDS reads of FU flows first, then the chain
                const bool lead = (i == seg_b);
                int4 e[FU];
                double tau[FU], P[FU][PPL];
#pragma unroll
                for (int u = 0; u < FU; ++u) {
                    e[u] = make_int4(lf[4 * (i + u)], lf[4 * (i + u) + 1], lf[4 * (i + u) + 2], 0);
                    tau[u] = lt[i + u];
                }
#pragma unroll
                for (int u = 0; u < FU; ++u)
#pragma unroll
                    for (int pp = 0; pp < PPL; ++pp) P[u][pp] = sP[e[u].x * TW + lane + 64 * pp];
#pragma unroll
                for (int u = 0; u < FU; ++u) {
                    if (e[u].z != fam) {
                        fam = e[u].z;
                        family_start(fam);
                    }
#pragma unroll
                    for (int pp = 0; pp < PPL; ++pp) run[pp] = __builtin_fma(P[u][pp], tau[u], run[pp]);
                    if (e[u].y >= 0)
#pragma unroll
                        for (int pp = 0; pp < PPL; ++pp)
                            sC[(lane + 64 * pp) * CS + e[u].y] = make_double2(st[pp] - P[u][pp], run[pp]);
                    if (lead) {
                        lead_c[u] = e[u].y;
#pragma unroll
                        for (int pp = 0; pp < PPL; ++pp) lead_a[u][pp] = run[pp];
                    }
                }
            }
            if (i < seg_e) {  // fewer than FU flows left (the segment's leading flows when i == seg_b)
                const bool lead = (i == seg_b);
#pragma unroll
                for (int u = 0; u < FU - 1; ++u) {
                    if (i + u < seg_e) {
                        const int4 e = make_int4(lf[4 * (i + u)], lf[4 * (i + u) + 1], lf[4 * (i + u) + 2], 0);
                        if (e.z != fam) {
                            fam = e.z;
                            family_start(fam);
                        }
                        const double tau = lt[i + u];
#pragma unroll
                        for (int pp = 0; pp < PPL; ++pp) {
                            const double P = sP[e.x * TW + lane + 64 * pp];
                            run[pp] = __builtin_fma(P, tau, run[pp]);
                            if (e.y >= 0) sC[(lane + 64 * pp) * CS + e.y] = make_double2(st[pp] - P, run[pp]);
                        }
                        if (lead) {
                            lead_c[u] = e.y;
#pragma unroll
                            for (int pp = 0; pp < PPL; ++pp) lead_a[u][pp] = run[pp];
                        }
                    }
                }
            }
#pragma unroll
            for (int pp = 0; pp < PPL; ++pp) sSeg[wave * TW + lane + 64 * pp] = run[pp];
        }
        __syncthreads();
        if (tile > tile0) reduce_sets(tile - 1);
        if (n_carry > 0) {  // the totals of the earlier segments of this segment's first family
            constexpr int CB = 8 / PPL;  // totals read per batch, then added in order
            double carry[PPL];
#pragma unroll
            for (int pp = 0; pp < PPL; ++pp) carry[pp] = 0.0;
            // W totals from wave w2: the reads first, then the adds in wave order
            auto totals = [&](int w2, auto width) {
                constexpr int W = decltype(width)::value;
                double t[W][PPL];
#pragma unroll
                for (int u = 0; u < W; ++u)
#pragma unroll
                    for (int pp = 0; pp < PPL; ++pp) t[u][pp] = sSeg[(w2 + u) * TW + lane + 64 * pp];
#pragma unroll
                for (int u = 0; u < W; ++u)
#pragma unroll
                    for (int pp = 0; pp < PPL; ++pp) carry[pp] += t[u][pp];
            };
            int w2 = carry_lo;
            for (; w2 + CB <= wave; w2 += CB) totals(w2, std::integral_constant<int, CB>{});
            if (CB > 4 && w2 + 4 <= wave) {  // the rest in steps of four, two and one
                totals(w2, std::integral_constant<int, (CB > 4 ? 4 : 1)>{});
                w2 += 4;
            }
            if (CB > 2 && w2 + 2 <= wave) {
                totals(w2, std::integral_constant<int, (CB > 2 ? 2 : 1)>{});
                w2 += 2;
            }
            if (w2 < wave) totals(w2, std::integral_constant<int, 1>{});
            // the leading classes' annuities from pass 1 (the same values sC holds), then any further
            // carried flows by reading sC back
#pragma unroll
            for (int u = 0; u < FU; ++u)
                if (u < n_carry && lead_c[u] >= 0)
#pragma unroll
                    for (int pp = 0; pp < PPL; ++pp) sC[(lane + 64 * pp) * CS + lead_c[u]].y = carry[pp] + lead_a[u][pp];
            for (int i = seg_b + FU; i < seg_b + n_carry; i += FU) {  // reads of FU classes, then the writes
                int cls[FU];
                double a[FU][PPL];
#pragma unroll
                for (int u = 0; u < FU; ++u) cls[u] = (i + u < seg_b + n_carry) ? lf[4 * (i + u) + 1] : -1;
#pragma unroll
                for (int u = 0; u < FU; ++u)
#pragma unroll
                    for (int pp = 0; pp < PPL; ++pp) a[u][pp] = (cls[u] >= 0) ? sC[(lane + 64 * pp) * CS + cls[u]].y : 0.0;
#pragma unroll
                for (int u = 0; u < FU; ++u)
                    if (cls[u] >= 0)
#pragma unroll
                        for (int pp = 0; pp < PPL; ++pp) sC[(lane + 64 * pp) * CS + cls[u]].y = carry[pp] + a[u][pp];
            }
        }
        __syncthreads();
        if (tile + 1 < tile_end) stage_a(p0 + TW);
        // C: netting sets, V_s = sum over the set's live classes of alpha F - beta A, in list order
        if (pg->set_mode == 0) {
            for (int s = 0; s < pg->S; ++s) {
                // the set's classes alive at t (a prefix of its entries), split evenly over the waves
                int cb = cb_set0, ce = ce_set0;
                if (s > 0) {
                    const int beg = cst(pg->set_off)[s];
                    const int end = beg + cst(pg->set_nlive)[(size_t)k * pg->S + s];
                    const int per = (end - beg + kValueWaves - 1) / kValueWaves;
                    cb = beg + wave * per;
                    ce = min(end, cb + per);
                }
                const cptr<int> scl = cst(pg->set_cls);
                const cptr<double> sal = cst(pg->set_alpha), sbe = cst(pg->set_beta);
                double v[PPL];
#pragma unroll
                for (int pp = 0; pp < PPL; ++pp) v[pp] = 0.0;
                // W entries from entry e: the class-leg reads first, then the FMAs in entry order
                auto entries = [&](int e, auto width) {
                    constexpr int W = decltype(width)::value;
                    double2 c2[W][PPL];
#pragma unroll
                    for (int u = 0; u < W; ++u)
#pragma unroll
                        for (int pp = 0; pp < PPL; ++pp) c2[u][pp] = sC[(lane + 64 * pp) * CS + scl[e + u]];
#pragma unroll
                    for (int u = 0; u < W; ++u)
#pragma unroll
                        for (int pp = 0; pp < PPL; ++pp)
                            v[pp] = __builtin_fma(-sbe[e + u], c2[u][pp].y, __builtin_fma(sal[e + u], c2[u][pp].x, v[pp]));
                };
                int e = cb;
                for (; e + FU <= ce; e += FU) entries(e, std::integral_constant<int, FU>{});
                if (FU > 2 && e + 2 <= ce) {  // the rest in steps of two and one
                    entries(e, std::integral_constant<int, (FU > 2 ? 2 : 1)>{});
                    e += 2;
                }
                if (e < ce) entries(e, std::integral_constant<int, 1>{});
#pragma unroll
                for (int pp = 0; pp < PPL; ++pp) sPart[(s * kValueWaves + wave) * TW + lane + 64 * pp] = v[pp];
            }
        } else {
            // one wave per set, summing the same kValueWaves chunks in the same order as the split
            // path (each chunk from zero, then the chunk sums in chunk order): the same values bit for bit
            for (int s = wave; s < pg->S; s += kValueWaves) {
                const int beg = cst(pg->set_off)[s], end = beg + cst(pg->set_nlive)[(size_t)k * pg->S + s];
                const int per = (end - beg + kValueWaves - 1) / kValueWaves;
#pragma unroll
                for (int pp = 0; pp < PPL; ++pp) {
                    double v = 0.0;
                    for (int w = 0; w < kValueWaves; ++w) {
                        const int cb = beg + w * per, ce = min(end, cb + per);
                        double vw = 0.0;
                        for (int e = cb; e < ce; ++e) {
                            const double2 c2 = sC[(lane + 64 * pp) * CS + cst(pg->set_cls)[e]];
                            vw = __builtin_fma(-cst(pg->set_beta)[e], c2.y, __builtin_fma(cst(pg->set_alpha)[e], c2.x, vw));
                        }
                        v += vw;
                    }
                    const uint64_t gp = p0 + lane + 64 * pp;
                    if (gp < n_paths) store_set(s, gp, v);
                }
            }
        }
        // the legs of the classes alive at t_k, for the per-trade PFE rows (reads sC only)
        if constexpr (CUBE) {
            const gmut<double> cube =
                glb((double*)(pg->ccube + ((uint64_t)pg->cube_row[k] * pg->cube_C + pg->cube_cls0) * n_paths));
            for (int c = pg->cls_live0[k] + wave; c < pg->C; c += kValueWaves)
#pragma unroll
                for (int pp = 0; pp < PPL; ++pp) {
                    const uint64_t gp = p0 + lane + 64 * pp;
                    if (gp < n_paths) {
                        const double2 v = sC[(lane + 64 * pp) * CS + c];
                        cube[2 * ((uint64_t)c * n_paths + gp)] = v.x;
                        cube[2 * ((uint64_t)c * n_paths + gp) + 1] = v.y;
                    }
                }
        }
        // D: per-trade exposure sums (reads sC only); work unit = (item, slice of the tile's paths)
        if (pg->want_trade && icnt > 0) {
            const uint64_t left = n_paths - p0;
            const int pvalid = (left < (uint64_t)TW) ? (int)left : TW;
            const double2* col = sC + icls;
            const double sg = isgn * kPosScale;  // sign, scaled
            const int pb = islice_b;
            int pe = min(islice_e, pvalid);
            if (det) pe = (p0 == 0 && pb == 0) ? 1 : pb;  // grid index 0: path 0 only
            // paths in order, eight class-leg reads in flight, then four, two and one: a slice's
            // remainder costs no loop overhead per path
            auto paths = [&](int p, auto width) {
                constexpr int W = decltype(width)::value;
                double2 c2[W];
#pragma unroll
                for (int u = 0; u < W; ++u) c2[u] = col[(p + u) * CS];
#pragma unroll
                for (int u = 0; u < W; ++u) {
                    const double sfl = sg * c2[u].x;  // (sign * F - (sign * K) A) 2^-512, exact sign flips and scaling
#pragma unroll
                    for (int q = 0; q < TPI; ++q) acc[q] += fma_pos(tK[q], c2[u].y, sfl);
                }
            };
            int p = pb;
            for (; p + 8 <= pe; p += 8) paths(p, std::integral_constant<int, 8>{});
            if (p + 4 <= pe) {
                paths(p, std::integral_constant<int, 4>{});
                p += 4;
            }
            if (p + 2 <= pe) {
                paths(p, std::integral_constant<int, 2>{});
                p += 2;
            }
            if (p < pe) paths(p, std::integral_constant<int, 1>{});
            if (pg->path_mtm) {
                for (int q = 0; q < icnt; ++q) {
                    const int j = ifirst + q;
                    const double N = pg->tr_N[j], NK = pg->tr_NK[j];
                    const gmut<double> dst = glb(pg->path_mtm + ((uint64_t)k * pg->n_trades + pg->tr_orig[j]) * n_paths + p0);
                    for (int p = islice_b; p < min(islice_e, pvalid); ++p)
                        dst[p] = N * col[p * CS].x - NK * col[p * CS].y;
                }
            }
        }
        __syncthreads();
    }
    if (tile0 < tile_end) reduce_sets(tile_end - 1);  // the last tile's set partials
    if (pg->want_trade) {
        // the slots' sums through LDS, then one fixed-order sum over the slices of each trade
        __syncthreads();
        double* sAcc = smem;  // [kTradeSlots][TPI]
        if (tid < n_slots)
#pragma unroll
            for (int q = 0; q < TPI; ++q) sAcc[tid * TPI + q] = acc[q];
        __syncthreads();
        const gmut<double> part = glb(pg->partial + ((uint64_t)k * pg->n_chunks + chunk) * pg->n_trades);
        for (int w = tid; w < pg->n_items * TPI; w += kValueThreads) {
            const int it = w / TPI;
            const int q = w - it * TPI;
            const int4 item = pg->items[it];
            if (q >= item.z) continue;
            const int j = item.y + q;
            double e = 0.0;  // a matured class contributes exact zeros
            if (it >= item0) {
                const double* a = sAcc + (it - item0) * TPI + q;
                double sum = 0.0;
                for (int sl = 0; sl < ns; ++sl) sum += a[sl * n_li * TPI];
                e = pg->tr_absN[j] * (sum * kPosUnscale);
            }
            part[pg->tr_orig[j]] = e;
        }
    }
}

// The plan's (PPL, TPI) variant, a uniform switch: one launch can hold plans of several variants.
// Four paths per lane come with two trades per item (more would not fit the registers of a
// 1024-thread block); two-factor plans take one path per lane.
template <int NFAC>
__device__ __forceinline__ void value_dispatch(const ValuePlan& plan, int k, int chunk, double* smem) {
    switch (__builtin_amdgcn_readfirstlane(cst(&plan)->variant)) {
        case kVariant(1, 2): value_unit<NFAC, 1, 2>(plan, k, chunk, smem); break;
        case kVariant(1, 8): value_unit<NFAC, 1, 8>(plan, k, chunk, smem); break;
        case kVariant(2, 2): if constexpr (NFAC == 1) value_unit<1, 2, 2>(plan, k, chunk, smem); break;
        case kVariant(2, 4): if constexpr (NFAC == 1) value_unit<1, 2, 4>(plan, k, chunk, smem); break;
        case kVariant(2, 8): if constexpr (NFAC == 1) value_unit<1, 2, 8>(plan, k, chunk, smem); break;
        case kVariant(4, 2): if constexpr (NFAC == 1) value_unit<1, 4, 2>(plan, k, chunk, smem); break;
        default: value_unit<NFAC, 1, 4>(plan, k, chunk, smem); break;
    }
}

// The same dispatch as a call of its own, so that the persistent loop's state does not add to the
// registers of the largest variant.
__device__ __attribute__((noinline)) void value_dispatch_call(const ValuePlan* plan, int k, int chunk, double* smem) {
    value_dispatch<1>(*plan, k, chunk, smem);
}

// One block per work unit; the hardware dispatcher hands out units in list order.
__global__ void __launch_bounds__(kValueThreads) value_kernel(const ValuePlan* plans, const int4* units) {
    extern __shared__ double smem[];
    const int4 u = units[blockIdx.x];
    value_dispatch<1>(plans[u.x], u.y, u.z, smem);
}

// A plan's own launch at one path per lane, one kernel per trades-per-item variant (the plan's), so
// that no variant's registers weigh on another's.
template <int TPI>
__global__ void __launch_bounds__(kValueThreads) value_p1_kernel(const ValuePlan* plans, const int4* units) {
    extern __shared__ double smem[];
    const int4 u = units[blockIdx.x];
    value_unit<1, 1, TPI>(plans[u.x], u.y, u.z, smem);
}
template __global__ void value_p1_kernel<2>(const ValuePlan*, const int4*);
template __global__ void value_p1_kernel<4>(const ValuePlan*, const int4*);
template __global__ void value_p1_kernel<8>(const ValuePlan*, const int4*);

// The same, writing the legs of the live classes for the per-trade PFE rows (one path per lane).
template <int TPI>
__global__ void __launch_bounds__(kValueThreads) value_cube_kernel(const ValuePlan* plans, const int4* units) {
    extern __shared__ double smem[];
    const int4 u = units[blockIdx.x];
    value_unit<1, 1, TPI, true>(plans[u.x], u.y, u.z, smem);
}
template __global__ void value_cube_kernel<2>(const ValuePlan*, const int4*);
template __global__ void value_cube_kernel<4>(const ValuePlan*, const int4*);
template __global__ void value_cube_kernel<8>(const ValuePlan*, const int4*);

// The same for two-factor (G2++) plans (one path per lane), one kernel per trades-per-item variant.
template <int TPI>
__global__ void __launch_bounds__(kValueThreads) value2_kernel(const ValuePlan* plans, const int4* units) {
    extern __shared__ double smem[];
    const int4 u = units[blockIdx.x];
    value_unit<2, 1, TPI>(plans[u.x], u.y, u.z, smem);
}
template __global__ void value2_kernel<2>(const ValuePlan*, const int4*);
template __global__ void value2_kernel<4>(const ValuePlan*, const int4*);
template __global__ void value2_kernel<8>(const ValuePlan*, const int4*);

// ---------------------------------------------------------------------------------------------
// Path mode: plans without per-trade outputs whose schedule families and netting sets fit registers.
// One thread carries path_ppt(NS) paths (kPathThreads apart, so loads and stores are coalesced): no
// LDS, no barriers. Per tile the families' start values st_f (1, P(t, start), or P(t, T_next) times
// the fixing factor of the period in progress); then per live cash-flow date d in date order the bond
// P_d, every family's flow at d (annuity_f += P_d tau_f, tau_f = 0 where f has none), and the set
// entries of the classes ending at d. The loop modifies no register under a branch on the family,
// so it compiles to straight-line arithmetic. The bond, fixing-factor, leg and set expressions are
// those of value_unit; the set sums run in cash-flow-date order rather than in set-entry order.
// ---------------------------------------------------------------------------------------------
// fn(v) on netting set s's registers (s < NS, wave-uniform). Each case ends in an empty asm statement
// of its own text, so that the compiler cannot merge the cases into one access at a variable index
// (which would move the set values to scratch memory).
#define EQHIP_SET_CASE(i)                      \
    if constexpr (NS > i)                       \
        if (s == i) {                           \
            fn(V[i]);                           \
            asm volatile("; path-mode set " #i); \
        }
template <int NS, int U, class Fn>
__device__ __forceinline__ void on_set(int s, double (&V)[NS][U], Fn fn) {
    static_assert(NS <= 8, "one case per set register");
    EQHIP_SET_CASE(0)
    EQHIP_SET_CASE(1)
    EQHIP_SET_CASE(2)
    EQHIP_SET_CASE(3)
    EQHIP_SET_CASE(4)
    EQHIP_SET_CASE(5)
    EQHIP_SET_CASE(6)
    EQHIP_SET_CASE(7)
}
#undef EQHIP_SET_CASE

// NFM: families held in registers (>= the plan's NF); NS: netting-set registers (>= the plan's S).
// The arithmetic of a path does not depend on NFM and NS, only the register footprint does.
template <int NFAC, int NFM, int NS>
__device__ __forceinline__ void value_path_unit(const ValuePlan& plan, const int k, const int chunk) {
    constexpr int U = path_ppt(NS);
    constexpr int TWP = kPathThreads * U;
    const __attribute__((address_space(4))) ValuePlan& g = *cst(&plan);
    const uint64_t n_paths = g.n_paths;
    const uint64_t span = (uint64_t)g.tiles_per_block * TWP;
    const uint64_t p_beg = (uint64_t)chunk * span;
    const uint64_t p_end = (p_beg + span < n_paths) ? p_beg + span : n_paths;
    const int fbeg = g.fix_off[k];
    const int e0 = g.pe_off[k], ne = g.pe_off[k + 1] - e0;
    const cptr<PathBond> pb = cst(g.pe + e0);
    const cptr<double> abd = cst((const double*)(g.AB + (size_t)k * g.D));
    const cptr<double> b2d = cst(g.AB2 + (NFAC == 2 ? (size_t)k * g.D : 0));
    const cptr<double> pab = cst((const double*)g.pc_ab);
    const cptr<int> pset = cst(g.pc_set);
    const cptr<int> fdc = cst((const int*)(g.fdesc + (size_t)k * g.NF));  // int4 per family
    const gptr<double> Rk = glb(g.R + (uint64_t)g.date_slot[k] * g.r_stride);
    const gptr<double> Yk = glb((NFAC == 2 ? g.R2 : g.R) + (uint64_t)g.date_slot[k] * g.r_stride);
    const int x_store = g.set_mtm != nullptr;
    for (uint64_t p0 = p_beg + threadIdx.x; p0 < p_end; p0 += TWP) {
        double r[U], y[U], P[U], st[NFM][U], run[NFM][U], V[NS][U];
        bool live[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t p = p0 + (uint64_t)u * kPathThreads;
            live[u] = p < p_end;
            r[u] = live[u] ? Rk[p] : 0.0;
            y[u] = (NFAC == 2 && live[u]) ? Yk[p] : 0.0;
        }
        auto bond = [&](double A, double B, double B2, int u) {
            return (NFAC == 2) ? bond_exp((A - B * r[u]) - B2 * y[u]) : A * bond_exp(-(B * r[u]));
        };
        // family starts: a period starting at t_k (1), a forward start (P(t, start)), a period in
        // progress (P(t, T_next) / P(T_fix, T_next | r_fix))
#pragma unroll
        for (int f = 0; f < NFM; ++f) {
            const int mode = (f < g.NF) ? fdc[4 * f] : 0;
            const int dz = fdc[4 * f + 2];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                run[f][u] = 0.0;
                st[f][u] = (mode == 1) ? 1.0 : 0.0;
            }
            if (mode >= 2) {
                const double A = abd[2 * dz], B = abd[2 * dz + 1], B2 = (NFAC == 2) ? b2d[dz] : 0.0;
#pragma unroll
                for (int u = 0; u < U; ++u) st[f][u] = bond(A, B, B2, u);
            }
            if (mode == 3) {
                const int w = fbeg + fdc[4 * f + 3];
                const uint64_t row = (uint64_t)g.fix_slot[w] * g.r_stride;
                const double c = g.fix_AB[w].x, b = g.fix_AB[w].y;
                const double b2 = (NFAC == 2) ? g.fix_B2[w] : 0.0;
                const double inv_a = 1.0 / c;  // one factor: 1 / A, as value_unit's inv_a
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint64_t p = p0 + (uint64_t)u * kPathThreads;
                    double ff;
                    if (NFAC == 2) {  // 1 / P(T_fix, T_next | x_f, y_f)
                        const double xf = live[u] ? glb(g.R)[row + p] : 0.0;
                        const double yf = live[u] ? glb(g.R2)[row + p] : 0.0;
                        ff = bond_exp(-((c - b * xf) - b2 * yf));
                    } else {
                        const double rf = live[u] ? glb(g.R)[row + p] : 0.0;
                        ff = inv_a * bond_exp(b * rf);
                    }
                    st[f][u] = st[f][u] * ff;
                }
            }
        }
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int u = 0; u < U; ++u) V[s][u] = 0.0;

        for (int e = 0; e < ne; ++e) {
            const int d = pb[e].d;
            const double A = abd[2 * d], B = abd[2 * d + 1], B2 = (NFAC == 2) ? b2d[d] : 0.0;
#pragma unroll
            for (int u = 0; u < U; ++u) P[u] = bond(A, B, B2, u);
#pragma unroll
            for (int f = 0; f < NFM; ++f) {
                const double tau = pb[e].tau[f];
#pragma unroll
                for (int u = 0; u < U; ++u) run[f][u] = __builtin_fma(P[u], tau, run[f][u]);
            }
            int c = pb[e].e0;
#pragma unroll
            for (int f = 0; f < NFM; ++f) {
                const int n = pb[e].n[f];
                for (int i = 0; i < n; ++i, ++c) {  // the sets of family f's class ending at d
                    const double al = pab[2 * c], be = pab[2 * c + 1];
                    on_set<NS, U>((NS == 1) ? 0 : pset[c], V, [&](double (&v)[U]) {
#pragma unroll
                        for (int u = 0; u < U; ++u)
                            v[u] = __builtin_fma(-be, run[f][u], __builtin_fma(al, st[f][u] - P[u], v[u]));
                    });
                }
            }
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (s >= g.S) break;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t p = p0 + (uint64_t)u * kPathThreads;
                if (!live[u]) continue;
                const double v = V[s][u];
                glb(g.X)[((uint64_t)s * g.n_dates + k) * n_paths + p] = (v > 0.0) ? v : 0.0;
                if (x_store) glb(g.set_mtm)[((uint64_t)k * g.S + s) * n_paths + p] = v;
            }
        }
    }
}

// One block per work unit (plan, date, chunk of tiles of kPathThreads x path_ppt(NS) paths); the
// launch's NFM and NS cover every plan it serves, and the plans are tiled for its NS.
template <int NFAC, int NFM, int NS>
__global__ void __launch_bounds__(kPathThreads) value_path_kernel(const ValuePlan* plans, const int4* units) {
    const int4 u = units[blockIdx.x];
    value_path_unit<NFAC, NFM, NS>(plans[u.x], u.y, u.z);
}
template __global__ void value_path_kernel<1, 2, 1>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<1, 2, 2>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<1, 2, 4>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<1, 2, 8>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<1, 4, 1>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<1, 4, 2>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<1, 4, 4>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<1, 4, 8>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<2, 2, 1>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<2, 2, 2>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<2, 2, 4>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<2, 2, 8>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<2, 4, 1>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<2, 4, 2>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<2, 4, 4>(const ValuePlan*, const int4*);
template __global__ void value_path_kernel<2, 4, 8>(const ValuePlan*, const int4*);

// The valuation kernels' exp on n values, for its accuracy check (eqhip_bond_exp): bond_exp_v's value
// where bond_exp gives the same bits, NaN where the two forms differ.
__global__ void bond_exp_kernel(const double* x, double* y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double a = bond_exp(x[i]), b = bond_exp_v(x[i]);
        y[i] = (__builtin_bit_cast(uint64_t, a) == __builtin_bit_cast(uint64_t, b)) ? b : __builtin_nan("");
    }
}

// One thread per path, blockIdx.y the date. The bonds of the date's remaining flows are summed in
// schedule order (the last one is also the floating leg's end bond); the trades' stores are
// coalesced across the wave (consecutive paths).
__global__ void __launch_bounds__(kFlatThreads) value_flat_kernel(FlatArgs g) {
#pragma clang fp contract(off)
    const uint64_t p = (uint64_t)blockIdx.x * kFlatThreads + threadIdx.x;
    if (p >= g.n_paths) return;
    const int k = blockIdx.y, m = g.m, first = g.first[k];
    const size_t stride = (size_t)g.n_dates * g.n_paths;
    double* x = g.X + (size_t)k * g.n_paths + p;
    if (first >= m) {  // matured: every trade is worth zero
        for (int j = 0; j < g.n_trades; ++j) x[j * stride] = 0.0;
        return;
    }
    const double r = g.date_det[k] ? g.r0 : g.R[(uint64_t)g.date_slot[k] * g.r_stride + p];
    const double2* ab = g.AB + (size_t)k * (m + 1);
    const double p_start = ab[m].x * bond_exp(-(ab[m].y * r));
    double annuity = 0.0, p_end = 0.0;
    for (int i = first; i < m; ++i) {
        const double P = ab[i].x * bond_exp(-(ab[i].y * r));
        const double leg = P * g.acc[i];
        annuity = (i == first) ? leg : annuity + leg;
        p_end = P;
    }
    const double fl = p_start - p_end;
    for (int j = 0; j < g.n_trades; ++j) {
        const double v = g.N[j] * fl - g.NK[j] * annuity;
        x[j * stride] = v > 0.0 ? v : 0.0;
    }
}

// Persistent variant: a fixed grid of blocks pulls work units from a device counter until the
// list is exhausted (every block exits once the counter passes n_units).
__global__ void __launch_bounds__(kValueThreads) value_kernel_persistent(const ValuePlan* plans, const int4* units,
                                                                         int n_units, unsigned int* counter) {
    extern __shared__ double smem[];
    __shared__ int s_unit;
    while (true) {
        __syncthreads();  // every thread is done with the previous unit and with s_unit
        if (threadIdx.x == 0) s_unit = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int idx = s_unit;
        if (idx >= n_units) break;
        const int4 u = units[idx];
        value_dispatch_call(plans + u.x, u.y, u.z, smem);
    }
}

// Sets the persistent launch's work counter to zero (a kernel, so it can sit in a captured graph).
__global__ void reset_counter_kernel(unsigned int* counter) { *counter = 0u; }

// ---------------------------------------------------------------------------------------------
// EE + PFE95 of one (set, date) row. EE = (fixed-order sum) / n. PFE = the element of rank
// q = min(ceil(0.95 n) - 1, n - 1) of the ascending order (exposure.rs:42-49), found exactly on
// the IEEE-754 bit patterns (monotone for x >= +0.0). The first pass is fused with the EE sum and
// builds two histograms: the top 12 bits (sign and exponent: one bin per binade) and, for the
// kSelectWin binades around the order statistic of the row's first 8192 keys, the binade's top 10
// mantissa bits. When the target binade lies in that window, one pass fixes 22 bits of the key;
// otherwise 12-bit MSB radix passes narrow the prefix. Once at most kSelectCand candidates share
// the prefix they are compacted into LDS and ranked by a bitonic sort. The first pass also keeps in
// LDS the keys of a bracket around the sample's order statistic; when the prefix's bin lies inside
// it, the candidates come from there and the compaction pass over the row is skipped.
// ---------------------------------------------------------------------------------------------
namespace {

// Block-wide: the bin holding `rank` among hist[0..NB) (counts), as (bin, count before it).
template <int NB>
__device__ void locate_rank(const unsigned int* hist, unsigned int* scan, unsigned long long rank,
                            unsigned int* s_bin, unsigned long long* s_before) {
    static_assert(NB % kSelectThreads == 0, "bins per thread");
    constexpr int kWaves = kSelectThreads / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = NB / kSelectThreads;
    unsigned int local = 0;
    for (int b = 0; b < per; ++b) local += hist[tid * per + b];
    // inclusive scan of the threads' counts: within each wave by shuffles, then over the wave totals
    unsigned int incl = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned int v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) scan[wave] = incl;
    __syncthreads();
    if (tid < 64) {
        const unsigned int t = (lane < kWaves) ? scan[lane] : 0u;
        unsigned int x = t;
#pragma unroll
        for (int off = 1; off < kWaves; off <<= 1) {
            const unsigned int v = __shfl_up(x, off, 64);
            if (lane >= off) x += v;
        }
        if (lane < kWaves) scan[kWaves + lane] = x - t;  // the counts of the earlier waves
    }
    __syncthreads();
    const unsigned long long incl_all = (unsigned long long)scan[kWaves + wave] + incl;
    const unsigned long long excl = incl_all - local;
    if (rank >= excl && rank < incl_all) {
        unsigned long long cum = excl;
        int b = tid * per;
        for (;; ++b) {
            if (cum + hist[b] > rank) break;
            cum += hist[b];
        }
        *s_bin = (unsigned int)b;
        *s_before = cum;
    }
    __syncthreads();
}

}  // namespace

// Histogram increment of one key by the lanes active in the calling branch: the lanes whose bin equals
// the first active lane's add once for all of them, the others one by one. A row whose keys pile into
// one bin (ties, narrow distributions) would otherwise serialise on one LDS address.
__device__ __forceinline__ void hist_add(unsigned int* hist, unsigned int b) {
    const unsigned int b0 = __builtin_amdgcn_readfirstlane(b);
    const uint64_t same = __ballot(b == b0);
    if (b == b0) {
        const unsigned int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(same >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)same, 0u));
        if (below == 0u) atomicAdd(&hist[b0], (unsigned int)__popcll(same));
    } else {
        atomicAdd(&hist[b], 1u);
    }
}

// Appends key to buf (capacity cap) for the lanes where in holds, in lane order, with one LDS atomic
// per wave; the count keeps growing past the capacity, so an overflow shows.
__device__ __forceinline__ void wave_append(bool in, unsigned long long key, unsigned long long* buf, unsigned int cap,
                                            unsigned int* count) {
    const uint64_t m = __ballot(in);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int lead = __ffsll((long long)m) - 1;
    unsigned int base = 0u;
    if (lane == lead) base = atomicAdd(count, (unsigned int)__popcll(m));
    base = __shfl(base, lead, 64);
    if (in) {
        const unsigned int pos =
            base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (pos < cap) buf[pos] = key;
    }
}

// Visits x[i] for i = start, start + T, start + 2T, ... in that order, with four loads in flight
// per thread (two blocks of 16 waves share a CU: the select kernels are compiled for 8 waves per SIMD,
// 64 VGPRs, and a pass over a row is latency-bound with one load in flight).
constexpr int kSelectU = 4;
constexpr int kSelectS = 8;  // sample keys per thread: the row's first kSelectS * kSelectThreads keys
template <class Keys, class F>
__device__ __forceinline__ void select_pass(const Keys& keys, uint64_t n, uint64_t start, F f) {
    constexpr int U = kSelectU;
    uint64_t i = start;
    for (; i + (uint64_t)(U - 1) * kSelectThreads < n; i += (uint64_t)U * kSelectThreads) {
        unsigned long long k[U];
#pragma unroll
        for (int u = 0; u < U; ++u) k[u] = keys(i + (uint64_t)u * kSelectThreads);
#pragma unroll
        for (int u = 0; u < U; ++u) f(k[u]);
    }
    for (; i < n; i += kSelectThreads) f(keys(i));
}

// Key sources of a select row: the bits of non-negative doubles, which order as the doubles do.
struct RowKeys {  // a row of positive parts
    const unsigned long long* x;
    __device__ unsigned long long operator()(uint64_t i) const { return glb(x)[i]; }
    __device__ double out(double v) const { return v; }
};
struct TradeKeys {  // a trade's positive parts, formed from its class's legs as the per-trade EE forms them
    const double2* c;
    double s, sK, absN;
    __device__ unsigned long long operator()(uint64_t i) const {
        const gptr<double> legs = glb((const double*)c);
        const double F = legs[2 * i], A = legs[2 * i + 1];
        const double x = __builtin_fma(-sK, A, s * F);
        return (unsigned long long)__double_as_longlong(x > 0.0 ? x : 0.0);
    }
    __device__ double out(double v) const { return absN * v; }
};

// Sorts cand[0, nc) ascending in LDS (bitonic, padded with the largest key to a power of two).
__device__ void sort_candidates(unsigned long long* cand, unsigned int nc) {
    const int tid = threadIdx.x;
    unsigned int m = 1;
    while (m < nc) m <<= 1;
    for (unsigned int i = nc + tid; i < m; i += kSelectThreads) cand[i] = ~0ull;
    __syncthreads();
    for (unsigned int size = 2; size <= m; size <<= 1) {
        for (unsigned int stride = size >> 1; stride > 0; stride >>= 1) {
            for (unsigned int i = tid; i < m; i += kSelectThreads) {
                const unsigned int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const unsigned long long a = cand[i], b = cand[j];
                    if ((a > b) == up) {
                        cand[i] = b;
                        cand[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// EE (the mean, when ee is not null) and the rank-q order statistic of one row of n keys.
template <class Keys>
__device__ void select_row(const Keys& keys, const uint64_t n, const uint64_t q, double* ee, double* pfe) {
    __shared__ unsigned int hist[4096];
    __shared__ unsigned int scan[kSelectThreads];
    __shared__ double red[kSelectThreads];
    // the window histogram of pass 1 ([kSelectWin][1024]); the same LDS later holds the candidates
    __shared__ unsigned long long cand[kSelectCand];
    unsigned int* hist2 = (unsigned int*)cand;
    static_assert(kSelectWin * 1024 * sizeof(unsigned int) <= sizeof(cand), "window histogram fits");
    __shared__ unsigned long long s_prefix, s_rank, s_before, s_min, s_max;
    __shared__ unsigned int s_bin, s_count, s_ncand, s_low, s_nb;
    // the keys of the first pass's bracket (see below)
    __shared__ unsigned long long bkt[kSelectBracket];

    const int tid = threadIdx.x;

    for (int b = tid; b < 4096; b += kSelectThreads) hist[b] = 0u;
    for (int b = tid; b < kSelectWin * 1024; b += kSelectThreads) hist2[b] = 0u;
    __syncthreads();
    // pass 1: EE sum, histogram of bits 63..52 and the window histogram. The first kSelectS keys of
    // every thread (x[tid + u T], the row's first U T keys) are binned first; the binade of their
    // rank-scaled order statistic centres the window.
    constexpr uint64_t kSample = (uint64_t)kSelectS * kSelectThreads;
    unsigned long long k0[kSelectS];
#pragma unroll
    for (int u = 0; u < kSelectS; ++u) {
        const uint64_t i = (uint64_t)tid + (uint64_t)u * kSelectThreads;
        k0[u] = (i < n) ? keys(i) : 0ull;
    }
#pragma unroll
    for (int u = 0; u < kSelectS; ++u)
        if ((uint64_t)tid + (uint64_t)u * kSelectThreads < n) hist_add(hist, (unsigned int)(k0[u] >> 52));
    __syncthreads();
    const uint64_t ns = n < kSample ? n : kSample;
    locate_rank<4096>(hist, scan, (q * ns) / n, &s_bin, &s_before);
    const unsigned int est = s_bin;
    // window [wb, wb + kSelectWin) of binades, inside bins 1..4095 (bin 0 holds zeros and subnormals)
    unsigned int wb = est > (unsigned)(kSelectWin / 2) ? min(est - kSelectWin / 2, 4096u - kSelectWin) : 1u;
    if (est == 0u) {
        // the sample's order statistic is a zero: where the row's is positive (a row just past its zeros)
        // it is among the smallest positive keys, so the window starts one binade below the sample's
        // lowest positive binade
        if (tid == 0) s_low = 4096u;
        __syncthreads();
        for (int b = tid; b < 4096; b += kSelectThreads)
            if (b > 0 && hist[b] > 0u) atomicMin(&s_low, (unsigned int)b);
        __syncthreads();
        const unsigned int low = s_low;
        if (low < 4096u) wb = min(low > 1u ? low - 1u : 1u, 4096u - kSelectWin);
    }
    // The bracket: the keys between the sample's order statistics kr ranks either side of its scaled
    // rank (3.5 standard deviations of a sample quantile, fewer when the row is long enough that the
    // bracket would overfill), widened to their histogram bins, are kept in LDS during the pass. When
    // the bin the histograms locate the rank-q key in lies inside the bracket and the bracket kept every
    // key, the candidates come from LDS and the compaction pass over the row is skipped.
    const uint64_t rc = (q * ns) / n;
    const double pq = ((double)q + 0.5) / (double)n;
    const double sds = sqrt((double)ns * pq * (1.0 - pq));
    const double kr = fmin(3.5 * sds + 1.0, (0.6 * kSelectBracket * (double)ns) / (2.0 * (double)n));
    const bool use_bkt = kr >= sds + 1.0;
    const uint64_t r_lo = (use_bkt && rc > (uint64_t)kr) ? rc - (uint64_t)kr : 0;
    const uint64_t r_hi = use_bkt ? min(ns - 1, rc + (uint64_t)kr) : 0;
    unsigned int blo_bin = 0u, bhi_bin = 0u;
    unsigned long long blo_before = 0ull, bhi_before = 0ull;
    if (use_bkt) {
        __syncthreads();  // every thread has read est
        locate_rank<4096>(hist, scan, r_lo, &s_bin, &s_before);
        blo_bin = s_bin;
        blo_before = s_before;
        __syncthreads();
        locate_rank<4096>(hist, scan, r_hi, &s_bin, &s_before);
        bhi_bin = s_bin;
        bhi_before = s_before;
    }
    // A key inside the window is counted in the window histogram only (its binade's bin receives the
    // window's per-binade totals after the pass), so every key costs one LDS atomic; the sample keys
    // inside the window move there from the binade histogram.
    double sum = 0.0;
    unsigned int c0 = 0u;
#pragma unroll
    for (int u = 0; u < kSelectS; ++u) {
        if ((uint64_t)tid + (uint64_t)u * kSelectThreads < n) {
            sum += __longlong_as_double(k0[u]);
            const unsigned int b = (unsigned int)(k0[u] >> 52);
            const unsigned int w = b - wb;  // wraps for b < wb
            if (w < (unsigned)kSelectWin) {
                atomicSub(&hist[b], 1u);
                hist_add(hist2, (w << 10) | (unsigned int)((k0[u] >> 42) & 1023u));
            }
        }
    }
    // the bracket's bounds: the first key of r_lo's bin and the end of r_hi's, a window binade's bin
    // narrowed by the sample's window histogram
    unsigned long long b_lo = ~0ull, b_hi = 0ull;  // empty unless use_bkt
    if (use_bkt) {
        __syncthreads();  // the sample's window histogram is complete
        auto bin_of = [&](unsigned int b, unsigned long long r, bool end) {
            const unsigned int w = b - wb;
            if (w < (unsigned)kSelectWin) {
                locate_rank<1024>(hist2 + (w << 10), scan, r, &s_bin, &s_before);
                const unsigned long long st = ((unsigned long long)b << 52) | ((unsigned long long)s_bin << 42);
                __syncthreads();  // every thread has read s_bin
                return end ? st + (1ull << 42) : st;
            }
            return end ? ((unsigned long long)(b + 1u) << 52) : ((unsigned long long)b << 52);
        };
        b_lo = bin_of(blo_bin, r_lo - blo_before, false);
        b_hi = bin_of(bhi_bin, r_hi - bhi_before, true);
        if (b_lo == 0ull) b_lo = 1ull;  // zeros are counted, never kept
        if (tid == 0) s_nb = 0u;
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kSelectS; ++u) {
            const bool valid = (uint64_t)tid + (uint64_t)u * kSelectThreads < n;
            wave_append(valid && k0[u] >= b_lo && k0[u] < b_hi, k0[u], bkt, kSelectBracket, &s_nb);
        }
    }
    // keys in bin 0 (zero exposures, the common case) are counted per thread, not by LDS atomics
    // that would all hit one address
    select_pass(keys, n, (uint64_t)tid + kSample, [&](unsigned long long key) {
        sum += __longlong_as_double(key);
        const unsigned int b = (unsigned int)(key >> 52);
        const unsigned int w = b - wb;
        if (b == 0u) ++c0;
        else if (w < (unsigned)kSelectWin) hist_add(hist2, (w << 10) | (unsigned int)((key >> 42) & 1023u));
        else hist_add(hist, b);
        if (use_bkt) wave_append(key >= b_lo && key < b_hi, key, bkt, kSelectBracket, &s_nb);
    });
    if (c0) atomicAdd(&hist[0], c0);
    red[tid] = sum;
    __syncthreads();
    {  // per-binade totals of the window into the binade histogram (a thread's bins share one binade)
        constexpr int per = kSelectWin * 1024 / kSelectThreads;
        static_assert(1024 % per == 0, "a thread's window bins lie in one binade");
        unsigned int c = 0u;
#pragma unroll
        for (int j = 0; j < per; ++j) c += hist2[tid * per + j];
        if (c) atomicAdd(&hist[wb + (unsigned)(tid * per) / 1024u], c);
    }
    if (ee) {  // the EE sum's fixed-order tree (a row without EE skips its barriers)
        for (int w = kSelectThreads / 2; w > 0; w >>= 1) {
            if (tid < w) red[tid] = red[tid] + red[tid + w];
            __syncthreads();
        }
        if (tid == 0) *ee = keys.out(red[0] / (double)n);
    } else {
        __syncthreads();  // the window totals are in the binade histogram
    }
    locate_rank<4096>(hist, scan, q, &s_bin, &s_before);
    if (tid == 0) {
        s_prefix = (unsigned long long)s_bin << 52;
        s_rank = q - s_before;
        s_count = hist[s_bin];
    }
    __syncthreads();

    int hi = 52;  // bits >= hi are fixed
    {  // the target binade inside the window: its mantissa histogram fixes ten more bits
        const unsigned int e = s_bin;
        const unsigned int w = e - wb;
        if (s_count > (unsigned int)kSelectCand && e != 0u && w < (unsigned)kSelectWin) {
            const unsigned long long rank = s_rank;
            const unsigned int* h = hist2 + (w << 10);
            locate_rank<1024>(h, scan, rank, &s_bin, &s_before);
            if (tid == 0) {
                s_prefix |= (unsigned long long)s_bin << 42;
                s_rank = rank - s_before;
                s_count = h[s_bin];
            }
            __syncthreads();
            hi = 42;
        }
    }
    while (true) {
        const unsigned long long prefix = s_prefix;
        const unsigned int count = s_count;
        if (count <= (unsigned int)kSelectCand || hi == 0) break;
        const int wd = hi >= 12 ? 12 : hi;
        const int sh = hi - wd;
        for (int b = tid; b < 4096; b += kSelectThreads) hist[b] = 0u;
        if (tid == 0) {
            s_min = ~0ull;
            s_max = 0ull;
        }
        __syncthreads();
        // the smallest and largest key under the prefix as well: when they are equal, every candidate
        // is the same value (ties, e.g. rows of zero exposure) and the remaining passes are skipped
        unsigned int c0 = 0u;
        unsigned long long kmin = ~0ull, kmax = 0ull;
        select_pass(keys, n, tid, [&](unsigned long long key) {
            if ((key >> hi) == (prefix >> hi)) {
                const unsigned int b = (unsigned int)((key >> sh) & ((1u << wd) - 1u));
                if (b == 0u) ++c0;
                else hist_add(hist, b);
                kmin = key < kmin ? key : kmin;
                kmax = key > kmax ? key : kmax;
            }
        });
        if (c0) atomicAdd(&hist[0], c0);
        if (kmin <= kmax) {
            atomicMin(&s_min, kmin);
            atomicMax(&s_max, kmax);
        }
        __syncthreads();
        if (s_min == s_max) {  // uniform across the block
            if (tid == 0) *pfe = keys.out(__longlong_as_double(s_min));
            return;
        }
        locate_rank<4096>(hist, scan, s_rank, &s_bin, &s_before);
        if (tid == 0) {
            s_prefix = prefix | ((unsigned long long)s_bin << sh);
            s_rank = s_rank - s_before;
            s_count = hist[s_bin];
        }
        __syncthreads();
        hi = sh;
    }
    if (hi == 0) {  // every bit fixed: the prefix is the key
        if (tid == 0) *pfe = keys.out(__longlong_as_double(s_prefix));
        return;
    }
    // compact the candidates sharing the prefix and rank them exactly: from the bracket when the
    // prefix's bin lies inside it and the bracket kept every key, else by a pass over the row
    if (tid == 0) s_ncand = 0u;
    __syncthreads();
    const unsigned long long prefix = s_prefix;
    const unsigned long long pend = prefix + (1ull << hi);
    if (use_bkt && s_nb <= (unsigned)kSelectBracket && prefix >= b_lo && pend <= b_hi) {
        const unsigned int nb = s_nb;
        for (unsigned int j = tid; j < nb; j += kSelectThreads) {
            const unsigned long long key = bkt[j];
            if ((key >> hi) == (prefix >> hi)) cand[atomicAdd(&s_ncand, 1u)] = key;
        }
    } else {
        select_pass(keys, n, tid, [&](unsigned long long key) {
            if ((key >> hi) == (prefix >> hi)) cand[atomicAdd(&s_ncand, 1u)] = key;
        });
    }
    __syncthreads();
    sort_candidates(cand, s_ncand);
    if (tid == 0) *pfe = keys.out(__longlong_as_double(cand[s_rank]));
}

__global__ void __launch_bounds__(kSelectThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) select_kernel(const SelectRow* rows) {
    const SelectRow row = rows[blockIdx.x];
    select_row(RowKeys{(const unsigned long long*)row.x}, row.n, row.q, row.ee, row.pfe);
}

// Per-trade PFE95 rows of one date group (each row's class legs written by the valuation).
__global__ void __launch_bounds__(kSelectThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) trade_select_kernel(const TradeSelectRow* rows) {
    const TradeSelectRow row = rows[blockIdx.x];
    select_row(TradeKeys{row.c, row.s, row.sK, row.absN}, row.n, row.q, nullptr, row.pfe);
}

// Per-trade EE: partial sums over chunks in chunk order, divided by the effective path count.
// partial is [date][part][trade], so the lanes of a wave (consecutive trades) read contiguously;
// trade_ee is [trade][date].
__global__ void trade_ee_kernel(const double* partial, int n_trades, int n_dates, int n_chunks,
                                const uint64_t* n_eff, double* trade_ee) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_trades * n_dates) return;
    const int k = i / n_trades;
    const int j = i - k * n_trades;
    const double* p = partial + (uint64_t)k * n_chunks * n_trades + j;
    double s = 0.0;
    for (int c = 0; c < n_chunks; ++c) s += p[(uint64_t)c * n_trades];
    trade_ee[(uint64_t)j * n_dates + k] = s / (double)n_eff[k];
}

// CVA = sum_k (1-R) EE_k (S_{k-1} - S_k) P(0,t_k | r0), S = exp(-h t), exposure.rs:242-275.
// The per-date factors of that sum, shared by every profile: (S_{k-1} - S_k, P(0,t_k | r0)).
__global__ void cva_weights_kernel(const double* times, int n_dates, double a, double b, double sigma,
                                   double r0, double hazard, double2* w) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_dates) return;
    const double t = times[k];
    const double survival = exp(-hazard * t);
    const double prev_survival = (k == 0) ? 1.0 : exp(-hazard * times[k - 1]);
    const double default_prob = prev_survival - survival;
    const double tau = t - 0.0;
    const double bf = hw_b(a, tau);
    const double disc = hw_a(a, b, sigma, tau, bf) * exp(-(bf * r0));
    w[k] = make_double2(default_prob, disc);
}

// ---------------------------------------------------------------------------------------------
// XVA reducer rows (quant-engine/src/xva.rs:48-64, 150-183). One block per time row; the values are
// read once (eight loads in flight per thread), the positive parts of the adjusted Q values are
// written for the row's quantile, and the negative parts are summed per thread in index order, then
// across threads by a fixed tree.
// ---------------------------------------------------------------------------------------------
namespace {
__device__ __forceinline__ double xva_adjust(double v, const XvaRowArgs& g) {
    if (!g.csa_vm) return v;
    const double excess = fabs(v) - g.threshold;
    const double collateral = (excess > 0.0) ? excess : 0.0;
    if (collateral < g.mta) return v;
    return (v >= 0.0) ? (v - collateral) - g.ia : (v + collateral) + g.ia;
}
__device__ __forceinline__ double block_sum(double x, double* red) {
    const int tid = threadIdx.x;
    red[tid] = x;
    __syncthreads();
    for (int w = kSelectThreads / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] = red[tid] + red[tid + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}
}  // namespace

__global__ void __launch_bounds__(kSelectThreads) xva_rows_kernel(XvaRowArgs g) {
#pragma clang fp contract(off)
    __shared__ double red[kSelectThreads];
    const int tid = threadIdx.x;
    const uint64_t k = blockIdx.x;
    const uint64_t n = g.n;
    const gptr<double> q = glb(g.q + k * g.q_stride);
    const gmut<double> pos = glb(g.pos + k * n);
    constexpr int U = 8;
    double neg_q = 0.0;
    uint64_t i = tid;
    for (; i + (uint64_t)(U - 1) * kSelectThreads < n; i += (uint64_t)U * kSelectThreads) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = q[i + (uint64_t)u * kSelectThreads];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double a = xva_adjust(v[u], g);
            pos[i + (uint64_t)u * kSelectThreads] = (a > 0.0) ? a : 0.0;
            neg_q += (a < 0.0) ? -a : 0.0;
        }
    }
    for (; i < n; i += kSelectThreads) {
        const double a = xva_adjust(q[i], g);
        pos[i] = (a > 0.0) ? a : 0.0;
        neg_q += (a < 0.0) ? -a : 0.0;
    }
    const double tq = block_sum(neg_q, red);
    if (tid == 0) g.ene_q[k] = tq;
    if (g.p) {
        const gptr<double> pr = glb(g.p + k * g.p_stride);
        double neg_p = 0.0;
        uint64_t j = tid;
        for (; j + (uint64_t)(U - 1) * kSelectThreads < n; j += (uint64_t)U * kSelectThreads) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = pr[j + (uint64_t)u * kSelectThreads];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const double a = xva_adjust(v[u], g);
                neg_p += (a < 0.0) ? -a : 0.0;
            }
        }
        for (; j < n; j += kSelectThreads) {
            const double a = xva_adjust(pr[j], g);
            neg_p += (a < 0.0) ? -a : 0.0;
        }
        const double tp = block_sum(neg_p, red);
        if (tid == 0) g.ene_p[k] = tp;
    }
}

// G2++ CVA factors: the discount is the model bond at the initial state (0, 0),
// exposure.rs:470-479 (unilateral_cva_2f), i.e. exp(c(t)).
__global__ void cva_weights2_kernel(const double* times, int n_dates, double a, double b, double sigma, double eta,
                                    double rho, double r0, double hazard, double2* w) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_dates) return;
    const double t = times[k];
    const double survival = exp(-hazard * t);
    const double prev_survival = (k == 0) ? 1.0 : exp(-hazard * times[k - 1]);
    const double default_prob = prev_survival - survival;
    const double disc = exp(hw2_c(a, b, sigma, eta, rho, r0, t - 0.0));
    w[k] = make_double2(default_prob, disc);
}

// One profile per thread, the dates in order, each term as (1-R) * EE_k * dS_k * P_k.
__global__ void __launch_bounds__(256) cva_kernel(const double* ee, int n_profiles, int n_dates, const double2* w,
                                                  double recovery, double* cva) {
#pragma clang fp contract(off)
    // one wave per profile: the lanes load and form 64 terms at a time, then every lane adds them in
    // date order (the reference's loop order, so the sum rounds exactly as one sequential loop)
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (j >= n_profiles) return;  // uniform across the wave
    const gptr<double> e = glb(ee + (uint64_t)j * n_dates);
    const gptr<double> wd = glb((const double*)w);
    const double q = 1.0 - recovery;
    double acc = 0.0;
    for (int base = 0; base < n_dates; base += 64) {
        const int k = base + lane;
        const double term = (k < n_dates) ? q * e[k] * wd[2 * k] * wd[2 * k + 1] : 0.0;
        const int cnt = min(64, n_dates - base);
        for (int i = 0; i < cnt; ++i) acc += __shfl(term, i);
    }
    if (lane == 0) cva[j] = acc;
}

// ---------------------------------------------------------------------------------------------
// Market-curve PV / DV01 (measure.cpp:160-222, market.cpp:26-41). Thread (trade, variant):
// variant 0 = PV, 1 = parallel bump, 2+i = pillar i bumped.
// ---------------------------------------------------------------------------------------------
// The first pillar at or after t, for pillars[0] < t < pillars[np - 1]: the index market.cpp:26-41's
// linear scan stops at, found by bisection.
__device__ __forceinline__ int pillar_at_or_after(const double* pillars, int np, double t) {
    int lo = 1, hi = np - 1;  // pillars[lo - 1] < t <= pillars[hi]
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (pillars[mid] < t) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

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
        const int i = pillar_at_or_after(pillars, np, t);
        double t0 = pillars[i - 1], t1 = pillars[i];
        double z0 = zr(i - 1), z1 = zr(i);
        double w = (t - t0) / (t1 - t0);
        z = z0 + w * (z1 - z0);
    }
    return exp(-z * t);
}

// measure.cpp:146-184 at valuation time v: each discount factor over DF(v) (DF(0) = 1 exactly, so v = 0
// is the t = 0 form bit for bit); the caller has refused v > start.
__device__ double npv_from_curve(const double* pillars, const double* zeros, int np,
                                 int bump_pillar, double bump, double notional, double fixed_rate,
                                 double start, const double* pay, const double* acc, int m, double valuation_time) {
#pragma clang fp contract(off)
    double dv = curve_df(pillars, zeros, np, bump_pillar, bump, valuation_time);
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
    // the curve in LDS: every discount factor bisects it
    __shared__ double s_pillars[kPvLdsPillars], s_zeros[kPvLdsPillars];
    const bool in_lds = g.np <= kPvLdsPillars;
    if (in_lds)
        for (int p = threadIdx.x; p < g.np; p += blockDim.x) {
            s_pillars[p] = g.pillars[p];
            s_zeros[p] = g.zeros[p];
        }
    __syncthreads();
    const double* pillars = in_lds ? s_pillars : g.pillars;
    const double* zeros = in_lds ? s_zeros : g.zeros;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int nv = 2 + g.np;
    if (i >= g.n_trades * nv) return;
    int j = i / nv, v = i - j * nv;
    int f0 = g.flow_off[j], m = g.flow_off[j + 1] - f0;
    const double* pay = g.pay + f0;
    const double* acc = g.acc + f0;
    // PV and the parallel DV01 at the valuation time; the bucketed ladder at t = 0 (measure.cpp:211-222)
    const double vt = (v < 2) ? g.valuation_time : 0.0;
    double base = npv_from_curve(pillars, zeros, g.np, -1, 0.0, g.notional[j], g.fixed[j],
                                 g.start[j], pay, acc, m, vt);
    if (v == 0) {
        g.pv[j] = base;
        return;
    }
    int bp = (v == 1) ? -2 : v - 2;
    double bumped = npv_from_curve(pillars, zeros, g.np, bp, g.bump, g.notional[j], g.fixed[j],
                                   g.start[j], pay, acc, m, vt);
    if (v == 1) g.dv01[j] = bumped - base;
    else if (g.bucketed) g.bucketed[(uint64_t)j * g.np + (v - 2)] = bumped - base;
}

// Interpolation weights of the zero rate at t on the pillars (the adjoint of curve_df's
// interpolation): weight w1 on pillar i1 and 1 - w1 on pillar i0 (i0 == i1 outside the pillars).
__device__ void curve_weights(const double* pillars, int np, double t, int& i0, int& i1, double& w1) {
#pragma clang fp contract(off)
    if (t <= pillars[0]) {
        i0 = i1 = 0;
        w1 = 1.0;
    } else if (t >= pillars[np - 1]) {
        i0 = i1 = np - 1;
        w1 = 1.0;
    } else {
        const int i = pillar_at_or_after(pillars, np, t);
        i0 = i - 1;
        i1 = i;
        w1 = (t - pillars[i0]) / (pillars[i1] - pillars[i0]);
    }
}

// Analytic sensitivities of one trade's PV (one thread per trade):
//   dPV/dz_p = sum_t c_t (-t DF(t)) dz(t)/dz_p over the cash-flow dates t with PV coefficients c_t
//   (N at the start, -N at the last payment, -N K tau_i at every payment; DF(0) = 1);
//   model NPV at t = 0 and its derivatives in a, b, sigma, r0 through P(0,t) = A(t) exp(-B(t) r0).
__global__ void pv_sens_kernel(PvArgs g, double a, double b, double sigma, double r0, int want_model,
                               double* pillar_delta, double* model_npv, double* model_greeks) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.n_trades) return;
    const int f0 = g.flow_off[j], m = g.flow_off[j + 1] - f0;
    const double N = g.notional[j], K = g.fixed[j];
    double* dz = pillar_delta + (size_t)j * g.np;
    for (int p = 0; p < g.np; ++p) dz[p] = 0.0;
    double npv = 0.0, ga = 0.0, gb = 0.0, gs = 0.0, gr = 0.0;
    auto add_flow = [&](double t, double c) {
        if (t > 0.0) {
            int i0, i1;
            double w1;
            curve_weights(g.pillars, g.np, t, i0, i1, w1);
            const double d = c * (-t * curve_df(g.pillars, g.zeros, g.np, -1, 0.0, t));
            if (i1 != i0) {
                dz[i0] += d * (1.0 - w1);
                dz[i1] += d * w1;
            } else {
                dz[i0] += d;  // flat extrapolation: the whole weight on the end pillar
            }
        }
        if (want_model) {
            if (t <= 0.0) {
                npv += c;  // P(0,0) = 1, no parameter dependence
                return;
            }
            const ZcbSens z = zcb_sens(a, b, sigma, t);
            const double P = z.A * exp(-(z.B * r0));
            const double cP = c * P;
            npv += cP;
            ga += cP * (z.lnA_a - r0 * z.B_a);
            gb += cP * z.lnA_b;
            gs += cP * z.lnA_s;
            gr -= cP * z.B;
        }
    };
    add_flow(g.start[j], N);
    add_flow(g.pay[f0 + m - 1], -N);
    for (int i = 0; i < m; ++i) add_flow(g.pay[f0 + i], -(N * K * g.acc[f0 + i]));
    if (want_model) {
        model_npv[j] = npv;
        model_greeks[(size_t)j * 4 + 0] = ga;
        model_greeks[(size_t)j * 4 + 1] = gb;
        model_greeks[(size_t)j * 4 + 2] = gs;
        model_greeks[(size_t)j * 4 + 3] = gr;
    }
}

}  // namespace eqhip
