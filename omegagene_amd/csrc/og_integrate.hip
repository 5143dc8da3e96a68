// og_integrate.hip -- FP64 integration on the device: leapfrog (presto),
// position Langevin, SHAKE pair/trio, analytic SETTLE, COM-motion removal,
// grid-coordinate refresh, PBC wrap and kinetic-energy reduction.
//
// Compiled with -ffp-contract=off: each expression rounds like the reference
// host code it follows, so these kernels match the CPU oracle bit for bit on
// identical inputs (reductions aside, which use a fixed tree order).
#include "og_terms.h"

#define OG_FORCE_VEL 4.184e-4                          // CelesteObject.cpp:37
#define OG_GAS_CONST 8.31451                           // CelesteObject.cpp:38
#define OG_KINETIC_COEFF ((1e7 / (4.184 * 1e3)) * 0.5) // CelesteObject.cpp:40

// ------------------------------------------------------------- Philox --
// Philox4x32-10 counter-based generator (J. K. Salmon, M. A. Moraes,
// R. O. Dror, D. E. Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11): ten rounds of two 32x32->64 multiplies with Weyl-sequence key bumps.
__device__ __forceinline__ uint4 og_philox(uint4 c, uint2 k) {
    for (int r = 0; r < 10; r++) {
        const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}
// Box-Muller in FP64 on two 32-bit uniforms in (0,1)
__device__ __forceinline__ double2 og_gauss2(uint32_t a, uint32_t b) {
    const double u1 = ((double)a + 0.5) * 2.3283064365386963e-10;
    const double u2 = ((double)b + 0.5) * 2.3283064365386963e-10;
    const double r  = sqrt(-2.0 * log(u1));
    double       sn, cs;
    sincos(6.283185307179586 * u2, &sn, &cs);
    return make_double2(r * cs, r * sn);
}

// --------------------------------------------------- leapfrog velocities --
// SubBox::cpy_vel_prev + update_velocities (SubBox.cpp:1427-1454)
__global__ void __launch_bounds__(OG_THREADS) k_leapfrog_vel(const og_dev_sys *__restrict__ S,
                                                             const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const int         i  = bm.y * OG_THREADS + threadIdx.x;
    if (i >= s.n_atoms || s.integrator != 0) return;
    const double minv = (double)s.mass_inv[i];
    for (int d = 0; d < 3; d++) {
        const double v    = s.vel_next[3 * i + d];
        s.vel[3 * i + d]  = v;
        s.vel_next[3 * i + d] = v + (((s.dt * -OG_FORCE_VEL) * s.work[3 * i + d]) * minv);
    }
}

// ------------------------------------------------------- COM cancel --
// COMMotion::cancel_translation (COMMotion.cpp:30-55) with disjoint groups:
// one block per system sums the momentum of each group in a fixed tree order.
__global__ void __launch_bounds__(OG_THREADS) k_com_sum(const og_dev_sys *__restrict__ S, const int2 *__restrict__ map) {
    const og_dev_sys &s = S[map[blockIdx.x].x];
    if (og_frozen(s)) return;
    __shared__ double sh[3 * OG_THREADS];
    if (!s.com_cancel || s.integrator != 0) return;
    for (int g = 0; g < s.n_com_groups; g++) {
        double m[3] = {0.0, 0.0, 0.0};
        for (int k = s.com_off[g] + threadIdx.x; k < s.com_off[g + 1]; k += blockDim.x) {
            const int    a  = s.com_atoms[k];
            const double ms = (double)s.mass[a];
            for (int d = 0; d < 3; d++) m[d] += s.vel_next[3 * a + d] * ms;
        }
        for (int d = 0; d < 3; d++) sh[3 * threadIdx.x + d] = m[d];
        __syncthreads();
        for (int off = blockDim.x >> 1; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off)
                for (int d = 0; d < 3; d++) sh[3 * threadIdx.x + d] += sh[3 * (threadIdx.x + off) + d];
            __syncthreads();
        }
        if (threadIdx.x == 0)
            for (int d = 0; d < 3; d++) s.com_mom[3 * g + d] = sh[d] * (double)s.com_minv[g];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(OG_THREADS) k_com_apply(const og_dev_sys *__restrict__ S,
                                                          const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const int         t  = bm.y * OG_THREADS + threadIdx.x;
    if (!s.com_cancel || s.integrator != 0 || s.n_com_groups == 0 || t >= s.com_off[s.n_com_groups]) return;
    int g = 0;
    while (t >= s.com_off[g + 1]) g++;
    const int a = s.com_atoms[t];
    for (int d = 0; d < 3; d++) s.vel_next[3 * a + d] -= s.com_mom[3 * g + d];
}

// ------------------------------------------------- scaling thermostat --
// ThermostatScaling::apply_thermostat (Thermostat.cpp:57-84) after the
// leapfrog velocity update and COM removal, for unconstrained leapfrog runs:
// the kinetic energy of the on-step velocities vel - dt/2 w FORCE_VEL m^-1
// sets one factor s = sqrt(1 + dt/tau (T0/T - 1)) and
// vel_next = (2 s - 1) vel + s dt (-FORCE_VEL) w m^-1.  T0 follows the
// heating ramp of SubBox::update_thermostat (SubBox.cpp:1745-1757).  Block
// partials in a fixed tree, then every block sums them in the same order.
__global__ void __launch_bounds__(OG_THREADS) k_thermo_sum(const og_dev_sys *__restrict__ S,
                                                           const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s) || s.thermostat != OG_TH_SCALING || s.integrator != 0) return;
    const int i  = bm.y * OG_THREADS + threadIdx.x;
    double    ke = 0.0;
    if (i < s.n_atoms) {
        double vel_norm = 0.0;
        for (int d = 0; d < 3; d++) {
            const double vel_tmp =
                s.vel[3 * i + d] - 0.5 * (s.dt * s.work[3 * i + d] * OG_FORCE_VEL * (double)s.mass_inv[i]);
            vel_norm += vel_tmp * vel_tmp;
        }
        ke = (double)s.mass[i] * vel_norm;
    }
    __shared__ double sh[OG_THREADS];
    sh[threadIdx.x] = ke;
    __syncthreads();
    for (int off = OG_THREADS / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) s.th_part[bm.y] = sh[0];
}

__global__ void __launch_bounds__(OG_THREADS) k_thermo_apply(const og_dev_sys *__restrict__ S,
                                                             const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s) || s.thermostat != OG_TH_SCALING || s.integrator != 0) return;
    __shared__ double sh[OG_THREADS];
    double            t = 0.0;
    for (int k = threadIdx.x; k < s.n_atom_blocks; k += OG_THREADS) t += s.th_part[k];
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int off = OG_THREADS / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    const long long step = s.cnt->step;
    double          T0   = s.temperature;
    if (step < s.heating_steps)
        T0 = s.th_temp_init + ((s.temperature - s.th_temp_init) / (double)s.heating_steps) * (double)step;
    const double dt_temperature = sh[0] * s.th_temp_coeff;
    const double scale          = sqrt(1 + s.dt * s.th_tau_inv * (T0 / dt_temperature - 1));
    const int    i              = bm.y * OG_THREADS + threadIdx.x;
    if (i >= s.n_atoms) return;
    for (int d = 0; d < 3; d++) {
        const double vel_diff = s.dt * -OG_FORCE_VEL * s.work[3 * i + d] * (double)s.mass_inv[i];
        s.vel_next[3 * i + d] = (2.0 * scale - 1.0) * s.vel[3 * i + d] + scale * vel_diff;
    }
}

// ---------------------------------------------------------- positions --
// SubBox::cpy_crd_prev + update_coordinates(dt, crd, vel_next) (SubBox.cpp:1409-1412,1562-1580)
__global__ void __launch_bounds__(OG_THREADS) k_positions(const og_dev_sys *__restrict__ S,
                                                          const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const int         i  = bm.y * OG_THREADS + threadIdx.x;
    if (i >= s.n_atoms || s.integrator != 0) return;
    for (int d = 0; d < 3; d++) {
        const double x        = s.crd[3 * i + d];
        s.crd_prev[3 * i + d] = x;
        const double diff     = s.vel_next[3 * i + d] * s.dt;
        s.crd[3 * i + d]      = x + diff;
    }
}

// ----------------------------------------------------- Langevin positions --
// DynamicsModeLangevin::calc_in_each_step (DynamicsMode.cpp:798-899):
// crd_prev2 <- crd_prev <- crd; first step x1 = x0 + dt v0; later steps
// SubBox::update_coordinates_langevin (SubBox.cpp:2077-2114) with N(0,1)
// noise from Philox keyed by (seed, step, atom).
// Core on values: x = current position, xp = previous position, vn = initial
// velocity (first step only), wk = gradient; the new position goes to xn.
__device__ __forceinline__ void og_langevin_core(const og_dev_sys &s, const int i, const long long step,
                                                 const double *x, const double *xp, const double *vn,
                                                 const double *wk, double mass, double minv, double *xn) {
    if (step == 0) {
        for (int d = 0; d < 3; d++) xn[d] = x[d] + vn[d] * s.dt;
        return;
    }
    double zeta[3];
    if (s.noise && s.cnt->noise_step == step) {
        // draws injected by the caller for this step (og_set_langevin_noise)
        for (int d = 0; d < 3; d++) zeta[d] = s.noise[3 * i + d];
    } else {
        const uint4   r  = og_philox(make_uint4((uint32_t)step, (uint32_t)(step >> 32), (uint32_t)i, 0x4c414e47u),
                                     make_uint2(s.seed_lo, s.seed_hi));
        const double2 z0 = og_gauss2(r.x, r.y);
        const double2 z1 = og_gauss2(r.z, r.w);
        zeta[0] = z0.x;
        zeta[1] = z0.y;
        zeta[2] = z1.x;
    }
    const double  dth     = s.dt * 0.5;
    const double  gm      = s.gamma;
    const double  sig     = sqrt(OG_GAS_CONST * s.temperature * gm * mass / dth * 1e-7);
    for (int d = 0; d < 3; d++) {
        const double det_f = -OG_FORCE_VEL * wk[d];
        const double stc_f = zeta[d] * sig;
        const double p1    = x[d];
        double       p2    = xp[d];
        p2 += nearbyint((p1 - p2) / (double)s.L[d]) * (double)s.L[d];
        xn[d] = 1.0 / (1.0 + gm * dth) * (2 * p1 - p2 + gm * dth * p2 + 4 * dth * dth * minv * (det_f + stc_f));
    }
}

__device__ __forceinline__ void og_langevin_atom(const og_dev_sys &s, const int i, const long long step,
                                                 const double *wk) {
    double x[3], xp[3], vn[3], xn[3];
    for (int d = 0; d < 3; d++) {
        x[d]  = s.crd[3 * i + d];
        xp[d] = s.crd_prev[3 * i + d];
        vn[d] = step == 0 ? s.vel_next[3 * i + d] : 0.0;
    }
    og_langevin_core(s, i, step, x, xp, vn, wk, (double)s.mass[i], (double)s.mass_inv[i], xn);
    for (int d = 0; d < 3; d++) {
        s.crd_prev2[3 * i + d] = xp[d];
        s.crd_prev[3 * i + d]  = x[d];
        if (step == 0) s.vel[3 * i + d] = vn[d];
        s.crd[3 * i + d] = xn[d];
    }
}

__global__ void __launch_bounds__(OG_THREADS) k_langevin_positions(const og_dev_sys *__restrict__ S,
                                                                   const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const int         i  = bm.y * OG_THREADS + threadIdx.x;
    if (i >= s.n_atoms || s.integrator != 1) return;
    const double wk[3] = {s.work[3 * i], s.work[3 * i + 1], s.work[3 * i + 2]};
    og_langevin_atom(s, i, s.cnt->step, wk);
}

// --------------------------------------------------------- constraints --
// ConstraintShake::shake_pair (ConstraintShake.cpp:18-49), one thread per constraint.
__device__ __forceinline__ void og_shake_pair(const og_dev_sys &s, const int t) {
    const int a = s.pairc_ids[2 * t], b = s.pairc_ids[2 * t + 1];
    double    d[3], dp[3];
    for (int k = 0; k < 3; k++) {
        d[k]  = og_mimg(s.crd[3 * b + k] - s.crd[3 * a + k], s.L[k]);
        dp[k] = og_mimg(s.crd_prev[3 * b + k] - s.crd_prev[3 * a + k], s.L[k]);
    }
    const double m1 = s.mass_inv[a], m2 = s.mass_inv[b], m12 = m1 + m2;
    const double dsq  = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double dsqp = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
    const double dip  = d[0] * dp[0] + d[1] * dp[1] + d[2] * dp[2];
    const double va = m12 * m12 * dsqp, vb = m12 * dip, vc = dsq - s.pairc_d2[t];
    const double disc = vb * vb - vc * va;
    const double vg   = (vb - sqrt(disc > 0.0 ? disc : 0.0)) / va;
    for (int k = 0; k < 3; k++) {
        s.crd[3 * a + k] += m1 * vg * dp[k];
        s.crd[3 * b + k] -= m2 * vg * dp[k];
    }
}

__global__ void __launch_bounds__(OG_THREADS) k_shake_pair(const og_dev_sys *__restrict__ S,
                                                           const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const int         t  = bm.y * OG_THREADS + threadIdx.x;
    if (s.constraint == 0 || t >= s.n_pairs_c) return;
    og_shake_pair(s, t);
}

// ConstraintShake::shake_trio (ConstraintShake.cpp:51-149), one thread per trio.
// inlined: as a called function it takes 248 VGPRs, which would set every caller's count
__device__ __forceinline__ void og_shake_trio(const og_dev_sys &s, int a1, int a2, int a3, const double *D) {
    double d12[3], d23[3], d31[3], p12[3], p23[3], p31[3];
    double g[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 3; k++) {
        d12[k] = og_mimg(s.crd[3 * a1 + k] - s.crd[3 * a2 + k], s.L[k]);
        p12[k] = og_mimg(s.crd_prev[3 * a1 + k] - s.crd_prev[3 * a2 + k], s.L[k]);
        d23[k] = og_mimg(s.crd[3 * a2 + k] - s.crd[3 * a3 + k], s.L[k]);
        p23[k] = og_mimg(s.crd_prev[3 * a2 + k] - s.crd_prev[3 * a3 + k], s.L[k]);
        d31[k] = og_mimg(s.crd[3 * a3 + k] - s.crd[3 * a1 + k], s.L[k]);
        p31[k] = og_mimg(s.crd_prev[3 * a3 + k] - s.crd_prev[3 * a1 + k], s.L[k]);
    }
    const double m1 = s.mass_inv[a1], m2 = s.mass_inv[a2], m3 = s.mass_inv[a3];
    const double c11 = m1 + m2, c12 = -m2, c13 = -m1;
    const double c21 = -m2, c22 = m2 + m3, c23 = -m3;
    const double c31 = -m1, c32 = -m3, c33 = m3 + m1;
    bool conv = false;
    for (int it = 0; it < s.max_loops; it++) {
        for (int k = 0; k < 3; k++) {
            d12[k] += (c11 * p12[k] * g[0] + c12 * p23[k] * g[1] + c13 * p31[k] * g[2]);
            d23[k] += (c21 * p12[k] * g[0] + c22 * p23[k] * g[1] + c23 * p31[k] * g[2]);
            d31[k] += (c31 * p12[k] * g[0] + c32 * p23[k] * g[1] + c33 * p31[k] * g[2]);
        }
#define OG_DOT(u, v) (u[0] * v[0] + u[1] * v[1] + u[2] * v[2])
        const double f11 = 2.0 * c11 * OG_DOT(d12, p12), f12 = 2.0 * c12 * OG_DOT(d12, p23),
                     f13 = 2.0 * c13 * OG_DOT(d12, p31);
        const double f21 = 2.0 * c21 * OG_DOT(d23, p12), f22 = 2.0 * c22 * OG_DOT(d23, p23),
                     f23 = 2.0 * c23 * OG_DOT(d23, p31);
        const double f31 = 2.0 * c31 * OG_DOT(d31, p12), f32 = 2.0 * c32 * OG_DOT(d31, p23),
                     f33 = 2.0 * c33 * OG_DOT(d31, p31);
        const double df0 = D[0] - OG_DOT(d12, d12), df1 = D[1] - OG_DOT(d23, d23), df2 = D[2] - OG_DOT(d31, d31);
#undef OG_DOT
        if (fabs(df0 / D[0]) < s.tol && fabs(df1 / D[1]) < s.tol && fabs(df2 / D[2]) < s.tol) {
            conv = true;
            break;
        }
        const double det = 1.0 / (f11 * f22 * f33 + f12 * f23 * f31 + f13 * f21 * f32 - f11 * f23 * f32
                                  - f12 * f21 * f33 - f13 * f22 * f31);
        const double b11 = f22 * f33 - f23 * f32, b12 = f32 * f13 - f33 * f12, b13 = f12 * f23 - f13 * f22;
        const double b21 = f23 * f31 - f21 * f33, b22 = f33 * f11 - f31 * f13, b23 = f13 * f21 - f11 * f23;
        const double b31 = f21 * f32 - f22 * f31, b32 = f31 * f12 - f32 * f11, b33 = f11 * f22 - f12 * f21;
        g[0] = (b11 * df0 + b12 * df1 + b13 * df2) * det;
        g[1] = (b21 * df0 + b22 * df1 + b23 * df2) * det;
        g[2] = (b31 * df0 + b32 * df1 + b33 * df2) * det;
        g[3] += g[0];
        g[4] += g[1];
        g[5] += g[2];
    }
    if (!conv) atomicAdd(&s.cnt->shake_fail, 1);
    for (int k = 0; k < 3; k++) {
        s.crd[3 * a1 + k] += m1 * (g[3] * p12[k] - g[5] * p31[k]);
        s.crd[3 * a2 + k] += m2 * (-g[3] * p12[k] + g[4] * p23[k]);
        s.crd[3 * a3 + k] += m3 * (-g[4] * p23[k] + g[5] * p31[k]);
    }
}

// Analytic SETTLE (S. Miyamoto, P. A. Kollman, J. Comput. Chem. 13 (1992) 952-962)
// for a rigid 3-site water (O, H1, H2); same arithmetic as the oracle's orc_settle.
// Core on values: xO/xH1/xH2 are the unconstrained positions (moved to the
// constrained ones), pO/pH1/pH2 the previous positions.  Returns false, with
// the positions untouched, when there is no rigid solution.
__device__ __forceinline__ bool og_settle_core(const og_dev_sys &s, double mO, double mH, double dOH, double dHH,
                                               double *xO, double *xH1, double *xH2, const double *pO,
                                               const double *pH1, const double *pH2) {
    const double wt = mO + 2.0 * mH;
    const double rc = 0.5 * dHH;
    const double h  = sqrt(dOH * dOH - rc * rc);
    const double ra = 2.0 * mH * h / wt;
    const double rb = h - ra;
    double b0[3], c0[3], xb[3], xc[3], com[3], a1[3], b1[3], c1[3];
    for (int d = 0; d < 3; d++) {
        b0[d] = og_mimg(pH1[d] - pO[d], s.L[d]);
        c0[d] = og_mimg(pH2[d] - pO[d], s.L[d]);
        xb[d] = og_mimg(xH1[d] - xO[d], s.L[d]);
        xc[d] = og_mimg(xH2[d] - xO[d], s.L[d]);
    }
    for (int d = 0; d < 3; d++) {
        com[d] = (mH * xb[d] + mH * xc[d]) / wt;
        a1[d]  = -com[d];
        b1[d]  = xb[d] - com[d];
        c1[d]  = xc[d] - com[d];
    }
    double Z[3] = {b0[1] * c0[2] - b0[2] * c0[1], b0[2] * c0[0] - b0[0] * c0[2], b0[0] * c0[1] - b0[1] * c0[0]};
    double X[3] = {a1[1] * Z[2] - a1[2] * Z[1], a1[2] * Z[0] - a1[0] * Z[2], a1[0] * Z[1] - a1[1] * Z[0]};
    double Y[3] = {Z[1] * X[2] - Z[2] * X[1], Z[2] * X[0] - Z[0] * X[2], Z[0] * X[1] - Z[1] * X[0]};
    const double nx = 1.0 / sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    const double ny = 1.0 / sqrt(Y[0] * Y[0] + Y[1] * Y[1] + Y[2] * Y[2]);
    const double nz = 1.0 / sqrt(Z[0] * Z[0] + Z[1] * Z[1] + Z[2] * Z[2]);
    for (int d = 0; d < 3; d++) {
        X[d] *= nx;
        Y[d] *= ny;
        Z[d] *= nz;
    }
#define OG_P(A, v) (A[0] * v[0] + A[1] * v[1] + A[2] * v[2])
    const double xb0 = OG_P(X, b0), yb0 = OG_P(Y, b0), xc0 = OG_P(X, c0), yc0 = OG_P(Y, c0);
    const double za1 = OG_P(Z, a1);
    const double xb1 = OG_P(X, b1), yb1 = OG_P(Y, b1), zb1 = OG_P(Z, b1);
    const double xc1 = OG_P(X, c1), yc1 = OG_P(Y, c1), zc1 = OG_P(Z, c1);
#undef OG_P
    // A displacement too large for a rigid solution makes one of these square
    // roots negative: the water is then left unconstrained and counted as a
    // constraint failure (OG_E_NOCONV) instead of turning into NaN.
    const double sinphi = za1 / ra;
    const double q1     = 1.0 - sinphi * sinphi;
    const double cosphi = sqrt(q1);
    const double sinpsi = (zb1 - zc1) / (2.0 * rc * cosphi);
    const double q2     = 1.0 - sinpsi * sinpsi;
    const double cospsi = sqrt(q2);
    const double ya2 = ra * cosphi;
    const double xb2 = -rc * cospsi;
    const double t1 = -rb * cosphi, t2 = rc * sinpsi * sinphi;
    const double yb2 = t1 - t2, yc2 = t1 + t2;
    const double alpa   = xb2 * (xb0 - xc0) + yb0 * yb2 + yc0 * yc2;
    const double beta   = xb2 * (yc0 - yb0) + xb0 * yb2 + xc0 * yc2;
    const double gama   = xb0 * yb1 - xb1 * yb0 + xc0 * yc1 - xc1 * yc0;
    const double al2be2 = alpa * alpa + beta * beta;
    const double q3     = al2be2 - gama * gama;
    const double sinthe = (alpa * gama - beta * sqrt(q3)) / al2be2;
    const double q4     = 1.0 - sinthe * sinthe;
    const double costhe = sqrt(q4);
    if (!(q1 > 0.0 && q2 >= 0.0 && q3 >= 0.0 && q4 >= 0.0)) return false;
    const double a3[3] = {-ya2 * sinthe, ya2 * costhe, za1};
    const double b3[3] = {xb2 * costhe - yb2 * sinthe, xb2 * sinthe + yb2 * costhe, zb1};
    const double c3[3] = {-xb2 * costhe - yc2 * sinthe, -xb2 * sinthe + yc2 * costhe, zc1};
    // Each atom moves by its correction (constrained minus unconstrained
    // position, both relative to the oxygen), so a hydrogen wrapped to the
    // other side of the box stays in its own periodic frame.
    for (int d = 0; d < 3; d++) {
        const double ca = com[d] + X[d] * a3[0] + Y[d] * a3[1] + Z[d] * a3[2];
        const double cb = com[d] + X[d] * b3[0] + Y[d] * b3[1] + Z[d] * b3[2];
        const double cc = com[d] + X[d] * c3[0] + Y[d] * c3[1] + Z[d] * c3[2];
        xO[d] += ca;
        xH1[d] += cb - xb[d];
        xH2[d] += cc - xc[d];
    }
    return true;
}


__device__ void og_settle(const og_dev_sys &s, int iO, int iH1, int iH2, double dOH, double dHH) {
    double    x[3][3], p[3][3];
    const int ia[3] = {iO, iH1, iH2};
    for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++) {
            x[a][d] = s.crd[3 * ia[a] + d];
            p[a][d] = s.crd_prev[3 * ia[a] + d];
        }
    if (!og_settle_core(s, s.mass[iO], s.mass[iH1], dOH, dHH, x[0], x[1], x[2], p[0], p[1], p[2])) {
        atomicAdd(&s.cnt->shake_fail, 1);
        return;
    }
    for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++) s.crd[3 * ia[a] + d] = x[a][d];
}

__global__ void __launch_bounds__(OG_TRIO_THREADS) k_trio(const og_dev_sys *__restrict__ S,
                                                          const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const int         t  = bm.y * OG_TRIO_THREADS + threadIdx.x;
    if (s.constraint == 0) return;
    if (t < s.n_trios) og_shake_trio(s, s.trio_ids[3 * t], s.trio_ids[3 * t + 1], s.trio_ids[3 * t + 2],
                                     s.trio_d2 + 3 * t);
    if (t < s.n_settle) og_settle(s, s.settle_ids[3 * t], s.settle_ids[3 * t + 1], s.settle_ids[3 * t + 2],
                                  s.settle_dOH, s.settle_dHH);
}

// ------------------------------------------------------------- finish --
// After constraints: velocity from positions (SubBox::set_velocity_from_crd,
// SubBox.cpp:1462-1481) or the Langevin central difference
// (SubBox::set_velocities_just_langevin, SubBox.cpp:2054-2066); grid
// coordinates advanced by the FP64 displacement (MiniCell::move_atom,
// MiniCell.cpp:1312-1325) and republished as float; PBC wrap
// (SubBox::revise_coordinates_pbc, SubBox.cpp:1482-1502); kinetic-energy
// partials (DynamicsMode::cal_kinetic_energy, DynamicsMode.cpp:358-372).
// Kinetic-energy partial of one block (fixed order) into its e_atom slot.
__device__ __forceinline__ void og_ke_partial(const og_dev_sys &s, const int slot, double ke, double *sh) {
    for (int off = 32; off > 0; off >>= 1) ke += __shfl_down(ke, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = ke;
    __syncthreads();
    if (threadIdx.x == 0 && slot < s.n_ke_parts) {
        double t = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += sh[k];
        s.e_atom[slot] = t;
    }
}

// Langevin velocity of one atom by the central difference over two steps
// (first step: the initial velocity vn); returns sum_d v_d^2.
__device__ __forceinline__ double og_langevin_vel(const og_dev_sys &s, const long long step, const double *x,
                                                  const double *xp2, const double *vn, double *v) {
    double ka = 0.0;
    if (step > 0) {
        const double dt2inv = 1.0 / (2 * s.dt);
        for (int d = 0; d < 3; d++) {
            double p2 = xp2[d];
            p2 += nearbyint((x[d] - p2) / (double)s.L[d]) * (double)s.L[d];
            v[d] = (x[d] - p2) * dt2inv;
            ka += v[d] * v[d];
        }
    } else {
        for (int d = 0; d < 3; d++) ka += vn[d] * vn[d];
    }
    return ka;
}

// Grid coordinates gc advanced by the step's displacement x - xp (float copy
// in f) and the position wrapped once into the box (xw).
__device__ __forceinline__ void og_advance(const og_dev_sys &s, const double *x, const double *xp, double *gc,
                                           float *f, double *xw) {
    for (int d = 0; d < 3; d++) {
        const double diff = x[d] - xp[d];
        gc[d]             = gc[d] + diff;
        f[d]              = (float)gc[d];
        xw[d]             = x[d];
        if (xw[d] >= (double)s.ub[d]) xw[d] -= (double)s.L[d];
        else if (xw[d] < (double)s.lb[d]) xw[d] += (double)s.L[d];
    }
}

__device__ __forceinline__ double og_finish_atom(const og_dev_sys &s, const int i, const long long step) {
    double ke;
    const int       g    = s.rev[i];
    double          x[3], xp[3];
    for (int d = 0; d < 3; d++) {
        x[d]  = s.crd[3 * i + d];
        xp[d] = s.crd_prev[3 * i + d];
    }
    if (s.integrator == 0) {
        if (s.constraint) {
            const double dt_inv = 1.0 / s.dt;
            for (int d = 0; d < 3; d++) s.vel_next[3 * i + d] = (x[d] - xp[d]) * dt_inv;
        }
        double ka = 0.0;
        for (int d = 0; d < 3; d++) {
            const double vj = (s.vel[3 * i + d] + s.vel_next[3 * i + d]) * 0.5;
            ka += vj * vj;
        }
        ke = ka * (double)s.mass[i];
    } else {
        double xp2[3], vn[3], v[3];
        for (int d = 0; d < 3; d++) {
            xp2[d] = s.crd_prev2[3 * i + d];
            vn[d]  = s.vel_next[3 * i + d];
        }
        const double ka = og_langevin_vel(s, step, x, xp2, vn, v);
        if (step > 0)
            for (int d = 0; d < 3; d++) s.vel[3 * i + d] = v[d];
        ke = ka * (double)s.mass[i];
    }
    double gc[3], xw[3];
    float  f[3];
    for (int d = 0; d < 3; d++) gc[d] = s.gcrd[3 * g + d];
    og_advance(s, x, xp, gc, f, xw);
    for (int d = 0; d < 3; d++) {
        s.gcrd[3 * g + d] = gc[d];
        s.crd[3 * i + d]  = xw[d];
    }
    const float4 old = s.gxyzq[g];
    s.gxyzq[g]       = make_float4(f[0], f[1], f[2], old.w);
    return ke;
}

__global__ void __launch_bounds__(OG_THREADS) k_finish(const og_dev_sys *__restrict__ S, const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const int         i  = bm.y * OG_THREADS + threadIdx.x;
    double            ke = 0.0;
    if (i < s.n_atoms) ke = og_finish_atom(s, i, s.cnt->step);
    if (bm.y == 0)  // slots beyond this kernel's blocks (used by the fused step) count zero
        for (int k = s.n_atom_blocks + threadIdx.x; k < s.n_ke_parts; k += OG_THREADS) s.e_atom[k] = 0.0;
    __shared__ double sh[OG_THREADS / 64];
    og_ke_partial(s, bm.y, ke, sh);
}

#define OG_STEP_THREADS 256

// End of an MD step for one system, by one block of OG_STEP_THREADS threads:
// pair/excluded-pair/bond energies (on logged or requested steps), kinetic
// energy from the per-block partials, the energy log entry, and the step
// counter.  c: the system's counters as read at the start of the block.
__device__ __forceinline__ void og_close_step(const og_dev_sys &s, const og_dev_counters &c) {
    // the kinetic partials are read first (loads issued four strides at a
    // time, added in stride order) so they overlap the counter reads
    double t = 0.0;
    for (int k = threadIdx.x; k < s.n_ke_parts; k += 4 * OG_STEP_THREADS) {
        double v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = k + j * OG_STEP_THREADS < s.n_ke_parts ? s.e_atom[k + j * OG_STEP_THREADS] : 0.0;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (k + j * OG_STEP_THREADS < s.n_ke_parts) t += v[j];
    }
    // read with the partials, before the counters are known (one memory
    // round trip less): the accumulator slots and, in thread 0, the energy
    // row, whose pair, excluded-pair, bonded and kinetic terms this kernel
    // replaces (the others come from earlier kernels of the step)
    const long long ev = og_eacc_load<OG_STEP_THREADS>(s);
    double row[OG_E_NTERMS];
    if (threadIdx.x == 0)
        for (int k = 0; k < OG_E_NTERMS; k++) row[k] = s.energy[k];
    if (og_frozen_c(c)) return;
    __shared__ double    sh[OG_STEP_THREADS / 64];
    __shared__ long long she[16 * OG_EACC_STRIDE];
    if (og_energy_step_c(s, c)) {
        og_eacc_clear<OG_STEP_THREADS>(s);
        og_eacc_reduce<OG_STEP_THREADS>(s, ev, she, row);
    }
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double ke = 0.0;
        for (int w = 0; w < OG_STEP_THREADS / 64; w++) ke += sh[w];
        row[OG_E_KIN]       = ke * OG_KINETIC_COEFF;
        s.energy[OG_E_KIN]  = row[OG_E_KIN];
        const long long stp = c.step;
        if (s.log_cap > 0 && (stp % s.log_stride) == 0) {
            double *e = s.elog + (size_t)((stp / s.log_stride) % s.log_cap) * OG_E_NTERMS;
            for (int k = 0; k < OG_E_NTERMS; k++) e[k] = row[k];
        }
        s.cnt->rebuild_now = 0;
        s.cnt->step        = stp + 1;
    }
}

// ------------------------------------------------- fused Langevin step --
// One launch for the whole integration step of a Langevin system whose
// SETTLE waters interact with nothing but their own atoms through exclusions
// and carry no bonds, VcMD groups or SHAKE groups:
//  * block 0 of a system takes every other ("free") atom through the same
//    sequence as the separate kernels -- gradient conversion, VcMD bias,
//    Langevin positions, SHAKE pairs then trios, finish -- with a block
//    barrier between the stages;
//  * blocks 1.. take the waters, three lanes per water (og_water_lanes):
//    gradient conversion, Langevin positions, SETTLE and finish.
// Each atom goes through the same device functions, in the same order, as in
// the separate kernels, so the trajectory is bitwise the same; only the
// FP64 block partials of the kinetic, excluded-pair and bond energies are
// grouped differently (block 0, then the water blocks; the remaining e_atom
// slots are zeroed).
// SETTLE waters of the fused Langevin step, three lanes per water (lane
// 3k + a holds atom a of the wave's water k; lane 63 is an idle copy): each
// lane loads its atom, takes the other two atoms' positions by shuffles for
// the excluded pairs and for SETTLE (solved identically in all three lanes),
// runs the value cores shared with the separate kernels and stores its atom
// once.  e_ex and ke collect the lane's contributions.
__device__ __forceinline__ void og_water_lanes(const og_dev_sys &s, const int blk, const long long step,
                                               double &e_ex, double &ke) {
    const int  lane = threadIdx.x & 63;
    const int  grp  = min(lane / 3, OG_WATERS_PER_WAVE - 1);
    const int  ac   = min(lane - 3 * grp, 2);
    const int  base = 3 * grp;
    const int  tr   = blk * OG_WATERS_PER_BLOCK + (threadIdx.x >> 6) * OG_WATERS_PER_WAVE + grp;
    const bool live = lane < 3 * OG_WATERS_PER_WAVE && tr < s.n_settle;
    const int  t    = min(tr, s.n_settle - 1);
    const int  cap  = s.cap_atom_array;
    int        ia[3];
    for (int k = 0; k < 3; k++) ia[k] = s.settle_ids[3 * t + k];
    const int   i  = ia[ac];
    const int   g  = s.rev[i];
    const int   e0 = s.exc_off[i], e1 = s.exc_off[i + 1];
    const float q  = s.charge[i];
    const float m  = s.mass[i];
    const float mi = s.mass_inv[i];
    double      x[3], xp[3], vn[3], w[3], gc[3];
    for (int d = 0; d < 3; d++) {
        x[d]  = s.crd[3 * i + d];
        xp[d] = s.crd_prev[3 * i + d];
        vn[d] = step == 0 ? s.vel_next[3 * i + d] : 0.0;
        gc[d] = s.gcrd[3 * g + d];
    }
    og_fx_read(s.fx, cap, og_fx_slot(g), w);
    const float gw = s.gxyzq[g].w;
    int4        ent[2];
    for (int j = 0; j < 2; j++)
        if (e0 + j < e1) ent[j] = s.exc_at[e0 + j];
    double X[3][3];
    float  Q[3];
    for (int k = 0; k < 3; k++) {
        for (int d = 0; d < 3; d++) X[k][d] = __shfl(x[d], base + k, 64);
        Q[k] = __shfl(q, base + k, 64);
    }
    // excluded pairs of the atom, in its list order (og_atom_work_calc)
    auto excl = [&](const int4 en) {
        const int   la = en.x == ia[0] ? 0 : (en.x == ia[1] ? 1 : 2);
        const int   lb = en.y == ia[0] ? 0 : (en.y == ia[1] ? 1 : 2);
        float       we[3];
        const float ene = og_excess_pair(s, X[la], X[lb], Q[la], Q[lb], we);
        if (la == ac) {
            for (int d = 0; d < 3; d++) w[d] += (double)we[d];
            if (live) e_ex += (double)ene;
        } else {
            for (int d = 0; d < 3; d++) w[d] -= (double)we[d];
        }
    };
    for (int j = 0; j < 2; j++)
        if (e0 + j < e1) excl(ent[j]);
    for (int e = e0 + 2; e < e1; e++) excl(s.exc_at[e]);
    double xn[3];
    og_langevin_core(s, i, step, x, xp, vn, w, (double)m, (double)mi, xn);
    double XN[3][3];
    for (int k = 0; k < 3; k++)
        for (int d = 0; d < 3; d++) XN[k][d] = __shfl(xn[d], base + k, 64);
    const double mO = __shfl((double)m, base, 64), mH = __shfl((double)m, base + 1, 64);
    if (!og_settle_core(s, mO, mH, s.settle_dOH, s.settle_dHH, XN[0], XN[1], XN[2], X[0], X[1], X[2]) && live &&
        ac == 0)
        atomicAdd(&s.cnt->shake_fail, 1);
    double v[3], xw[3];
    float  f[3];
    const double ka = og_langevin_vel(s, step, XN[ac], xp, vn, v);
    og_advance(s, XN[ac], x, gc, f, xw);
    if (!live) return;
    ke += ka * (double)m;
    og_fx_consume(s.fx, cap, og_fx_slot(g));
    for (int d = 0; d < 3; d++) {
        s.work[3 * i + d]      = w[d];
        s.crd_prev2[3 * i + d] = xp[d];
        s.crd_prev[3 * i + d]  = x[d];
        s.crd[3 * i + d]       = xw[d];
        s.vel[3 * i + d]       = step == 0 ? vn[d] : v[d];
        s.gcrd[3 * g + d]      = gc[d];
    }
    s.gxyzq[g] = make_float4(f[0], f[1], f[2], gw);
}

__device__ void og_vcmd_block(const og_dev_sys &s, double *bias, int &active);  // VcMD section below

__global__ void __launch_bounds__(OG_WATER_THREADS) k_langevin_step(const og_dev_sys *__restrict__ S,
                                                                    const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const long long step = s.cnt->step;
    const int       tid  = threadIdx.x;
    double          e_ex = 0.0, e_b = 0.0, ke = 0.0;
    if (bm.y == 0) {
        const int nf = s.n_free;
        for (int k = tid; k < nf; k += OG_WATER_THREADS) og_atom_work(s, s.free_ids[k], e_ex, e_b);
        __syncthreads();
        if (s.vc_n_states || s.mc_mode == OG_MC_VAUS) {
            __shared__ double bias[3];
            __shared__ int    active;
            og_vcmd_block(s, bias, active);
            __syncthreads();
        }
        for (int k = tid; k < nf; k += OG_WATER_THREADS) {
            const int    i     = s.free_ids[k];
            const double wk[3] = {s.work[3 * i], s.work[3 * i + 1], s.work[3 * i + 2]};
            og_langevin_atom(s, i, step, wk);
        }
        __syncthreads();
        if (s.constraint) {
            for (int t = tid; t < s.n_pairs_c; t += OG_WATER_THREADS) og_shake_pair(s, t);
            __syncthreads();
            for (int t = tid; t < s.n_trios; t += OG_WATER_THREADS)
                og_shake_trio(s, s.trio_ids[3 * t], s.trio_ids[3 * t + 1], s.trio_ids[3 * t + 2], s.trio_d2 + 3 * t);
            __syncthreads();
        }
        for (int k = tid; k < nf; k += OG_WATER_THREADS) ke += og_finish_atom(s, s.free_ids[k], step);
        for (int k = s.n_water_blocks + 1 + tid; k < s.n_ke_parts; k += OG_WATER_THREADS) s.e_atom[k] = 0.0;
    } else {
        og_water_lanes(s, bm.y - 1, step, e_ex, ke);
    }
    __shared__ double sh[2 * (OG_WATER_THREADS / 64)];
    og_ke_partial(s, bm.y, ke, sh);
    if (og_energy_step(s)) {
        __syncthreads();  // sh is reused
        og_add_misc_energies(s, bm.y, e_ex, e_b, sh);
    }
}

// One block per system, the step's last kernel.
__global__ void __launch_bounds__(OG_STEP_THREADS) k_step_close(const og_dev_sys *__restrict__ S,
                                                               const int2 *__restrict__ map) {
    const og_dev_sys &s = S[map[blockIdx.x].x];
    // one batch of loads for the counters this kernel reads (a single block
    // runs per system, so every dependent round trip adds to the step)
    const og_dev_counters c = *s.cnt;
    og_close_step(s, c);
}

// ----------------------------------------------------------------- VcMD --
// One wave per system, after the forces and before the velocity update
// (DynamicsModePresto/Langevin::calc_in_each_step, DynamicsMode.cpp:504-512,
// :824-834):
//  * ExtendedVcMD::set_crd_centers (Extend.cpp:753-797): group coordinates
//    unwrapped against their previous values, mass centres;
//  * set_struct_parameters_mass_center (:668-692): lambda = |minimum-image
//    difference of the centres|, unit vector;
//  * scale_force (:1035-1102): outside [min - sigma, max + sigma) of the
//    current state the gradient of every group atom gets
//    +-k_B T * coef * excess * u * m_i / M_group, and the reported energy is
//    coef * excess^2 (the reference's energy/force pair, kept as is);
//  * vs_step + trial_transition (:808-845, :907-1033): the state counts a
//    visit while lambda is in its range, and every vc_interval steps moves
//    to a neighbouring state whose range holds lambda with probability
//    proportional to 1/q_cano (uniform from Philox keyed by seed and step).
// V-AUS (mc_mode OG_MC_VAUS, ExtendedVAUS, Extend.cpp:366-540) takes the same
// groups and lambda; its state moves first (og_mc_transition) and its force is
// og_vaus_lane0's.
// Called by every thread of a block (any size); bias and active live in LDS.
__device__ void og_vaus_lane0(const og_dev_sys &s, const double *diff, double lam, double *bias, int &active);
__device__ void og_vcmd_block(const og_dev_sys &s, double *bias, int &active) {
    const int lane = threadIdx.x, nt = blockDim.x;
    const int n    = s.vc_na + s.vc_nb;
    for (int k = lane; k < n; k += nt) {
        const int a = s.vc_atoms[k];
        for (int d = 0; d < 3; d++) {
            const double Lh   = (double)(s.L[d] * 0.5f);
            double       tmp  = s.crd[3 * a + d];
            double       diff = tmp - s.vc_crdg[3 * k + d];
            while (diff > Lh) {
                diff -= (double)s.L[d];
                tmp -= (double)s.L[d];
            }
            while (-diff > Lh) {
                diff += (double)s.L[d];
                tmp += (double)s.L[d];
            }
            s.vc_crdg[3 * k + d] = tmp;
        }
    }
    __syncthreads();
    if (lane == 0) {
        // mass centres summed in group order, as the reference loop
        double ca[3] = {0.0, 0.0, 0.0}, cb[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < n; k++) {
            const double ms = (double)s.mass[s.vc_atoms[k]];
            for (int d = 0; d < 3; d++) {
                if (k < s.vc_na) ca[d] += s.vc_crdg[3 * k + d] * ms;
                else cb[d] += s.vc_crdg[3 * k + d] * ms;
            }
        }
        double diff[3];
        for (int d = 0; d < 3; d++) diff[d] = og_mimg(ca[d] * s.vc_minv_a - cb[d] * s.vc_minv_b, s.L[d]);
        const double lam = sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
        if (s.mc_mode == OG_MC_VAUS) {
            og_vaus_lane0(s, diff, lam, bias, active);
        } else {
        const int    vs  = s.cnt->vs;
        const double lo = s.vc_lo[vs], hi = s.vc_hi[vs];
        double       ex = 0.0, rec = 0.0;
        if (lam < lo - s.vc_sigma) {
            ex  = lam - (lo - s.vc_sigma);
            rec = s.vc_coef * ex;
        } else if (lam >= hi + s.vc_sigma) {
            ex  = lam - (hi + s.vc_sigma);
            rec = s.vc_coef * ex;
        }
        const double dew = s.vc_constk * rec;
        for (int d = 0; d < 3; d++) bias[d] = diff[d] / lam;  // unit vector, scaled below
        for (int d = 0; d < 3; d++) bias[d] = dew * bias[d];
        active                = rec != 0.0;
        s.energy[OG_E_EXTEND] = rec * ex;
        s.energy[OG_E_LAMBDA] = lam;
        s.energy[OG_E_VSTATE] = (double)vs;
        // visit count and transition trial
        const long long step     = s.cnt->step;
        const bool      in_range = lam >= lo && lam < hi;
        if (in_range) s.vc_hist[vs] += 1.0;
        if (step > 0 && step % s.vc_interval == 0 && in_range) {
            int cand[2] = {vs, -1};
            if (vs > 0) {
                if (lam < s.vc_hi[vs - 1]) cand[1] = vs - 1;
                else if (vs < s.vc_n_states - 1 && lam >= s.vc_lo[vs + 1]) cand[1] = vs + 1;
            } else if (s.vc_n_states > 1 && lam >= s.vc_lo[1]) {
                cand[1] = 1;
            }
            if (cand[1] >= 0) {
                const uint4  r   = og_philox(make_uint4((uint32_t)step, (uint32_t)(step >> 32), 0u, 0x5643u),
                                             make_uint2(s.seed_lo, s.seed_hi));
                const double u   = ((double)r.x + 0.5) * 2.3283064365386963e-10;
                const double w0  = s.vc_qinv[cand[0]], w1 = s.vc_qinv[cand[1]];
                s.cnt->vs        = u * (w0 + w1) < w0 ? cand[0] : cand[1];
            }
        }
        }
    }
    __syncthreads();
    if (!active) return;
    // gradient ("work") of the group atoms; the groups are disjoint
    for (int k = lane; k < n; k += nt) {
        const int    a    = s.vc_atoms[k];
        const bool   ga   = k < s.vc_na;
        const double minv = ga ? s.vc_minv_a : s.vc_minv_b;
        const double dir  = ga ? 1.0 : -1.0;
        for (int d = 0; d < 3; d++) s.work[3 * a + d] += dir * (bias[d] * minv) * (double)s.mass[a];
    }
}

// ------------------------------------------------------- V-McMD, V-AUS --
// ExtendedVMcMD::trial_transition + set_current_vstate (Extend.cpp:143-176):
// a move up is tried first, then down; a move needs lambda inside both
// (closed) ranges and its uniform draw above 1 - the state's probability
// for that direction.  The two draws come from Philox keyed by seed and step.
__device__ int og_mc_trial(const og_dev_sys &s, int src, int rel, double lam, double u) {
    if (!(lam >= s.mc_lo[src] && lam <= s.mc_hi[src])) return src;
    if (src == 0 && rel == -1) return src;
    if (src == s.mc_n_states - 1 && rel == 1) return src;
    const int dst = src + rel;
    if (lam >= s.mc_lo[dst] && lam <= s.mc_hi[dst]) {
        const double p = rel == -1 ? s.mc_pdown[src] : s.mc_pup[src];
        if (u > 1.0 - p) return dst;
    }
    return src;
}
__device__ int og_mc_transition(const og_dev_sys &s, int vs, double lam, long long step) {
    const uint4  r    = og_philox(make_uint4((uint32_t)step, (uint32_t)(step >> 32), 1u, 0x4D43u),
                                  make_uint2(s.seed_lo, s.seed_hi));
    const double u_up = ((double)r.x + 0.5) * 2.3283064365386963e-10;
    const double u_dn = ((double)r.y + 0.5) * 2.3283064365386963e-10;
    const int    up   = og_mc_trial(s, vs, 1, lam, u_up);
    return up != vs ? up : og_mc_trial(s, vs, -1, lam, u_dn);
}
// the state's d ln P / d lambda at lambda clamped to its range
__device__ double og_mc_dlnp(const og_dev_sys &s, int vs, double param) {
    const double *c = s.mc_poly + (size_t)vs * OG_MC_NPOLY;
    double        d = c[0], t = 1.0;
    for (int i = 1; i <= s.mc_order[vs]; i++) {
        t *= param;
        d += c[i] * t;
    }
    return d;
}

// V-AUS on lane 0 of the group block (ExtendedVAUS::scale_force,
// Extend.cpp:460-540): the state moves first (apply_bias, :126-141); outside
// [lo - sigma, hi + sigma] a harmonic recovery, and k_B T (d ln P(lambda
// clamped) + recovery) along the unit vector.  The reference reports no
// bias energy for V-AUS.
__device__ void og_vaus_lane0(const og_dev_sys &s, const double *diff, double lam, double *bias, int &active) {
    const long long step = s.cnt->step;
    int             vs   = s.cnt->vs;
    if (step > 0 && step % s.mc_interval == 0) {
        vs        = og_mc_transition(s, vs, lam, step);
        s.cnt->vs = vs;
    }
    const double lo = s.mc_lo[vs], hi = s.mc_hi[vs];
    double       param = lam, recovery = 0.0;
    if (param <= lo) {
        if (param <= lo - s.vc_sigma) recovery = s.vc_coef * (param - (lo - s.vc_sigma));
        param = lo;
    } else if (param >= hi) {
        if (param >= hi + s.vc_sigma) recovery = s.vc_coef * (param - (hi + s.vc_sigma));
        param = hi;
    }
    const double dew = s.mc_constk * (og_mc_dlnp(s, vs, param) + recovery);
    for (int d = 0; d < 3; d++) bias[d] = diff[d] / lam;
    for (int d = 0; d < 3; d++) bias[d] = dew * bias[d];
    active                = 1;
    s.energy[OG_E_EXTEND] = 0.0;
    s.energy[OG_E_LAMBDA] = lam;
    s.energy[OG_E_VSTATE] = (double)vs;
}

// V-McMD (ExtendedVMcMD::apply_bias + scale_force, Extend.cpp:126-200), one
// wave per system after the gradient is complete: lambda = the potential
// energy of the step from the fixed-point accumulators (read, not cleared:
// the step's closing kernel takes them), the state move every mc_interval
// steps, then the factor k_B T d ln P(lambda clamped) that k_vmcmd_scale
// applies to every atom's gradient.
__global__ void __launch_bounds__(64) k_vmcmd(const og_dev_sys *__restrict__ S, const int2 *__restrict__ map) {
    const og_dev_sys &s = S[map[blockIdx.x].x];
    if (og_frozen(s) || s.mc_mode != OG_MC_VMCMD) return;
    long long v = 0;
    for (int slot = threadIdx.x; slot < OG_EACC_SLOTS; slot += 64) {
        const long long *a = s.e_acc + (size_t)slot * OG_EACC_STRIDE;
        v += a[OG_EACC_VDW] + a[OG_EACC_ELE] + a[OG_EACC_EXCESS] + a[OG_EACC_BOND] + a[OG_EACC_ANGLE] +
             a[OG_EACC_TORSION] + a[OG_EACC_IMPRO] + a[OG_EACC_VDW14] + a[OG_EACC_ELE14];
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (threadIdx.x != 0) return;
    const double    lam  = (double)v * OG_FX_INV + s.e_self;
    const long long step = s.cnt->step;
    int             vs   = s.cnt->vs;
    if (step > 0 && step % s.mc_interval == 0) {
        vs        = og_mc_transition(s, vs, lam, step);
        s.cnt->vs = vs;
    }
    const double lo = s.mc_lo[vs], hi = s.mc_hi[vs];
    const double param = lam <= lo ? lo : (lam >= hi ? hi : lam);
    s.cnt->mc_scale       = s.mc_constk * og_mc_dlnp(s, vs, param);
    s.energy[OG_E_EXTEND] = 0.0;
    s.energy[OG_E_LAMBDA] = lam;
    s.energy[OG_E_VSTATE] = (double)vs;
}
__global__ void __launch_bounds__(OG_THREADS) k_vmcmd_scale(const og_dev_sys *__restrict__ S,
                                                            const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s) || s.mc_mode != OG_MC_VMCMD) return;
    const int i = bm.y * OG_THREADS + threadIdx.x;
    if (i >= s.n_atoms) return;
    const double f = s.cnt->mc_scale;
    for (int k = 0; k < 3; k++) s.work[3 * i + k] *= f;
}
void og_launch_vmcmd(const og_launch &sys, const og_launch &atoms) {
    hipLaunchKernelGGL(k_vmcmd, dim3(sys.n_blocks), dim3(64), 0, sys.stream, sys.sys, sys.map);
    hipLaunchKernelGGL(k_vmcmd_scale, dim3(atoms.n_blocks), dim3(OG_THREADS), 0, atoms.stream, atoms.sys, atoms.map);
}

__global__ void __launch_bounds__(64) k_vcmd(const og_dev_sys *__restrict__ S, const int2 *__restrict__ map) {
    const og_dev_sys &s = S[map[blockIdx.x].x];
    if (og_frozen(s)) return;
    if (s.vc_n_states == 0 && s.mc_mode != OG_MC_VAUS) return;
    __shared__ double bias[3];
    __shared__ int    active;
    og_vcmd_block(s, bias, active);
}

// ------------------------------------------------------------- launchers --
void og_launch_vcmd(const og_launch &a) {
    hipLaunchKernelGGL(k_vcmd, dim3(a.n_blocks), dim3(64), 0, a.stream, a.sys, a.map);
}
void og_launch_leapfrog_vel(const og_launch &a) {
    hipLaunchKernelGGL(k_leapfrog_vel, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_com_sum(const og_launch &a) {
    hipLaunchKernelGGL(k_com_sum, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_com_apply(const og_launch &a) {
    hipLaunchKernelGGL(k_com_apply, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_thermostat(const og_launch &a) {
    hipLaunchKernelGGL(k_thermo_sum, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
    hipLaunchKernelGGL(k_thermo_apply, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_positions(const og_launch &a) {
    hipLaunchKernelGGL(k_positions, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_langevin_positions(const og_launch &a) {
    hipLaunchKernelGGL(k_langevin_positions, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_constraints(const og_launch &trio, const og_launch &misc) {
    if (misc.n_blocks)
        hipLaunchKernelGGL(k_shake_pair, dim3(misc.n_blocks), dim3(OG_THREADS), 0, misc.stream, misc.sys, misc.map);
    if (trio.n_blocks)
        hipLaunchKernelGGL(k_trio, dim3(trio.n_blocks), dim3(OG_TRIO_THREADS), 0, trio.stream, trio.sys, trio.map);
}
void og_launch_langevin_step(const og_launch &a) {
    hipLaunchKernelGGL(k_langevin_step, dim3(a.n_blocks), dim3(OG_WATER_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_finish(const og_launch &a) {
    hipLaunchKernelGGL(k_finish, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_step_close(const og_launch &a) {
    hipLaunchKernelGGL(k_step_close, dim3(a.n_blocks), dim3(OG_STEP_THREADS), 0, a.stream, a.sys, a.map);
}
