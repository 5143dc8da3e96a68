// og_bonded.hip -- bonded terms beyond harmonic bonds (angles, torsions,
// impropers, scaled 1-4 pairs) and SHAKE on four-atom groups, in FP64.
//
// The reference evaluates these on the host for every step
// (SubBox::calc_energy_angles / _torsions / _impros / _14nb,
// SubBox.cpp:1250-1346, with ForceField::calc_angle / calc_torsion /
// calc_14pair, ForceField.cpp:138-331) and runs ConstraintShake::shake_quad
// (ConstraintShake.cpp:151-276) after the pair and trio constraints.  Here one
// thread takes one term: it writes each atom's gradient share to its own
// contribution slot (no atomics; each atom sums its slots in a fixed order in
// og_atom_work_calc), and the block's energy sums go to the fixed-point
// accumulators.  Compiled with -ffp-contract=off: every expression rounds as
// the reference's double code.
#include "og_internal.h"

namespace {

constexpr double kChargeCoeff = 332.06378;  // CelesteObject.cpp:36
constexpr double kPiRef       = 3.14159265; // CelesteObject.cpp:26 (the reference's PI)
constexpr double kEps         = 1e-10;      // CelesteObject.cpp:31
constexpr double kEps3        = 1e-30;      // CelesteObject.cpp:31

// minimum-image difference xa - xb (og_mimg per axis)
__device__ __forceinline__ void mimg3(const og_dev_sys &s, const double *xa, const double *xb, double *d) {
    for (int k = 0; k < 3; k++) d[k] = og_mimg(xa[k] - xb[k], s.L[k]);
}

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ void load3(const og_dev_sys &s, int a, double *x) {
    for (int k = 0; k < 3; k++) x[k] = s.crd[3 * a + k];
}

__device__ __forceinline__ void store3(double *c, int slot, const double *v) {
    for (int k = 0; k < 3; k++) c[3 * (size_t)slot + k] = v[k];
}

// Angle t (ForceField::calc_angle, ForceField.cpp:138-177): E = k (theta - theta0)^2
// with theta the angle at the middle atom; shares w1, -(w1 + w2), w2.
__device__ double angle_term(const og_dev_sys &s, int t) {
    const int a1 = s.ang_ids[3 * t], a2 = s.ang_ids[3 * t + 1], a3 = s.ang_ids[3 * t + 2];
    const double2 par = s.ang_par[t];
    double x1[3], x2[3], x3[3], d12[3], d32[3];
    load3(s, a1, x1);
    load3(s, a2, x2);
    load3(s, a3, x3);
    mimg3(s, x1, x2, d12);
    mimg3(s, x3, x2, d32);
    const double r12_2      = d12[0] * d12[0] + d12[1] * d12[1] + d12[2] * d12[2];
    const double r32_2      = d32[0] * d32[0] + d32[1] * d32[1] + d32[2] * d32[2];
    const double r12r32     = sqrt(r12_2 * r32_2);
    const double r12r32_inv = 1.0 / r12r32;
    double       cost       = (d12[0] * d32[0] + d12[1] * d32[1] + d12[2] * d32[2]) * r12r32_inv;
    cost                    = fmax(fmin(cost, 1.0), -1.0);
    const double theta      = acos(cost);
    const double dif        = theta - par.y;
    const double coef       = par.x * dif;
    const double r12inv_2   = 1.0 / r12_2;
    const double r32inv_2   = 1.0 / r32_2;
    const double sin_theta  = fmax(kEps, sin(theta));
    const double grad_coef  = -(2.0 * coef) / sin_theta;
    double       w1[3], w2[3], wm[3];
    for (int k = 0; k < 3; k++) {
        w1[k] = grad_coef * (d32[k] * r12r32_inv - d12[k] * cost * r12inv_2);
        w2[k] = grad_coef * (d12[k] * r12r32_inv - d32[k] * cost * r32inv_2);
        wm[k] = -(w1[k] + w2[k]);
    }
    store3(s.bcontrib, 3 * t, w1);
    store3(s.bcontrib, 3 * t + 1, wm);
    store3(s.bcontrib, 3 * t + 2, w2);
    return coef * dif;
}

// Torsion or improper u (ForceField::calc_torsion, ForceField.cpp:179-299):
// E = V (1 + cos(n phi - phase)) / overlaps; shares w1, w2, -(w1 + w2 + w3), w3.
__device__ double torsion_term(const og_dev_sys &s, int u, int slot0) {
    const int   *a   = s.tor_ids + 4 * u;
    const double4 par = s.tor_par[u];  // {V, overlaps, symmetry, phase}
    double w1[3] = {0, 0, 0}, w2[3] = {0, 0, 0}, w3[3] = {0, 0, 0};
    double ene = 0.0;
    if (par.x > kEps) {
        double x1[3], x2[3], x3[3], x4[3], d21[3], d32[3], d43[3];
        load3(s, a[0], x1);
        load3(s, a[1], x2);
        load3(s, a[2], x3);
        load3(s, a[3], x4);
        mimg3(s, x2, x1, d21);
        mimg3(s, x3, x2, d32);
        mimg3(s, x4, x3, d43);
        double p12[3], p23[3];
        cross3(d21, d32, p12);
        cross3(d32, d43, p23);
        const double p12_2  = p12[0] * p12[0] + p12[1] * p12[1] + p12[2] * p12[2];
        const double p23_2  = p23[0] * p23[0] + p23[1] * p23[1] + p23[2] * p23[2];
        const double p12p23 = sqrt(p12_2 * p23_2);
        if (p12p23 > 0.0) {
            const double p12p23inv = 1.0 / p12p23;
            const double s1223     = p12[0] * p23[0] + p12[1] * p23[1] + p12[2] * p23[2];
            double       cosp      = s1223 * p12p23inv;
            cosp                   = fmax(-1.0, fmin(1.0, cosp));
            double p123[3];
            cross3(p12, p23, p123);
            double       phi    = acos(cosp);
            const double s32123 = d32[0] * p123[0] + d32[1] * p123[1] + d32[2] * p123[2];
            if (s32123 < 0.0) phi = 2.0 * kPiRef - phi;
            const double symmetry = par.z, phase = par.w;
            const double phimod   = symmetry * phi - phase;
            const double div      = 1.0 / par.y;
            ene                   = par.x * (1.0 + cos(phimod)) * div;
            const double sinp     = sin(phi);
            const int    nrot     = (int)symmetry;
            // the multiplicity table of ForceField.cpp:7-17: n for even n, 0 for odd (n in 0..10, checked by the host)
            const double gmul     = (nrot & 1) ? 0.0 : (double)nrot;
            const double sin_fphs = sin(phase);
            const double cos_fphs = cos(phase);
            const double sin_rot  = sin(symmetry * phi);
            const double cos_rot  = cos(symmetry * phi);
            const double dflim    = cos_fphs * (symmetry - gmul + gmul * cosp);
            const double df0      = cos_fphs * sin_rot - sin_fphs * cos_rot;
            double       df1      = dflim;
            if (kEps <= fabs(sinp)) df1 = df0 / sinp;
            const double work_coeff = symmetry * par.x * div * df1 * p12p23inv;
            const double r12        = 1.0 / p12_2 * s1223;
            const double r23        = 1.0 / p23_2 * s1223;
            double op_p23d32[3], op_p12d32[3];
            cross3(p23, d32, op_p23d32);
            cross3(p12, d32, op_p12d32);
            for (int k = 0; k < 3; k++) w1[k] = work_coeff * (op_p23d32[k] - r12 * op_p12d32[k]);
            const double d2132[3] = {d21[0] + d32[0], d21[1] + d32[1], d21[2] + d32[2]};
            double op_d2132_p23[3], op_d2132_p12[3], op_p12_d43[3], op_p23_d43[3];
            cross3(d2132, p23, op_d2132_p23);
            cross3(d2132, p12, op_d2132_p12);
            cross3(p12, d43, op_p12_d43);
            cross3(p23, d43, op_p23_d43);
            for (int k = 0; k < 3; k++)
                w2[k] = work_coeff * (op_d2132_p23[k] + op_p12_d43[k] - r12 * op_d2132_p12[k] - r23 * op_p23_d43[k]);
            for (int k = 0; k < 3; k++) w3[k] = work_coeff * (op_p12d32[k] - r23 * op_p23d32[k]);
        }
    }
    double wm[3];
    for (int k = 0; k < 3; k++) wm[k] = -(w1[k] + w2[k] + w3[k]);
    store3(s.bcontrib, slot0, w1);
    store3(s.bcontrib, slot0 + 1, w2);
    store3(s.bcontrib, slot0 + 2, wm);
    store3(s.bcontrib, slot0 + 3, w3);
    return ene;
}

// 1-4 pair t (ForceField::calc_14pair, ForceField.cpp:301-331): LJ of the
// entry's type pair and Coulomb, each scaled; shares w, -w.
__device__ void nb14_term(const og_dev_sys &s, int t, int slot0, double &e_vdw, double &e_ele) {
    const int a1 = s.nb14_ids[2 * t], a2 = s.nb14_ids[2 * t + 1];
    const int tp = s.nb14_tp[t];
    const double2 cf = s.nb14_coef[t];  // {LJ scale, Coulomb scale}
    double x1[3], x2[3], d14[3];
    load3(s, a1, x1);
    load3(s, a2, x2);
    mimg3(s, x1, x2, d14);
    const double r14_2      = d14[0] * d14[0] + d14[1] * d14[1] + d14[2] * d14[2];
    const double r14_2_inv  = 1.0 / r14_2;
    const double r14_inv    = sqrt(r14_2_inv);
    const double r14_inv_6  = r14_2_inv * r14_2_inv * r14_2_inv;
    const double r14_inv_12 = r14_inv_6 * r14_inv_6;
    const double term6      = (double)s.lj6[tp] * r14_inv_6;
    const double term12     = (double)s.lj12[tp] * r14_inv_12;
    e_vdw                   = (-term6 + term12) * cf.x;
    const double coef_work  = r14_2_inv * cf.x * (-12.0 * term12 + 6.0 * term6);
    const double cc         = (double)s.charge[a1] * (double)s.charge[a2] * kChargeCoeff * cf.y;
    e_ele                   = cc * r14_inv;
    const double coef_ele   = -cc * r14_2_inv * r14_inv;
    double       w[3], wm[3];
    for (int k = 0; k < 3; k++) {
        w[k]  = coef_work * d14[k] + coef_ele * d14[k];
        wm[k] = -w[k];
    }
    store3(s.bcontrib, slot0, w);
    store3(s.bcontrib, slot0 + 1, wm);
}

// block-wide sums of the five bonded energies (fixed tree order), left in thread 0
__device__ __forceinline__ void block_sum5(double *e, double *sh) {
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 5; q++) e[q] += __shfl_down(e[q], off, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 5; q++) sh[5 * w + q] = e[q];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < 5; q++) {
            e[q] = 0.0;
            for (int k = 0; k < (int)(blockDim.x >> 6); k++) e[q] += sh[5 * k + q];
        }
}

}  // namespace

// One thread per term: angles, then torsions, impropers, 1-4 pairs.
__global__ void __launch_bounds__(OG_THREADS) k_bonded(const og_dev_sys *__restrict__ S,
                                                       const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const int t    = bm.y * OG_THREADS + threadIdx.x;
    double    e[5] = {0, 0, 0, 0, 0};  // angle, torsion, improper, 1-4 LJ, 1-4 Coulomb
    const int nt0 = s.n_angles, nt1 = nt0 + s.n_torsions, nt2 = nt1 + s.n_impros;
    if (t < nt0) {
        e[0] = angle_term(s, t);
    } else if (t < nt1) {
        e[1] = torsion_term(s, t - nt0, 3 * nt0 + 4 * (t - nt0));
    } else if (t < nt2) {
        e[2] = torsion_term(s, t - nt0, 3 * nt0 + 4 * (t - nt0));
    } else if (t < s.n_bterms) {
        nb14_term(s, t - nt2, 3 * nt0 + 4 * (nt2 - nt0) + 2 * (t - nt2), e[3], e[4]);
    }
    if (!og_energy_step(s)) return;
    __shared__ double sh[5 * (OG_THREADS / 64)];
    block_sum5(e, sh);
    if (threadIdx.x == 0) {
        OG_GLOBAL(long long) *A = (OG_GLOBAL(long long) *)s.e_acc + (bm.y % OG_EACC_SLOTS) * OG_EACC_STRIDE;
        og_atomic_add_fx_g(A + OG_EACC_ANGLE, og_to_fixed(e[0]));
        og_atomic_add_fx_g(A + OG_EACC_TORSION, og_to_fixed(e[1]));
        og_atomic_add_fx_g(A + OG_EACC_IMPRO, og_to_fixed(e[2]));
        og_atomic_add_fx_g(A + OG_EACC_VDW14, og_to_fixed(e[3]));
        og_atomic_add_fx_g(A + OG_EACC_ELE14, og_to_fixed(e[4]));
    }
}

// ------------------------------------------------------- four-atom SHAKE --
namespace {

// ConstraintObject::calc_linear_eq (Constraint.cpp:141-203): LU with column
// pivoting in the reference's index order; false when a pivot vanishes.
__device__ bool linear_eq6(double a[6][6], double x[6], const double b[6]) {
    int pivot[6];
    for (int i = 0; i < 6; i++) pivot[i] = i;
    for (int i = 0; i < 6; i++) {
        int    index = i;
        double max_a = fabs(a[i][pivot[i]]);
        for (int j = i + 1; j < 6; j++) {
            const double tmp_a = fabs(a[i][pivot[j]]);
            if (tmp_a > max_a) {
                index = j;
                max_a = tmp_a;
            }
        }
        if (index != i) {
            const int tmp = pivot[i];
            pivot[i]      = pivot[index];
            pivot[index]  = tmp;
        }
        if (fabs(a[i][pivot[i]]) < kEps3) return false;
        a[i][pivot[i]] = 1.0 / a[i][pivot[i]];
        for (int j = i + 1; j < 6; j++) {
            a[i][pivot[j]] *= a[i][pivot[i]];
            for (int k = i + 1; k < 6; k++) a[k][pivot[j]] -= a[i][pivot[j]] * a[k][pivot[i]];
        }
    }
    for (int i = 0; i < 6; i++) {
        double v = b[pivot[i]];
        for (int j = 0; j < i; j++) v -= a[j][pivot[i]] * x[j];
        x[i] = v;
    }
    for (int i = 5; i >= 0; i--) {
        double v = x[i];
        for (int j = i + 1; j < 6; j++) v -= a[j][pivot[i]] * x[j];
        x[i] = v * a[i][pivot[i]];
    }
    return true;
}

}  // namespace

// One thread per group: Newton iteration on the six squared distances (pairs
// 01 12 20 30 23 13, Constraint.cpp:4) with a 6x6 solve per iteration, then
// the displacement of the four atoms along the previous-step bond vectors.
__global__ void __launch_bounds__(OG_QUAD_THREADS) k_shake_quad(const og_dev_sys *__restrict__ S,
                                                                const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const int q = bm.y * OG_QUAD_THREADS + threadIdx.x;
    if (s.constraint == 0 || q >= s.n_quads) return;
    const int   *a  = s.quad_ids + 4 * q;
    const double *D = s.quad_d2 + 6 * q;
    const int pi[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {2, 3}, {1, 3}};
    double m[4], cur[4][3], prev[4][3];
    for (int i = 0; i < 4; i++) {
        m[i] = (double)s.mass_inv[a[i]];
        load3(s, a[i], cur[i]);
        for (int k = 0; k < 3; k++) prev[i][k] = s.crd_prev[3 * a[i] + k];
    }
    const double weight[6][6] = {{m[0] + m[1], -m[1], -m[0], -m[0], 0.0, -m[1]},
                                 {-m[1], m[1] + m[2], -m[2], 0.0, -m[2], m[1]},
                                 {-m[0], -m[2], m[2] + m[0], m[0], m[2], 0.0},
                                 {-m[0], 0.0, m[0], m[3] + m[0], -m[3], -m[3]},
                                 {0.0, -m[2], m[2], -m[3], m[2] + m[3], m[3]},
                                 {-m[1], m[1], 0.0, -m[3], m[3], m[1] + m[3]}};
    double d_cur[6][3], d_prev[6][3];
    for (int p = 0; p < 6; p++) {
        mimg3(s, cur[pi[p][0]], cur[pi[p][1]], d_cur[p]);
        mimg3(s, prev[pi[p][0]], prev[pi[p][1]], d_prev[p]);
    }
    double coef[6] = {0, 0, 0, 0, 0, 0}, coef_post[6] = {0, 0, 0, 0, 0, 0};
    bool   conv = false;
    for (int it = 0; it < s.max_loops; it++) {
        double dv[6][3];
        for (int i = 0; i < 6; i++) coef[i] = coef_post[i];
        for (int i = 0; i < 6; i++) {
            for (int k = 0; k < 3; k++) dv[i][k] = d_cur[i][k];
            for (int j = 0; j < 6; j++)
                for (int k = 0; k < 3; k++) dv[i][k] += weight[j][i] * d_prev[j][k] * coef[j];
        }
        double grad[6][6];
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < 6; j++)
                grad[i][j] = 2.0 * weight[j][i] * (dv[j][0] * d_prev[i][0] + dv[j][1] * d_prev[i][1] +
                                                   dv[j][2] * d_prev[i][2]);
        conv = true;
        double sd[6];
        for (int i = 0; i < 6; i++) {
            const double norm = dv[i][0] * dv[i][0] + dv[i][1] * dv[i][1] + dv[i][2] * dv[i][2];
            sd[i]             = D[i] - norm;
            conv              = conv && (fabs(sd[i] / D[i]) < s.tol);
        }
        if (conv) break;
        double result[6];
        for (int i = 0; i < 6; i++) {
            result[i] = sd[i];
            for (int j = 0; j < 6; j++) result[i] += grad[j][i] * coef[j];
        }
        // a vanishing pivot leaves coef_post as it was (the reference ignores the solver's status)
        linear_eq6(grad, coef_post, result);
    }
    if (!conv) atomicAdd(&s.cnt->shake_fail, 1);
    double *x = s.crd;
    for (int k = 0; k < 3; k++) {
        x[3 * a[0] + k] += m[0] * (coef[0] * d_prev[0][k] - coef[2] * d_prev[2][k] - coef[3] * d_prev[3][k]);
        x[3 * a[1] + k] += m[1] * (-coef[0] * d_prev[0][k] + coef[1] * d_prev[1][k] + coef[5] * d_prev[5][k]);
        x[3 * a[2] + k] += m[2] * (-coef[1] * d_prev[1][k] + coef[2] * d_prev[2][k] + coef[4] * d_prev[4][k]);
        x[3 * a[3] + k] += m[3] * (coef[3] * d_prev[3][k] - coef[4] * d_prev[4][k] - coef[5] * d_prev[5][k]);
    }
}

void og_launch_bonded(const og_launch &a) {
    if (a.n_blocks) hipLaunchKernelGGL(k_bonded, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_shake_quad(const og_launch &a) {
    if (a.n_blocks)
        hipLaunchKernelGGL(k_shake_quad, dim3(a.n_blocks), dim3(OG_QUAD_THREADS), 0, a.stream, a.sys, a.map);
}
