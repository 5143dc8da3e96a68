/*
 * oracle_bonded.c -- CPU restatement of the reference's bonded terms beyond
 * bonds (angles, torsions, impropers, 1-4 pairs) and of SHAKE on four-atom
 * groups.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by the product).
 * Compiled with -ffp-contract=off.  The reference evaluates these terms in
 * double (real / real_bp, CelesteObject.h:11-24) on double coordinates, with
 * PBC's float box lengths in the minimum image.
 */
#include "oracle.h"
#include <math.h>

static const double CHARGE_COEFF = 332.06378; /* CelesteObject.cpp:36 */
static const double PI_REF       = 3.14159265; /* CelesteObject.cpp:26 */
static const double EPS_REF      = 1e-10;      /* CelesteObject.cpp:31 */
static const double EPS3_REF     = 1e-30;      /* CelesteObject.cpp:31 */

/* PBC::diff_crd_minim_image, double overload (PBC.cpp:62-82) */
static void mimg(double d[3], const double *x1, const double *x2, const float L[3]) {
    for (int i = 0; i < 3; i++) {
        float L_half     = (float)(L[i] * 0.5);
        float L_half_inv = (float)(1.0 / L_half);
        d[i]             = x1[i] - x2[i];
        d[i]             = d[i] - (double)L[i] * (int)(d[i] * (double)L_half_inv);
        d[i]             = d[i] - (double)L[i] * (int)(d[i] * (double)L_half_inv);
        /* within rounding of +-L the two passes swap -L and +L: one nearest-image step more */
        if (fabs(d[i]) > 0.75 * (double)L[i]) d[i] -= (double)L[i] * nearbyint(d[i] / (double)L[i]);
    }
}

static void cross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* SubBox::calc_energy_angles (SubBox.cpp:1250-1270) + ForceField::calc_angle
 * (ForceField.cpp:138-177): E = k (theta - theta0)^2. */
double orc_angles(const float L[3], int n, const int *ids, const double *k_eps, const double *theta0,
                  const double *crd, double *work) {
    double e_sum = 0.0;
    for (int t = 0; t < n; t++) {
        const int a1 = ids[3 * t], a2 = ids[3 * t + 1], a3 = ids[3 * t + 2];
        double    d12[3], d32[3];
        mimg(d12, crd + 3 * a1, crd + 3 * a2, L);
        mimg(d32, crd + 3 * a3, crd + 3 * a2, L);
        double r12_2      = d12[0] * d12[0] + d12[1] * d12[1] + d12[2] * d12[2];
        double r32_2      = d32[0] * d32[0] + d32[1] * d32[1] + d32[2] * d32[2];
        double r12r32     = sqrt(r12_2 * r32_2);
        double r12r32_inv = 1.0 / r12r32;
        double cost       = (d12[0] * d32[0] + d12[1] * d32[1] + d12[2] * d32[2]) * r12r32_inv;
        cost              = fmax(fmin(cost, 1.0), -1.0);
        double theta      = acos(cost);
        double dif        = theta - theta0[t];
        double coef       = k_eps[t] * dif;
        e_sum += coef * dif;
        double r12inv_2  = 1.0 / r12_2;
        double r32inv_2  = 1.0 / r32_2;
        double sin_theta = sin(theta);
        sin_theta        = fmax(EPS_REF, sin_theta);
        double grad_coef = -(2.0 * coef) / sin_theta;
        double w1[3], w2[3];
        for (int i = 0; i < 3; i++) {
            w1[i] = grad_coef * (d32[i] * r12r32_inv - d12[i] * cost * r12inv_2);
            w2[i] = grad_coef * (d12[i] * r12r32_inv - d32[i] * cost * r32inv_2);
        }
        for (int d = 0; d < 3; d++) {
            work[3 * a1 + d] += w1[d];
            work[3 * a2 + d] -= w1[d] + w2[d];
            work[3 * a3 + d] += w2[d];
        }
    }
    return e_sum;
}

/* ForceField::calc_torsion (ForceField.cpp:179-299), for torsions and
 * impropers (SubBox.cpp:1272-1321): E = V (1 + cos(n phi - phase)) / overlaps. */
static void torsion_term(const float L[3], const double *c1, const double *c2, const double *c3, const double *c4,
                         double ene_p, double overlaps, double symmetry, double phase, double *ene, double w1[3],
                         double w2[3], double w3[3]) {
    static const double gmul[11] = {0.0, 0.0, 2.0, 0.0, 4.0, 0.0, 6.0, 0.0, 8.0, 0.0, 10.0}; /* ForceField.cpp:7-17 */
    *ene = 0.0;
    for (int d = 0; d < 3; d++) w1[d] = w2[d] = w3[d] = 0.0;
    if (ene_p <= EPS_REF) return;
    double d21[3], d32[3], d43[3];
    mimg(d21, c2, c1, L);
    mimg(d32, c3, c2, L);
    mimg(d43, c4, c3, L);
    double p12[3], p23[3];
    cross(d21, d32, p12);
    cross(d32, d43, p23);
    double p12_2  = p12[0] * p12[0] + p12[1] * p12[1] + p12[2] * p12[2];
    double p23_2  = p23[0] * p23[0] + p23[1] * p23[1] + p23[2] * p23[2];
    double p12p23 = sqrt(p12_2 * p23_2);
    if (p12p23 <= 0.0) return;
    double p12p23inv = 1.0 / p12p23;
    double s1223     = p12[0] * p23[0] + p12[1] * p23[1] + p12[2] * p23[2];
    double cosp      = s1223 * p12p23inv;
    cosp             = fmax(-1.0, fmin(1.0, cosp));
    double p123[3];
    cross(p12, p23, p123);
    double phi    = acos(cosp);
    double s32123 = d32[0] * p123[0] + d32[1] * p123[1] + d32[2] * p123[2];
    if (s32123 < 0.0) phi = 2.0 * PI_REF - phi;
    double phimod = symmetry * phi - phase;
    double div    = 1.0 / overlaps;
    *ene          = ene_p * (1.0 + cos(phimod)) * div;
    double sinp     = sin(phi);
    int    nrot     = (int)symmetry;
    double sin_fphs = sin(phase);
    double cos_fphs = cos(phase);
    double sin_rot  = sin(symmetry * phi);
    double cos_rot  = cos(symmetry * phi);
    double dums     = sinp;
    double dflim    = cos_fphs * (symmetry - gmul[nrot] + gmul[nrot] * cosp);
    double df0      = cos_fphs * sin_rot - sin_fphs * cos_rot;
    double df1      = dflim;
    if (EPS_REF <= fabs(dums)) df1 = df0 / dums;
    double work_coeff = symmetry * ene_p * div * df1 * p12p23inv;
    double r12        = 1.0 / p12_2 * s1223;
    double r23        = 1.0 / p23_2 * s1223;
    double op_p23d32[3], op_p12d32[3];
    cross(p23, d32, op_p23d32);
    cross(p12, d32, op_p12d32);
    for (int i = 0; i < 3; i++) w1[i] = work_coeff * (op_p23d32[i] - r12 * op_p12d32[i]);
    double d2132[3] = {d21[0] + d32[0], d21[1] + d32[1], d21[2] + d32[2]};
    double op_d2132_p23[3], op_d2132_p12[3], op_p12_d43[3], op_p23_d43[3];
    cross(d2132, p23, op_d2132_p23);
    cross(d2132, p12, op_d2132_p12);
    cross(p12, d43, op_p12_d43);
    cross(p23, d43, op_p23_d43);
    for (int i = 0; i < 3; i++)
        w2[i] = work_coeff * (op_d2132_p23[i] + op_p12_d43[i] - r12 * op_d2132_p12[i] - r23 * op_p23_d43[i]);
    double op_p12_d32[3], op_p23_d32[3];
    cross(p12, d32, op_p12_d32);
    cross(p23, d32, op_p23_d32);
    for (int i = 0; i < 3; i++) w3[i] = work_coeff * (op_p12_d32[i] - r23 * op_p23_d32[i]);
}

double orc_torsions(const float L[3], int n, const int *ids, const double *ene, const int *overlaps,
                    const int *symmetry, const double *phase, const double *crd, double *work) {
    double e_sum = 0.0;
    for (int t = 0; t < n; t++) {
        const int *a = ids + 4 * t;
        double     e, w1[3], w2[3], w3[3];
        torsion_term(L, crd + 3 * a[0], crd + 3 * a[1], crd + 3 * a[2], crd + 3 * a[3], ene[t], (double)overlaps[t],
                     (double)symmetry[t], phase[t], &e, w1, w2, w3);
        e_sum += e;
        for (int d = 0; d < 3; d++) {
            work[3 * a[0] + d] += w1[d];
            work[3 * a[1] + d] += w2[d];
            work[3 * a[2] + d] -= w1[d] + w2[d] + w3[d];
            work[3 * a[3] + d] += w3[d];
        }
    }
    return e_sum;
}

/* SubBox::calc_energy_14nb (SubBox.cpp:1323-1346) + ForceField::calc_14pair
 * (ForceField.cpp:301-331): the entry's own type pair, float LJ table and
 * charges promoted to double. */
void orc_nb14(const float L[3], int n, const int *ids, const int *types, const double *coeff_vdw,
              const double *coeff_ele, const float *lj6, const float *lj12, int n_types, const float *charge,
              const double *crd, double *work, double *e_vdw, double *e_ele) {
    *e_vdw = *e_ele = 0.0;
    for (int t = 0; t < n; t++) {
        const int a1 = ids[2 * t], a2 = ids[2 * t + 1];
        const int tp = types[2 * t] * n_types + types[2 * t + 1];
        double    d14[3];
        mimg(d14, crd + 3 * a1, crd + 3 * a2, L);
        double r14_2      = d14[0] * d14[0] + d14[1] * d14[1] + d14[2] * d14[2];
        double r14_2_inv  = 1.0 / r14_2;
        double r14_inv    = sqrt(r14_2_inv);
        double r14_inv_6  = r14_2_inv * r14_2_inv * r14_2_inv;
        double r14_inv_12 = r14_inv_6 * r14_inv_6;
        double term6      = (double)lj6[tp] * r14_inv_6;
        double term12     = (double)lj12[tp] * r14_inv_12;
        *e_vdw += (-term6 + term12) * coeff_vdw[t];
        double coef_work = r14_2_inv * coeff_vdw[t] * (-12.0 * term12 + 6.0 * term6);
        double cc        = (double)charge[a1] * (double)charge[a2] * CHARGE_COEFF * coeff_ele[t];
        *e_ele += cc * r14_inv;
        double coef_ele = -cc * r14_2_inv * r14_inv;
        for (int d = 0; d < 3; d++) {
            double w = coef_work * d14[d] + coef_ele * d14[d];
            work[3 * a1 + d] += w;
            work[3 * a2 + d] -= w;
        }
    }
}

/* ConstraintObject::calc_linear_eq (Constraint.cpp:141-203): LU with column
 * pivoting on the transposed index order the reference uses. */
static int linear_eq6(double a[6][6], double x[6], const double b[6]) {
    int pivot[6];
    for (int i = 0; i < 6; i++) pivot[i] = i;
    for (int i = 0; i < 6; i++) {
        int    index = i;
        double max_a = fabs(a[i][pivot[i]]);
        for (int j = i + 1; j < 6; j++) {
            double tmp_a = fabs(a[i][pivot[j]]);
            if (tmp_a > max_a) {
                index = j;
                max_a = tmp_a;
            }
        }
        if (index != i) {
            int tmp      = pivot[i];
            pivot[i]     = pivot[index];
            pivot[index] = tmp;
        }
        if (fabs(a[i][pivot[i]]) < EPS3_REF) return 1;
        a[i][pivot[i]] = 1.0 / a[i][pivot[i]];
        for (int j = i + 1; j < 6; j++) {
            a[i][pivot[j]] *= a[i][pivot[i]];
            for (int k = i + 1; k < 6; k++) a[k][pivot[j]] -= a[i][pivot[j]] * a[k][pivot[i]];
        }
    }
    for (int i = 0; i < 6; i++) {
        double new_x = b[pivot[i]];
        for (int j = 0; j < i; j++) new_x -= a[j][pivot[i]] * x[j];
        x[i] = new_x;
    }
    for (int i = 5; i >= 0; i--) {
        double new_x = x[i];
        for (int j = i + 1; j < 6; j++) new_x -= a[j][pivot[i]] * x[j];
        x[i] = new_x * a[i][pivot[i]];
    }
    return 0;
}

/* ConstraintShake::shake_quad (ConstraintShake.cpp:151-276): Newton iteration
 * on the six squared distances of a four-atom group (pairs 01 12 20 30 23 13,
 * Constraint.cpp:4) with a 6x6 linear solve per iteration.  Returns the number
 * of groups that did not converge within max_loops (the reference prints them). */
int orc_shake_quad(const float L[3], int n, const int *ids, const double *dist2, double *crd, const double *crd_prev,
                   const float *mass_inv_f, double tol, int max_loops) {
    static const int pairs_idx[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {2, 3}, {1, 3}};
    int n_fail = 0;
    for (int q = 0; q < n; q++) {
        const int *a = ids + 4 * q;
        double     m[4];
        double     cur[4][3], prev[4][3];
        for (int i = 0; i < 4; i++) {
            m[i] = (double)mass_inv_f[a[i]];
            for (int d = 0; d < 3; d++) {
                cur[i][d]  = crd[3 * a[i] + d];
                prev[i][d] = crd_prev[3 * a[i] + d];
            }
        }
        const double weight[6][6] = {{m[0] + m[1], -m[1], -m[0], -m[0], 0.0, -m[1]},
                                     {-m[1], m[1] + m[2], -m[2], 0.0, -m[2], m[1]},
                                     {-m[0], -m[2], m[2] + m[0], m[0], m[2], 0.0},
                                     {-m[0], 0.0, m[0], m[3] + m[0], -m[3], -m[3]},
                                     {0.0, -m[2], m[2], -m[3], m[2] + m[3], m[3]},
                                     {-m[1], m[1], 0.0, -m[3], m[3], m[1] + m[3]}};
        double d_cur[6][3], d_prev[6][3];
        for (int p = 0; p < 6; p++) {
            mimg(d_cur[p], cur[pairs_idx[p][0]], cur[pairs_idx[p][1]], L);
            mimg(d_prev[p], prev[pairs_idx[p][0]], prev[pairs_idx[p][1]], L);
        }
        double coef[6] = {0, 0, 0, 0, 0, 0}, coef_post[6] = {0, 0, 0, 0, 0, 0};
        int    converge = 0;
        for (int it = 0; it < max_loops; it++) {
            double dv[6][3];
            for (int i = 0; i < 6; i++) coef[i] = coef_post[i];
            for (int i = 0; i < 6; i++) {
                for (int d = 0; d < 3; d++) dv[i][d] = d_cur[i][d];
                for (int j = 0; j < 6; j++)
                    for (int d = 0; d < 3; d++) dv[i][d] += weight[j][i] * d_prev[j][d] * coef[j];
            }
            double grad[6][6];
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++)
                    grad[i][j] = 2.0 * weight[j][i] *
                                 (dv[j][0] * d_prev[i][0] + dv[j][1] * d_prev[i][1] + dv[j][2] * d_prev[i][2]);
            converge = 1;
            double sd[6];
            for (int i = 0; i < 6; i++) {
                double norm = dv[i][0] * dv[i][0] + dv[i][1] * dv[i][1] + dv[i][2] * dv[i][2];
                sd[i]       = dist2[6 * q + i] - norm;
                converge    = converge && (fabs(sd[i] / dist2[6 * q + i]) < tol);
            }
            if (converge) break;
            double result[6];
            for (int i = 0; i < 6; i++) {
                result[i] = sd[i];
                for (int j = 0; j < 6; j++) result[i] += grad[j][i] * coef[j];
            }
            linear_eq6(grad, coef_post, result);
        }
        if (!converge) n_fail++;
        for (int d = 0; d < 3; d++) {
            crd[3 * a[0] + d] += m[0] * (coef[0] * d_prev[0][d] - coef[2] * d_prev[2][d] - coef[3] * d_prev[3][d]);
            crd[3 * a[1] + d] += m[1] * (-coef[0] * d_prev[0][d] + coef[1] * d_prev[1][d] + coef[5] * d_prev[5][d]);
            crd[3 * a[2] + d] += m[2] * (-coef[1] * d_prev[1][d] + coef[2] * d_prev[2][d] + coef[4] * d_prev[4][d]);
            crd[3 * a[3] + d] += m[3] * (coef[3] * d_prev[3][d] - coef[4] * d_prev[4][d] - coef[5] * d_prev[5][d]);
        }
    }
    return n_fail;
}
