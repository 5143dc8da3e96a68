/*
 * oracle_integ.c -- CPU restatement of the reference FP64 integration,
 * constraint and thermostat primitives, plus a textbook analytic SETTLE.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by the product).
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

static const double FORCE_VEL     = 4.184e-4;                       /* CelesteObject.cpp:37 */
static const double GAS_CONST     = 8.31451;                        /* CelesteObject.cpp:38 */
static const double KINETIC_COEFF = (1e7 / (4.184 * 1e3)) * 0.5;    /* CelesteObject.cpp:40 */

/* SubBox::update_velocities (SubBox.cpp:1445-1454) */
void orc_update_velocities(int n, double dt, const double *vel, const double *work, const float *mass_inv,
                           double *vel_next) {
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++)
            vel_next[i * 3 + d] = vel[i * 3 + d] + (dt * -FORCE_VEL * work[i * 3 + d] * (double)mass_inv[i]);
}

/* SubBox::update_coordinates (SubBox.cpp:1567-1580) */
void orc_update_coordinates(int n, double dt, const double *crd_prev, const double *vel, double *crd) {
    for (int i = 0; i < 3 * n; i++) {
        double diff = vel[i] * dt;
        crd[i]      = crd_prev[i] + diff;
    }
}

/* PBC::diff_crd_minim_image, double overload (PBC.cpp:62-82) */
static void mimg(double d[3], const double *x1, const double *x2, const float L[3]) {
    for (int i = 0; i < 3; i++) {
        float L_half     = (float)(L[i] * 0.5);
        float L_half_inv = (float)(1.0 / L_half);
        d[i]             = x1[i] - x2[i];
        d[i]             = d[i] - (double)(L[i] * (float)(int)(d[i] * (double)L_half_inv));
        d[i]             = d[i] - (double)(L[i] * (float)(int)(d[i] * (double)L_half_inv));
        /* within rounding of +-L the two passes swap -L and +L: one nearest-image step more */
        if (fabs(d[i]) > 0.75 * (double)L[i]) d[i] -= (double)L[i] * nearbyint(d[i] / (double)L[i]);
    }
}

/* ConstraintShake::shake_pair (ConstraintShake.cpp:18-49) */
void orc_shake_pair(const float L[3], int n, const int *ids, const double *dist2, double *crd,
                    const double *crd_prev, const float *mass_inv) {
    for (int k = 0; k < n; k++) {
        int    a = ids[2 * k], b = ids[2 * k + 1];
        double d12[3], d12p[3];
        mimg(d12, crd + b * 3, crd + a * 3, L);
        mimg(d12p, crd_prev + b * 3, crd_prev + a * 3, L);
        double m1 = mass_inv[a], m2 = mass_inv[b], m12 = m1 + m2;
        double dsq  = d12[0] * d12[0] + d12[1] * d12[1] + d12[2] * d12[2];
        double dsqp = d12p[0] * d12p[0] + d12p[1] * d12p[1] + d12p[2] * d12p[2];
        double dip  = d12[0] * d12p[0] + d12[1] * d12p[1] + d12[2] * d12p[2];
        double va = m12 * m12 * dsqp, vb = m12 * dip, vc = dsq - dist2[k];
        double disc = vb * vb - vc * va;
        double vg   = (vb - sqrt(disc > 0.0 ? disc : 0.0)) / va;
        for (int d = 0; d < 3; d++) {
            crd[a * 3 + d] += m1 * vg * d12p[d];
            crd[b * 3 + d] -= m2 * vg * d12p[d];
        }
    }
}

/* ConstraintShake::shake_trio (ConstraintShake.cpp:51-149): Newton iteration on
 * the three squared-distance constraints with a Cramer-rule 3x3 solve. */
int orc_shake_trio(const float L[3], int n, const int *ids, const double *dist2, double *crd,
                   const double *crd_prev, const float *mass_inv, double tol, int max_loops) {
    int n_fail = 0;
    for (int k = 0; k < n; k++) {
        int     a1 = ids[3 * k], a2 = ids[3 * k + 1], a3 = ids[3 * k + 2];
        double  d12[3], d23[3], d31[3], p12[3], p23[3], p31[3];
        double  g[6] = {0, 0, 0, 0, 0, 0};
        const double *D = dist2 + 3 * k;
        mimg(d12, crd + a1 * 3, crd + a2 * 3, L);
        mimg(p12, crd_prev + a1 * 3, crd_prev + a2 * 3, L);
        mimg(d23, crd + a2 * 3, crd + a3 * 3, L);
        mimg(p23, crd_prev + a2 * 3, crd_prev + a3 * 3, L);
        mimg(d31, crd + a3 * 3, crd + a1 * 3, L);
        mimg(p31, crd_prev + a3 * 3, crd_prev + a1 * 3, L);
        double m1 = mass_inv[a1], m2 = mass_inv[a2], m3 = mass_inv[a3];
        double c11 = m1 + m2, c12 = -m2, c13 = -m1;
        double c21 = -m2, c22 = m2 + m3, c23 = -m3;
        double c31 = -m1, c32 = -m3, c33 = m3 + m1;
        int conv = 0;
        for (int it = 0; it < max_loops; it++) {
            for (int d = 0; d < 3; d++) {
                d12[d] += (c11 * p12[d] * g[0] + c12 * p23[d] * g[1] + c13 * p31[d] * g[2]);
                d23[d] += (c21 * p12[d] * g[0] + c22 * p23[d] * g[1] + c23 * p31[d] * g[2]);
                d31[d] += (c31 * p12[d] * g[0] + c32 * p23[d] * g[1] + c33 * p31[d] * g[2]);
            }
#define DOT(u, v) (u[0] * v[0] + u[1] * v[1] + u[2] * v[2])
            double f11 = 2.0 * c11 * DOT(d12, p12), f12 = 2.0 * c12 * DOT(d12, p23), f13 = 2.0 * c13 * DOT(d12, p31);
            double f21 = 2.0 * c21 * DOT(d23, p12), f22 = 2.0 * c22 * DOT(d23, p23), f23 = 2.0 * c23 * DOT(d23, p31);
            double f31 = 2.0 * c31 * DOT(d31, p12), f32 = 2.0 * c32 * DOT(d31, p23), f33 = 2.0 * c33 * DOT(d31, p31);
            double diff[3] = {D[0] - DOT(d12, d12), D[1] - DOT(d23, d23), D[2] - DOT(d31, d31)};
#undef DOT
            if (fabs(diff[0] / D[0]) < tol && fabs(diff[1] / D[1]) < tol && fabs(diff[2] / D[2]) < tol) {
                conv = 1;
                break;
            }
            double det = 1.0 / (f11 * f22 * f33 + f12 * f23 * f31 + f13 * f21 * f32 - f11 * f23 * f32
                                - f12 * f21 * f33 - f13 * f22 * f31);
            double b11 = f22 * f33 - f23 * f32, b12 = f32 * f13 - f33 * f12, b13 = f12 * f23 - f13 * f22;
            double b21 = f23 * f31 - f21 * f33, b22 = f33 * f11 - f31 * f13, b23 = f13 * f21 - f11 * f23;
            double b31 = f21 * f32 - f22 * f31, b32 = f31 * f12 - f32 * f11, b33 = f11 * f22 - f12 * f21;
            g[0] = (b11 * diff[0] + b12 * diff[1] + b13 * diff[2]) * det;
            g[1] = (b21 * diff[0] + b22 * diff[1] + b23 * diff[2]) * det;
            g[2] = (b31 * diff[0] + b32 * diff[1] + b33 * diff[2]) * det;
            g[3] += g[0];
            g[4] += g[1];
            g[5] += g[2];
        }
        if (!conv) n_fail++;
        for (int d = 0; d < 3; d++) {
            crd[a1 * 3 + d] += m1 * (g[3] * p12[d] - g[5] * p31[d]);
            crd[a2 * 3 + d] += m2 * (-g[3] * p12[d] + g[4] * p23[d]);
            crd[a3 * 3 + d] += m3 * (-g[4] * p23[d] + g[5] * p31[d]);
        }
    }
    return n_fail;
}

/* Analytic SETTLE (S. Miyamoto and P. A. Kollman, J. Comput. Chem. 13, 952 (1992)).
 * Works on minimum-image unwrapped copies of each water, relative to its oxygen. */
int orc_settle(const float L[3], int n, const int *ids, double dOH, double dHH, double *crd,
               const double *crd_prev, const float *mass) {
    int n_fail = 0;
    for (int k = 0; k < n; k++) {
        int    iO = ids[3 * k], iH1 = ids[3 * k + 1], iH2 = ids[3 * k + 2];
        double mO = mass[iO], mH = mass[iH1];
        double wt = mO + 2.0 * mH;
        double rc = 0.5 * dHH;
        double h  = sqrt(dOH * dOH - rc * rc);
        double ra = 2.0 * mH * h / wt;
        double rb = h - ra;
        /* old geometry relative to O (minimum image) */
        double b0[3], c0[3], a1[3], b1[3], c1[3];
        mimg(b0, crd_prev + iH1 * 3, crd_prev + iO * 3, L);
        mimg(c0, crd_prev + iH2 * 3, crd_prev + iO * 3, L);
        /* new unconstrained positions relative to new O, then to their centre of mass */
        double xb[3], xc[3];
        mimg(xb, crd + iH1 * 3, crd + iO * 3, L);
        mimg(xc, crd + iH2 * 3, crd + iO * 3, L);
        double com[3];
        for (int d = 0; d < 3; d++) {
            com[d] = (mH * xb[d] + mH * xc[d]) / wt; /* relative to new O */
            a1[d]  = -com[d];
            b1[d]  = xb[d] - com[d];
            c1[d]  = xc[d] - com[d];
        }
        /* frame: Z normal to the old plane, X perpendicular to a1 and Z, Y = Z x X */
        double Z[3] = {b0[1] * c0[2] - b0[2] * c0[1], b0[2] * c0[0] - b0[0] * c0[2], b0[0] * c0[1] - b0[1] * c0[0]};
        double X[3] = {a1[1] * Z[2] - a1[2] * Z[1], a1[2] * Z[0] - a1[0] * Z[2], a1[0] * Z[1] - a1[1] * Z[0]};
        double Y[3] = {Z[1] * X[2] - Z[2] * X[1], Z[2] * X[0] - Z[0] * X[2], Z[0] * X[1] - Z[1] * X[0]};
        double nx = 1.0 / sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
        double ny = 1.0 / sqrt(Y[0] * Y[0] + Y[1] * Y[1] + Y[2] * Y[2]);
        double nz = 1.0 / sqrt(Z[0] * Z[0] + Z[1] * Z[1] + Z[2] * Z[2]);
        for (int d = 0; d < 3; d++) {
            X[d] *= nx;
            Y[d] *= ny;
            Z[d] *= nz;
        }
#define PX(v) (X[0] * v[0] + X[1] * v[1] + X[2] * v[2])
#define PY(v) (Y[0] * v[0] + Y[1] * v[1] + Y[2] * v[2])
#define PZ(v) (Z[0] * v[0] + Z[1] * v[1] + Z[2] * v[2])
        double xb0 = PX(b0), yb0 = PY(b0), xc0 = PX(c0), yc0 = PY(c0);
        double za1 = PZ(a1);
        double xb1 = PX(b1), yb1 = PY(b1), zb1 = PZ(b1);
        double xc1 = PX(c1), yc1 = PY(c1), zc1 = PZ(c1);
#undef PX
#undef PY
#undef PZ
        double sinphi = za1 / ra;
        double q1     = 1.0 - sinphi * sinphi;
        double cosphi = sqrt(q1);
        double sinpsi = (zb1 - zc1) / (2.0 * rc * cosphi);
        double q2     = 1.0 - sinpsi * sinpsi;
        double cospsi = sqrt(q2);
        double ya2 = ra * cosphi;
        double xb2 = -rc * cospsi;
        double t1 = -rb * cosphi, t2 = rc * sinpsi * sinphi;
        double yb2 = t1 - t2, yc2 = t1 + t2;
        double alpa   = xb2 * (xb0 - xc0) + yb0 * yb2 + yc0 * yc2;
        double beta   = xb2 * (yc0 - yb0) + xb0 * yb2 + xc0 * yc2;
        double gama   = xb0 * yb1 - xb1 * yb0 + xc0 * yc1 - xc1 * yc0;
        double al2be2 = alpa * alpa + beta * beta;
        double q3     = al2be2 - gama * gama;
        double sinthe = (alpa * gama - beta * sqrt(q3)) / al2be2;
        double q4     = 1.0 - sinthe * sinthe;
        double costhe = sqrt(q4);
        /* no rigid solution: leave the water unconstrained and count it */
        if (!(q1 > 0.0 && q2 >= 0.0 && q3 >= 0.0 && q4 >= 0.0)) {
            n_fail++;
            continue;
        }
        double a3[3] = {-ya2 * sinthe, ya2 * costhe, za1};
        double b3[3] = {xb2 * costhe - yb2 * sinthe, xb2 * sinthe + yb2 * costhe, zb1};
        double c3[3] = {-xb2 * costhe - yc2 * sinthe, -xb2 * sinthe + yc2 * costhe, zc1};
        /* each atom moves by its correction, so it stays in its own periodic frame */
        for (int d = 0; d < 3; d++) {
            double ca = com[d] + X[d] * a3[0] + Y[d] * a3[1] + Z[d] * a3[2];
            double cb = com[d] + X[d] * b3[0] + Y[d] * b3[1] + Z[d] * b3[2];
            double cc = com[d] + X[d] * c3[0] + Y[d] * c3[1] + Z[d] * c3[2];
            crd[iO * 3 + d] += ca;
            crd[iH1 * 3 + d] += cb - xb[d];
            crd[iH2 * 3 + d] += cc - xc[d];
        }
    }
    return n_fail;
}

/* SubBox::set_velocity_from_crd (SubBox.cpp:1462-1481) */
void orc_velocity_from_crd(int n, double dt_inv, const double *crd, const double *crd_prev, double *vel_next) {
    for (int i = 0; i < 3 * n; i++) vel_next[i] = (crd[i] - crd_prev[i]) * dt_inv;
}

/* SubBox::revise_coordinates_pbc (SubBox.cpp:1482-1502) */
void orc_revise_pbc(int n, const float L[3], const float lb[3], double *crd) {
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            float ub = lb[d] + L[d];
            if (crd[i * 3 + d] >= ub) crd[i * 3 + d] -= L[d];
            else if (crd[i * 3 + d] < lb[d]) crd[i * 3 + d] += L[d];
        }
}

/* COMMotion::cancel_translation (COMMotion.cpp:30-55) */
void orc_cancel_com(int n_groups, const int *off, const int *atoms, const float *minv_g, const float *mass,
                    double *vel) {
    for (int g = 0; g < n_groups; g++) {
        double mom[3] = {0, 0, 0};
        for (int k = off[g]; k < off[g + 1]; k++)
            for (int d = 0; d < 3; d++) mom[d] += vel[atoms[k] * 3 + d] * mass[atoms[k]];
        for (int d = 0; d < 3; d++) mom[d] *= minv_g[g];
        for (int k = off[g]; k < off[g + 1]; k++)
            for (int d = 0; d < 3; d++) vel[atoms[k] * 3 + d] -= mom[d];
    }
}

/* DynamicsMode::cal_kinetic_energy (DynamicsMode.cpp:358-372) on vel_just from
 * SubBox::copy_vel_just (SubBox.cpp:1504-1516). */
double orc_kinetic(int n, const float *mass, const double *vel, const double *vel_next) {
    double kpre = 0.0;
    for (int i = 0; i < n; i++) {
        double ka = 0.0;
        for (int d = 0; d < 3; d++) {
            double vj = (vel[i * 3 + d] + vel_next[i * 3 + d]) * 0.5;
            ka += vj * vj;
        }
        kpre += ka * mass[i];
    }
    return kpre * KINETIC_COEFF;
}

/* SubBox::update_coordinates_langevin (SubBox.cpp:2077-2114) */
void orc_langevin_positions(int n, double dt_half, double gamma, double T, const float L[3],
                            const double *crd_prev, const double *crd_prev2, const double *work,
                            const float *mass, const float *mass_inv, const double *zeta, double *crd) {
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double det_f = -FORCE_VEL * work[i * 3 + d];
            double stc_f = zeta[i * 3 + d] * sqrt(GAS_CONST * T * gamma * (double)mass[i] / dt_half * 1e-7);
            double p1 = crd_prev[i * 3 + d], p2 = crd_prev2[i * 3 + d];
            p2 += nearbyint((p1 - p2) / (double)L[d]) * (double)L[d];
            crd[i * 3 + d] = 1.0 / (1.0 + gamma * dt_half)
                             * (2 * p1 - p2 + gamma * dt_half * p2
                                + 4 * dt_half * dt_half * (double)mass_inv[i] * (det_f + stc_f));
        }
}

/* ThermostatScaling::apply_thermostat (Thermostat.cpp:57-84): the kinetic
 * energy of the on-step velocities v(t-dt/2) + dt/2 a(t) sets one scale
 * factor, then v(t+dt/2) = (2 s - 1) v(t-dt/2) + s dt a(t).  `real` is double,
 * mass and mass_inv are the float arrays (SubBox.cpp:691). */
double orc_thermostat_scaling(int n, double dt, double tau_inv, double temperature, double temperature_coeff,
                              const double *work, const double *vel, const float *mass, const float *mass_inv,
                              double *vel_next) {
    double kine_pre = 0.0;
    for (int i = 0; i < n; i++) {
        double vel_norm = 0;
        for (int d = 0; d < 3; d++) {
            double vel_tmp = vel[i * 3 + d] - 0.5 * (dt * work[i * 3 + d] * FORCE_VEL * mass_inv[i]);
            vel_norm += vel_tmp * vel_tmp;
        }
        kine_pre += mass[i] * vel_norm;
    }
    double dt_temperature = kine_pre * temperature_coeff;
    double scale          = sqrt(1 + dt * tau_inv * (temperature / dt_temperature - 1));
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double vel_diff      = dt * -FORCE_VEL * work[i * 3 + d] * mass_inv[i];
            vel_next[i * 3 + d] = (2.0 * scale - 1.0) * vel[i * 3 + d] + scale * vel_diff;
        }
    return scale;
}

/* SubBox::set_velocities_just_langevin (SubBox.cpp:2054-2066) */
void orc_langevin_velocity(int n, double dt, const float L[3], const double *crd, const double *crd_prev2,
                           double *vel) {
    const double dt2inv = 1.0 / (2 * dt);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double p1 = crd[i * 3 + d], p2 = crd_prev2[i * 3 + d];
            p2 += nearbyint((p1 - p2) / (double)L[d]) * (double)L[d];
            vel[i * 3 + d] = (p1 - p2) * dt2inv;
        }
}

/* ---- VcMD 1-D virtual-state bias ---------------------------------------
 * ExtendedVcMD::set_crd_centers (Extend.cpp:753-797), set_struct_parameters_mass_center
 * (:668-692), scale_force (:1035-1102) and the visit count / transition
 * candidates of vs_step + set_vs_next (:808-832, :846-880), for one
 * dimension with group a = enhance pair[0] and b = pair[1].
 * crdg[(na+nb)*3] holds the unwrapped group coordinates between calls.
 * minv_a/minv_b are the float group inverse masses (MmSystem.cpp:572-582).
 * Returns the bias energy coef*excess^2 and writes lambda, whether lambda is
 * inside the current state's range, and the transition candidates
 * cand[0] = vs, cand[1] = neighbour or -1. */
/* Group coordinates unwrapped against their previous values, the two mass
 * centres and their minimum-image difference (ExtendedVcMD::set_crd_centers
 * Extend.cpp:753-797 and ExtendedVAUS::set_crd_centers :366-413 are the same
 * loop); returns |diff|. */
static double group_diff(const float L[3], int na, const int *atoms_a, int nb, const int *atoms_b, double *crdg,
                         const double *crd, const float *mass, float minv_a, float minv_b, double diff[3]) {
    double c[2][3] = {{0, 0, 0}, {0, 0, 0}};
    for (int g = 0; g < 2; g++) {
        const int  n   = g == 0 ? na : nb;
        const int *ids = g == 0 ? atoms_a : atoms_b;
        double    *cg  = crdg + (g == 0 ? 0 : 3 * na);
        for (int i = 0; i < n; i++) {
            const int a = ids[i];
            for (int d = 0; d < 3; d++) {
                const float Lh   = L[d] * 0.5f;
                double      tmp  = crd[3 * a + d];
                double      dif  = tmp - cg[3 * i + d];
                while (dif > Lh) {
                    dif -= L[d];
                    tmp -= L[d];
                }
                while (-dif > Lh) {
                    dif += L[d];
                    tmp += L[d];
                }
                cg[3 * i + d] = tmp;
                c[g][d] += tmp * mass[a];
            }
        }
    }
    for (int d = 0; d < 3; d++) {
        c[0][d] *= minv_a;
        c[1][d] *= minv_b;
        double      v   = c[0][d] - c[1][d];
        const float Lhi = 1.0f / (L[d] * 0.5f);
        /* PBC::diff_crd_minim_image, double overload (PBC.cpp:62-82) */
        v       = v - L[d] * (float)(int)(v * Lhi);
        v       = v - L[d] * (float)(int)(v * Lhi);
        if (fabs(v) > 0.75 * (double)L[d]) v -= (double)L[d] * nearbyint(v / (double)L[d]);
        diff[d] = v;
    }
    return sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
}

double orc_vcmd_bias(const float L[3], int na, const int *atoms_a, int nb, const int *atoms_b, double *crdg,
                     const double *crd, const float *mass, float minv_a, float minv_b, int n_states,
                     const double *lo, const double *hi, int vs, double sigma, double coef, double temperature,
                     double *work, double *lambda_out, int *in_range, int cand[2]) {
    double       diff[3];
    const double lam = group_diff(L, na, atoms_a, nb, atoms_b, crdg, crd, mass, minv_a, minv_b, diff);
    double       u[3];
    for (int d = 0; d < 3; d++) u[d] = diff[d] / lam;
    double ex = 0.0, rec = 0.0;
    if (lam < lo[vs] - sigma) {
        ex  = lam - (lo[vs] - sigma);
        rec = coef * ex;
    } else if (lam >= hi[vs] + sigma) {
        ex  = lam - (hi[vs] + sigma);
        rec = coef * ex;
    }
    const double const_k = (8.31451 / 4.184) * 1e-3 * temperature; /* Extend.cpp:573-576 */
    const double dew     = const_k * rec;
    double       dir     = 1.0;
    for (int g = 0; g < 2; g++) {
        const int   n    = g == 0 ? na : nb;
        const int  *ids  = g == 0 ? atoms_a : atoms_b;
        const float minv = g == 0 ? minv_a : minv_b;
        double      bias[3];
        for (int d = 0; d < 3; d++) bias[d] = dew * u[d] * (double)minv;
        for (int i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) work[3 * ids[i] + d] += dir * bias[d] * (double)mass[ids[i]];
        dir *= -1.0;
    }
    *lambda_out = lam;
    *in_range   = lam >= lo[vs] && lam < hi[vs];
    cand[0]     = vs;
    cand[1]     = -1;
    if (vs > 0) {
        if (lam < hi[vs - 1]) cand[1] = vs - 1;
        else if (vs < n_states - 1 && lam >= lo[vs + 1]) cand[1] = vs + 1;
    } else if (n_states > 1 && lam >= lo[1]) {
        cand[1] = 1;
    }
    return rec * ex;
}

/* ---- V-McMD / V-AUS virtual states --------------------------------------
 * ExtendedVMcMD::trial_transition + set_current_vstate (Extend.cpp:143-176):
 * from state vs, a move up is tried first, then a move down; a move needs
 * lambda inside both ranges (closed intervals) and its uniform draw above
 * 1 - the state's probability for that direction.  u_up / u_down are the
 * draws of the two trials.  Returns the new state. */
static int vs_trial(int n, const double *lo, const double *hi, const double *p_down, const double *p_up, int src,
                    int rel, double lambda, double u) {
    if (!(lambda >= lo[src] && lambda <= hi[src])) return src;
    if (src == 0 && rel == -1) return src;
    if (src == n - 1 && rel == 1) return src;
    const int dst = src + rel;
    if (lambda >= lo[dst] && lambda <= hi[dst]) {
        const double p = rel == -1 ? p_down[src] : p_up[src];
        if (u > 1.0 - p) return dst;
    }
    return src;
}
int orc_vs_transition(int n, const double *lo, const double *hi, const double *p_down, const double *p_up, int vs,
                      double lambda, double u_up, double u_down) {
    const int up = vs_trial(n, lo, hi, p_down, p_up, vs, 1, lambda, u_up);
    if (up != vs) return up;
    return vs_trial(n, lo, hi, p_down, p_up, vs, -1, lambda, u_down);
}

/* ExtendedVMcMD::scale_force (Extend.cpp:177-200): the factor the whole
 * gradient is multiplied by, k_B T (kcal/mol) times the state's polynomial
 * d ln P / d lambda at lambda clamped to the state's range (set_temperature,
 * :119-122). */
double orc_vmcmd_factor(const double *poly, int order, double lo, double hi, double temperature, double lambda) {
    double param = lambda;
    if (lambda <= lo) param = lo;
    else if (lambda >= hi) param = hi;
    double d_ln_p = poly[0], t = 1.0;
    for (int i = 1; i < order + 1; i++) {
        t *= param;
        d_ln_p += poly[i] * t;
    }
    const double const_k = (8.31451 / 4.184) * 1e-3 * temperature;
    return const_k * d_ln_p;
}

/* ExtendedVAUS::cal_struct_parameters + scale_force (Extend.cpp:427-540) for
 * one pair of groups: lambda = distance between the mass centres; outside
 * [lo - sigma, hi + sigma] a harmonic recovery term; the force
 * k_B T (d ln P(lambda clamped) + recovery) along the unit vector, weighted by
 * m_i / M_group, + on group a and - on group b.  Returns lambda. */
double orc_vaus_bias(const float L[3], int na, const int *atoms_a, int nb, const int *atoms_b, double *crdg,
                     const double *crd, const float *mass, float minv_a, float minv_b, const double *poly,
                     int order, double lo, double hi, double sigma, double recov_coef, double temperature,
                     double *work) {
    double       diff[3];
    const double lam = group_diff(L, na, atoms_a, nb, atoms_b, crdg, crd, mass, minv_a, minv_b, diff);
    double       u[3];
    for (int d = 0; d < 3; d++) u[d] = diff[d] / lam;
    double param = lam, recovery = 0.0;
    if (param <= lo) {
        if (param <= lo - sigma) recovery = recov_coef * (param - (lo - sigma));
        param = lo;
    } else if (param >= hi) {
        if (param >= hi + sigma) recovery = recov_coef * (param - (hi + sigma));
        param = hi;
    }
    double t = 1.0, d_ln_p = poly[0];
    for (int i = 1; i <= order; i++) {
        t *= param;
        d_ln_p += poly[i] * t;
    }
    const double const_k = (8.31451 / 4.184) * 1e-3 * temperature;
    const double dew     = const_k * (d_ln_p + recovery);
    double       dir     = 1.0;
    for (int g = 0; g < 2; g++) {
        const int   n    = g == 0 ? na : nb;
        const int  *ids  = g == 0 ? atoms_a : atoms_b;
        const float minv = g == 0 ? minv_a : minv_b;
        double      bias[3];
        for (int d = 0; d < 3; d++) bias[d] = dew * u[d] * (double)minv;
        for (int i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) work[3 * ids[i] + d] += dir * bias[d] * (double)mass[ids[i]];
        dir *= -1.0;
    }
    return lam;
}
