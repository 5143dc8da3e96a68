/*
 * oracle_pair.c -- CPU restatement of the reference LJ + zero-dipole pair
 * path, its constants, self energy, excluded-pair ("excess") terms and bonds.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by the product).
 * Compiled with -ffp-contract=off so every float/double rounding happens
 * where the reference host build rounds it.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NAC ORC_N_ATOM_CELL
static const double CHARGE_COEFF = 332.06378; /* CelesteObject.cpp:36 */
static const double PI_REF       = 3.14159265; /* CelesteObject.cpp:26 */
static const double EPS_REF      = 1e-10;      /* CelesteObject.cpp:31 */

void orc_zd_setup(orc_zd_t *zd, double cutoff, double alpha) { orc_zm_setup(zd, cutoff, alpha, ORC_ZM_DIPOLE, 1); }

/* ZeroMultipoleSum::set_zms_params (ZeroMultipoleSum.cpp:108-190).  Every
 * member is real_pw (float); the double literals promote their expressions,
 * which round to float on assignment. */
void orc_zm_setup(orc_zd_t *zd, double cutoff, double alpha, int mode, int zo_grad_bug) {
    memset(zd, 0, sizeof *zd);
    zd->mode        = mode;
    zd->zo_grad_bug = zo_grad_bug;
    zd->cutoff      = (float)cutoff;
    zd->alpha       = (float)alpha;
    float cutoff_2  = zd->cutoff * zd->cutoff;
    float cutoff_3  = cutoff_2 * zd->cutoff;
    float tmp       = zd->cutoff * zd->alpha;
    float errorfc   = erfcf(tmp);
    float expterm   = expf(-tmp * tmp);
    zd->piewald     = 0.0f;
    if (zd->alpha < EPS_REF) {
        zd->fcoeff = (float)(1.0 / cutoff_3);
        zd->bcoeff = (float)(0.5 * zd->fcoeff);
    } else {
        float root_pi = (float)sqrt(PI_REF);
        zd->piewald   = (float)(2.0 / root_pi * zd->alpha);
        zd->fcoeff    = errorfc / cutoff_3 + zd->piewald * expterm / cutoff_2;
        zd->bcoeff    = (float)(0.5 * zd->fcoeff);
    }
    zd->zcore = errorfc / zd->cutoff + zd->bcoeff * cutoff_2;
    /* the higher orders: derivative coefficients of erfc(a r)/r at the cutoff, then the polynomial
     * coefficients and cores */
    const float c = zd->cutoff, c4 = cutoff_2 * cutoff_2, c5 = cutoff_3 * cutoff_2, c6 = cutoff_3 * cutoff_3,
                c7 = cutoff_3 * c4;
    float hzqrc, d1, d2, d3, d4;
    if (zd->alpha < EPS_REF) {
        hzqrc = (float)(2.0 / cutoff_3);
        d1    = (float)(1.0 / cutoff_2);
        d2    = (float)(2.0 / cutoff_3);
        d3    = (float)(6.0 / c4);
        d4    = (float)(24.0 / c5);
    } else {
        float root_pi = (float)sqrt(PI_REF);
        float t3      = tmp * tmp * tmp;
        float t5      = t3 * tmp * tmp;
        float t7      = t5 * tmp * tmp;
        hzqrc         = (float)((errorfc + expterm * (tmp + t3) * 2.0 / root_pi) * 2.0 / cutoff_3);
        d1            = (float)((errorfc + expterm * tmp * 2.0 / root_pi) / cutoff_2);
        d2            = (float)((errorfc + expterm * (tmp + t3) * 2.0 / root_pi) * 2.0 / cutoff_3);
        d3 = (float)((errorfc + expterm * (tmp + 2.0 / 3.0 * t3 + 2.0 / 3.0 * t5) * 2.0 / root_pi) * 6.0 / c4);
        /* the reference's series (ZeroMultipoleSum.cpp:152-154), its "1/3 + t7" term as written */
        d4 = (float)((errorfc + expterm * (tmp + 2.0 / 3.0 * t3 + 1.0 / 6.0 * t5 + 1.0 / 3.0 + t7) * 2.0 / root_pi) *
                     24.0 / c5);
    }
    zd->zq2  = (float)(-3.0 / 4.0 * zd->fcoeff - hzqrc / 4.0);
    zd->zq4  = (float)(zd->fcoeff / 8.0 / cutoff_2 + hzqrc / 8.0 / cutoff_2);
    zd->zq22 = (float)(2.0 * zd->zq2);
    zd->zq44 = (float)(4.0 * zd->zq4);
    zd->zo2  = (float)(-15.0 / 16.0 * d1 / c - 7.0 / 16.0 * d2 - 1.0 / 16.0 * d3 * c);
    zd->zo4  = (float)(5.0 / 16.0 * d1 / cutoff_3 + 5.0 / 16.0 * d2 / cutoff_2 + 1.0 / 16.0 * d3 / c);
    zd->zo6  = (float)(-1.0 / 16.0 * d1 / c5 - 1.0 / 16.0 * d2 / c4 - 1.0 / 48.0 * d3 / cutoff_3);
    zd->zo22 = (float)(2.0 * zd->zo2);
    zd->zo44 = (float)(4.0 * zd->zo4);
    zd->zo66 = (float)(6.0 * zd->zo6);
    zd->zh2  = (float)(-35.0 / 32.0 * d1 / c - 19.0 / 32.0 * d2 - 1.0 / 96.0 * d4 * cutoff_2 - 1.0 / 8.0 * d3 * c);
    zd->zh4  = (float)(1.0 / 64.0 * d4 + 35.0 / 64.0 * d1 / cutoff_3 + 35.0 / 64.0 * d2 / cutoff_2 +
                      5.0 / 32.0 * d3 / c);
    zd->zh6  = (float)(-7.0 / 32.0 * d1 / c5 - 7.0 / 32.0 * d2 / c4 - 1.0 / 12.0 * d3 / cutoff_3 -
                      1.0 / 96.0 * d4 / cutoff_2);
    zd->zh8  = (float)(5.0 / 128.0 * d2 / c6 + 5.0 / 128.0 * d1 / c7 + 1.0 / 64.0 * d3 / c5 + 1.0 / 384.0 * d4 / c4);
    zd->zh22 = (float)(2.0 * zd->zh2);
    zd->zh44 = (float)(4.0 * zd->zh4);
    zd->zh66 = (float)(6.0 * zd->zh6);
    zd->zh88 = (float)(8.0 * zd->zh8);
    switch (mode) {
    case ORC_ZM_QUADRUPOLE: zd->zmcore = errorfc / c - zd->zq2 * cutoff_2 - zd->zq4 * c4; break;
    case ORC_ZM_OCTUPOLE: zd->zmcore = errorfc / c - zd->zo2 * cutoff_2 - zd->zo4 * c4 - zd->zo6 * c6; break;
    case ORC_ZM_HEXADECAPOLE:
        zd->zmcore = errorfc / c - zd->zh2 * cutoff_2 - zd->zh4 * c4 - zd->zh6 * c6 - zd->zh8 * c6 * cutoff_2;
        break;
    default: zd->zmcore = zd->zcore; break;
    }
}

/* The electrostatic term of ForceField::calc_pairwise for the selected order:
 * calc_zero02pole(_alpha0) ... calc_zero16pole(_alpha0) (ZeroMultipoleSum.cpp:309-442),
 * float locals. */
static void zm_pair(const orc_zd_t *zd, float r12, float r12_2, float r12_inv, float r12_2_inv, float r12_3_inv,
                    float cc, float *e, float *g) {
    const int   a0    = zd->alpha < EPS_REF;
    const float r12_4 = r12_2 * r12_2;
    const float r12_6 = r12_4 * r12_2;
    const float r12_8 = r12_4 * r12_4;
    float       tmp = 0.0f, errorfc = 1.0f;
    if (!a0) {
        tmp     = r12 * zd->alpha;
        errorfc = erfcf(tmp);
    }
    switch (zd->mode) {
    case ORC_ZM_QUADRUPOLE:
        if (a0) {
            *e = cc * (r12_inv - zd->zmcore - zd->zq2 * r12_2 - zd->zq4 * r12_4);
            *g = -cc * (r12_3_inv + zd->zq22 + zd->zq44 * r12_2);
        } else {
            *e = cc * (r12_inv * errorfc - zd->zmcore - zd->zq2 * r12_2 - zd->zq4 * r12_4);
            *g = -cc * (r12_3_inv * errorfc + zd->piewald * expf(-tmp * tmp) * r12_2_inv + zd->zq22 + zd->zq44 * r12_2);
        }
        break;
    case ORC_ZM_OCTUPOLE:
        if (a0) {
            *e = cc * (r12_inv - zd->zmcore - zd->zo2 * r12_2 - zd->zo4 * r12_4 - zd->zo6 * r12_6);
            *g = -cc * (r12_3_inv + zd->zo22 + zd->zo44 * r12_2 + zd->zo66 * r12_4);
        } else {
            const float g2 = zd->zo_grad_bug ? zd->zh22 : zd->zo22, g4 = zd->zo_grad_bug ? zd->zh44 : zd->zo44,
                        g6 = zd->zo_grad_bug ? zd->zh66 : zd->zo66;
            *e = cc * (r12_inv * errorfc - zd->zmcore - zd->zo2 * r12_2 - zd->zo4 * r12_4 - zd->zo6 * r12_6);
            *g = -cc * (r12_3_inv * errorfc + zd->piewald * expf(-tmp * tmp) * r12_2_inv + g2 + g4 * r12_2 + g6 * r12_4);
        }
        break;
    case ORC_ZM_HEXADECAPOLE:
        if (a0) {
            *e = cc * (r12_inv - zd->zmcore - zd->zh2 * r12_2 - zd->zh4 * r12_4 - zd->zh6 * r12_6 - zd->zh8 * r12_8);
            *g = -cc * (r12_3_inv + zd->zh22 + zd->zh44 * r12_2 + zd->zh66 * r12_4 + zd->zh88 * r12_6);
        } else {
            *e = cc * (r12_inv * errorfc - zd->zmcore - zd->zh2 * r12_2 - zd->zh4 * r12_4 - zd->zh6 * r12_6 -
                       zd->zh8 * r12_8);
            *g = -cc * (r12_3_inv * errorfc + zd->piewald * expf(-tmp * tmp) * r12_2_inv + zd->zh22 +
                        zd->zh44 * r12_2 + zd->zh66 * r12_4 + zd->zh88 * r12_6);
        }
        break;
    default:
        if (a0) {
            *e = cc * (r12_inv - zd->zcore + zd->bcoeff * r12_2);
            *g = -cc * (r12_3_inv - zd->fcoeff);
        } else {
            *e = cc * (r12_inv * errorfc - zd->zcore + zd->bcoeff * r12_2);
            *g = -cc * (errorfc * r12_3_inv + zd->piewald * expf(-tmp * tmp) * r12_2_inv - zd->fcoeff);
        }
        break;
    }
}

/* ZeroMultipoleSum::cal_self_energy (ZeroMultipoleSum.cpp:43-106), dipole case. */
double orc_self_energy(const orc_zd_t *zd, int n, const float *q, int n_excess, const int *ex, double *e_self) {
    double *dwork = (double *)calloc((size_t)n, sizeof(double));
    for (int i = 0; i < n_excess; i++) {
        dwork[ex[2 * i]] += q[ex[2 * i + 1]];
        dwork[ex[2 * i + 1]] += q[ex[2 * i]];
    }
    double zmcore = zd->zmcore; /* the selected order's core (ZeroMultipoleSum.cpp:72-80) */
    float  mon    = 0.0f;       /* d_self_mon is a real_pw member */
    for (int i = 0; i < n; i++) {
        float  pq = -zd->piewald * q[i] * q[i];
        double e  = ((double)pq - zmcore * q[i] * (q[i] + dwork[i])) * 0.5 * CHARGE_COEFF;
        if (e_self) e_self[i] = e;
        mon = (float)((double)mon + e);
    }
    free(dwork);
    return (double)mon;
}

/* MmSystem::set_excess_pairs (MmSystem.cpp:700-736): i<j from nb15off order. */
int orc_excess_pairs(int n, const int *nb15off, int *out, int cap) {
    int k = 0;
    for (int a = 0; a < n; a++)
        for (int j = 0; j < ORC_MAX_N_NB15OFF; j++) {
            int b = nb15off[a * ORC_MAX_N_NB15OFF + j];
            if (b == -1) break;
            if (a >= b) continue;
            if (k >= cap) return -1;
            out[2 * k]     = a;
            out[2 * k + 1] = b;
            k++;
        }
    return k;
}

/* ForceField::calc_pairwise (ForceField.cpp:333-418) with the ZD kernels
 * calc_zero02pole_alpha0 (ZeroMultipoleSum.cpp:309-332) / calc_zero02pole (:333-346).
 * float locals, double literals promote then round on assignment. */
int orc_pair_fp32(const orc_zd_t *zd, const float d12[3], float a6, float a12, float q1, float q2,
                  float *e_vdw, float *e_ele, float work[3]) {
    float r12_2 = d12[0] * d12[0] + d12[1] * d12[1] + d12[2] * d12[2];
    float r12   = sqrtf(r12_2);
    *e_vdw = *e_ele = 0.0f;
    work[0] = work[1] = work[2] = 0.0f;
    if (r12 >= zd->cutoff) return 0;
    float r12_inv    = (float)(1.0 / (double)r12);
    float r12_2_inv  = (float)(1.0 / (double)r12_2);
    float r12_3_inv  = r12_inv * r12_2_inv;
    float r12_6_inv  = r12_3_inv * r12_3_inv;
    float r12_12_inv = r12_6_inv * r12_6_inv;
    float term6      = a6 * r12_6_inv;
    float term12     = a12 * r12_12_inv;
    *e_vdw           = -term6 + term12;
    float work_coef  = (float)((double)r12_2_inv * (-12.0 * (double)term12 + 6.0 * (double)term6));
    float cc         = (float)((double)(q1 * q2) * CHARGE_COEFF);
    float g_ele;
    zm_pair(zd, r12, r12_2, r12_inv, r12_2_inv, r12_3_inv, cc, e_ele, &g_ele);
    for (int d = 0; d < 3; d++) {
        float wv = work_coef * d12[d];
        float we = g_ele * d12[d];
        work[d]  = wv + we;
    }
    return 1;
}

/* CelesteObject.cpp:32-40 */
static const double ELEM_CHARGE  = 1.60217657e-19;
static const double AVOGADRO     = 6.0221413e23;
static const double PERMITTIVITY = 8.85418782e-12;
static const double BOLTZMAN     = 1.380658e-23;

void orc_dh_setup(orc_ff_t *ff, double dielectric, double ionic_strength, double temperature) {
    ff->dielectric_inv = (float)(1.0 / dielectric);
    ff->debye_inv      = (float)(1.0 / (sqrt(PERMITTIVITY * dielectric * BOLTZMAN * temperature /
                                             (2 * AVOGADRO * ELEM_CHARGE * ELEM_CHARGE * ionic_strength)) *
                                        1e10));
}

/* ForceField::calc_pairwise (ForceField.cpp:333-418) with the HPS switch
 * (:380-391) and ZeroMultipoleSum::calc_debye_huckel (ZeroMultipoleSum.cpp:443-455):
 * float locals; exp of a float is the float overload. */
int orc_pair_ff_fp32(const orc_zd_t *zd, const orc_ff_t *ff, const float d12[3], int t, float a6, float a12,
                     float q1, float q2, float *e_vdw, float *e_ele, float work[3]) {
    if (!ff || (!ff->hps && !ff->dh)) return orc_pair_fp32(zd, d12, a6, a12, q1, q2, e_vdw, e_ele, work);
    float r12_2 = d12[0] * d12[0] + d12[1] * d12[1] + d12[2] * d12[2];
    float r12   = sqrtf(r12_2);
    *e_vdw = *e_ele = 0.0f;
    work[0] = work[1] = work[2] = 0.0f;
    if (r12 >= zd->cutoff) return 0;
    float r12_inv    = (float)(1.0 / (double)r12);
    float r12_2_inv  = (float)(1.0 / (double)r12_2);
    float r12_3_inv  = r12_inv * r12_2_inv;
    float r12_6_inv  = r12_3_inv * r12_3_inv;
    float r12_12_inv = r12_6_inv * r12_6_inv;
    float term6      = a6 * r12_6_inv;
    float term12     = a12 * r12_12_inv;
    float ev         = -term6 + term12;
    float work_coef  = (float)((double)r12_2_inv * (-12.0 * (double)term12 + 6.0 * (double)term6));
    if (ff->hps) {
        const float hc = ff->hps_cutoff[t], hl = ff->hps_lambda[t];
        if (r12 > hc) {
            ev *= hl;
            work_coef *= hl;
        } else {
            ev += (1 - hl) * ff->hps_eps;
        }
    }
    *e_vdw   = ev;
    float cc = (float)((double)(q1 * q2) * CHARGE_COEFF);
    float g_ele;
    if (ff->dh) {
        float ld_exp = expf(-r12 * ff->debye_inv);
        *e_ele       = cc * r12_inv * ff->dielectric_inv * ld_exp;
        g_ele        = -*e_ele * (r12_inv + ff->debye_inv);
        if (ff->dh_grad_fix) g_ele *= r12_inv;
    } else {
        zm_pair(zd, r12, r12_2, r12_inv, r12_2_inv, r12_3_inv, cc, e_ele, &g_ele);
    }
    for (int d = 0; d < 3; d++) {
        float wv = work_coef * d12[d];
        float we = g_ele * d12[d];
        work[d]  = wv + we;
    }
    return 1;
}

/* PBC::fix_pbc_image on a float coordinate (PBC.cpp:106-121). */
static void fix_image(float *x, int image, const float L[3]) {
    if (image == 0) return;
    if ((image & 1) == 1) x[0] -= L[0];
    else if ((image & 2) == 2) x[0] += L[0];
    if ((image & 4) == 4) x[1] -= L[1];
    else if ((image & 8) == 8) x[1] += L[1];
    if ((image & 16) == 16) x[2] -= L[2];
    else if ((image & 32) == 32) x[2] += L[2];
}

/* SubBox::calc_energy_pairwise (SubBox.cpp:956-1125) */
int64_t orc_pairs_celllist(const orc_zd_t *zd, const orc_grid_t *g, const orc_cells_t *c,
                           const orc_cellpair_t *cps, int n_cp, const int *atom_type, int n_types,
                           const float *lj6, const float *lj12, const float *charge, double *work,
                           double energy[2]) {
    return orc_pairs_celllist_ff(zd, NULL, g, c, cps, n_cp, atom_type, n_types, lj6, lj12, charge, work, energy);
}
int64_t orc_pairs_celllist_ff(const orc_zd_t *zd, const orc_ff_t *ff, const orc_grid_t *g, const orc_cells_t *c,
                              const orc_cellpair_t *cps, int n_cp, const int *atom_type, int n_types,
                              const float *lj6, const float *lj12, const float *charge, double *work,
                              double energy[2]) {
    int64_t n_in = 0;
    memset(work, 0, sizeof(double) * (size_t)c->n_atom_array * 3);
    energy[0] = energy[1] = 0.0;
    for (int cp = 0; cp < n_cp; cp++) {
        int c1 = cps[cp].cell_id1, c2 = cps[cp].cell_id2;
        for (int i2 = 0; i2 < NAC; i2++) {
            int   g2 = c2 * NAC + i2;
            int   id2 = c->atomids[g2];
            float x2[3] = {(float)c->crd[g2 * 3], (float)c->crd[g2 * 3 + 1], (float)c->crd[g2 * 3 + 2]};
            fix_image(x2, cps[cp].image, g->L);
            for (int i1 = 0; i1 < NAC; i1++) {
                int bit = i2 * NAC + i1;
                if (((uint32_t)cps[cp].pair_mask[bit / 32] >> (bit % 32)) & 1u) continue;
                int   g1 = c1 * NAC + i1;
                int   id1 = c->atomids[g1];
                float x1[3] = {(float)c->crd[g1 * 3], (float)c->crd[g1 * 3 + 1], (float)c->crd[g1 * 3 + 2]};
                float d12[3] = {x1[0] - x2[0], x1[1] - x2[1], x1[2] - x2[2]};
                int   t = atom_type[id1] * n_types + atom_type[id2];
                float ev, ee, w[3];
                n_in += orc_pair_ff_fp32(zd, ff, d12, t, lj6[t], lj12[t], charge[id1], charge[id2], &ev, &ee, w);
                energy[0] += ev;
                energy[1] += ee;
                for (int d = 0; d < 3; d++) {
                    work[g1 * 3 + d] += w[d];
                    work[g2 * 3 + d] += -(double)w[d];
                }
            }
        }
    }
    return n_in;
}

/* SubBox::calc_energy_pairwise_wo_neighborsearch (SubBox.cpp:1127-1220) */
int64_t orc_pairs_allpairs_fp32(const orc_zd_t *zd, const float L[3], int n, const double *crd,
                                const int *atom_type, int n_types, const float *lj6, const float *lj12,
                                const float *charge, const int *nb15off, double *work, double energy[2]) {
    return orc_pairs_allpairs_ff(zd, NULL, L, n, crd, atom_type, n_types, lj6, lj12, charge, nb15off, work, energy);
}
int64_t orc_pairs_allpairs_ff(const orc_zd_t *zd, const orc_ff_t *ff, const float L[3], int n, const double *crd,
                              const int *atom_type, int n_types, const float *lj6, const float *lj12,
                              const float *charge, const int *nb15off, double *work, double energy[2]) {
    int64_t n_in = 0;
    float   L_half[3];
    for (int d = 0; d < 3; d++) L_half[d] = (float)(L[d] * 0.5);
    memset(work, 0, sizeof(double) * (size_t)n * 3);
    energy[0] = energy[1] = 0.0;
    for (int a1 = 0; a1 < n; a1++) {
        float x1[3] = {(float)crd[a1 * 3], (float)crd[a1 * 3 + 1], (float)crd[a1 * 3 + 2]};
        for (int a2 = 0; a2 < a1; a2++) {
            int excl = 0;
            for (int k = 0; k < ORC_MAX_N_NB15OFF; k++)
                if (nb15off[a1 * ORC_MAX_N_NB15OFF + k] == a2) excl = 1;
            if (excl) continue;
            float x2[3] = {(float)crd[a2 * 3], (float)crd[a2 * 3 + 1], (float)crd[a2 * 3 + 2]};
            for (int d = 0; d < 3; d++) {
                if (x2[d] - x1[d] >= L_half[d]) x2[d] -= L[d];
                else if (x2[d] - x1[d] <= -L_half[d]) x2[d] += L[d];
            }
            float d12[3] = {x1[0] - x2[0], x1[1] - x2[1], x1[2] - x2[2]};
            int   t = atom_type[a1] * n_types + atom_type[a2];
            float ev, ee, w[3];
            n_in += orc_pair_ff_fp32(zd, ff, d12, t, lj6[t], lj12[t], charge[a1], charge[a2], &ev, &ee, w);
            energy[0] += ev;
            energy[1] += ee;
            for (int d = 0; d < 3; d++) {
                work[a1 * 3 + d] += w[d];
                work[a2 * 3 + d] -= w[d];
            }
        }
    }
    return n_in;
}

/* FP64 version of the same physics for error statements (double math,
 * exact minimum image on double coordinates, erfc in double). */
int64_t orc_pairs_celllist_fp64(const orc_zd_t *zd, const orc_grid_t *g, const orc_cells_t *c,
                                const orc_cellpair_t *cps, int n_cp, const double *crd_atom,
                                const int *atom_type, int n_types, const float *lj6, const float *lj12,
                                const float *charge, double *work, double energy[3]) {
    int64_t n_in = 0;
    int     n_atoms = 0;
    for (int i = 0; i < c->n_atom_array; i++)
        if (c->atomids[i] + 1 > n_atoms) n_atoms = c->atomids[i] + 1;
    memset(work, 0, sizeof(double) * (size_t)n_atoms * 3);
    energy[0] = energy[1] = energy[2] = 0.0;
    const double rc = zd->cutoff, alpha = zd->alpha;
    /* constants in double from the same closed forms */
    double fcoeff, bcoeff, zcore, piewald = 0.0;
    if (alpha < EPS_REF) {
        fcoeff = 1.0 / (rc * rc * rc);
        bcoeff = 0.5 * fcoeff;
        zcore  = 1.0 / rc + bcoeff * rc * rc;
    } else {
        piewald = 2.0 / sqrt(M_PI) * alpha;
        fcoeff  = erfc(alpha * rc) / (rc * rc * rc) + piewald * exp(-alpha * alpha * rc * rc) / (rc * rc);
        bcoeff  = 0.5 * fcoeff;
        zcore   = erfc(alpha * rc) / rc + bcoeff * rc * rc;
    }
    for (int cp = 0; cp < n_cp; cp++) {
        int c1 = cps[cp].cell_id1, c2 = cps[cp].cell_id2;
        for (int i2 = 0; i2 < NAC; i2++) {
            int id2 = c->atomids[c2 * NAC + i2];
            for (int i1 = 0; i1 < NAC; i1++) {
                int bit = i2 * NAC + i1;
                if (((uint32_t)cps[cp].pair_mask[bit / 32] >> (bit % 32)) & 1u) continue;
                int    id1 = c->atomids[c1 * NAC + i1];
                double d[3];
                /* the image is the cell pair's (a small box lists one cell pair
                 * under several images): take the grid-coordinate separation
                 * with that image and carry the FP64 atom separation to the
                 * same periodic copy */
                const int img = cps[cp].image;
                for (int k = 0; k < 3; k++) {
                    const int    lo = 1 << (2 * k), hi = 2 << (2 * k);
                    const double sh = (img & lo) ? -(double)g->L[k] : (img & hi) ? (double)g->L[k] : 0.0;
                    const double dg = c->crd[(c1 * NAC + i1) * 3 + k] - (c->crd[(c2 * NAC + i2) * 3 + k] + sh);
                    d[k]            = crd_atom[id1 * 3 + k] - crd_atom[id2 * 3 + k];
                    d[k] -= (double)g->L[k] * nearbyint((d[k] - dg) / (double)g->L[k]);
                }
                double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                double r  = sqrt(r2);
                if (r >= rc) continue;
                n_in++;
                int    t = atom_type[id1] * n_types + atom_type[id2];
                double ri = 1.0 / r, r2i = ri * ri, r6i = r2i * r2i * r2i;
                double t6 = (double)lj6[t] * r6i, t12 = (double)lj12[t] * r6i * r6i;
                double cc = (double)charge[id1] * (double)charge[id2] * CHARGE_COEFF;
                double ev = t12 - t6, ee, gcoef = r2i * (-12.0 * t12 + 6.0 * t6);
                if (alpha < EPS_REF) {
                    ee = cc * (ri - zcore + bcoeff * r2);
                    gcoef += -cc * (ri * r2i - fcoeff);
                } else {
                    double ef = erfc(alpha * r);
                    ee        = cc * (ri * ef - zcore + bcoeff * r2);
                    gcoef += -cc * (ef * ri * r2i + piewald * exp(-alpha * alpha * r2) * r2i - fcoeff);
                }
                energy[0] += ev;
                energy[1] += ee;
                energy[2] += gcoef * r2; /* virial: r_ij . dE/dr_i */
                for (int k = 0; k < 3; k++) {
                    work[id1 * 3 + k] += gcoef * d[k];
                    work[id2 * 3 + k] -= gcoef * d[k];
                }
            }
        }
    }
    return n_in;
}

/* PBC::diff_crd_minim_image, float overload (PBC.cpp:44-51). */
static void minim_image_f(float d[3], const float *x1, const float *x2, const float L[3]) {
    for (int i = 0; i < 3; i++) {
        float L_half     = (float)(L[i] * 0.5);
        float L_half_inv = (float)(1.0 / L_half);
        d[i]             = x1[i] - x2[i];
        d[i]             = d[i] - L[i] * (float)(int)(d[i] * L_half_inv);
        d[i]             = d[i] - L[i] * (float)(int)(d[i] * L_half_inv);
        /* within rounding of +-L the two passes swap -L and +L: one nearest-image step more */
        if (fabsf(d[i]) > 0.75f * L[i]) d[i] -= L[i] * rintf(d[i] / L[i]);
    }
}
/* double overload (PBC.cpp:62-82) */
static void minim_image_d(double d[3], const double *x1, const double *x2, const float L[3]) {
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

/* The excluded-pair term of the selected order: calc_zero02pole_excess(_alpha0) ...
 * calc_zero16pole_excess(_alpha0) (ZeroMultipoleSum.cpp:191-307).  bug_compat
 * keeps the reference's forms where they differ from the order's polynomial:
 * the dipole alpha > 0 energy -erf(ar) r (:214), the octupole r^5 in place of
 * r^6 (:255, :271), the hexadecapole alpha = 0 gradient c66 r^2 and -1/r^3
 * (:288) and alpha > 0 energy without the r^6 and r^8 terms (:304). */
static void excess_term(const orc_zd_t *zd, float r12, float r12_2, float r12_inv, float r12_2_inv, float r12_3_inv,
                        float cc, int bug, float *ene, float *gc) {
    const int a0 = zd->alpha < EPS_REF;
    float     tmp = 0.0f, errorfc = 1.0f, errorfn = 0.0f, pe = 0.0f;
    if (!a0) {
        tmp     = zd->alpha * r12;
        errorfc = erfcf(tmp);
        errorfn = (float)(1.0 - (double)errorfc);
        pe      = zd->piewald * expf(-tmp * tmp) * r12_2_inv;
    }
    switch (zd->mode) {
    case ORC_ZM_QUADRUPOLE:
        if (a0) {
            *ene = cc * (-zd->zq2 * r12_2 - zd->zq4 * r12_2 * r12_2);
            *gc  = -cc * (zd->zq22 + zd->zq44 * r12_2);
        } else {
            *ene = cc * (-errorfn * r12_inv - zd->zq2 * r12_2 - zd->zq4 * r12_2 * r12_2);
            *gc  = -cc * (errorfc * r12_3_inv + pe + zd->zq22 + zd->zq44 * r12_2 - r12_3_inv);
        }
        break;
    case ORC_ZM_OCTUPOLE:
        if (a0) {
            const double r12_4 = r12_2 * r12_2; /* a double local in the reference (:254) */
            *ene = (float)(cc * (-zd->zo2 * r12_2 - zd->zo4 * r12_4 -
                                 (bug ? zd->zo6 * r12_4 * r12 : zd->zo6 * r12_4 * r12_2)));
            *gc  = (float)(-cc * (zd->zo22 + zd->zo44 * r12_2 + zd->zo66 * r12_4));
        } else {
            const float r12_4 = r12_2 * r12_2;
            *ene = cc * (-errorfn * r12_inv - zd->zo2 * r12_2 - zd->zo4 * r12_4 -
                         (bug ? zd->zo6 * r12_4 * r12 : zd->zo6 * r12_4 * r12_2));
            *gc  = -cc * (errorfc * r12_3_inv + pe + zd->zo22 + zd->zo44 * r12_2 + zd->zo66 * r12_4 - r12_3_inv);
        }
        break;
    case ORC_ZM_HEXADECAPOLE: {
        const float r12_4 = r12_2 * r12_2, r12_6 = r12_4 * r12_2, r12_8 = r12_4 * r12_4;
        if (a0) {
            *ene = cc * (-zd->zh2 * r12_2 - zd->zh4 * r12_4 - zd->zh6 * r12_6 - zd->zh8 * r12_8);
            if (bug) *gc = -cc * (zd->zh22 + zd->zh44 * r12_2 + zd->zh66 * r12_2 + zd->zh88 * r12_6 - r12_3_inv);
            else *gc = -cc * (zd->zh22 + zd->zh44 * r12_2 + zd->zh66 * r12_4 + zd->zh88 * r12_6);
        } else {
            if (bug) *ene = cc * (-errorfn * r12_inv - zd->zh2 * r12_2 - zd->zh4 * r12_4);
            else *ene = cc * (-errorfn * r12_inv - zd->zh2 * r12_2 - zd->zh4 * r12_4 - zd->zh6 * r12_6 - zd->zh8 * r12_8);
            *gc = -cc * (errorfc * r12_3_inv + pe + zd->zh22 + zd->zh44 * r12_2 + zd->zh66 * r12_4 + zd->zh88 * r12_6 -
                         r12_3_inv);
        }
        break;
    }
    default:
        if (a0) {
            *ene = cc * zd->bcoeff * r12_2;
            *gc  = cc * zd->fcoeff;
        } else {
            if (bug) *ene = cc * (-errorfn * r12 + zd->bcoeff * r12_2);
            else *ene = cc * (-errorfn * r12_inv + zd->bcoeff * r12_2);
            *gc = -cc * (errorfc * r12_3_inv + pe - zd->fcoeff - r12_3_inv);
        }
        break;
    }
}

/* SubBox::calc_energy_ele_excess (SubBox.cpp:1347-1367) + ForceField::calc_zms_excess
 * (ForceField.cpp:420-437) + calc_zero02pole_excess(_alpha0) (ZeroMultipoleSum.cpp:191-217). */
double orc_excess(const orc_zd_t *zd, const float L[3], int n_excess, const int *pairs, const double *crd,
                  const float *q, int bug_compat, double *work) {
    double e_sum = 0.0;
    for (int i = 0; i < n_excess; i++) {
        int   a = pairs[2 * i], b = pairs[2 * i + 1];
        float c1[3] = {(float)crd[a * 3], (float)crd[a * 3 + 1], (float)crd[a * 3 + 2]};
        float c2[3] = {(float)crd[b * 3], (float)crd[b * 3 + 1], (float)crd[b * 3 + 2]};
        float d12[3];
        minim_image_f(d12, c1, c2, L);
        float r12_2     = d12[0] * d12[0] + d12[1] * d12[1] + d12[2] * d12[2];
        float r12       = sqrtf(r12_2);
        float r12_inv   = (float)(1.0 / (double)r12);
        float r12_2_inv = r12_inv * r12_inv;
        float r12_3_inv = r12_2_inv * r12_inv;
        float cc        = (float)((double)(q[a] * q[b]) * CHARGE_COEFF);
        float ene, gc;
        excess_term(zd, r12, r12_2, r12_inv, r12_2_inv, r12_3_inv, cc, bug_compat, &ene, &gc);
        e_sum += ene;
        for (int d = 0; d < 3; d++) {
            float w = gc * d12[d];
            work[a * 3 + d] += w;
            work[b * 3 + d] -= w;
        }
    }
    return e_sum;
}

/* SubBox::calc_energy_bonds (SubBox.cpp:1229-1249) + ForceField::calc_bond (ForceField.cpp:108-136). */
double orc_bonds(const float L[3], int n_bonds, const int *pairs, const double *k, const double *r0,
                 const double *crd, double *work) {
    double e = 0.0;
    for (int i = 0; i < n_bonds; i++) {
        int    a = pairs[2 * i], b = pairs[2 * i + 1];
        double d[3];
        minim_image_d(d, crd + a * 3, crd + b * 3, L);
        double r    = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double dif  = r - r0[i];
        double coef = k[i] * dif;
        e += coef * dif;
        double gcoef = 2.0 * coef / r;
        for (int q = 0; q < 3; q++) {
            double w = d[q] * gcoef;
            work[a * 3 + q] += w;
            work[b * 3 + q] -= w;
        }
    }
    return e;
}
