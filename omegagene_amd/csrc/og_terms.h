// og_terms.h -- per-atom terms shared by the force-conversion kernel and the
// fused Langevin step: excluded-pair ZD terms, harmonic bonds, and the
// conversion of an atom's fixed-point pair gradient into its FP64 "work".
// Every function evaluates in the reference rounding (no FMA contraction),
// whichever source file includes it.
#pragma once
#include "og_internal.h"

// Block-wide sums of two doubles in a fixed order; the totals are left in thread 0.
__device__ __forceinline__ void og_block_sum2(double &a, double &b, double *sh) {
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        sh[2 * w]     = a;
        sh[2 * w + 1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = 0.0;
        b = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
            a += sh[2 * k];
            b += sh[2 * k + 1];
        }
    }
}

// The excluded-pair term of the zero-quadrupole, -octupole and -hexadecapole
// orders (calc_zero04pole_excess ... calc_zero16pole_excess(_alpha0),
// ZeroMultipoleSum.cpp:218-307), float as the reference; excess_bug keeps the
// reference's forms where they misplace or drop a power (:255, :271, :288,
// :304).  w = gradient coefficient times d; returns the energy.
// Written without branches on the order: every form is evaluated (the unused
// coefficients are zeros, which leave a float sum's bits unchanged) and the
// order's result selected.  (An if/else-if chain over the order here was
// compiled by this toolchain with the hexadecapole alpha > 0 gradient scaled
// by a register never assigned on that path.)
__device__ __forceinline__ float og_excess_zm(const og_dev_sys &s, float r, float r2, float rinv, float r2inv,
                                              float r3inv, float cc, const float *d, float *w) {
#pragma clang fp contract(off)
    const float c2 = s.zm_c[0], c4 = s.zm_c[1], c6 = s.zm_c[2], c8 = s.zm_c[3];
    const float g2 = s.zm_cc[0], g4 = s.zm_cc[1], g6 = s.zm_cc[2], g8 = s.zm_cc[3];
    const int   ord = s.zm_order;
    const bool  bug = s.excess_bug != 0;
    const float r4 = r2 * r2, r6 = r4 * r2, r8 = r4 * r4;
    // alpha = 0
    const float  e1a = cc * (-c2 * r2 - c4 * r2 * r2);
    const float  g1a = -cc * (g2 + g4 * r2);
    const double r4d = (double)r4;  // a double local in the reference's octupole (:254)
    const float  e2a = (float)(cc * (-c2 * r2 - c4 * r4d - (bug ? c6 * r4d * r : c6 * r4d * r2)));
    const float  g2a = (float)(-cc * (g2 + g4 * r2 + g6 * r4d));
    const float  e3a = cc * (-c2 * r2 - c4 * r4 - c6 * r6 - c8 * r8);
    const float  g3a = bug ? -cc * (g2 + g4 * r2 + g6 * r2 + g8 * r6 - r3inv) : -cc * (g2 + g4 * r2 + g6 * r4 + g8 * r6);
    // alpha > 0 (the gradient takes the same form in every order)
    const float x     = s.alpha * r;
    const float erfcv = erfcf(x);
    const float erfv  = (float)(1.0 - (double)erfcv);
    const float pe    = s.piewald * expf(-x * x) * r2inv;
    const float e1b   = cc * (-erfv * rinv - c2 * r2 - c4 * r2 * r2);
    const float e2b   = cc * (-erfv * rinv - c2 * r2 - c4 * r4 - (bug ? c6 * r4 * r : c6 * r4 * r2));
    const float e3b   = bug ? cc * (-erfv * rinv - c2 * r2 - c4 * r4)
                            : cc * (-erfv * rinv - c2 * r2 - c4 * r4 - c6 * r6 - c8 * r8);
    const float gb    = -cc * (erfcv * r3inv + pe + g2 + g4 * r2 + g6 * r4 + g8 * r6 - r3inv);
    const bool  a0    = s.alpha == 0.0f;
    const float ene   = a0 ? (ord == 1 ? e1a : ord == 2 ? e2a : e3a) : (ord == 1 ? e1b : ord == 2 ? e2b : e3b);
    const float gc    = a0 ? (ord == 1 ? g1a : ord == 2 ? g2a : g3a) : gb;
    for (int k = 0; k < 3; k++) w[k] = gc * d[k];
    return ene;
}

// Excluded pair (a, b), a < b in the excess list, from the atoms' FP64
// coordinates and charges: gradient w added to a and subtracted from b,
// energy returned, in the reference float arithmetic
// (ForceField::calc_zms_excess ForceField.cpp:420-437,
// calc_zero02pole_excess(_alpha0) ZeroMultipoleSum.cpp:191-217).
__device__ __forceinline__ float og_excess_pair(const og_dev_sys &s, const double *xa, const double *xb, float qa,
                                                float qb, float *w) {
#pragma clang fp contract(off)
    float d[3];
    for (int k = 0; k < 3; k++) d[k] = og_mimg_f((float)xa[k] - (float)xb[k], s.L[k]);
    const float r2    = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const float r     = sqrtf(r2);
    const float rinv  = 1.0f / r;
    const float r2inv = rinv * rinv;
    const float r3inv = r2inv * rinv;
    const float cc    = (float)((double)(qa * qb) * 332.06378);
    float       ene, gc;
    if (s.zm_order) return og_excess_zm(s, r, r2, rinv, r2inv, r3inv, cc, d, w);
    if (s.alpha == 0.0f) {
        ene = cc * s.bcoeff * r2;
        gc  = cc * s.fcoeff;
    } else {
        const float x     = s.alpha * r;
        const float erfcv = erfcf(x);
        const float erfv  = (float)(1.0 - (double)erfcv);
        ene = s.excess_bug ? cc * (-erfv * r + s.bcoeff * r2) : cc * (-erfv * rinv + s.bcoeff * r2);
        gc  = -cc * (erfcv * r3inv + s.piewald * expf(-x * x) * r2inv - s.fcoeff - r3inv);
    }
    for (int k = 0; k < 3; k++) w[k] = gc * d[k];
    return ene;
}

// Harmonic bond (a, b) with constants k, r0: gradient w added to a and
// subtracted from b, energy returned, in FP64 (ForceField::calc_bond
// ForceField.cpp:108-136).
__device__ __forceinline__ double og_bond_pair(const og_dev_sys &s, double bk, double r0, const double *xa,
                                               const double *xb, double *w) {
#pragma clang fp contract(off)
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = og_mimg(xa[k] - xb[k], s.L[k]);
    const double r    = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double dif  = r - r0;
    const double coef = bk * dif;
    const double gcf  = 2.0 * coef / r;
    for (int k = 0; k < 3; k++) w[k] = d[k] * gcf;
    return coef * dif;
}

#ifndef OG_TERM_BATCH
#define OG_TERM_BATCH 8  // list entries whose loads are issued together
#endif

// Atom i: fixed-point pair gradient (grid slot) -> FP64 "work" in atom order
// (SubBox::add_work_from_minicell, SubBox.cpp:1398-1407), plus the atom's
// excluded-pair ZD terms (SubBox::calc_energy_ele_excess SubBox.cpp:1347-1367)
// and bonds through per-atom lists: every pair term is evaluated for both of
// its atoms in the orientation of the reference list, so no atomics are
// needed and the sums are deterministic.  Energies are counted on the pair's
// first atom.  The lists are walked in batches: the entries, then the
// partners' coordinates, are loaded for a whole batch before the terms are
// added in list order.  Returns the work in w; stores nothing.
__device__ __forceinline__ void og_atom_work_calc(const og_dev_sys &s, const int i, double *w, double &e_ex,
                                                  double &e_b) {
#pragma clang fp contract(off)
    const int g   = s.rev[i];
    const int cap = s.cap_atom_array;
    const int x0 = s.exc_off[i], x1 = s.exc_off[i + 1];
    const int b0 = s.bond_off[i], b1 = s.bond_off[i + 1];
    double    xi[3];
    for (int k = 0; k < 3; k++) xi[k] = s.crd[3 * i + k];
    const float qi = s.charge[i];
    og_fx_read(s.fx, cap, og_fx_slot(g), w);
    for (int eb = x0; eb < x1; eb += OG_TERM_BATCH) {
        const int n = min(OG_TERM_BATCH, x1 - eb);
        int       other[OG_TERM_BATCH];
        bool      first[OG_TERM_BATCH];
#pragma unroll
        for (int j = 0; j < OG_TERM_BATCH; j++)
            if (j < n) {
                const int4 ent = s.exc_at[eb + j];
                first[j]       = ent.x == i;
                other[j]       = first[j] ? ent.y : ent.x;
            }
        double xo[OG_TERM_BATCH][3];
        float  qo[OG_TERM_BATCH];
#pragma unroll
        for (int j = 0; j < OG_TERM_BATCH; j++)
            if (j < n) {
                for (int k = 0; k < 3; k++) xo[j][k] = s.crd[3 * other[j] + k];
                qo[j] = s.charge[other[j]];
            }
#pragma unroll
        for (int j = 0; j < OG_TERM_BATCH; j++)
            if (j < n) {
                float we[3];
                if (first[j]) {
                    const float ene = og_excess_pair(s, xi, xo[j], qi, qo[j], we);
                    for (int k = 0; k < 3; k++) w[k] += (double)we[k];
                    e_ex += (double)ene;
                } else {
                    og_excess_pair(s, xo[j], xi, qo[j], qi, we);
                    for (int k = 0; k < 3; k++) w[k] -= (double)we[k];
                }
            }
    }
    for (int eb = b0; eb < b1; eb += OG_TERM_BATCH) {
        const int n = min(OG_TERM_BATCH, b1 - eb);
        int       other[OG_TERM_BATCH], t[OG_TERM_BATCH];
        bool      first[OG_TERM_BATCH];
#pragma unroll
        for (int j = 0; j < OG_TERM_BATCH; j++)
            if (j < n) {
                const int4 ent = s.bond_at[eb + j];
                first[j]       = ent.x == i;
                other[j]       = first[j] ? ent.y : ent.x;
                t[j]           = ent.z;
            }
        double xo[OG_TERM_BATCH][3], bk[OG_TERM_BATCH], r0[OG_TERM_BATCH];
#pragma unroll
        for (int j = 0; j < OG_TERM_BATCH; j++)
            if (j < n) {
                for (int k = 0; k < 3; k++) xo[j][k] = s.crd[3 * other[j] + k];
                bk[j] = s.bond_k[t[j]];
                r0[j] = s.bond_r0[t[j]];
            }
#pragma unroll
        for (int j = 0; j < OG_TERM_BATCH; j++)
            if (j < n) {
                double wb[3];
                if (first[j]) {
                    const double ene = og_bond_pair(s, bk[j], r0[j], xi, xo[j], wb);
                    for (int k = 0; k < 3; k++) w[k] += wb[k];
                    e_b += ene;
                } else {
                    og_bond_pair(s, bk[j], r0[j], xo[j], xi, wb);
                    for (int k = 0; k < 3; k++) w[k] -= wb[k];
                }
            }
    }
    // angles, torsions, impropers and 1-4 pairs: the atom's shares, written
    // by k_bonded earlier in the evaluation, in ascending slot order
    if (s.n_bterms) {
        const int c0 = s.bc_off[i], c1 = s.bc_off[i + 1];
        for (int e = c0; e < c1; e++) {
            const double *c = s.bcontrib + 3 * (size_t)s.bc_at[e];
            for (int k = 0; k < 3; k++) w[k] += c[k];
        }
    }
}

// Stores atom i's work and clears its fixed-point accumulators for the next
// evaluation.
__device__ __forceinline__ void og_atom_work_store(const og_dev_sys &s, const int i, const double *w) {
    og_fx_consume(s.fx, s.cap_atom_array, og_fx_slot(s.rev[i]));
    for (int k = 0; k < 3; k++) s.work[3 * i + k] = w[k];
}

__device__ __forceinline__ void og_atom_work(const og_dev_sys &s, const int i, double &e_ex, double &e_b) {
    double w[3];
    og_atom_work_calc(s, i, w, e_ex, e_b);
    og_atom_work_store(s, i, w);
}

// Adds a block's excluded-pair and bond energy sums to the system's
// fixed-point accumulators (slot chosen by the block), on energy steps.
__device__ __forceinline__ void og_add_misc_energies(const og_dev_sys &s, int slot, double e_ex, double e_b,
                                                     double *sh) {
    og_block_sum2(e_ex, e_b, sh);
    if (threadIdx.x == 0) {
        OG_GLOBAL(long long) *A = (OG_GLOBAL(long long) *)s.e_acc + (slot % OG_EACC_SLOTS) * OG_EACC_STRIDE;
        og_atomic_add_fx_g(A + OG_EACC_EXCESS, og_to_fixed(e_ex));
        og_atomic_add_fx_g(A + OG_EACC_BOND, og_to_fixed(e_b));
    }
}
