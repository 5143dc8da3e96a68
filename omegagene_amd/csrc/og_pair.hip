// og_pair.hip -- LJ + zero-dipole cell-pair kernel, excluded-pair (excess)
// and bond terms, fixed-point gradient reduction and energy reductions.
//
// Pair kernel (reference: kernel_pairwise_ljzd + cal_pair, cuda_setup.cu:487-675;
// CPU: SubBox::calc_energy_pairwise, SubBox.cpp:956-1125).  It evaluates the
// same pairs as the reference -- every unmasked atom pair of every listed
// cell pair closer than the cutoff -- but schedules them for a wave64:
//
// * one wave per cell1 (the last cells of a large system: two waves, each
//   with half of the cell's records); its 8 atoms ("i") sit in scalar registers;
// * filter: the wave reads its cell pairs 8 at a time (lane = 8*pair + j) and
//   keeps the cell2 atoms ("rows") within the cutoff of the cell1 bounding
//   box, appending them (shifted to the pair's periodic image) to a per-wave
//   LDS queue.  About half of the rows of a 13 A list survive at 12 A;
// * rows: every 64 queued rows are one pass with lane = row, looping over
//   the 8 i atoms.  The row's j-gradient stays in its lane (no cross-lane
//   reduction) and is added once as 64-bit fixed point; each lane keeps 24
//   partial i-gradients, reduced over the wave once per cell1.
//
// FP32 pair math as the reference, fixed-point atomics (order-independent,
// so every run gives the same bits), energies summed per lane in FP64 and
// reduced per cell in a fixed order.
#include "og_terms.h"

#define OG_CHARGE_COEFF 332.06378f   // CelesteObject.cpp:36
#define OG_QCAP 128                  // per-wave row queue (two passes of 64)
#define OG_LJ_LDS_ENTRIES 512        // LJ table entries (whole leading rows) held in LDS per block

typedef float og_f4 __attribute__((ext_vector_type(4)));
typedef float og_f2 __attribute__((ext_vector_type(2)));

// Pair forms (the FORM template argument): 0 = LJ + zero-dipole at alpha 0,
// 1 = LJ + zero-dipole at alpha > 0 (cal_pair, cuda_setup.cu:487-521),
// 2 = hydrophobicity-scale LJ + Debye-Hueckel (cal_pair_hps_dh,
// cuda_setup.cu:1273-1332), 3 / 4 = LJ + zero-quadrupole, -octupole or
// -hexadecapole at alpha 0 / > 0 (the reference's host forms,
// ZeroMultipoleSum.cpp:347-442).  The LJ table entry is {A6, A12} for forms 0 and
// 1 and {A6, A12, HPS switch distance, HPS lambda} for form 2.
template <int FORM> struct og_ljt { typedef og_f2 type; };
template <> struct og_ljt<2> { typedef og_f4 type; };

// erfc(x) for x >= 0 from Abramowitz & Stegun 7.1.26 (|abs error| <= 1.5e-7),
// sharing exp(-x^2) with the force term.
__device__ __forceinline__ float og_erfc_as(float x, float expmx2) {
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, x, 1.0f));
    float       p = fmaf(t, 1.061405429f, -1.453152027f);
    p             = fmaf(t, p, 1.421413741f);
    p             = fmaf(t, p, -0.284496736f);
    p             = fmaf(t, p, 0.254829592f);
    return p * t * expmx2;
}

__device__ __forceinline__ float og_rlane(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// Cell1 data held by the wave: atom coordinates, charge * 332.06378, LJ row
// offsets (type * n_types, the same in the global table and in the block's
// LDS copy of its leading rows) and the bounding box.
struct og_cell1 {
    float    x[OG_NAC], y[OG_NAC], z[OG_NAC], q[OG_NAC];
    unsigned row[OG_NAC];
    float    lo[3], hi[3];
};

// j-gradient of the last pass, not yet added: its fixed-point atomics are
// issued after the next pass has read its rows, so they complete while that
// pass computes.  (A vector-memory wait counts stores and atomics too, so an
// atomic issued just before a load that is waited on stalls the wave until
// the memory-side add has completed.)
struct og_jpend {
    float x, y, z;
    int   a;
    bool  live;
};

// Fixed-point add of a lane's gradient (x, y, z) at grid slot a (the
// j-gradient enters with a minus sign).  Zero contributions are skipped.
__device__ __forceinline__ void og_add_fx3(OG_GLOBAL(long long) *FX, int cap, int a, double sx, double sy,
                                           double sz) {
    const long long vx = og_fx_fast(sx), vy = og_fx_fast(sy), vz = og_fx_fast(sz);
    if (vx != 0) og_atomic_add_fx_g(FX + a, vx);
    if (vy != 0) og_atomic_add_fx_g(FX + cap + a, vy);
    if (vz != 0) og_atomic_add_fx_g(FX + 2 * cap + a, vz);
}

// j-gradient of a row: (x, y) packed into one 64-bit atomic; z at 2^19 units
// into the 32-bit field of its slot in the slot pair's word, and when the
// next lane holds the pair's other slot (rows of one cell2 queue in atom
// order) the lower lane adds both fields with one atomic.  The memory-side
// atomics cost per element, so about 1.6 per row instead of three.  Called
// by the whole wave.
__device__ __forceinline__ void og_flush_j(OG_GLOBAL(long long) *FX, int cap, og_jpend &p) {
    const int  lane = threadIdx.x & 63;
    const int  f    = og_fx_slot(p.a);
    const int  fz   = (int)__builtin_rintf(-p.z * 524288.0f);  // |z| < 4096: exact scaling, half-even rounding
    const int  fn = __shfl_down(f, 1, 64), zn = __shfl_down(fz, 1, 64), ln = __shfl_down((int)p.live, 1, 64);
    const int  fp = __shfl_up(f, 1, 64), lp = __shfl_up((int)p.live, 1, 64);
    const bool lower = !(f & 1) && lane < 63 && ln && fn == (f ^ 1);  // this lane adds the pair's both fields
    const bool upper = (f & 1) && lane > 0 && lp && fp == (f ^ 1);    // the lane below adds this field
    if (p.live) {
        const unsigned long long xy = og_fxj_pack_f(-p.x, -p.y);
        if (xy != 0) og_atomic_add_fx_g(FX + 3 * cap + f, (long long)xy);
        const long long vz = lower ? (long long)fz + ((long long)zn << 32) : (long long)fz << (32 * (f & 1));
        if (!upper && vz != 0) og_atomic_add_fx_g(FX + OG_FX_ZJ(cap) + (f >> 1), vz);
    }
    p.live = false;
}

// One pass over up to 64 queued rows (lane = row; lanes >= cnt idle).
// m8: the row's excluded pairs (bit i set = skip atom i).  LJ_LDS: the
// cell1's types are rows of the block's LDS copy of the LJ table.
template <int FORM, bool ENERGY, bool LJ_LDS>
__device__ __forceinline__ void og_rows(const og_dev_sys &s, const og_cell1 &c, const og_f4 pj, const unsigned tj,
                                        const unsigned m8, const int a2, const bool live,
                                        const typename og_ljt<FORM>::type *__restrict__ ljl,
                                        const OG_GLOBAL(typename og_ljt<FORM>::type) *__restrict__ LJP,
                                        float (&fi)[3 * OG_NAC], int &npair, float &vmax, float &vsum, double &ev,
                                        double &ee, double &vir, og_jpend &pend) {
    const float rc2 = s.cutoff2;
    const float zcore = s.zcore, bcoeff = s.bcoeff, fcoeff = s.fcoeff;
    const float alpha = s.alpha, piewald = s.piewald;
    const float hps_eps = s.hps_eps, kappa = s.dh_kappa;
    const float zmc = s.zm_core, ze0 = s.zm_c[0], ze1 = s.zm_c[1], ze2 = s.zm_c[2], ze3 = s.zm_c[3];
    const float zg0 = s.zm_pg[0], zg1 = s.zm_pg[1], zg2 = s.zm_pg[2], zg3 = s.zm_pg[3];
    const bool  dh_fix_on = s.dh_grad_fix != 0;
    float fjx = 0.0f, fjy = 0.0f, fjz = 0.0f, fe = 0.0f, fq = 0.0f, fv = 0.0f;
    float px = pj.x, py = pj.y, pz = pj.z, pw = pj.w;
    // the row's eight LJ pairs {A6, A12} load together, ahead of their use
    typename og_ljt<FORM>::type ljv[OG_NAC];
#pragma unroll
    for (int i = 0; i < OG_NAC; i++) ljv[i] = LJ_LDS ? ljl[c.row[i] + tj] : LJP[c.row[i] + tj];
    // the tables hold {6 A6, 12 A12}; a caller's table in the reference's {A6, A12} is scaled here
    if (s.lj_unscaled) {
#pragma unroll
        for (int i = 0; i < OG_NAC; i++) {
            ljv[i].x *= 6.0f;
            ljv[i].y *= 12.0f;
        }
    }
#pragma unroll
    for (int i = 0; i < OG_NAC; i++) {
        const float dx = c.x[i] - px;
        const float dy = c.y[i] - py;
        const float dz = c.z[i] - pz;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const bool  in = r2 < rc2, unmasked = !((m8 >> i) & 1u);
        const bool  ok = in && unmasked;
        // the pair count on the scalar unit: each ballot is its compare's lane mask as it stands
        if constexpr (ENERGY)
            npair += __builtin_popcountll(__builtin_amdgcn_ballot_w64(in) & __builtin_amdgcn_ballot_w64(unmasked));
        const typename og_ljt<FORM>::type lj = ljv[i];  // {6 A6, 12 A12(, HPS switch, HPS lambda)}
        // ZSEL (energy steps, forms 0, 1, 3, 4): a pair that is not evaluated
        // gets 1/r = 0 and charge product 0, which zero its gradient and both
        // energies (one select instead of three); the HPS form's short-range
        // constant needs the selects
        constexpr bool ZSEL = ENERGY && FORM != 2;
        const float rinv  = ZSEL && !ok ? 0.0f : __builtin_amdgcn_rsqf(r2);
        const float r2inv = rinv * rinv;
        const float r6inv = r2inv * r2inv * r2inv;
        // r^2 times the LJ gradient coefficient, 6 A6 r^-6 - 12 A12 r^-12, and (energy steps) the LJ energy
        const float lji   = fmaf(-lj.y, r6inv, lj.x);
        const float ljf   = r6inv * lji;
        const float cc    = ZSEL && !ok ? 0.0f : c.q[i] * pw;
        float       g, eel, elj = 0.0f;
        // LJ energy: forms 0, 1, 3, 4 keep -12 times it, r^-6 (6 A6 + lji) = r^-6 (12 A6 - 12 A12 r^-6), and the
        // cell's sum is scaled once (two fewer instructions per pair slot); the HPS form keeps the energy itself
        if constexpr (ENERGY) {
            if constexpr (FORM == 2) elj = r6inv * fmaf(lj.y * (1.0f / 12.0f), r6inv, lj.x * (-1.0f / 6.0f));
            else elj = r6inv * (lj.x + lji);
        }
        if constexpr (FORM == 0) {
            eel = cc * (rinv - zcore + bcoeff * r2);
            g   = r2inv * ljf - cc * (rinv * r2inv - fcoeff);
        } else if constexpr (FORM == 2) {
            // HPS: beyond the pair's switch distance lambda * LJ, inside it LJ + (1 - lambda) eps;
            // Debye-Hueckel: cc (charges x 332.06378 / eps_r) exp(-kappa r) / r, gradient coefficient
            // -E (1/r + kappa) as the reference (times 1/r with the gradient fix)
            const float r   = r2 * rinv;
            const bool  far = r > lj.z;
            const float sc  = far ? lj.w : 1.0f;
            elj             = fmaf(elj, sc, far ? 0.0f : (1.0f - lj.w) * hps_eps);
            eel             = cc * rinv * __expf(-r * kappa);
            g = sc * r2inv * ljf - eel * (rinv + kappa) * (dh_fix_on ? rinv : 1.0f);
        } else if constexpr (FORM == 3 || FORM == 4) {
            // zero-quadrupole/-octupole/-hexadecapole: the order's polynomials in r^2 by Horner's rule
            const float pe = r2 * fmaf(r2, fmaf(r2, fmaf(r2, ze3, ze2), ze1), ze0);
            const float pg = fmaf(r2, fmaf(r2, fmaf(r2, zg3, zg2), zg1), zg0);
            if constexpr (FORM == 3) {
                eel = cc * (rinv - zmc - pe);
                g   = r2inv * ljf - cc * (rinv * r2inv + pg);
            } else {
                const float r   = r2 * rinv;
                const float ar  = alpha * r;
                const float ex  = __expf(-ar * ar);
                const float erc = og_erfc_as(ar, ex);
                eel = cc * (rinv * erc - zmc - pe);
                g   = r2inv * ljf - cc * (erc * rinv * r2inv + piewald * ex * r2inv + pg);
            }
        } else {
            const float r   = r2 * rinv;
            const float ar  = alpha * r;
            const float ex  = __expf(-ar * ar);
            const float erc = og_erfc_as(ar, ex);
            eel = cc * (rinv * erc - zcore + bcoeff * r2);
            g   = r2inv * ljf - cc * (erc * rinv * r2inv + piewald * ex * r2inv - fcoeff);
        }
        // lanes without a pair (masked, out of range, idle) contribute exact zeros
        if (!ZSEL) g = ok ? g : 0.0f;
        fjx               = fmaf(g, dx, fjx);
        fjy               = fmaf(g, dy, fjy);
        fjz               = fmaf(g, dz, fjz);
        fi[3 * i]         = fmaf(g, dx, fi[3 * i]);
        fi[3 * i + 1]     = fmaf(g, dy, fi[3 * i + 1]);
        fi[3 * i + 2]     = fmaf(g, dz, fi[3 * i + 2]);
        if (ENERGY) {
            fe += ZSEL || ok ? elj : 0.0f;
            fq += ZSEL || ok ? eel : 0.0f;
            fv = fmaf(g, r2, fv);
        }
        // finish atom i here: without this the compiler defers the
        // accumulations and keeps all 8 atoms' distances live
        asm volatile("" : "+v"(fjx), "+v"(fjy), "+v"(fjz), "+v"(fi[3 * i]), "+v"(fi[3 * i + 1]), "+v"(fi[3 * i + 2]),
                     "+v"(px), "+v"(py), "+v"(pz), "+v"(pw));
        if constexpr (ENERGY) asm volatile("" : "+v"(fe), "+v"(fq), "+v"(fv));
    }
    pend.x    = fjx;
    pend.y    = fjy;
    pend.z    = fjz;
    pend.a    = a2;
    pend.live = live;
    // range check of the converted values; fmaxf drops NaN, the sum keeps it
    vmax = fmaxf(vmax, fmaxf(fabsf(fjx), fmaxf(fabsf(fjy), fabsf(fjz))));
    vsum += fjx + fjy + fjz;
    if (ENERGY) {
        ev += (double)fe;
        ee += (double)fq;
        vir += (double)fv;
    }
}

// Pass over queue entries [q0, q0 + cnt): reads the rows, adds the previous
// pass's j-gradient, then evaluates the rows.
template <int FORM, bool ENERGY, bool LJ_LDS>
__device__ __forceinline__ void og_pass(const og_dev_sys &s, const og_cell1 &c, const og_f4 *qp, const int2 *qa,
                                        const int q0, const int cnt, const typename og_ljt<FORM>::type *ljl,
                                        const OG_GLOBAL(typename og_ljt<FORM>::type) *__restrict__ LJP,
                                        OG_GLOBAL(long long) *FX,
                                        float (&fi)[3 * OG_NAC], int &npair, float &vmax, float &vsum, double &ev,
                                        double &ee, double &vir, og_jpend &pend) {
    const int  lane = threadIdx.x & 63;
    const int  slot = (q0 + lane) & (OG_QCAP - 1);
    const bool live = lane < cnt;
    og_f4      pj   = qp[slot];
    const int2 ax   = qa[slot];
    // idle lanes (stale queue slots): finite coordinates, LJ column 0, every pair masked
    if (!live) pj = og_f4{0.0f, 0.0f, 0.0f, 0.0f};
    const unsigned m8  = live ? ((unsigned)ax.y >> 24) : 0xffu;
    const unsigned tj  = live ? ((unsigned)ax.y & 0xffffffu) : 0u;
    // every outstanding load of the filter is complete before the atomics go
    // out, so nothing in the pass waits on the vector-memory counter
    // (s_waitcnt vmcnt(0); expcnt and lgkmcnt left at their maxima)
    __builtin_amdgcn_s_waitcnt(0x0F70);
    og_flush_j(FX, s.cap_atom_array, pend);
    og_rows<FORM, ENERGY, LJ_LDS>(s, c, pj, tj, m8, ax.x, live, ljl, LJP, fi, npair, vmax, vsum, ev, ee, vir, pend);
}

// a <- the kept half of (a, b) summed with the partner lane's other half
// (partner = lane ^ off, off = 32 or 16); "upper" lanes keep b.
__device__ __forceinline__ float og_keep_add(float a, float b, bool upper, int off) {
    const float keep = upper ? b : a;
    const float send = upper ? a : b;
    return keep + __shfl_xor(send, off, 64);
}

// One cell1 (one wave), its cell-pair records [p0, p1) (all of them, or
// half of them for a split cell).  ENERGY = false on steps whose energies
// are not logged: forces and the pair count only.  LJ_LDS: the cell1's LJ rows are in
// the block's LDS copy ljl of the table's leading rows.
template <int FORM, bool ENERGY, bool LJ_LDS>
__device__ __forceinline__ void og_pair_cell(const og_dev_sys &s, const int c1, const int p0, const int p1,
                                             const og_f4 *__restrict__ shift,
                                             const typename og_ljt<FORM>::type *__restrict__ ljl, og_f4 *qp,
                                             int2 *qa) {
    const int lane = threadIdx.x & 63;
    const int jl   = lane & 7;

    const OG_GLOBAL(og_f4) *__restrict__ X      = (const OG_GLOBAL(og_f4) *)s.gxyzq;
    const OG_GLOBAL(int) *__restrict__ TY       = (const OG_GLOBAL(int) *)s.gtype;
    typedef typename og_ljt<FORM>::type ljt;
    const OG_GLOBAL(ljt) *__restrict__ LJP =
        (const OG_GLOBAL(ljt) *)(FORM == 2 ? (const void *)s.lj_hps : (const void *)s.lj_pair);
    const OG_GLOBAL(unsigned) *__restrict__ CP  = (const OG_GLOBAL(unsigned) *)s.cp_c2;
    const OG_GLOBAL(uint64_t) *__restrict__ CM  = (const OG_GLOBAL(uint64_t) *)s.cp_mask;
    OG_GLOBAL(long long) *__restrict__ FX       = (OG_GLOBAL(long long) *)s.fx;
    const int   cap = s.cap_atom_array;
    const int   nt  = s.n_types;
    // keep-test radius: the cutoff plus 10 ppm, so float rounding of the box
    // distance can never drop a row that holds an in-range pair
    const float rcb2 = s.cutoff2 * 1.00001f;

    // the first group's cell-pair records load with the cell1 atoms
    // (lane = 8 * (cell pair in the group) + j atom)
    unsigned cp = 0u;
    uint64_t mk = ~0ull;
    bool     valid = false;
    if (p0 < p1) {
        const int t = p0 + (lane >> 3);
        valid       = t < p1;
        const int tc = valid ? t : p1 - 1;
        cp = CP[tc];
        mk = CM[tc];
    }

    og_cell1 c;
    {
        const og_f4 p  = X[c1 * OG_NAC + jl];
        const int   ty = TY[c1 * OG_NAC + jl];
        // charge x 332.06378 (x 1/eps_r under Debye-Hueckel)
        const float qc = FORM == 2 ? p.w * OG_CHARGE_COEFF * s.dh_diel_inv : p.w * OG_CHARGE_COEFF;
#pragma unroll
        for (int i = 0; i < OG_NAC; i++) {
            c.x[i]   = og_rlane(p.x, i);
            c.y[i]   = og_rlane(p.y, i);
            c.z[i]   = og_rlane(p.z, i);
            c.q[i]   = og_rlane(qc, i);
            const unsigned t = (unsigned)__builtin_amdgcn_readlane(ty, i) * (unsigned)nt;
            c.row[i] = t;
        }
        c.lo[0] = c.hi[0] = c.x[0];
        c.lo[1] = c.hi[1] = c.y[0];
        c.lo[2] = c.hi[2] = c.z[0];
#pragma unroll
        for (int i = 1; i < OG_NAC; i++) {
            c.lo[0] = fminf(c.lo[0], c.x[i]), c.hi[0] = fmaxf(c.hi[0], c.x[i]);
            c.lo[1] = fminf(c.lo[1], c.y[i]), c.hi[1] = fmaxf(c.hi[1], c.y[i]);
            c.lo[2] = fminf(c.lo[2], c.z[i]), c.hi[2] = fmaxf(c.hi[2], c.z[i]);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {  // wave-uniform: keep in scalar registers
            c.lo[k] = og_rlane(c.lo[k], 0);
            c.hi[k] = og_rlane(c.hi[k], 0);
        }
    }

    float  fi[3 * OG_NAC];
#pragma unroll
    for (int k = 0; k < 3 * OG_NAC; k++) fi[k] = 0.0f;
    double ev = 0.0, ee = 0.0, vir = 0.0;  // ev: -12 x the LJ energy except under HPS (scaled at the end)
    int    npair = 0;                    // in-cutoff unmasked pairs of this cell1 (wave-uniform)
    float  vmax = 0.0f, vsum = 0.0f;     // j-gradient range check
    int    qn = 0, qd = 0;               // rows queued / rows evaluated (wave-uniform)
    og_jpend pend{0.0f, 0.0f, 0.0f, 0, false};

    for (int base = p0; base < p1; base += OG_NAC) {
        const int   a2 = (int)(cp & 0xffffffu) * OG_NAC + jl;
        const og_f4 pj = X[a2];
        const int   ty = TY[a2];
        // the next group's records load while this group's atoms are in flight
        // (past the end: a valid record that the validity flag masks)
        const int      tn     = base + OG_NAC + (lane >> 3);
        const bool     validn = tn < p1;
        const int      tcn    = validn ? tn : p1 - 1;
        const unsigned cpn    = CP[tcn];
        const uint64_t mkn    = CM[tcn];
        // wait for this group's atoms (and types) here, with the next records still in flight
        asm volatile("" ::"v"(pj.x), "v"(ty));
        const og_f4    sh = shift[(cp >> 24) & 63u];  // periodic image of the cell pair
        const float    px = pj.x + sh.x, py = pj.y + sh.y, pz = pj.z + sh.z;
        const float    gx = fmaxf(fmaxf(c.lo[0] - px, px - c.hi[0]), 0.0f);
        const float    gy = fmaxf(fmaxf(c.lo[1] - py, py - c.hi[1]), 0.0f);
        const float    gz = fmaxf(fmaxf(c.lo[2] - pz, pz - c.hi[2]), 0.0f);
        const unsigned m8 = valid ? (unsigned)(mk >> (8 * jl)) & 0xffu : 0xffu;
        const bool     use = (gx * gx + gy * gy + gz * gz < rcb2) && m8 != 0xffu;
        const uint64_t ub  = __ballot(use);
        const int2     ent = make_int2(a2, (int)((unsigned)ty | (m8 << 24)));
        if (use) {
            const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(ub >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)ub, 0u));
            const int slot = (qn + rank) & (OG_QCAP - 1);
            qp[slot] = og_f4{px, py, pz, pj.w};
            qa[slot] = ent;
        }
        qn += __builtin_popcountll(ub);
        if (qn - qd >= 64) {
            og_pass<FORM, ENERGY, LJ_LDS>(s, c, qp, qa, qd, 64, ljl, LJP, FX, fi, npair, vmax, vsum, ev, ee, vir, pend);
            qd += 64;
        }
        cp    = cpn;
        mk    = mkn;
        valid = validn;
    }
    if (qn > qd)
        og_pass<FORM, ENERGY, LJ_LDS>(s, c, qp, qa, qd, qn - qd, ljl, LJP, FX, fi, npair, vmax, vsum, ev, ee, vir, pend);
    og_flush_j(FX, cap, pend);

    // i-gradients: reduce-scatter the 24 partial sums over the wave.  After
    // the xor-32/16/8 steps lane L holds atom L>>3's (x, y, z) summed over 8
    // lanes; an 8-lane butterfly completes the sum.
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
    float      h12[12], h6[6], h3[3];
#pragma unroll
    for (int k = 0; k < 12; k++) h12[k] = og_keep_add(fi[k], fi[12 + k], b5, 32);
#pragma unroll
    for (int k = 0; k < 6; k++) h6[k] = og_keep_add(h12[k], h12[6 + k], b4, 16);
#pragma unroll
    for (int k = 0; k < 3; k++) h3[k] = og_dpp_sum8(og_keep_add(h6[k], h6[3 + k], b3, 8));
    if (!(vmax <= OG_FXJ_ROW_LIMIT) || !(fabsf(vsum) < INFINITY)) atomicOr(&s.cnt->flags, OG_F_NAN);
    if (jl == 0) {
        const int a1 = c1 * OG_NAC + (lane >> 3);
        if (!(fabsf(h3[0]) + fabsf(h3[1]) + fabsf(h3[2]) <= (float)OG_FX_LIMIT)) atomicOr(&s.cnt->flags, OG_F_NAN);
        og_add_fx3(FX, cap, og_fx_slot(a1), (double)h3[0], (double)h3[1], (double)h3[2]);
    }
    if (!ENERGY) return;
    for (int off = 1; off < 64; off <<= 1) {
        ev += __shfl_xor(ev, off, 64);
        ee += __shfl_xor(ee, off, 64);
        vir += __shfl_xor(vir, off, 64);
    }
    if (lane == 0) {
        OG_GLOBAL(long long) *A = (OG_GLOBAL(long long) *)s.e_acc + (c1 % OG_EACC_SLOTS) * OG_EACC_STRIDE;
        og_atomic_add_fx_g(A + OG_EACC_VDW, og_to_fixed(FORM == 2 ? ev : ev * (-1.0 / 12.0)));
        og_atomic_add_fx_g(A + OG_EACC_ELE, og_to_fixed(ee));
        og_atomic_add_fx_g(A + OG_EACC_VIR, og_to_fixed(vir));
        og_atomic_add_fx_g(A + OG_EACC_PAIRS, (long long)npair);
    }
}

// WAVES waves (one work unit each: a cell, or half of one of the last
// cells) per block sharing the image table and an LDS
// copy of the leading rows of the LJ table (as many whole rows as fit in
// OG_LJ_LDS_ENTRIES; the engine numbers types by frequency, so for water the
// oxygen and hydrogen rows against every type).  A cell1 whose types are all
// in it reads LJ pairs from LDS, otherwise from the global table.  The block
// size is a launch shape the autotuner chooses (results do not depend on it).
template <int FORM, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) k_pair(const og_dev_sys *__restrict__ S, const int2 *__restrict__ map) {
    typedef typename og_ljt<FORM>::type ljt;
    __shared__ og_f4 shift[64];              // image code -> shift of cell2
    __shared__ ljt   ljl[OG_LJ_LDS_ENTRIES];
    __shared__ og_f4 qpos[WAVES][OG_QCAP];  // row queue: shifted x, y, z, charge
    __shared__ int2  qaux[WAVES][OG_QCAP];  // grid slot, LJ column | mask byte << 24
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    // block-uniform exits (one system per block; this kernel never sets the overflow flag)
    if (og_frozen(s) || (s.cnt->flags & OG_F_CP_OVERFLOW)) return;
    const int nt   = s.n_types;
    const int rows = min(nt, OG_LJ_LDS_ENTRIES / nt);  // table rows held in LDS
    if (threadIdx.x < 64) {
        // image bits of MiniCell::add_cell_pair (MiniCell.cpp:1174-1187): neg bit -> -L, pos bit -> +L
        const unsigned img = threadIdx.x;
        const float    L0 = s.L[0], L1 = s.L[1], L2 = s.L[2];
        shift[img] = og_f4{(img & 1u) ? -L0 : ((img & 2u) ? L0 : 0.0f), (img & 4u) ? -L1 : ((img & 8u) ? L1 : 0.0f),
                           (img & 16u) ? -L2 : ((img & 32u) ? L2 : 0.0f), 0.0f};
    }
    const ljt *tab = (const ljt *)(FORM == 2 ? (const void *)s.lj_hps : (const void *)s.lj_pair);
    for (int k = threadIdx.x; k < rows * nt; k += 64 * WAVES) ljl[k] = tab[k];
    __syncthreads();
    // work unit u of the system: cells 0 .. n-m-1 whole, then the last m
    // cells as two units each (first and second half of their records)
    const int w  = threadIdx.x >> 6;
    const int n  = s.cnt->n_cells;
    const int m  = og_pair_split_cells(s.pair_split, n);
    const int u  = __builtin_amdgcn_readfirstlane(bm.y * WAVES + w);
    if (u >= n + m) return;
    int c1 = u, half = -1;
    if (u >= n - m) {
        c1   = n - m + ((u - (n - m)) >> 1);
        half = (u - (n - m)) & 1;
    }
    int p0 = __builtin_amdgcn_readfirstlane(s.cp_head[c1]);
    int p1 = __builtin_amdgcn_readfirstlane(s.cp_head[c1 + 1]);
    if (half >= 0) {
        const int pm = p0 + ((p1 - p0 + 1) >> 1);
        if (half == 0) p1 = pm;
        else p0 = pm;
    }
    const bool lds = __builtin_amdgcn_readfirstlane(__all(s.gtype[c1 * OG_NAC + (threadIdx.x & 7)] < rows)) != 0;
    if (og_energy_step(s)) {
        if (lds) og_pair_cell<FORM, true, true>(s, c1, p0, p1, shift, ljl, qpos[w], qaux[w]);
        else og_pair_cell<FORM, true, false>(s, c1, p0, p1, shift, ljl, qpos[w], qaux[w]);
    } else {
        if (lds) og_pair_cell<FORM, false, true>(s, c1, p0, p1, shift, ljl, qpos[w], qaux[w]);
        else og_pair_cell<FORM, false, false>(s, c1, p0, p1, shift, ljl, qpos[w], qaux[w]);
    }
}

// Per atom: fixed-point pair gradient -> FP64 "work" with the atom's
// excluded-pair and bond terms (og_atom_work).
__global__ void __launch_bounds__(OG_THREADS) k_forces_to_work(const og_dev_sys *__restrict__ S,
                                                               const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const int i    = bm.y * OG_THREADS + threadIdx.x;
    double    e_ex = 0.0, e_b = 0.0;
    if (i < s.n_atoms) og_atom_work(s, i, e_ex, e_b);
    if (!og_energy_step(s)) return;
    __shared__ double sh[2 * (OG_THREADS / 64)];
    og_add_misc_energies(s, bm.y, e_ex, e_b, sh);
}

// Energies of a single-point force call from the fixed-point accumulators
// (an MD step takes them in k_step_close).
__global__ void __launch_bounds__(64) k_energy_reduce(const og_dev_sys *__restrict__ S,
                                                                  const int2 *__restrict__ map) {
    const og_dev_sys &s = S[map[blockIdx.x].x];
    if (threadIdx.x == 0) s.cnt->rebuild_now = 0;  // a forced rebuild of this evaluation is done
    if (og_frozen(s) || !og_energy_step(s)) return;
    __shared__ long long sh[16 * OG_EACC_STRIDE];
    og_take_pair_energies<64>(s, sh);
}

// ------------------------------------------------------------- launchers --
template <int WAVES>
static void og_launch_pair_w(const og_launch &a, int form) {
    if (form == OG_FORM_HPS_DH)
        hipLaunchKernelGGL((k_pair<2, WAVES>), dim3(a.n_blocks), dim3(64 * WAVES), 0, a.stream, a.sys, a.map);
    else if (form == OG_FORM_ZM)
        hipLaunchKernelGGL((k_pair<3, WAVES>), dim3(a.n_blocks), dim3(64 * WAVES), 0, a.stream, a.sys, a.map);
    else if (form == OG_FORM_ZM_ALPHA)
        hipLaunchKernelGGL((k_pair<4, WAVES>), dim3(a.n_blocks), dim3(64 * WAVES), 0, a.stream, a.sys, a.map);
    else if (form == OG_FORM_ZD_ALPHA)
        hipLaunchKernelGGL((k_pair<1, WAVES>), dim3(a.n_blocks), dim3(64 * WAVES), 0, a.stream, a.sys, a.map);
    else
        hipLaunchKernelGGL((k_pair<0, WAVES>), dim3(a.n_blocks), dim3(64 * WAVES), 0, a.stream, a.sys, a.map);
}

void og_launch_pair(const og_launch &a, int form, int waves) {
    switch (waves) {
    case 1: og_launch_pair_w<1>(a, form); break;
    case 2: og_launch_pair_w<2>(a, form); break;
    case 8: og_launch_pair_w<8>(a, form); break;
    default: og_launch_pair_w<4>(a, form); break;
    }
}
void og_launch_forces_to_work(const og_launch &a) {
    hipLaunchKernelGGL(k_forces_to_work, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_energy_reduce(const og_launch &a) {
    hipLaunchKernelGGL(k_energy_reduce, dim3(a.n_blocks), dim3(64), 0, a.stream, a.sys, a.map);
}
