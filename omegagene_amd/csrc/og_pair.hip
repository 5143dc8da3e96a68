// og_pair.hip -- LJ + zero-dipole cell-pair kernel, excluded-pair (excess)
// and bond terms, fixed-point gradient reduction and energy reductions.
//
// Pair kernel (reference: kernel_pairwise_ljzd + cal_pair, cuda_setup.cu:487-675;
// CPU: SubBox::calc_energy_pairwise, SubBox.cpp:956-1125): one wave64 per
// cell1; every cell pair is one 8x8 tile with lane = a1 + 8*a2, so one pass
// over 64 lanes covers the tile (the reference needs two half passes of a
// warp32).  FP32 pair math, j-gradients reduced over the 8 i-lanes with DPP
// and added as 64-bit fixed point, i-gradients kept in registers across the
// cell1's pairs, energies accumulated per lane in FP64 and reduced per cell
// in a fixed order (deterministic, no floating-point atomics).
#include "og_internal.h"

#define OG_CHARGE_COEFF 332.06378f   // CelesteObject.cpp:36

typedef float og_f4 __attribute__((ext_vector_type(4)));
typedef float og_f2 __attribute__((ext_vector_type(2)));

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

// Adds one tile's j-gradient.  After the 8-lane butterfly every lane of a
// j row holds the row's (x, y, z) sums bit for bit, so lane k < 3 of the row
// converts and adds component k: one conversion and one atomic instruction
// serve all three components of the 8 j atoms (24 active lanes).
// The largest converted magnitude is kept in vmax and checked once per cell
// against the conversion range; fmaxf drops a NaN operand, so a NaN is
// caught instead by the running sum vsum.
__device__ __forceinline__ void og_flush_j(OG_GLOBAL(long long) *FX, int cap, int a2, int i, float rx, float ry,
                                           float rz, float &vmax, float &vsum) {
    const float     v  = i == 0 ? rx : (i == 1 ? ry : rz);
    const long long fv = og_fx_fast(-(double)v);
    if (i < 3 && fv != 0) og_atomic_add_fx_g(FX + i * cap + a2, fv);
    vmax = fmaxf(vmax, fabsf(v));
    vsum += v;
}

// Periodic shift of a tile's j cell (image bits of MiniCell::add_cell_pair,
// MiniCell.cpp:1174-1187: neg bit -> -L, pos bit -> +L) as integer selects on
// wave-uniform values, so it stays on the scalar unit.
__device__ __forceinline__ float og_img_shift(unsigned img, unsigned neg, unsigned pos, float L) {
    const int Lb = __builtin_amdgcn_readfirstlane(__float_as_int(L));
    const int b  = (img & neg) ? (Lb ^ (int)0x80000000) : ((img & pos) ? Lb : 0);
    return __int_as_float(__builtin_amdgcn_readfirstlane(b));
}

// One cell1 (one wave).  ENERGY = false on steps whose energies are not
// logged: forces and the pair count only.
template <int ALPHA, bool ENERGY> __device__ __forceinline__ void og_pair_cell(const og_dev_sys &s, const int c1) {
    const int lane = threadIdx.x & 63;
    const int i    = lane & 7;
    const int jl   = lane >> 3;

    const OG_GLOBAL(og_f4) *__restrict__ X      = (const OG_GLOBAL(og_f4) *)s.gxyzq;
    const OG_GLOBAL(int) *__restrict__ TY       = (const OG_GLOBAL(int) *)s.gtype;
    const OG_GLOBAL(og_f2) *__restrict__ LJP    = (const OG_GLOBAL(og_f2) *)s.lj_pair;
    const OG_GLOBAL(unsigned) *__restrict__ CP  = (const OG_GLOBAL(unsigned) *)s.cp_c2;
    const OG_GLOBAL(uint64_t) *__restrict__ CM  = (const OG_GLOBAL(uint64_t) *)s.cp_mask;
    OG_GLOBAL(long long) *__restrict__ FX       = (OG_GLOBAL(long long) *)s.fx;
    const int   cap    = s.cap_atom_array;
    const int   nt     = s.n_types;
    const float rc2    = s.cutoff2;
    const float zcore  = s.zcore, bcoeff = s.bcoeff, fcoeff = s.fcoeff;
    const float alpha  = s.alpha, piewald = s.piewald;
    const float L0 = s.L[0], L1 = s.L[1], L2 = s.L[2];
    const int   p0 = __builtin_amdgcn_readfirstlane(s.cp_head[c1]);
    const int   p1 = __builtin_amdgcn_readfirstlane(s.cp_head[c1 + 1]);

    const og_f4  pi   = X[c1 * OG_NAC + i];
    const unsigned trow = (unsigned)(TY[c1 * OG_NAC + i] * nt);
    const float  qi   = pi.w * OG_CHARGE_COEFF;

    double dfx = 0.0, dfy = 0.0, dfz = 0.0, ev = 0.0, ee = 0.0, vir = 0.0;
    int    npair = 0;  // in-cutoff unmasked pairs of this cell1 (wave-uniform)
    float  vmax = 0.0f, vsum = 0.0f;  // j-gradient range check (see og_flush_j)

    for (int base = p0; base < p1; base += 64) {
        const int nb = min(64, p1 - base);
        // 64 cell-pair records per vector load; broadcast with readlane
        const unsigned cpv = lane < nb ? CP[base + lane] : 0u;
        const uint64_t mkv = lane < nb ? CM[base + lane] : ~0ull;
        const unsigned mlo = (unsigned)mkv, mhi = (unsigned)(mkv >> 32);
        float fx = 0.0f, fy = 0.0f, fz = 0.0f, fe = 0.0f, fq = 0.0f, fv = 0.0f;
        unsigned c2n = __builtin_amdgcn_readlane(cpv, 0);
        int      a2n = (int)(c2n & 0xffffffu) * OG_NAC + jl;
        og_f4    pjn = X[a2n];
        int      tjn = TY[a2n];
        for (int k = 0; k < nb; k++) {
            const unsigned c2i  = c2n;
            const int      a2   = a2n;
            const og_f4    pj   = pjn;
            const int      tj   = tjn;
            if (k + 1 < nb) {  // prefetch the next tile's j atoms
                c2n = __builtin_amdgcn_readlane(cpv, k + 1);
                a2n = (int)(c2n & 0xffffffu) * OG_NAC + jl;
                pjn = X[a2n];
                tjn = TY[a2n];
            }
            // readlane returns int: widen through uint32_t so the low word is not sign-extended
            const uint64_t mask = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(mhi, k) << 32) |
                                  (uint64_t)(uint32_t)__builtin_amdgcn_readlane(mlo, k);
            // image shift of the tile: wave-uniform, selected on the scalar unit
            const unsigned img = c2i >> 24;
            const float    sx  = og_img_shift(img, 1u, 2u, L0);
            const float    sy  = og_img_shift(img, 4u, 8u, L1);
            const float    sz  = og_img_shift(img, 16u, 32u, L2);
            const float    dx  = pi.x - (pj.x + sx);
            const float    dy  = pi.y - (pj.y + sy);
            const float    dz  = pi.z - (pj.z + sz);
            const float    r2  = dx * dx + dy * dy + dz * dz;
            const bool     excl = __builtin_amdgcn_inverse_ballot_w64(mask);
            const bool     ok   = !excl && (r2 < rc2);
            const uint64_t okb  = __ballot(ok);
            npair += __builtin_popcountll(okb);
            // About 40% of the tiles of a 13 A list hold no pair inside the
            // 12 A cutoff: skip them.  Their contributions are exact zeros, so
            // the result is bit-identical to evaluating them.
            if (okb == 0) continue;
            const og_f2  lj  = LJP[trow + (unsigned)tj];  // {A6, A12}, one 8-byte load
            const float  a6  = lj.x;
            const float  a12 = lj.y;
            const float    r2s  = ok ? r2 : 1.0f;
            const float    rinv  = __builtin_amdgcn_rsqf(r2s);
            const float    r2inv = rinv * rinv;
            const float    r6inv = r2inv * r2inv * r2inv;
            const float    t6    = a6 * r6inv;
            const float    t12   = a12 * r6inv * r6inv;
            const float    cc    = qi * pj.w;
            float          g, eel;
            if (ALPHA == 0) {
                eel = cc * (rinv - zcore + bcoeff * r2s);
                g   = r2inv * (6.0f * t6 - 12.0f * t12) - cc * (rinv * r2inv - fcoeff);
            } else {
                const float r   = r2s * rinv;
                const float ar  = alpha * r;
                const float ex  = __expf(-ar * ar);
                const float erc = og_erfc_as(ar, ex);
                eel = cc * (rinv * erc - zcore + bcoeff * r2s);
                g   = r2inv * (6.0f * t6 - 12.0f * t12) - cc * (erc * rinv * r2inv + piewald * ex * r2inv - fcoeff);
            }
            g              = ok ? g : 0.0f;
            const float wx = g * dx, wy = g * dy, wz = g * dz;
            fx += wx;
            fy += wy;
            fz += wz;
            if (ENERGY) {
                fe += ok ? (t12 - t6) : 0.0f;
                fq += ok ? eel : 0.0f;
                fv += g * r2s;
            }
            const float rjx = og_dpp_sum8(wx);
            const float rjy = og_dpp_sum8(wy);
            const float rjz = og_dpp_sum8(wz);
            og_flush_j(FX, cap, a2, i, rjx, rjy, rjz, vmax, vsum);
        }
        dfx += (double)fx;
        dfy += (double)fy;
        dfz += (double)fz;
        if (ENERGY) {
            ev += (double)fe;
            ee += (double)fq;
            vir += (double)fv;
        }
    }
    for (int off = 8; off < 64; off <<= 1) {
        dfx += __shfl_xor(dfx, off, 64);
        dfy += __shfl_xor(dfy, off, 64);
        dfz += __shfl_xor(dfz, off, 64);
    }
    if (!(vmax <= (float)OG_FX_LIMIT) || !(fabsf(vsum) < INFINITY)) atomicOr(&s.cnt->flags, OG_F_NAN);
    if (jl == 0) {
        const int a1 = c1 * OG_NAC + i;
        if (!(fabs(dfx) + fabs(dfy) + fabs(dfz) <= OG_FX_LIMIT)) atomicOr(&s.cnt->flags, OG_F_NAN);
        const long long vx = og_fx_fast(dfx), vy = og_fx_fast(dfy), vz = og_fx_fast(dfz);
        if ((vx | vy | vz) != 0) {
            og_atomic_add_fx_g(FX + a1, vx);
            og_atomic_add_fx_g(FX + cap + a1, vy);
            og_atomic_add_fx_g(FX + 2 * cap + a1, vz);
        }
    }
    if (!ENERGY) return;
    for (int off = 1; off < 64; off <<= 1) {
        ev += __shfl_xor(ev, off, 64);
        ee += __shfl_xor(ee, off, 64);
        vir += __shfl_xor(vir, off, 64);
    }
    if (lane == 0) {
        s.e_cell[4 * c1]     = ev;
        s.e_cell[4 * c1 + 1] = ee;
        s.e_cell[4 * c1 + 2] = vir;
        s.e_cell[4 * c1 + 3] = (double)npair;
    }
}

template <int ALPHA>
__global__ void __launch_bounds__(OG_THREADS) k_pair(const og_dev_sys *__restrict__ S, const int2 *__restrict__ map) {
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const int c1 = __builtin_amdgcn_readfirstlane(bm.y * (OG_THREADS / 64) + (threadIdx.x >> 6));
    if (c1 >= s.cnt->n_cells || (s.cnt->flags & OG_F_CP_OVERFLOW)) return;
    if (og_energy_step(s)) og_pair_cell<ALPHA, true>(s, c1);
    else og_pair_cell<ALPHA, false>(s, c1);
}

// ------------------------------------------------------------ misc terms --
// Excluded-pair ZD terms (SubBox::calc_energy_ele_excess SubBox.cpp:1347-1367,
// ForceField::calc_zms_excess ForceField.cpp:420-437,
// calc_zero02pole_excess(_alpha0) ZeroMultipoleSum.cpp:191-217) in the
// reference float arithmetic, and harmonic bonds E = k (r - r0)^2 in FP64
// (ForceField::calc_bond ForceField.cpp:108-136).  Both add into the
// fixed-point gradient at the atoms' grid slots.
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

// Excluded pair (a, b), a < b in the excess list: gradient w added to a and
// subtracted from b, energy returned (reference float arithmetic).
__device__ __forceinline__ float og_excess_pair(const og_dev_sys &s, int a, int b, float *w) {
#pragma clang fp contract(off)
    float d[3];
    for (int k = 0; k < 3; k++) {
        const float Lh  = s.L[k] * 0.5f;
        const float Lhi = 1.0f / Lh;
        float       v   = (float)s.crd[3 * a + k] - (float)s.crd[3 * b + k];
        v               = v - s.L[k] * (float)(int)(v * Lhi);
        v               = v - s.L[k] * (float)(int)(v * Lhi);
        d[k]            = v;
    }
    const float r2    = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const float r     = sqrtf(r2);
    const float rinv  = 1.0f / r;
    const float r2inv = rinv * rinv;
    const float r3inv = r2inv * rinv;
    const float cc    = (float)((double)(s.charge[a] * s.charge[b]) * 332.06378);
    float       ene, gc;
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

// Harmonic bond (a, b): gradient w added to a and subtracted from b, energy returned (FP64).
__device__ __forceinline__ double og_bond_pair(const og_dev_sys &s, int t, int a, int b, double *w) {
#pragma clang fp contract(off)
    double d[3];
    for (int k = 0; k < 3; k++) {
        const float Lh  = s.L[k] * 0.5f;
        const float Lhi = 1.0f / Lh;
        double      v   = s.crd[3 * a + k] - s.crd[3 * b + k];
        v               = v - (double)(s.L[k] * (float)(int)(v * (double)Lhi));
        v               = v - (double)(s.L[k] * (float)(int)(v * (double)Lhi));
        d[k]            = v;
    }
    const double r    = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double dif  = r - s.bond_r0[t];
    const double coef = s.bond_k[t] * dif;
    const double gcf  = 2.0 * coef / r;
    for (int k = 0; k < 3; k++) w[k] = d[k] * gcf;
    return coef * dif;
}

// Per atom: fixed-point pair gradient (grid order) -> FP64 "work" in atom order
// (SubBox::add_work_from_minicell, SubBox.cpp:1398-1407), plus the atom's
// excluded-pair ZD terms (SubBox::calc_energy_ele_excess SubBox.cpp:1347-1367,
// ForceField::calc_zms_excess ForceField.cpp:420-437,
// calc_zero02pole_excess(_alpha0) ZeroMultipoleSum.cpp:191-217) and bonds
// E = k (r - r0)^2 (ForceField::calc_bond ForceField.cpp:108-136) through
// per-atom lists: every pair term is evaluated for both of its atoms in the
// orientation of the reference list, so no atomics are needed and the sums
// are deterministic.  Energies are counted on the pair's first atom.
// Clears the fixed-point accumulators for the next evaluation.
__global__ void __launch_bounds__(OG_THREADS) k_forces_to_work(const og_dev_sys *__restrict__ S,
                                                               const int2 *__restrict__ map) {
#pragma clang fp contract(off)
    const int2        bm = map[blockIdx.x];
    const og_dev_sys &s  = S[bm.x];
    if (og_frozen(s)) return;
    const int i   = bm.y * OG_THREADS + threadIdx.x;
    double    e_ex = 0.0, e_b = 0.0;
    if (i < s.n_atoms) {
        const int g   = s.rev[i];
        const int cap = s.cap_atom_array;
        double    w[3];
        for (int k = 0; k < 3; k++) {
            const long long v = s.fx[k * cap + g];
            s.fx[k * cap + g] = 0;
            w[k]              = (double)v * OG_FX_INV;
        }
        for (int e = s.exc_off[i]; e < s.exc_off[i + 1]; e++) {
            const int t = s.exc_at[e];
            const int a = s.excess[2 * t], b = s.excess[2 * t + 1];
            float     we[3];
            const float ene = og_excess_pair(s, a, b, we);
            if (i == a) {
                for (int k = 0; k < 3; k++) w[k] += (double)we[k];
                e_ex += (double)ene;
            } else {
                for (int k = 0; k < 3; k++) w[k] -= (double)we[k];
            }
        }
        for (int e = s.bond_off[i]; e < s.bond_off[i + 1]; e++) {
            const int t = s.bond_at[e];
            const int a = s.bond_ids[2 * t], b = s.bond_ids[2 * t + 1];
            double    wb[3];
            const double ene = og_bond_pair(s, t, a, b, wb);
            if (i == a) {
                for (int k = 0; k < 3; k++) w[k] += wb[k];
                e_b += ene;
            } else {
                for (int k = 0; k < 3; k++) w[k] -= wb[k];
            }
        }
        for (int k = 0; k < 3; k++) s.work[3 * i + k] = w[k];
    }
    __shared__ double sh[2 * (OG_THREADS / 64)];
    og_block_sum2(e_ex, e_b, sh);
    if (threadIdx.x == 0 && bm.y < s.n_misc_blocks) {
        s.e_misc[2 * bm.y]     = e_ex;
        s.e_misc[2 * bm.y + 1] = e_b;
    }
}

// Deterministic per-system energy reduction: 1024 threads, each with a fixed
// strided share of the per-cell and per-block partials (16-byte loads, four in
// flight), then a fixed wave and block tree.
#define OG_RED_THREADS 1024
__global__ void __launch_bounds__(OG_RED_THREADS) k_energy_reduce(const og_dev_sys *__restrict__ S,
                                                                  const int2 *__restrict__ map) {
    const og_dev_sys &s = S[map[blockIdx.x].x];
    if (og_frozen(s) || !og_energy_step(s)) return;
    const int n = s.cnt->n_cells;
    double    v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // vdw, ele, virial, pairs, excess, bond
    const double2 *E = (const double2 *)s.e_cell;
#pragma unroll 4
    for (int k = threadIdx.x; k < n; k += OG_RED_THREADS) {
        const double2 u = E[2 * k], w = E[2 * k + 1];
        v[0] += u.x;
        v[1] += u.y;
        v[2] += w.x;
        v[3] += w.y;
    }
    const double2 *M = (const double2 *)s.e_misc;
    for (int k = threadIdx.x; k < s.n_misc_blocks; k += OG_RED_THREADS) {
        const double2 u = M[k];
        v[4] += u.x;
        v[5] += u.y;
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 6; q++) v[q] += __shfl_down(v[q], off, 64);
    __shared__ double sh[6 * (OG_RED_THREADS / 64)];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 6; q++) sh[6 * wv + q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < OG_RED_THREADS / 64; w++)
            for (int q = 0; q < 6; q++) t[q] += sh[6 * w + q];
        s.energy[OG_E_VDW]    = t[0];
        s.energy[OG_E_ELE]    = t[1];
        s.energy[OG_E_VIRIAL] = t[2];
        s.energy[OG_E_NPAIRS] = t[3];
        s.energy[OG_E_EXCESS] = t[4];
        s.energy[OG_E_BOND]   = t[5];
        s.energy[OG_E_SELF]   = s.e_self;
    }
}

// ------------------------------------------------------------- launchers --
void og_launch_pair(const og_launch &a, int alpha_mode) {
    if (alpha_mode)
        hipLaunchKernelGGL(k_pair<1>, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
    else
        hipLaunchKernelGGL(k_pair<0>, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_forces_to_work(const og_launch &a) {
    hipLaunchKernelGGL(k_forces_to_work, dim3(a.n_blocks), dim3(OG_THREADS), 0, a.stream, a.sys, a.map);
}
void og_launch_energy_reduce(const og_launch &a) {
    hipLaunchKernelGGL(k_energy_reduce, dim3(a.n_blocks), dim3(OG_RED_THREADS), 0, a.stream, a.sys, a.map);
}
