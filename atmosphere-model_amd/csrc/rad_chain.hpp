// rad_chain.hpp -- the column recurrences of the two-stream solver on
// one wavefront, shared by the reference-table radiation (k_rad.hip) and the
// general two-stream columns (k_ts.hip).
//
// src/mod_radiation.f90: the direct beam down the column (221-249) and the
// adding method of calculate_fluxes (349-381).  One wavefront holds RG
// g-points of one column; CG = 16 lanes run one g-point's recurrences over
// blocks of layers, combined across the group with DPP row scans.
#pragma once

#include "atm_common.hpp"

namespace radchain {

constexpr int RG = 4;       // g-points per pass (64 lanes = RG groups of 16)
constexpr int RMAXL = 4;    // layers per lane at most: L <= 64 * RMAXL
constexpr int RARR = 5;     // LDS arrays of L + 1 per g-point

// v / d with z the correctly rounded 1 / d: the quotient q0 = v z is
// within an ulp of v / d, and one correction with the exact residual
// v - q0 d gives the correctly rounded quotient (Markstein's theorem, round
// to nearest, no overflow or underflow), as the division v / d would.  A
// position's divisor is shared by the 16 g-points and both tables, so its
// reciprocal is formed once.
__device__ __forceinline__ double div_by(double v, double d, double z) {
    const double q0 = v * z;
    return fma(fma(-q0, d, v), z, q0);
}

// ---------------------------------------------------------------------------
// The column recurrences on 16 lanes per g-point.  Lane j of the group owns
// the layers [lo, hi) of block j (ceil(L/16) layers each).  A block is
// first reduced to one composite layer, then the composites are combined
// across the group with 4-step scans over DPP row moves (from the surface
// up and from the top down); each lane then runs the reference's recurrences over its
// own layers from the boundary values of its block.  Values agree with the
// sequential column to rounding (the grouping of the products differs).
// ---------------------------------------------------------------------------
constexpr int CG = 16;  // lanes per g-point
static_assert(RG * CG == 64, "one wavefront holds the RG g-points' lane groups");

// a stack of layers: reflectance for light from above (rt) and from below
// (rb), transmittance, emission up out of its top and down out of its bottom
struct Comp {
    double rt, rb, t, su, sd;
};

__device__ __forceinline__ Comp comp_identity() { return Comp{0., 0., 1., 0., 0.}; }

// stack u on top of stack w (the adding of two layers, 349-381).  REFL
// false: no layer reflects light from below (u.rb = 0), so the multiple-
// reflection factor d is exactly 1.
template <bool REFL = true>
__device__ __forceinline__ Comp comp_over(const Comp &u, const Comp &w) {
    const double d = REFL ? 1. / (1. - u.rb * w.rt) : 1.;
    Comp c;
    c.rt = u.rt + u.t * u.t * w.rt * d;
    c.rb = w.rb + w.t * w.t * u.rb * d;
    c.t = u.t * w.t * d;
    c.su = u.su + u.t * (w.su + w.rt * u.sd) * d;
    c.sd = w.sd + w.t * (u.sd + u.rb * w.su) * d;
    return c;
}

// Moves within a row of 16 lanes (one g-point's group) by the DPP modifier
// of a VALU move, two 32-bit halves per double: row_up<D> gives lane j the
// value of lane j - D (row_shr), row_down<D> that of lane j + D (row_shl).
// A lane whose source lies outside its row receives 0 (bound_ctrl); every
// caller uses the moved value only where the source is inside the row.
template <int CTRL>
__device__ __forceinline__ double dpp_row(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
    const int rlo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
    const int rhi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)rhi << 32) | (unsigned)rlo));
}
template <int D>
__device__ __forceinline__ double row_up(double x) {
    static_assert(D >= 1 && D < CG, "row shift");
    return dpp_row<0x110 + D>(x);
}
template <int D>
__device__ __forceinline__ double row_down(double x) {
    static_assert(D >= 1 && D < CG, "row shift");
    return dpp_row<0x100 + D>(x);
}

template <int D>
__device__ __forceinline__ Comp comp_row_up(const Comp &c) {
    return Comp{row_up<D>(c.rt), row_up<D>(c.rb), row_up<D>(c.t), row_up<D>(c.su), row_up<D>(c.sd)};
}

template <int D>
__device__ __forceinline__ Comp comp_row_down(const Comp &c) {
    return Comp{row_down<D>(c.rt), row_down<D>(c.rb), row_down<D>(c.t), row_down<D>(c.su), row_down<D>(c.sd)};
}

// one step of the inclusive scans over the group: v = blocks 0..j over the
// surface (from below), u = blocks j..15 (from above)
template <bool REFL, int D>
__device__ __forceinline__ void scan_up(Comp &v, int j) {
    const Comp w = comp_row_up<D>(v);
    if (j >= D) v = comp_over<REFL>(v, w);
}
template <bool REFL, int D>
__device__ __forceinline__ void scan_down(Comp &u, int j) {
    const Comp x = comp_row_down<D>(u);
    if (j + D < CG) u = comp_over<REFL>(x, u);
}

// calculate_fluxes (349-381) of one g-point on lane j of its group.  On
// entry R, T, Su, Sd hold the layer properties; on return R[k] and T[k]
// hold the upward and downward fluxes of level k (the direct beam B added
// to the latter when B != nullptr).  A lane reads and writes only the
// entries of its own block (and the top block entry L); the recurrence
// values stay in registers (NBMAX >= ceil(L / CG)).  REFL false: every R is
// 0 (the longwave), so each 1 / (1 - R a) is exactly 1.
template <bool REFL, int NBMAX>
__device__ __forceinline__ void adding_group(int L, int j, double *R, double *T, const double *Su, const double *Sd,
                             const double *B, double toa, double sfc_up, double sfc_alb) {
    const int nb = (L + CG - 1) / CG;
    const int lo = min(j * nb, L), hi = min(lo + nb, L), n = hi - lo;
    // the block as one layer, built from its bottom
    Comp c = comp_identity();
#pragma unroll
    for (int i = 0; i < NBMAX; i++)
        if (i < n) {
            const int k = lo + i;
            c = comp_over<REFL>(Comp{R[k], R[k], T[k], Su[k], Sd[k]}, c);
        }
    // the surface is a layer that transmits nothing
    const Comp sfc{sfc_alb, 0., 0., sfc_up, 0.};
    Comp v = j == 0 ? comp_over<REFL>(c, sfc) : c;  // blocks 0..j over the surface
    static_assert(CG == 16, "four scan steps");
    scan_up<REFL, 1>(v, j);
    scan_up<REFL, 2>(v, j);
    scan_up<REFL, 4>(v, j);
    scan_up<REFL, 8>(v, j);
    Comp u = c;  // blocks j..15
    scan_down<REFL, 1>(u, j);
    scan_down<REFL, 2>(u, j);
    scan_down<REFL, 4>(u, j);
    scan_down<REFL, 8>(u, j);
    Comp below = comp_row_up<1>(v), above = comp_row_down<1>(u);
    if (j == 0) below = sfc;
    if (j == CG - 1) above = comp_identity();
    // albedo a and emission e of everything below each level of the block,
    // up the block (the reference's loop body)
    double a[NBMAX + 1], e[NBMAX + 1], bet[NBMAX];
    a[0] = below.rt;
    e[0] = below.su;
#pragma unroll
    for (int i = 0; i < NBMAX; i++) {
        if (i < n) {
            const int k = lo + i;
            const double Rk = R[k], Tk = T[k];
            const double b = REFL ? 1. / (1. - Rk * a[i]) : 1.;
            bet[i] = b;
            a[i + 1] = Rk + (Tk * Tk) * b * a[i];
            e[i + 1] = Su[k] + Tk * b * (e[i] + Sd[k] * a[i]);
        }
    }
    double an = a[0], en = e[0];  // level hi
#pragma unroll
    for (int i = 1; i <= NBMAX; i++)
        if (i == n) {
            an = a[i];
            en = e[i];
        }
    // downward flux at the top of the block from the stack above it (the
    // top block: toa itself)
    double fd = above.sd + above.t * toa + above.rb * en;
    if (REFL) fd = fd / (1. - above.rb * an);
    if (n > 0 && hi == L) {
        R[L] = an * toa + en;
        T[L] = B ? fd + B[L] : fd;
    }
#pragma unroll
    for (int i = NBMAX - 1; i >= 0; i--) {
        if (i < n) {
            const int k = lo + i;
            const double fdn = bet[i] * (Sd[k] + T[k] * fd + R[k] * e[i]);
            R[k] = a[i] * fdn + e[i];
            T[k] = B ? fdn + B[k] : fdn;
            fd = fdn;
        }
    }
}

// the direct beam down the column (221-249) on lane j of a g-point's group:
// on entry B holds the layers' direct transmittance, Su and Sd the direct-
// to-diffuse reflectance and transmittance; on return B[k] is the beam at
// level k and Su, Sd the layers' diffuse sources.  Returns B[L].
__device__ __forceinline__ double beam_group(int L, int j, double toa_dir, double *B, double *Su, double *Sd) {
    const int nb = (L + CG - 1) / CG;
    const int lo = min(j * nb, L), hi = min(lo + nb, L);
    double p = 1.;
    for (int k = hi - 1; k >= lo; k--) p = p * B[k];
    {  // product of blocks j..15
        double w = row_down<1>(p);
        if (j + 1 < CG) p = p * w;
        w = row_down<2>(p);
        if (j + 2 < CG) p = p * w;
        w = row_down<4>(p);
        if (j + 4 < CG) p = p * w;
        w = row_down<8>(p);
        if (j + 8 < CG) p = p * w;
    }
    double above = row_down<1>(p);
    if (j == CG - 1) above = 1.;
    double bk1 = toa_dir * above;
    if (lo < hi && hi == L) B[L] = toa_dir;
    for (int k = hi - 1; k >= lo; k--) {
        const double bk = B[k] * bk1;
        Su[k] = Su[k] * bk1;
        Sd[k] = Sd[k] * bk1;
        B[k] = bk;
        bk1 = bk;
    }
    return toa_dir;
}

// Optics arithmetic type R: double on the FP64 path; float on the
// mixed-precision path (the g-point optics: extinction, two-stream
// coefficients, reflectance, transmittance and sources), whose results
// enter the LDS arrays and the adding method in double.
__device__ __forceinline__ double odiv(double a, double b) { return a / b; }
__device__ __forceinline__ float odiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

}  // namespace radchain
