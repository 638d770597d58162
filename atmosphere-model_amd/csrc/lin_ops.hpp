// lin_ops.hpp -- per-point bodies of one dynamics step for the forward,
// tangent-linear and adjoint models (host and device).
//
// The step (src/mod_sw_dyn.f90:80-416 then 424-773, with the halo
// exchanges of mod_sync) is written as a program of operations in single-
// assignment form: every intermediate array of the reference's level loops
// gets its own array, so the adjoint sweep finds each operation's inputs
// as the forward sweep left them.  Each operation is a per-point body, a
// template on the accessor that supplies its scalar type (lin_types.hpp):
// the same body, statement for statement the reference's, evaluates the
// forward value (double), the tangent (Jac<1>) or the local partials
// (Jac<NR>, one slot per read) from which the adjoint is scattered.
//
// Reads name a slot (compile-time), the operand (index into the operation's
// inputs) and the position; every body reads all its values before the
// arithmetic, with no read under a data-dependent branch, so the slots are
// constants after inlining.
#pragma once

#include "atm_common.hpp"
#include "lin_types.hpp"

constexpr int LIN_MAXIN = 8, LIN_MAXOUT = 4;

// One operation's launch: geometry, coefficients, two integers of the
// operation, and per operand the value array and the tangent or adjoint
// array (null for an operand the step does not differentiate, e.g. the net
// flux, which is a fixed forcing).
struct LinArgs {
    Geo g;
    DynCoef c;
    int e0, e1;
    const double *ak, *bk;  // hybrid coefficients (the remap's target grid)
    double top;             // Ak(1), the top pressure
    const double *iv[LIN_MAXIN];
    double *it[LIN_MAXIN];
    double *ov[LIN_MAXOUT];
    double *ot[LIN_MAXOUT];
};

#define RD(s, arr, I_, J_, K_) a.template rd<s>((arr), (I_), (J_), (K_))

// ---------------------------------------------------------------------------
// Accessors
// ---------------------------------------------------------------------------
struct AccFwd {
    using S = double;
    const LinArgs &x;
    template <int s>
    __host__ __device__ S rd(int arr, int I, int J, int K) const { return x.iv[arr][x.g.ix(I, J, K)]; }
    __host__ __device__ void wr(int o, int I, int J, int K, S v) const { x.ov[o][x.g.ix(I, J, K)] = v; }
};

struct AccTan {
    using S = Jac<1>;
    const LinArgs &x;
    template <int s>
    __host__ __device__ S rd(int arr, int I, int J, int K) const {
        const size_t i = x.g.ix(I, J, K);
        S r(x.iv[arr][i]);
        r.d[0] = x.it[arr] ? x.it[arr][i] : 0.;
        return r;
    }
    __host__ __device__ void wr(int o, int I, int J, int K, const S &v) const {
        const size_t i = x.g.ix(I, J, K);
        x.ov[o][i] = v.v;
        x.ot[o][i] = v.d[0];
    }
};

// Adjoint: slot s remembers the operand and position it read; a write
// takes the output's adjoint (and clears it: each single-assignment array
// has one writer, so its adjoint is complete when the sweep reaches the
// writer) and adds adjoint x partial to the adjoint of every value read.
// ADD is the accumulation: an atomic add on the device (several points
// read the same value), a plain add on the host.
template <int NR, class ADD>
struct AccAdj {
    using S = Jac<NR>;
    const LinArgs &x;
    int sarr[NR];
    size_t six[NR];
    double acc[NR];  // adjoint owed to each read, summed over the outputs
    __host__ __device__ explicit AccAdj(const LinArgs &a) : x(a) {
        for (int s = 0; s < NR; s++) acc[s] = 0.;
    }
    template <int s>
    __host__ __device__ S rd(int arr, int I, int J, int K) {
        static_assert(s < NR, "read slot");
        const size_t i = x.g.ix(I, J, K);
        sarr[s] = arr;
        six[s] = i;
        return S::seed(x.iv[arr][i], s);
    }
    __host__ __device__ void wr(int o, int I, int J, int K, const S &v) {
        const size_t i = x.g.ix(I, J, K);
        const double ab = x.ot[o][i];
        if (ab == 0.) return;
        x.ot[o][i] = 0.;
#pragma unroll
        for (int s = 0; s < NR; s++) acc[s] += ab * v.d[s];
    }
    // after the body: one addition per read, none where the adjoint owed
    // is zero (a partial that a frozen branch zeroes)
    __host__ __device__ void flush() {
#pragma unroll
        for (int s = 0; s < NR; s++) {
            double *t = x.it[sarr[s]];
            if (t && acc[s] != 0.) ADD::add(t + six[s], acc[s]);
        }
    }
};

// ---------------------------------------------------------------------------
// Operations of the C-grid step (mod_sw_dyn.f90:80-416)
// ---------------------------------------------------------------------------
// playkap and the heating rate (99-106; D grid 457-464).  in: pkap, net
// flux, dp; out: heating rate
struct OpHeat {
    static constexpr int NR = 5;
    static constexpr bool ADJ_DIRECT = true;  // neighbouring points share few reads: no LDS planes
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S p1 = RD(0, 0, I, J, K + 1), p0 = RD(1, 0, I, J, K);
        S n1 = RD(2, 1, I, J, K + 1), n0 = RD(3, 1, I, J, K);
        S d = RD(4, 2, I, J, K);
        S pk = (p1 - p0) / log(p1 / p0);
        a.wr(0, I, J, K, (n1 - n0) / (ATM_CP_DRY * d / x.c.gravity) / pk);
    }
};

// C-grid winds and Courant numbers (109-134).  in: ud, vd; out: uc, vc, cx, cy
struct OpCWinds {
    static constexpr int NR = 8;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S u00 = RD(0, 0, I, J, K), um0 = RD(1, 0, I - 1, J, K), u01 = RD(2, 0, I, J + 1, K),
          um1 = RD(3, 0, I - 1, J + 1, K);
        S v00 = RD(4, 1, I, J, K), v0m = RD(5, 1, I, J - 1, K), v10 = RD(6, 1, I + 1, J, K),
          v1m = RD(7, 1, I + 1, J - 1, K);
        S ub = .5 * (u00 + um0);
        S vb = .5 * (v00 + v0m);
        S u = .5 * (ub + .5 * (u01 + um1));
        S v = .5 * (vb + .5 * (v10 + v1m));
        a.wr(0, I, J, K, u);
        a.wr(1, I, J, K, v);
        a.wr(2, I, J, K, x.c.dtdx * u);
        a.wr(3, I, J, K, x.c.dtdx * v);  // Q2: y Courant number uses dt/dx
    }
};

// A-grid winds (121-125), used by the kinetic energy at every level with
// the values of the last level (k = nlay).  in: ud, vd (level K = nlay-1);
// out: ua, va (planes)
struct OpCUaVa {
    static constexpr int NR = 12;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S u00 = RD(0, 0, I, J, K), um0 = RD(1, 0, I - 1, J, K), u01 = RD(2, 0, I, J + 1, K),
          um1 = RD(3, 0, I - 1, J + 1, K), u10 = RD(4, 0, I + 1, J, K), u11 = RD(5, 0, I + 1, J + 1, K);
        S v00 = RD(6, 1, I, J, K), v0m = RD(7, 1, I, J - 1, K), v10 = RD(8, 1, I + 1, J, K),
          v1m = RD(9, 1, I + 1, J - 1, K), v01 = RD(10, 1, I, J + 1, K), v11 = RD(11, 1, I + 1, J + 1, K);
        S ub = .5 * (u00 + um0);
        S vb = .5 * (v00 + v0m);
        S u = .5 * (ub + .5 * (u01 + um1));
        S v = .5 * (vb + .5 * (v10 + v1m));
        a.wr(0, I, J, 0, .5 * (u + .25 * (u10 + u00 + u11 + u01)));
        a.wr(1, I, J, 0, .5 * (v + .25 * (v01 + v00 + v11 + v10)));
    }
};

// denominators (136-139; D grid 451-454).  in: uc, vc; out: denx, deny
struct OpDenoms {
    static constexpr int NR = 4;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S u1 = RD(0, 0, I + 1, J, K), u0 = RD(1, 0, I, J, K);
        S v1 = RD(2, 1, I, J + 1, K), v0 = RD(3, 1, I, J, K);
        a.wr(0, I, J, K, 1. / (1. - x.c.dtdx * (u1 - u0)));
        a.wr(1, I, J, K, 1. / (1. - x.c.dtdy * (v1 - v0)));
    }
};

// inner y step: PCM y fluxes, tmp and Laplacian (142-155, 209-222; D grid
// 470-482, 535-547).  in: q, cy, deny; out: tmp, lap
struct OpInnerY {
    static constexpr int NR = 8;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S q0 = RD(0, 0, I, J, K), qs = RD(1, 0, I, J - 1, K), qn = RD(2, 0, I, J + 1, K);
        S qe = RD(3, 0, I + 1, J, K), qw = RD(4, 0, I - 1, J, K);
        S c0 = RD(5, 1, I, J, K), c1 = RD(6, 1, I, J + 1, K);
        S dn = RD(7, 2, I, J, K);
        S f0 = pcm_face(qs, q0, c0);
        S f1 = pcm_face(q0, qn, c1);
        a.wr(0, I, J, K, .5 * (q0 + dn * (q0 - x.c.rA * (f1 - f0))));
        a.wr(1, I, J, K, -2. * q0 * x.c.rdx2rdy2 + (qe + qw) * x.c.rdx2 + (qn + qs) * x.c.rdy2);
    }
};

// inner x step: PCM x fluxes, tmp (176-185, 241-250; D grid 503-511,
// 566-574).  in: q, cx, denx; out: tmp
struct OpInnerX {
    static constexpr int NR = 6;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S q0 = RD(0, 0, I, J, K), qw = RD(1, 0, I - 1, J, K), qe = RD(2, 0, I + 1, J, K);
        S c0 = RD(3, 1, I, J, K), c1 = RD(4, 1, I + 1, J, K);
        S dn = RD(5, 2, I, J, K);
        S f0 = pcm_face(qw, q0, c0);
        S f1 = pcm_face(q0, qe, c1);
        a.wr(0, I, J, K, .5 * (q0 + dn * (q0 - x.c.rA * (f1 - f0))));
    }
};

// outer PPM face value along x (DIR 0) or y (DIR 1); MASS: times the
// Courant number (the mass flux dpfx / dpfy).  in: tmp, Courant; out: face
template <int V, int DIR, bool MASS>
struct OpOuter {
    static constexpr int NR = 7;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        constexpr int sI = DIR == 0 ? 1 : 0, sJ = DIR == 0 ? 0 : 1;
        S t0 = RD(0, 0, I - 3 * sI, J - 3 * sJ, K), t1 = RD(1, 0, I - 2 * sI, J - 2 * sJ, K);
        S t2 = RD(2, 0, I - sI, J - sJ, K), t3 = RD(3, 0, I, J, K);
        S t4 = RD(4, 0, I + sI, J + sJ, K), t5 = RD(5, 0, I + 2 * sI, J + 2 * sJ, K);
        S cf = RD(6, 1, I, J, K);
        S f = ppm_face<V>(t0, t1, t2, t3, t4, t5, cf);
        if (MASS)
            a.wr(0, I, J, K, f * cf);
        else
            a.wr(0, I, J, K, f);
    }
};

// dpc, x part (165-171; D grid 492-498).  in: dp, dpfx, lap; out: dpc
struct OpDpcX {
    static constexpr int NR = 6;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S d = RD(0, 0, I, J, K), f1 = RD(1, 1, I + 1, J, K), f0 = RD(2, 1, I, J, K);
        S l1 = RD(3, 2, I + 1, J, K), l0 = RD(4, 2, I, J, K), lm = RD(5, 2, I - 1, J, K);
        a.wr(0, I, J, K, d - (f1 - f0 + x.c.dtdx * x.c.vx * (l1 - 2. * l0 + lm)));
    }
};

// dpc, y part (195-201; D grid 521-527).  in: dpc (x part), dpfy, lap; out: dpc
struct OpDpcY {
    static constexpr int NR = 6;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S d = RD(0, 0, I, J, K), f1 = RD(1, 1, I, J + 1, K), f0 = RD(2, 1, I, J, K);
        S l1 = RD(3, 2, I, J + 1, K), l0 = RD(4, 2, I, J, K), lm = RD(5, 2, I, J - 1, K);
        a.wr(0, I, J, K, d - (f1 - f0 + x.c.dtdy * x.c.vy * (l1 - 2. * l0 + lm)));
    }
};

// ptc, x part (229-236; D grid 554-561).  in: dp, pt, fx, dpfx, lap; out: ptc
struct OpPtcX {
    static constexpr int NR = 9;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S dpl = RD(0, 0, I, J, K), ptl = RD(1, 1, I, J, K);
        S fx1 = RD(2, 2, I + 1, J, K), fx0 = RD(3, 2, I, J, K);
        S d1 = RD(4, 3, I + 1, J, K), d0 = RD(5, 3, I, J, K);
        S l1 = RD(6, 4, I + 1, J, K), l0 = RD(7, 4, I, J, K), lm = RD(8, 4, I - 1, J, K);
        a.wr(0, I, J, K, dpl * ptl - (fx1 * d1 - fx0 * d0 + x.c.vx * ((l1 - l0) * d1 - (l0 - lm) * d0)));
    }
};

// ptc, y part and division by dpc (257-266; D grid 581-590).
// in: dp, fy, dpfy, lap, heating rate, dpc, ptc (x part); out: ptc
struct OpPtcY {
    static constexpr int NR = 11;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S dpl = RD(0, 0, I, J, K);
        S fy1 = RD(1, 1, I, J + 1, K), fy0 = RD(2, 1, I, J, K);
        S d1 = RD(3, 2, I, J + 1, K), d0 = RD(4, 2, I, J, K);
        S l1 = RD(5, 3, I, J + 1, K), l0 = RD(6, 3, I, J, K), lm = RD(7, 3, I, J - 1, K);
        S h = RD(8, 4, I, J, K), dpc = RD(9, 5, I, J, K), ptc = RD(10, 6, I, J, K);
        a.wr(0, I, J, K,
             (ptc - (fy1 * d1 - fy0 * d0 + x.c.vy * ((l1 - l0) * d1 - (l0 - lm) * d0)) + x.c.dt * dpl * h) / dpc);
    }
};

// kinetic energy (297-317): Q4 shifted pencils, last-level ua/va.
// in: uc, vc, ua, va; out: ke
struct OpCKe {
    static constexpr int NR = 14;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S va = RD(0, 3, I, J, 0), ua = RD(1, 2, I, J, 0);
        S v0 = RD(2, 1, I, J - 2, K), v1 = RD(3, 1, I, J - 1, K), v2 = RD(4, 1, I, J, K);
        S v3 = RD(5, 1, I, J + 1, K), v4 = RD(6, 1, I, J + 2, K), v5 = RD(7, 1, I, J + 3, K);
        S u0 = RD(8, 0, I - 2, J, K), u1 = RD(9, 0, I - 1, J, K), u2 = RD(10, 0, I, J, K);
        S u3 = RD(11, 0, I + 1, J, K), u4 = RD(12, 0, I + 2, J, K), u5 = RD(13, 0, I + 3, J, K);
        S fy = ppm_face<V_PPM_U>(v0, v1, v2, v3, v4, v5, x.c.dtdy * va);
        S fx = ppm_face<V_PPM_U>(u0, u1, u2, u3, u4, u5, x.c.dtdx * ua);
        S k = va * fy;
        a.wr(0, I, J, K, .5 * (k + ua * fx));
    }
};

// C-grid vorticity (319-323).  in: uc, vc; out: vorticity
struct OpCVort {
    static constexpr int NR = 4;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S u0 = RD(0, 0, I, J, K), us = RD(1, 0, I, J - 1, K);
        S v0 = RD(2, 1, I, J, K), vw = RD(3, 1, I - 1, J, K);
        a.wr(0, I, J, K, x.c.coriolis - (u0 - us) * x.c.rdy + (v0 - vw) * x.c.rdx);
    }
};

// C-grid pressure gradient force (325-339).  in: pkap, gz; out: pgfx, pgfy
struct OpCPgf {
    static constexpr int NR = 12;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S pa = RD(0, 0, I, J, K), pd = RD(1, 0, I, J, K + 1);
        S pbx = RD(2, 0, I - 1, J, K + 1), pcx = RD(3, 0, I - 1, J, K);
        S pby = RD(4, 0, I, J - 1, K + 1), pcy = RD(5, 0, I, J - 1, K);
        S gw0 = RD(6, 1, I - 1, J, K), g01 = RD(7, 1, I, J, K + 1), gw1 = RD(8, 1, I - 1, J, K + 1);
        S g00 = RD(9, 1, I, J, K), gs0 = RD(10, 1, I, J - 1, K), gs1 = RD(11, 1, I, J - 1, K + 1);
        a.wr(0, I, J, K, ((pa - pbx) * (gw0 - g01) + (pcx - pd) * (gw1 - g00)) / (pa - pbx + pcx - pd));
        a.wr(1, I, J, K, ((pa - pby) * (gs0 - g01) + (pcy - pd) * (gs1 - g00)) / (pa - pby + pcy - pd));
    }
};

// denominators and Courant numbers from the D winds (344-350).
// in: ud, vd; out: denx, deny, cx, cy
struct OpCDenoms2 {
    static constexpr int NR = 4;
    static constexpr bool ADJ_DIRECT = true;  // neighbouring points share few reads: no LDS planes
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S u0 = RD(0, 0, I, J, K), uw = RD(1, 0, I - 1, J, K);
        S v0 = RD(2, 1, I, J, K), vs = RD(3, 1, I, J - 1, K);
        a.wr(0, I, J, K, 1. / (1. - x.c.dtdx * (u0 - uw)));
        a.wr(1, I, J, K, 1. / (1. - x.c.dtdy * (v0 - vs)));
        a.wr(2, I, J, K, x.c.dtdx * u0);
        a.wr(3, I, J, K, x.c.dtdx * v0);  // Q2
    }
};

// inner x PCM of vorticity with the Courant numbers shifted by e0 (C grid
// -1, Q4; D grid 0) and taken from level e1 (D grid: nlay-1, Q22; C grid:
// e1 < 0, the current level), tmp and Laplacian (353-367; D grid 708-722).
// in: vorticity, cx, denx; out: tmp, lap
struct OpVortInnerX {
    static constexpr int NR = 8;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        const int sh = x.e0, Kc = x.e1 >= 0 ? x.e1 : K;
        S w0 = RD(0, 0, I, J, K), ww = RD(1, 0, I - 1, J, K), we = RD(2, 0, I + 1, J, K);
        S wn = RD(3, 0, I, J + 1, K), ws = RD(4, 0, I, J - 1, K);
        S c0 = RD(5, 1, I + sh, J, Kc), c1 = RD(6, 1, I + 1 + sh, J, Kc);
        S dn = RD(7, 2, I, J, Kc);
        S f0 = pcm_face(ww, w0, c0);
        S f1 = pcm_face(w0, we, c1);
        a.wr(0, I, J, K, .5 * (w0 + dn * (w0 - x.c.rA * (f1 - f0))));
        a.wr(1, I, J, K, -2. * w0 * x.c.rdx2rdy2 + (we + ww) * x.c.rdx2 + (wn + ws) * x.c.rdy2);
    }
};

// inner y PCM of vorticity (388-397; D grid 744-753), shifts as above.
// in: vorticity, cy, deny; out: tmp
struct OpVortInnerY {
    static constexpr int NR = 6;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        const int sh = x.e0, Kc = x.e1 >= 0 ? x.e1 : K;
        S w0 = RD(0, 0, I, J, K), ws = RD(1, 0, I, J - 1, K), wn = RD(2, 0, I, J + 1, K);
        S c0 = RD(3, 1, I, J + sh, Kc), c1 = RD(4, 1, I, J + 1 + sh, Kc);
        S dn = RD(5, 2, I, J, Kc);
        S f0 = pcm_face(ws, w0, c0);
        S f1 = pcm_face(w0, wn, c1);
        a.wr(0, I, J, K, .5 * (w0 + dn * (w0 - x.c.rA * (f1 - f0))));
    }
};

// uc update (369-382): outer PPM_U along y on tmp shifted by +1 (Q4).
// in: tmp, cy, vd, lap, ke, pgfx, uc; out: uc
struct OpCUc {
    static constexpr int NR = 14;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S t0 = RD(0, 0, I, J - 2, K), t1 = RD(1, 0, I, J - 1, K), t2 = RD(2, 0, I, J, K);
        S t3 = RD(3, 0, I, J + 1, K), t4 = RD(4, 0, I, J + 2, K), t5 = RD(5, 0, I, J + 3, K);
        S cy = RD(6, 1, I, J, K), vd = RD(7, 2, I, J, K);
        S l1 = RD(8, 3, I, J + 1, K), l0 = RD(9, 3, I, J, K);
        S k0 = RD(10, 4, I, J, K), kw = RD(11, 4, I - 1, J, K);
        S pg = RD(12, 5, I, J, K), uc = RD(13, 6, I, J, K);
        S fy = ppm_face<V_PPM_U>(t0, t1, t2, t3, t4, t5, cy);
        a.wr(0, I, J, K, uc + x.c.dt * (vd * fy + x.c.vy * (l1 - l0)) - x.c.dtdx * (k0 - kw - pg));
    }
};

// vc update (399-412): outer PPM_U along x on tmp shifted by +1 (Q4).
// in: tmp, cx, ud, lap, ke, pgfy, vc; out: vc
struct OpCVc {
    static constexpr int NR = 14;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S t0 = RD(0, 0, I - 2, J, K), t1 = RD(1, 0, I - 1, J, K), t2 = RD(2, 0, I, J, K);
        S t3 = RD(3, 0, I + 1, J, K), t4 = RD(4, 0, I + 2, J, K), t5 = RD(5, 0, I + 3, J, K);
        S cx = RD(6, 1, I, J, K), ud = RD(7, 2, I, J, K);
        S l1 = RD(8, 3, I + 1, J, K), l0 = RD(9, 3, I, J, K);
        S k0 = RD(10, 4, I, J, K), ks = RD(11, 4, I, J - 1, K);
        S pg = RD(12, 5, I, J, K), vc = RD(13, 6, I, J, K);
        S fx = ppm_face<V_PPM_U>(t0, t1, t2, t3, t4, t5, cx);
        a.wr(0, I, J, K, vc - x.c.dt * (ud * fx + x.c.vx * (l1 - l0)) - x.c.dtdy * (k0 - ks - pg));
    }
};

// ---------------------------------------------------------------------------
// Operations of the D-grid step (mod_sw_dyn.f90:424-773)
// ---------------------------------------------------------------------------
// Courant numbers from the C winds (446-449).  in: uc, vc; out: cx, cy
struct OpDCourant {
    static constexpr int NR = 2;
    static constexpr bool ADJ_DIRECT = true;  // neighbouring points share few reads: no LDS planes
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S u0 = RD(0, 0, I, J, K), v0 = RD(1, 1, I, J, K);
        a.wr(0, I, J, K, x.c.dtdx * u0);
        a.wr(1, I, J, K, x.c.dtdx * v0);  // Q2
    }
};

// B-grid winds (441-444), used by the kinetic energy with the values of the
// last level (Q21).  in: uc, vc (level nlay-1); out: ub, vb (planes)
struct OpDUbVb {
    static constexpr int NR = 4;
    static constexpr bool ADJ_DIRECT = true;  // neighbouring points share few reads: no LDS planes
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S u0 = RD(0, 0, I, J, K), us = RD(1, 0, I, J - 1, K);
        S v0 = RD(2, 1, I, J, K), vw = RD(3, 1, I - 1, J, K);
        a.wr(0, I, J, 0, .5 * (u0 + us));
        a.wr(1, I, J, 0, .5 * (v0 + vw));
    }
};

// interpolate_to_corners (921-956) of level K + e0 of q into the same
// level of the output
struct OpCorners {
    static constexpr int NR = 16;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        const int Kq = K + x.e0;
        // q(I + i - 2, J + j - 2) in slot 4 i + j
        S q00 = RD(0, 0, I - 2, J - 2, Kq), q01 = RD(1, 0, I - 2, J - 1, Kq), q02 = RD(2, 0, I - 2, J, Kq),
          q03 = RD(3, 0, I - 2, J + 1, Kq);
        S q10 = RD(4, 0, I - 1, J - 2, Kq), q11 = RD(5, 0, I - 1, J - 1, Kq), q12 = RD(6, 0, I - 1, J, Kq),
          q13 = RD(7, 0, I - 1, J + 1, Kq);
        S q20 = RD(8, 0, I, J - 2, Kq), q21 = RD(9, 0, I, J - 1, Kq), q22 = RD(10, 0, I, J, Kq),
          q23 = RD(11, 0, I, J + 1, Kq);
        S q30 = RD(12, 0, I + 1, J - 2, Kq), q31 = RD(13, 0, I + 1, J - 1, Kq), q32 = RD(14, 0, I + 1, J, Kq),
          q33 = RD(15, 0, I + 1, J + 1, Kq);
        // Q(a, b) = q(I + a, J + b): a, b = -2..1 -> qij with i = a + 2, j = b + 2
        // TX(b) = 7/12 (Q(0,b) + Q(-1,b)) - 1/12 (Q(1,b) + Q(-2,b))
        S tx0 = ATM_W7_12 * (q22 + q12) - ATM_W1_12 * (q32 + q02);   // b = 0
        S txm = ATM_W7_12 * (q21 + q11) - ATM_W1_12 * (q31 + q01);   // b = -1
        S txp = ATM_W7_12 * (q23 + q13) - ATM_W1_12 * (q33 + q03);   // b = 1
        S txm2 = ATM_W7_12 * (q20 + q10) - ATM_W1_12 * (q30 + q00);  // b = -2
        // TY(a) = 7/12 (Q(a,0) + Q(a,-1)) - 1/12 (Q(a,1) + Q(a,-2))
        S ty0 = ATM_W7_12 * (q22 + q21) - ATM_W1_12 * (q23 + q20);   // a = 0
        S tym = ATM_W7_12 * (q12 + q11) - ATM_W1_12 * (q13 + q10);   // a = -1
        S typ = ATM_W7_12 * (q32 + q31) - ATM_W1_12 * (q33 + q30);   // a = 1
        S tym2 = ATM_W7_12 * (q02 + q01) - ATM_W1_12 * (q03 + q00);  // a = -2
        S o = (9. / 32.) * (tx0 + txm) - (1. / 32.) * (txp + txm2);
        o = o + (9. / 32.) * (ty0 + tym) - (1. / 32.) * (typ + tym2);
        a.wr(0, I, J, Kq, o);
    }
};

// D-grid kinetic energy (639-657), PPM_CONSTRAINED, last-level ub/vb.
// in: ud, vd, ub, vb; out: ke
struct OpDKe {
    static constexpr int NR = 14;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S vb = RD(0, 3, I, J, 0), ub = RD(1, 2, I, J, 0);
        S v0 = RD(2, 1, I, J - 3, K), v1 = RD(3, 1, I, J - 2, K), v2 = RD(4, 1, I, J - 1, K);
        S v3 = RD(5, 1, I, J, K), v4 = RD(6, 1, I, J + 1, K), v5 = RD(7, 1, I, J + 2, K);
        S u0 = RD(8, 0, I - 3, J, K), u1 = RD(9, 0, I - 2, J, K), u2 = RD(10, 0, I - 1, J, K);
        S u3 = RD(11, 0, I, J, K), u4 = RD(12, 0, I + 1, J, K), u5 = RD(13, 0, I + 2, J, K);
        S fy = ppm_face<V_PPM_C>(v0, v1, v2, v3, v4, v5, x.c.dtdy * vb);
        S fx = ppm_face<V_PPM_C>(u0, u1, u2, u3, u4, u5, x.c.dtdx * ub);
        S k = vb * fy;
        a.wr(0, I, J, K, .5 * (k + ub * fx));
    }
};

// D-grid vorticity (659-663).  in: ud, vd; out: vorticity
struct OpDVort {
    static constexpr int NR = 4;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S un = RD(0, 0, I, J + 1, K), u0 = RD(1, 0, I, J, K);
        S ve = RD(2, 1, I + 1, J, K), v0 = RD(3, 1, I, J, K);
        a.wr(0, I, J, K, x.c.coriolis - (un - u0) * x.c.rdy + (ve - v0) * x.c.rdx);
    }
};

// D-grid pressure gradient force on the B grid (665-679).
// in: pkapb, gzb; out: pgfx, pgfy
struct OpDPgf {
    static constexpr int NR = 12;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S pe0 = RD(0, 0, I + 1, J, K), p01 = RD(1, 0, I, J, K + 1), p00 = RD(2, 0, I, J, K);
        S pe1 = RD(3, 0, I + 1, J, K + 1), pn0 = RD(4, 0, I, J + 1, K), pn1 = RD(5, 0, I, J + 1, K + 1);
        S g00 = RD(6, 1, I, J, K), ge1 = RD(7, 1, I + 1, J, K + 1), g01 = RD(8, 1, I, J, K + 1);
        S ge0 = RD(9, 1, I + 1, J, K), gn1 = RD(10, 1, I, J + 1, K + 1), gn0 = RD(11, 1, I, J + 1, K);
        a.wr(0, I, J, K, ((pe0 - p01) * (g00 - ge1) + (p00 - pe1) * (g01 - ge0)) / (pe0 - p01 + p00 - pe1));
        a.wr(1, I, J, K, ((pn0 - p01) * (g00 - gn1) + (p00 - pn1) * (g01 - gn0)) / (pn0 - p01 + p00 - pn1));
    }
};

// divergence of the D winds (682-685).  in: ud, vd; out: divergence
struct OpDDiv {
    static constexpr int NR = 4;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S u0 = RD(0, 0, I, J, K), uw = RD(1, 0, I - 1, J, K);
        S v0 = RD(2, 1, I, J, K), vs = RD(3, 1, I, J - 1, K);
        a.wr(0, I, J, K, x.c.rdx * (u0 - uw) + x.c.rdy * (v0 - vs));
    }
};

// Laplacian (688-699); e0 != 0 scales it by the damping coefficient (701-702)
struct OpDel2 {
    static constexpr int NR = 5;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S q0 = RD(0, 0, I, J, K), qe = RD(1, 0, I + 1, J, K), qw = RD(2, 0, I - 1, J, K);
        S qn = RD(3, 0, I, J + 1, K), qs = RD(4, 0, I, J - 1, K);
        S v = -2. * q0 * x.c.rdx2rdy2 + (qe + qw) * x.c.rdx2 + (qn + qs) * x.c.rdy2;
        if (x.e0)
            a.wr(0, I, J, K, x.c.divdamp * v);
        else
            a.wr(0, I, J, K, v);
    }
};

// ud update (724-738); the Courant numbers of the last level (Q22).
// in: tmp, cy, vc, lap, ke, divergence, pgfx, ud; out: ud
struct OpDUd {
    static constexpr int NR = 16;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        const int Lm = x.g.L - 1;
        S t0 = RD(0, 0, I, J - 3, K), t1 = RD(1, 0, I, J - 2, K), t2 = RD(2, 0, I, J - 1, K);
        S t3 = RD(3, 0, I, J, K), t4 = RD(4, 0, I, J + 1, K), t5 = RD(5, 0, I, J + 2, K);
        S cy = RD(6, 1, I, J, Lm), vc = RD(7, 2, I, J, K);
        S l0 = RD(8, 3, I, J, K), ls = RD(9, 3, I, J - 1, K);
        S ke = RD(10, 4, I + 1, J, K), k0 = RD(11, 4, I, J, K);
        S de = RD(12, 5, I + 1, J, K), d0 = RD(13, 5, I, J, K);
        S pg = RD(14, 6, I, J, K), ud = RD(15, 7, I, J, K);
        S fy = ppm_face<V_PPM_C>(t0, t1, t2, t3, t4, t5, cy);
        a.wr(0, I, J, K, ud + x.c.dt * (vc * fy + x.c.vy * (l0 - ls)) - x.c.dtdx * (ke - k0 - de + d0 - pg));
    }
};

// vd update (755-769); the Courant numbers of the last level (Q22).
// in: tmp, cx, uc, lap, ke, divergence, pgfy, vd; out: vd
struct OpDVd {
    static constexpr int NR = 16;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        const int Lm = x.g.L - 1;
        S t0 = RD(0, 0, I - 3, J, K), t1 = RD(1, 0, I - 2, J, K), t2 = RD(2, 0, I - 1, J, K);
        S t3 = RD(3, 0, I, J, K), t4 = RD(4, 0, I + 1, J, K), t5 = RD(5, 0, I + 2, J, K);
        S cx = RD(6, 1, I, J, Lm), uc = RD(7, 2, I, J, K);
        S l0 = RD(8, 3, I, J, K), lw = RD(9, 3, I - 1, J, K);
        S kn = RD(10, 4, I, J + 1, K), k0 = RD(11, 4, I, J, K);
        S dn = RD(12, 5, I, J + 1, K), d0 = RD(13, 5, I, J, K);
        S pg = RD(14, 6, I, J, K), vd = RD(15, 7, I, J, K);
        S fx = ppm_face<V_PPM_C>(t0, t1, t2, t3, t4, t5, cx);
        a.wr(0, I, J, K, vd - x.c.dt * (uc * fx + x.c.vx * (l0 - lw)) - x.c.dtdy * (kn - k0 - dn + d0 - pg));
    }
};

#undef RD

// ---------------------------------------------------------------------------
// Column operations: pressure, p**kappa and geopotential
// ---------------------------------------------------------------------------
// plev(k) = plev(k+1) + src(k), pkap = plev**kappa from the top down (cgrid
// 89-92, dgrid 433-437), then, with GZ, gz(k+1) = gz(k) + cp th(k)
// (pkap(k) - pkap(k+1)) from the surface up (cgrid 276-292, dgrid 606-626).
// Level nlay of plev/pkap (the top) and level 0 of gz (the surface) are
// constants.  S = double: forward; S = Jac<1>: tangent (T* the tangents).
template <class S, bool GZ>
__host__ __device__ inline void scan_column(const Geo &g, size_t c0, const double *src, const double *tsrc,
                                            const double *th, const double *tth, double *plev, double *tplev,
                                            double *pkap, double *tpkap, double *gz, double *tgz) {
    const size_t pl = g.plane();
    const int L = g.L;
    auto ld = [&](const double *v, const double *t, size_t i) {
        S r(v[i]);
        if constexpr (!std::is_same<S, double>::value) r.d[0] = t ? t[i] : 0.;
        return r;
    };
    auto st = [&](double *v, double *t, size_t i, const S &x) {
        if constexpr (std::is_same<S, double>::value) {
            v[i] = x;
        } else {
            v[i] = x.v;
            t[i] = x.d[0];
        }
    };
    S p = S(plev[c0 + (size_t)L * pl]);
    for (int K = L - 1; K >= 0; K--) {
        const size_t i = c0 + (size_t)K * pl;
        p = p + ld(src, tsrc, i);
        st(plev, tplev, i, p);
        st(pkap, tpkap, i, pow_kappa(p));
    }
    if (GZ) {
        S z = S(gz[c0]);
        S pk = ld(pkap, tpkap, c0);
        for (int K = 0; K < L; K++) {
            const size_t i = c0 + (size_t)K * pl;
            S pk1 = K + 1 < L ? ld(pkap, tpkap, i + pl) : S(pkap[i + pl]);
            z = z + ATM_CP_DRY * ld(th, tth, i) * (pk - pk1);
            st(gz, tgz, i + pl, z);
            pk = pk1;
        }
    }
}

// Adjoint of scan_column: the adjoints of gz, pkap and plev (consumed and
// cleared) into those of src and th.  Each column is one thread's, so the
// additions need no atomics.
template <bool GZ>
__host__ __device__ inline void scan_column_adj(const Geo &g, size_t c0, double *asrc, const double *th, double *ath,
                                                const double *plev, double *aplev, const double *pkap, double *apkap,
                                                double *agz) {
    const size_t pl = g.plane();
    const int L = g.L;
    if (GZ) {
        double carry = 0.;  // adjoint of gz(K+1) from the levels above
        for (int K = L - 1; K >= 0; K--) {
            const size_t i = c0 + (size_t)K * pl;
            const double a = agz[i + pl] + carry;
            agz[i + pl] = 0.;
            const double pk = pkap[i], pk1 = pkap[i + pl];
            ath[i] += a * (ATM_CP_DRY * (pk - pk1));
            const double w = a * (ATM_CP_DRY * th[i]);
            apkap[i] += w;
            if (K + 1 < L) apkap[i + pl] -= w;
            carry = a;
        }
    }
    double carry = 0.;  // adjoint of plev(K) through plev(K-1) = plev(K) + src(K-1)
    for (int K = 0; K < L; K++) {
        const size_t i = c0 + (size_t)K * pl;
        const double a = carry + aplev[i] + apkap[i] * (ATM_KAPPA * pkap[i] / plev[i]);
        aplev[i] = 0.;
        apkap[i] = 0.;
        asrc[i] += a;
        carry = a;
    }
}

// Periodic halo fill of one point (mod_sync.f90:109-217 on one image): the
// source of a halo point is the interior point at the wrapped indices.
__host__ __device__ inline int wrap_index(int I, int n) {
    int r = (I - ATM_HALO) % n;
    if (r < 0) r += n;
    return ATM_HALO + r;
}
