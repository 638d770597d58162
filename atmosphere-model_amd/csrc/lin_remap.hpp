// lin_remap.hpp -- the vertical remap (src/mod_remap.f90:27-356) as
// operations of the linearised step: per-point bodies (lin_ops.hpp style)
// for the grids, face coordinates and energies, and column operations for
// remap_column and the potential-temperature reconstruction.
//
// A column operation runs in one thread per column.  Forward and tangent
// evaluate the column on their scalar type, keeping every intermediate (the
// eliminated edge values, gam, the back-substituted and limited level
// values, the edges and curvature) in the operation's own workspace, laid
// out [slot][level][column]; the adjoint sweep, which runs after the
// forward sweep of the same step, walks the column backwards over them.  Each backward step applies
// the partials of one small local function (an elimination step, a clamp,
// a Huynh limit, one source/destination overlap), obtained by evaluating
// the same template the forward pass used on Jac<K>, so the reverse pass
// takes exactly the forward pass's branches.
#pragma once

#include "lin_ops.hpp"

// ---------------------------------------------------------------------------
// Local functions of vertical_ppm (287-356) and remap_column (165-285),
// 1-based as in the reference.  Shared by forward, tangent and adjoint.
// ---------------------------------------------------------------------------
// top layer (292-296): X(L+1), X(L), X(L-1), V(L), V(L-1) -> level_var(L+1), gam(L)
template <class S>
__host__ __device__ inline void rm_elim_top(const S &xp, const S &x0, const S &xm, const S &v0, const S &vm, S &lv,
                                            S &gm) {
    S alpha = (x0 - xm) / (xp - x0);
    S beta = alpha * (.5 + alpha);
    lv = (2. * alpha * (1. + alpha) * v0 + vm) / beta;
    gm = (1. + alpha * (1.5 + alpha)) / beta;
}
// interior layer k (299-304): X(k+1), X(k), X(k-1), V(k), V(k-1),
// level_var(k+1), gam(k) -> level_var(k), gam(k-1)
template <class S>
__host__ __device__ inline void rm_elim_step(const S &xp, const S &x0, const S &xm, const S &v0, const S &vm,
                                             const S &lvn, const S &g, S &lv, S &gm) {
    S alpha = (xp - x0) / (x0 - xm);
    S beta = 2. + 2. * alpha - g;
    lv = (3. * (v0 + alpha * vm) - lvn) / beta;
    gm = alpha / beta;
}
// bottom layer (307-309): X(3), X(2), X(1), V(1), V(2), level_var(2), gam(1)
// -> level_var(1); alpha is that of k = 2
template <class S>
__host__ __device__ inline S rm_elim_bottom(const S &x3, const S &x2, const S &x1, const S &v1, const S &v2,
                                            const S &lv2, const S &g1) {
    S alpha = (x3 - x2) / (x2 - x1);
    S beta = 1. + alpha * (1.5 + alpha);
    return (2. * alpha * (1. + alpha) * v1 + v2 - beta * lv2) / (alpha * (.5 + alpha) - beta * g1);
}
// limits of level_var(j) (316-334): the boundary clamps at j = L and j = 2,
// the monotonicity constraint at j = 3..L-1 (its k = L-1 iteration, which
// reads gam(L+1), is a no-op after the boundary clamp: Q7); ge = gam(j+1),
// gc = gam(j-1) (gam(L+1) taken as 0) decide the branch
template <class S>
__host__ __device__ inline S rm_clamp(int j, int L, const S &bs, const S &vlo, const S &vhi, double ge, double gc) {
    S l = bs;
    if (j == L || j == 2) {
        l = dmin2(l, dmax2(vhi, vlo));
        l = dmax2(l, dmin2(vhi, vlo));
    } else if (j >= 3 && j <= L - 1) {
        const S lo = dmin2(vlo, vhi), hi = dmax2(vlo, vhi);
        if (LIN_DECIDE(ge * gc > 0)) {
            l = dmin2(l, hi);
            l = dmax2(l, lo);
        } else if (LIN_DECIDE(ge > 0)) {
            l = dmax2(l, lo);
        } else {
            l = dmin2(l, hi);
        }
    }
    return l;
}
// Huynh constraint (340-352, Q8: gam used as slopes) on the top edge of
// layer k >= 3: V(k), gam(k-1), gam(k-2)
template <class S>
__host__ __device__ inline S rm_huynh_top(const S &t, const S &v, const S &gb, const S &ga) {
    S pmp1 = v - 2. * gb;
    S lac1 = pmp1 + 1.5 * ga;
    return dmin2(dmax2(t, dmin3(v, pmp1, lac1)), dmax3(v, pmp1, lac1));
}
// ... and on its bottom edge: V(k), gam(k), gam(k-1)
template <class S>
__host__ __device__ inline S rm_huynh_btm(const S &b, const S &v, const S &gc, const S &gb) {
    S pmp1 = v + 2. * gc;
    S lac1 = pmp1 - 1.5 * gb;
    return dmin2(dmax2(b, dmin3(v, pmp1, lac1)), dmax3(v, pmp1, lac1));
}
// contribution of source layer k to destination layer m (193-281), zero
// when they do not overlap
template <class S>
__host__ __device__ inline S rm_overlap(const S &b, const S &t, const S &cv, const S &v, const S &xk, const S &xk1,
                                        const S &dm0, const S &dm1) {
    S acc(0.);
    if (xk1 >= dm0 || xk <= dm1) return acc;
    const S dsrc = xk1 - xk;
    const S s1 = (dm0 - xk) / dsrc;
    const S s2 = (dm1 - xk) / dsrc;
    const S factor = dsrc / (dm1 - dm0);
    const S tmp3 = 2. * (s1 * s1);
    if (dm1 <= xk1 && xk1 < dm0 && dm0 < xk) {  // case 3
        S tmp1 = s1 - 1., tmp2 = s1 + 1.;
        acc = acc + tmp1 / 6. * (3. * (b * tmp1 - t * tmp2) + cv * (tmp3 - tmp2)) * factor;
    }
    if (xk1 < dm1 && dm1 < xk && xk <= dm0) {  // case 4
        acc = acc + (s2 / 2.) * (b * (2. - s2) + s2 * (t + cv * (1. - ATM_W2_3 * s2))) * factor;
    }
    if (dm1 <= xk1 && xk <= dm0) {  // case 5
        acc = acc + v * factor;
    }
    if (xk1 < dm1 && dm0 < xk) {  // case 6
        S tmp1 = s1 + s2;
        acc = acc + (s1 - s2) / 6. * (3. * (b * (tmp1 - 2.) - t * tmp1) + cv * (tmp3 + tmp1 * (2. * s2 - 3.))) *
                        factor;
    }
    return acc;
}

// The source layers that overlap destination layer m (193-281 let a pair
// through when X(k+1) < D(m) and X(k) > D(m+1)).  Faces decrease upward in
// both grids, so for each m these are a contiguous run of k, and the run's
// start only moves up as m grows: the first layer of the run is the first k
// from the previous run's start with X(k+1) < D(m), and the run ends before
// the first k with X(k) <= D(m+1).  Walking the runs visits exactly the
// pairs the reference's loop over all (k, m) accumulates, in the same
// ascending order, with O(L) face reads per column instead of O(L^2).
template <class XF>
__host__ __device__ inline int rm_first_overlap(const XF &X, int k, int L, double dm0) {
    while (k <= L && LIN_DECIDE(val(X(k + 1)) >= dm0)) k++;
    return k;
}

// Column workspace: slot s, 1-based level k (0..L+1), column t.
struct RmWS {
    double *w;
    size_t ncol, t;
    int L;
    __host__ __device__ double &at(int s, int k) const { return w[((size_t)s * (L + 2) + k) * ncol + t]; }
};
// slots: values (and, for the tangent, their tangents at slot + RM_NV);
// the adjoint keeps its accumulators at slot + RM_NV
enum { RM_LV = 0, RM_GM, RM_BS, RM_LF, RM_T, RM_B, RM_CV, RM_NV };
constexpr int RM_SLOTS = 2 * RM_NV;
// Workspaces of the step's column operations: the u, v and energy columns
// of remap_column and the pt reconstruction each keep their own, so the
// forward sweep's intermediates are still there when the adjoint sweep
// reaches the operation, and the three remap columns can run together.
enum { RM_WS_U = 0, RM_WS_V, RM_WS_E, RM_WS_PT, RM_NWS };

template <class S>
__host__ __device__ inline S rm_get(const RmWS &w, int s, int k);
template <>
__host__ __device__ inline double rm_get<double>(const RmWS &w, int s, int k) { return w.at(s, k); }
template <>
__host__ __device__ inline Jac<1> rm_get<Jac<1>>(const RmWS &w, int s, int k) {
    Jac<1> r(w.at(s, k));
    r.d[0] = w.at(s + RM_NV, k);
    return r;
}
__host__ __device__ inline void rm_put(const RmWS &w, int s, int k, double v) { w.at(s, k) = v; }
__host__ __device__ inline void rm_put(const RmWS &w, int s, int k, const Jac<1> &v) {
    w.at(s, k) = v.v;
    w.at(s + RM_NV, k) = v.d[0];
}

// A column of LinArgs operands at (I, J): operand `a`, 1-based level k.
// S = double reads values, S = Jac<1> values and tangents.
template <class S>
__host__ __device__ inline S col_rd(const LinArgs &x, int a, size_t c0, int k) {
    const size_t i = c0 + (size_t)(k - 1) * x.g.plane();
    S r(x.iv[a][i]);
    if constexpr (!std::is_same<S, double>::value) r.d[0] = x.it[a] ? x.it[a][i] : 0.;
    return r;
}
template <class S>
__host__ __device__ inline void col_wr(const LinArgs &x, int a, size_t c0, int k, const S &v) {
    const size_t i = c0 + (size_t)(k - 1) * x.g.plane();
    if constexpr (std::is_same<S, double>::value) {
        x.ov[a][i] = v;
    } else {
        x.ov[a][i] = v.v;
        x.ot[a][i] = v.d[0];
    }
}

// remap_column (165-285) of one column forward (double) or forward +
// tangent (Jac<1>).  Operands: in 0 the layer values V (L), in 1 the
// source faces X (L+1), in 2 the destination faces D (L+1); out 0 the
// remapped values.  Every intermediate goes to the workspace.
template <class S>
__host__ __device__ inline void rm_column(const LinArgs &x, size_t c0, const RmWS &w) {
    const int L = x.g.L;
    auto V = [&](int k) { return col_rd<S>(x, 0, c0, k); };
    auto X = [&](int k) { return col_rd<S>(x, 1, c0, k); };
    auto D = [&](int k) { return col_rd<S>(x, 2, c0, k); };
    // elimination sweep; the recurrence (level_var(k+1), gam(k)) and the
    // faces and values of the current layers are carried in registers
    {
        S xp = X(L + 1), x0 = X(L), xm = X(L - 1), v0 = V(L), vm = V(L - 1);
        S lv, gm;
        rm_elim_top<S>(xp, x0, xm, v0, vm, lv, gm);
        rm_put(w, RM_LV, L + 1, lv);
        rm_put(w, RM_GM, L, gm);
        for (int k = L; k >= 2; k--) {
            // xp = X(k+1), x0 = X(k), xm = X(k-1), v0 = V(k), vm = V(k-1)
            S l2, g2;
            rm_elim_step<S>(xp, x0, xm, v0, vm, lv, gm, l2, g2);
            rm_put(w, RM_LV, k, l2);
            rm_put(w, RM_GM, k - 1, g2);
            lv = l2;
            gm = g2;
            if (k > 2) {
                xp = x0;
                x0 = xm;
                xm = X(k - 2);
                v0 = vm;
                vm = V(k - 2);
            }
        }
        // xp = X(3), x0 = X(2), xm = X(1), v0 = V(2), vm = V(1)
        rm_put(w, RM_LV, 1, rm_elim_bottom<S>(xp, x0, xm, vm, v0, lv, gm));
    }
    // back substitution (313-315) and limits of the level values
    auto G = [&](int i) { return (i >= 1 && i <= L) ? val(rm_get<S>(w, RM_GM, i)) : 0.; };
    {
        S bs = rm_get<S>(w, RM_LV, 1);
        rm_put(w, RM_BS, 1, bs);
        rm_put(w, RM_LF, 1, bs);
        for (int j = 2; j <= L + 1; j++) {
            bs = rm_get<S>(w, RM_LV, j) - rm_get<S>(w, RM_GM, j - 1) * bs;
            rm_put(w, RM_BS, j, bs);
            const S vlo = V(j - 1), vhi = j <= L ? V(j) : S(0.);
            rm_put(w, RM_LF, j, j <= L ? rm_clamp<S>(j, L, bs, vlo, vhi, G(j + 1), G(j - 1)) : bs);
        }
    }
    // edges and curvature of each layer (336-354)
    for (int k = 1; k <= L; k++) {
        const S v = V(k);
        S t = rm_get<S>(w, RM_LF, k + 1), b = rm_get<S>(w, RM_LF, k);
        if (k >= 3) {
            t = rm_huynh_top<S>(t, v, rm_get<S>(w, RM_GM, k - 1), rm_get<S>(w, RM_GM, k - 2));
            b = rm_huynh_btm<S>(b, v, k <= L ? rm_get<S>(w, RM_GM, k) : S(0.), rm_get<S>(w, RM_GM, k - 1));
        }
        rm_put(w, RM_T, k, t);
        rm_put(w, RM_B, k, b);
        rm_put(w, RM_CV, k, 6. * v - 3. * (t + b));
    }
    // overlaps, for each destination layer in ascending source layer
    int k0 = 1;
    for (int m = 1; m <= L; m++) {
        const S dm0 = D(m), dm1 = D(m + 1);
        S acc(0.);
        k0 = rm_first_overlap(X, k0, L, val(dm0));
        for (int k = k0; k <= L; k++) {
            const S xk = X(k), xk1 = X(k + 1);
            if (LIN_DECIDE(val(xk) <= val(dm1))) break;
            acc = acc + rm_overlap<S>(rm_get<S>(w, RM_B, k), rm_get<S>(w, RM_T, k), rm_get<S>(w, RM_CV, k), V(k), xk,
                                      xk1, dm0, dm1);
        }
        col_wr<S>(x, 0, c0, m, acc);
    }
}

// Adjoint of rm_column: the reverse pass over the intermediates that the
// forward sweep of the same step left in this operation's workspace.  The
// output's adjoint (consumed) is added into the adjoints of V, X and D;
// each column is one thread's.  Every sweep keeps the adjoints that
// consecutive steps of its loop add to (a layer's edges while the overlap
// walk stays on it, the back substitution's chain, a window of faces) in
// registers and writes each to memory once, when no later step of the
// sweep adds to it.
__host__ __device__ inline void rm_column_adj(const LinArgs &x, size_t c0, const RmWS &w) {
    const int L = x.g.L;
    auto V = [&](int k) { return col_rd<double>(x, 0, c0, k); };
    auto X = [&](int k) { return col_rd<double>(x, 1, c0, k); };
    auto D = [&](int k) { return col_rd<double>(x, 2, c0, k); };
    const size_t pl = x.g.plane();
    auto aV = [&](int k) -> double & { return x.it[0][c0 + (size_t)(k - 1) * pl]; };
    auto aX = [&](int k) -> double & { return x.it[1][c0 + (size_t)(k - 1) * pl]; };
    auto aD = [&](int k) -> double & { return x.it[2][c0 + (size_t)(k - 1) * pl]; };
    auto W = [&](int s, int k) { return w.at(s, k); };
    auto A = [&](int s, int k) -> double & { return w.at(s + RM_NV, k); };

    // overlaps (the pairs the forward pass visited).  The walk never returns
    // to a layer it left, so the sums of the current layer kc are written
    // when it moves on; a layer it never reaches keeps zero adjoints.
    for (int k = 1; k <= L; k++) A(RM_B, k) = A(RM_T, k) = A(RM_CV, k) = 0.;
    {
        int kc = 0;
        double pB = 0., pT = 0., pCV = 0., pV = 0., pX0 = 0., pX1 = 0.;  // of layer kc: B, T, CV, V, X(kc), X(kc+1)
        double pD0 = 0., pD1 = 0.;                                      // D(m), D(m+1)
        auto leave = [&](int knew) {
            if (kc > 0) {
                A(RM_B, kc) = pB;
                A(RM_T, kc) = pT;
                A(RM_CV, kc) = pCV;
                aV(kc) += pV;
                aX(kc) += pX0;
                if (knew == kc + 1) {
                    pX0 = pX1;
                } else {
                    aX(kc + 1) += pX1;
                    pX0 = 0.;
                }
            }
            pX1 = pB = pT = pCV = pV = 0.;
            kc = knew;
        };
        int k0 = 1;
        for (int m = 1; m <= L; m++) {
            if (m > 1) {
                aD(m - 1) += pD0;
                pD0 = pD1;
                pD1 = 0.;
            }
            double &ao = x.ot[0][c0 + (size_t)(m - 1) * pl];
            const double am = ao;
            ao = 0.;
            if (am == 0.) continue;
            const double dm0 = D(m), dm1 = D(m + 1);
            k0 = rm_first_overlap(X, k0, L, dm0);
            for (int k = k0; k <= L; k++) {
                const double xk = X(k), xk1 = X(k + 1);
                if (LIN_DECIDE(xk <= dm1)) break;
                using J8 = Jac<8>;
                const J8 r = rm_overlap<J8>(J8::seed(W(RM_B, k), 0), J8::seed(W(RM_T, k), 1),
                                            J8::seed(W(RM_CV, k), 2), J8::seed(V(k), 3), J8::seed(xk, 4),
                                            J8::seed(xk1, 5), J8::seed(dm0, 6), J8::seed(dm1, 7));
                if (k != kc) leave(k);
                pB += am * r.d[0];
                pT += am * r.d[1];
                pCV += am * r.d[2];
                pV += am * r.d[3];
                pX0 += am * r.d[4];
                pX1 += am * r.d[5];
                pD0 += am * r.d[6];
                pD1 += am * r.d[7];
            }
        }
        aD(L) += pD0;
        aD(L + 1) += pD1;
        if (kc > 0) leave(L + 2);
    }
    // edges and curvature, from the top layer down: level_var(k+1) and
    // gam(k) receive their last adjoint at layer k
    {
        double pLF = 0.;                // level_var(k+1), from layer k+1's bottom edge
        double g0 = 0., g1 = 0., g2 = 0.;  // gam(k), gam(k-1), gam(k-2)
        for (int k = L; k >= 1; k--) {
            const double acv = A(RM_CV, k);
            double av = 6. * acv;
            const double at = A(RM_T, k) - 3. * acv, ab = A(RM_B, k) - 3. * acv;
            double lt = at, lb = ab;
            if (k >= 3) {
                using J4 = Jac<4>;
                const J4 t = rm_huynh_top<J4>(J4::seed(W(RM_LF, k + 1), 0), J4::seed(V(k), 1),
                                              J4::seed(W(RM_GM, k - 1), 2), J4::seed(W(RM_GM, k - 2), 3));
                lt = at * t.d[0];
                av += at * t.d[1];
                g1 += at * t.d[2];
                g2 += at * t.d[3];
                const J4 b = rm_huynh_btm<J4>(J4::seed(W(RM_LF, k), 0), J4::seed(V(k), 1), J4::seed(W(RM_GM, k), 2),
                                              J4::seed(W(RM_GM, k - 1), 3));
                lb = ab * b.d[0];
                av += ab * b.d[1];
                g0 += ab * b.d[2];
                g1 += ab * b.d[3];
            }
            aV(k) += av;
            A(RM_LF, k + 1) = pLF + lt;
            pLF = lb;
            A(RM_GM, k) = g0;
            g0 = g1;
            g1 = g2;
            g2 = 0.;
        }
        A(RM_LF, 1) = pLF;
    }
    // limits and back substitution, from the top level down:
    // bs(j) = lv(j) - gam(j-1) bs(j-1)
    auto G = [&](int i) { return (i >= 1 && i <= L) ? W(RM_GM, i) : 0.; };
    {
        double pBS = 0., pAV = 0.;  // adjoints of bs(j) and of V(j)
        for (int j = L + 1; j >= 2; j--) {
            const double alf = A(RM_LF, j);
            double abs_ = alf, avj = pAV, avm = 0.;
            if (j <= L) {
                using J3 = Jac<3>;
                const J3 l = rm_clamp<J3>(j, L, J3::seed(W(RM_BS, j), 0), J3::seed(V(j - 1), 1), J3::seed(V(j), 2),
                                          G(j + 1), G(j - 1));
                abs_ = alf * l.d[0];
                avm = alf * l.d[1];
                avj += alf * l.d[2];
                aV(j) += avj;
            }
            pAV = avm;
            abs_ += pBS;
            A(RM_LV, j) = abs_;
            A(RM_GM, j - 1) -= abs_ * W(RM_BS, j - 1);
            pBS = -abs_ * W(RM_GM, j - 1);
        }
        aV(1) += pAV;
        A(RM_LV, 1) = A(RM_LF, 1) + pBS;
    }
    // elimination sweep backwards: bottom, then k = 2..L, then the top.
    // At step k the adjoints of X(k-1) and V(k-1) are complete.
    {
        using J7 = Jac<7>;
        const J7 l1 = rm_elim_bottom<J7>(J7::seed(X(3), 0), J7::seed(X(2), 1), J7::seed(X(1), 2), J7::seed(V(1), 3),
                                         J7::seed(V(2), 4), J7::seed(W(RM_LV, 2), 5), J7::seed(W(RM_GM, 1), 6));
        const double a = A(RM_LV, 1);
        double xm = a * l1.d[2], x0 = a * l1.d[1], xp = a * l1.d[0];  // X(k-1), X(k), X(k+1) at k = 2
        double vm = a * l1.d[3], v0 = a * l1.d[4];                    // V(k-1), V(k)
        double cLV = a * l1.d[5], cGM = a * l1.d[6];                  // level_var(k), gam(k-1)
        for (int k = 2; k <= L; k++) {
            J7 lv, gm;
            rm_elim_step<J7>(J7::seed(X(k + 1), 0), J7::seed(X(k), 1), J7::seed(X(k - 1), 2), J7::seed(V(k), 3),
                             J7::seed(V(k - 1), 4), J7::seed(W(RM_LV, k + 1), 5), J7::seed(W(RM_GM, k), 6), lv, gm);
            const double al = A(RM_LV, k) + cLV, ag = A(RM_GM, k - 1) + cGM;
            xp += al * lv.d[0] + ag * gm.d[0];
            x0 += al * lv.d[1] + ag * gm.d[1];
            xm += al * lv.d[2] + ag * gm.d[2];
            v0 += al * lv.d[3] + ag * gm.d[3];
            vm += al * lv.d[4] + ag * gm.d[4];
            cLV = al * lv.d[5] + ag * gm.d[5];
            cGM = al * lv.d[6] + ag * gm.d[6];
            aX(k - 1) += xm;
            aV(k - 1) += vm;
            xm = x0;
            x0 = xp;
            xp = 0.;
            vm = v0;
            v0 = 0.;
        }
        // xm = X(L), x0 = X(L+1), vm = V(L) pending
        using J5 = Jac<5>;
        J5 lv, gm;
        rm_elim_top<J5>(J5::seed(X(L + 1), 0), J5::seed(X(L), 1), J5::seed(X(L - 1), 2), J5::seed(V(L), 3),
                        J5::seed(V(L - 1), 4), lv, gm);
        const double al = A(RM_LV, L + 1) + cLV, ag = A(RM_GM, L) + cGM;
        aX(L + 1) += x0 + (al * lv.d[0] + ag * gm.d[0]);
        aX(L) += xm + (al * lv.d[1] + ag * gm.d[1]);
        aX(L - 1) += al * lv.d[2] + ag * gm.d[2];
        aV(L) += vm + (al * lv.d[3] + ag * gm.d[3]);
        aV(L - 1) += al * lv.d[4] + ag * gm.d[4];
    }
}

// Potential temperature from the remapped energy (129-155): in 0 the
// remapped energy, 1 the kinetic energy of the remapped winds, 2 target
// playkap, 3 target pkap (L+1), 4 gz (surface at level 1); out 0 pt.
template <class S>
__host__ __device__ inline S rm_pt_level(const S &e, const S &gzr, const S &ke, const S &tplk, const S &tk,
                                         const S &tk1, S &gzn) {
    const S p = (e - gzr - ke) / (ATM_CV_DRY * tplk + .5 * ATM_CP_DRY * (tk - tk1));
    gzn = gzr + ATM_CP_DRY * p * (tk - tk1);
    return p;
}

template <class S>
__host__ __device__ inline void rm_pt_column(const LinArgs &x, size_t c0, const RmWS &w) {
    const int L = x.g.L;
    S gzr = S(x.iv[4][c0]);  // surface geopotential: a constant
    for (int m = 1; m <= L; m++) {
        rm_put(w, RM_BS, m, gzr);
        S gzn;
        const S p = rm_pt_level<S>(col_rd<S>(x, 0, c0, m), gzr, col_rd<S>(x, 1, c0, m), col_rd<S>(x, 2, c0, m),
                                   col_rd<S>(x, 3, c0, m), m < L ? col_rd<S>(x, 3, c0, m + 1) : S(x.iv[3][c0 + (size_t)L * x.g.plane()]), gzn);
        col_wr<S>(x, 0, c0, m, p);
        gzr = gzn;
    }
}

__host__ __device__ inline void rm_pt_column_adj(const LinArgs &x, size_t c0, const RmWS &w) {
    const int L = x.g.L;
    // the geopotential chain: in the workspace from the forward sweep
    const size_t pl = x.g.plane();
    double agz = 0.;  // adjoint of gz_remap(m + 1)
    for (int m = L; m >= 1; m--) {
        const size_t i = c0 + (size_t)(m - 1) * pl;
        double &ao = x.ot[0][i];
        const double ap = ao;
        ao = 0.;
        using J6 = Jac<6>;
        J6 gzn;
        const double tk1 = x.iv[3][i + pl];
        const J6 p = rm_pt_level<J6>(J6::seed(x.iv[0][i], 0), J6::seed(w.at(RM_BS, m), 1), J6::seed(x.iv[1][i], 2),
                                     J6::seed(x.iv[2][i], 3), J6::seed(x.iv[3][i], 4), J6::seed(tk1, 5), gzn);
        double a[6];
        for (int s = 0; s < 6; s++) a[s] = ap * p.d[s] + agz * gzn.d[s];
        x.it[0][i] += a[0];
        x.it[1][i] += a[2];
        x.it[2][i] += a[3];
        x.it[3][i] += a[4];
        if (m < L) x.it[3][i + pl] += a[5];
        agz = a[1];  // gz_remap(1) is the surface: a constant
    }
}

// ---------------------------------------------------------------------------
// Per-point operations of the remap
// ---------------------------------------------------------------------------
#define RD(s, arr, I_, J_, K_) a.template rd<s>((arr), (I_), (J_), (K_))

// playkap = (pkap(k+1) - pkap(k)) / log(pkap(k+1) / pkap(k)) (48-53, 64-70)
struct OpRemapPlk {
    static constexpr int NR = 2;
    static constexpr bool ADJ_DIRECT = true;  // neighbouring points share few reads: no LDS planes
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S pk1 = RD(0, 0, I, J, K + 1), pk = RD(1, 0, I, J, K);
        a.wr(0, I, J, K, (pk1 - pk) / log(pk1 / pk));
    }
};

// target_plev = Ak + Bk (plev(1) - top) (58-59); ak, bk, ak0 of LinArgs
struct OpTargetPlev {
    static constexpr int NR = 1;
    static constexpr bool ADJ_DIRECT = true;  // neighbouring points share few reads: no LDS planes
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S ps = RD(0, 0, I, J, 0);
        a.wr(0, I, J, K, x.ak[K] + x.bk[K] * (ps - x.top));
    }
};

// x**kappa (61-62)
struct OpPowKappa {
    static constexpr int NR = 1;
    static constexpr bool ADJ_DIRECT = true;  // neighbouring points share few reads: no LDS planes
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S p = RD(0, 0, I, J, K);
        a.wr(0, I, J, K, pow_kappa(p));
    }
};

// face coordinates of the wind columns (84-90, 94-100): 4-point averages
// along J (DIR 1, u) or I (DIR 0, v)
template <int DIR>
struct OpFaceAvg {
    static constexpr int NR = 4;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        constexpr int sI = DIR == 0 ? 1 : 0, sJ = DIR == 0 ? 0 : 1;
        S q0 = RD(0, 0, I, J, K), qm = RD(1, 0, I - sI, J - sJ, K);
        S qp = RD(2, 0, I + sI, J + sJ, K), qm2 = RD(3, 0, I - 2 * sI, J - 2 * sJ, K);
        a.wr(0, I, J, K, ATM_W7_12 * (q0 + qm) - ATM_W1_12 * (qp + qm2));
    }
};

// the D winds the remap reads: the new interior, and outside it the wind
// halo of the start state (Q5).  in: the D step's winds, the start winds
struct OpWindSource {
    static constexpr int NR = 2;
    static constexpr bool ADJ_DIRECT = true;  // neighbouring points share few reads: no LDS planes
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S w1 = RD(0, 0, I, J, K), w0 = RD(1, 1, I, J, K);
        const bool inside = I >= ATM_HALO && I < ATM_HALO + x.g.nx && J >= ATM_HALO && J < ATM_HALO + x.g.ny;
        a.wr(0, I, J, K, inside ? w1 : w0);
    }
};

// total energy of a layer (111-128).  in: pt, playkap, gz, ud, vd
struct OpEnergyIn {
    static constexpr int NR = 16;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S pt = RD(0, 0, I, J, K), plk = RD(1, 1, I, J, K), g1 = RD(2, 2, I, J, K + 1), g0 = RD(3, 2, I, J, K);
        S u00 = RD(4, 3, I, J, K), u01 = RD(5, 3, I, J + 1, K), um0 = RD(6, 3, I - 1, J, K);
        S um1 = RD(7, 3, I - 1, J + 1, K), u10 = RD(8, 3, I + 1, J, K), u11 = RD(9, 3, I + 1, J + 1, K);
        S v00 = RD(10, 4, I, J, K), v10 = RD(11, 4, I + 1, J, K), v0m = RD(12, 4, I, J - 1, K);
        S v1m = RD(13, 4, I + 1, J - 1, K), v01 = RD(14, 4, I, J + 1, K), v11 = RD(15, 4, I + 1, J + 1, K);
        S ua = (.25 * (u00 + u01) + .125 * (um0 + um1 + u10 + u11));
        S va = (.25 * (v00 + v10) + .125 * (v0m + v1m + v01 + v11));
        a.wr(0, I, J, K, ATM_CV_DRY * pt * plk + .5 * (g1 + g0 + ua * ua + va * va));
    }
};

// kinetic energy of the remapped winds on the A grid (137-142)
struct OpKeRemap {
    static constexpr int NR = 12;
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S u00 = RD(0, 0, I, J, K), u01 = RD(1, 0, I, J + 1, K), um0 = RD(2, 0, I - 1, J, K);
        S um1 = RD(3, 0, I - 1, J + 1, K), u10 = RD(4, 0, I + 1, J, K), u11 = RD(5, 0, I + 1, J + 1, K);
        S v00 = RD(6, 1, I, J, K), v10 = RD(7, 1, I + 1, J, K), v0m = RD(8, 1, I, J - 1, K);
        S v1m = RD(9, 1, I + 1, J - 1, K), v01 = RD(10, 1, I, J + 1, K), v11 = RD(11, 1, I + 1, J + 1, K);
        S ua = (.25 * (u00 + u01) + .125 * (um0 + um1 + u10 + u11));
        S va = (.25 * (v00 + v10) + .125 * (v0m + v1m + v01 + v11));
        a.wr(0, I, J, K, .5 * (ua * ua + va * va));
    }
};

// dp = target_plev(k) - target_plev(k+1) (157-158)
struct OpDpOut {
    static constexpr int NR = 2;
    static constexpr bool ADJ_DIRECT = true;  // neighbouring points share few reads: no LDS planes
    template <class A>
    __host__ __device__ static void run(A &a, const LinArgs &x, int I, int J, int K) {
        using S = typename A::S;
        S p0 = RD(0, 0, I, J, K), p1 = RD(1, 0, I, J, K + 1);
        a.wr(0, I, J, K, p0 - p1);
    }
};

#undef RD

enum { RM_COL_REMAP = 0, RM_COL_PT = 1 };

// The remap program (perform_vertical_remapping 27-163).  Grids over
// isd-3..ied+3, winds over isd-1..ied+1, energy and output on the interior.
template <class E>
void lin_remap_program(E &e) {
    const Geo &g = e.g;
    const int L = g.L;
    auto R = [&](int a, int b) { return lin_box(g, a, b); };
    // current grid, p**kappa, geopotential (42-53, 72-78), playkap (Q17)
    e.scan(true, R(4, 4), A_DP1, A_PT1, A_RPLEV, A_RPKAP, A_RGZ);
    e.template op<OpRemapPlk>(R(4, 4), 0, L, {A_RPKAP}, {A_RPLK});
    // target grid (55-70)
    e.template op<OpTargetPlev>(R(4, 4), 0, L, {A_RPLEV}, {A_TPLEV});
    e.template op<OpPowKappa>(R(4, 4), 0, L, {A_TPLEV}, {A_TPKAP});
    e.template op<OpRemapPlk>(R(4, 4), 0, L, {A_TPKAP}, {A_TPLK});
    // winds (80-105): face coordinates, the wind source with the Q5 ring
    e.template op<OpFaceAvg<1>>(R(6, 6), 0, L + 1, {A_RPLEV}, {A_UFX});
    e.template op<OpFaceAvg<1>>(R(6, 6), 0, L + 1, {A_TPLEV}, {A_UFD});
    e.template op<OpFaceAvg<0>>(R(6, 6), 0, L + 1, {A_RPLEV}, {A_VFX});
    e.template op<OpFaceAvg<0>>(R(6, 6), 0, L + 1, {A_TPLEV}, {A_VFD});
    e.template op<OpWindSource>(R(6, 6), 0, L, {A_UD1, A_UD0}, {A_UDM});
    e.template op<OpWindSource>(R(6, 6), 0, L, {A_VD1, A_VD0}, {A_VDM});
    // total energy of each layer (111-128): from the winds before the remap
    e.template op<OpEnergyIn>(R(7, 7), 0, L, {A_PT1, A_RPLK, A_RGZ, A_UDM, A_VDM}, {A_ENER});
    // remap_column of the u and v winds and of the energy: independent
    // columns, one launch
    e.columns(RM_COL_REMAP, {e.col(R(6, 6), {A_UDM, A_UFX, A_UFD}, {A_UR}, RM_WS_U),
                             e.col(R(6, 6), {A_VDM, A_VFX, A_VFD}, {A_VR}, RM_WS_V),
                             e.col(R(7, 7), {A_ENER, A_RPLEV, A_TPLEV}, {A_EREM}, RM_WS_E)});
    // pt from the remapped energy and the remapped winds' kinetic energy (129-155)
    e.template op<OpKeRemap>(R(7, 7), 0, L, {A_UR, A_VR}, {A_KER});
    e.columns(RM_COL_PT, {e.col(R(7, 7), {A_EREM, A_KER, A_TPLK, A_TPKAP, A_RGZ}, {A_PT2}, RM_WS_PT)});
    // write-back (157-161): dp from the target grid; the winds are A_UR, A_VR
    e.template op<OpDpOut>(R(7, 7), 0, L, {A_TPLEV}, {A_DP2});
}
