// remap_column.hpp -- one column of the vertical PPM remap, as two passes
// (src/mod_remap.f90:165-356).  Host and device: the library's kernels run
// it per thread; a host build of the same code is checked against the
// oracle's remap_column on the CPU.
//
// A column source P provides X(k) (source face coordinates, k = 1..L+1),
// D(k) (destination faces) and V(k) (layer values, k = 1..L); lv and gm
// are strided (ns) scratch arrays of L+2 entries for the eliminated edge
// values and gam.  Face coordinates also come in two halves: xr(k) / dr(k)
// issue the loads a coordinate is formed from (type P::XR) and xv() forms
// it, so both passes load the next level's values one level before they
// use them: the loads are in flight while the current level is computed
// instead of stalling the column's serial chain every level.
#pragma once

#include "atm_common.hpp"

// vertical_ppm elimination sweep (297-311), 1-based k.
template <class P>
__host__ __device__ void ppm_forward(const P &p, int L, double *lv, double *gm, size_t ns) {
    double xp = p.X(L + 1), x0 = p.X(L), xm = p.X(L - 1);
    double v0 = p.V(L), vm = p.V(L - 1);
    // level k - 2 of the next iteration, loaded one iteration ahead
    int kn = L - 2 >= 1 ? L - 2 : 1;
    typename P::XR xn = p.xr(kn);
    double vn = p.V(kn);
    double alpha = (x0 - xm) / (xp - x0);
    double beta = alpha * (.5 + alpha);
    double lvn = (2. * alpha * (1. + alpha) * v0 + vm) / beta;  // level_var(L+1)
    lv[(size_t)(L + 1) * ns] = lvn;
    double g = (1. + alpha * (1.5 + alpha)) / beta;              // gam(L)
    gm[(size_t)L * ns] = g;
    for (int k = L; k >= 2; k--) {
        // xp = X(k+1), x0 = X(k), xm = X(k-1), v0 = V(k), vm = V(k-1),
        // lvn = level_var(k+1), g = gam(k)
        alpha = (xp - x0) / (x0 - xm);
        beta = 2. + 2. * alpha - g;
        const double l = (3. * (v0 + alpha * vm) - lvn) / beta;
        lv[(size_t)k * ns] = l;
        g = alpha / beta;
        gm[(size_t)(k - 1) * ns] = g;
        lvn = l;
        if (k > 2) {
            xp = x0;
            x0 = xm;
            xm = p.xv(xn);  // X(k - 2)
            v0 = vm;
            vm = vn;        // V(k - 2)
            kn = k - 3 >= 1 ? k - 3 : 1;
            xn = p.xr(kn);
            vn = p.V(kn);
        }
    }
    // alpha of k = 2; v0 = V(2), vm = V(1), lvn = level_var(2), g = gam(1)
    beta = 1. + alpha * (1.5 + alpha);
    lv[ns] = (2. * alpha * (1. + alpha) * vm + v0 - beta * lvn) / (alpha * (.5 + alpha) - beta * g);
}

// Ascending pass: PPM coefficients of each source layer plus the overlap
// walk.  sink(m, acc, dm0, dm1) receives destination layer m in order.
template <class P, class Sink>
__host__ __device__ void ppm_merge(const P &p, int L, const double *lv, const double *gm, size_t ns, Sink &sink) {
    auto G = [&](int i) { return i <= L ? gm[(size_t)i * ns] : 0.; };
    // limiter window of the current source layer k
    int k = 1;
    double gA = 0., gB = 0., gC = G(1), gD = G(2), gE = G(3);  // gam(k-2 .. k+2)
    double vC = p.V(1), vD = p.V(2);                            // var(k), var(k+1)
    double bsC = lv[ns];   // back-substituted level_var(k) before the clamps
    double lfC = bsC;      // final level_var(k)
    double bsN = 0., lfN = 0.;
    // current layer: btm, top, curvature, faces
    double b = 0., t = 0., cv = 0., sk = p.X(1), sk1 = p.X(2);
    // what the next advance (to layer k + 1) reads, loaded one layer ahead:
    // gam(k+3), var(k+2), X(k+2), level_var(k+2); indices past the column
    // are clamped into it and their values replaced as G and V do
    auto cl = [&](int i) { return i <= L + 1 ? i : L + 1; };
    double ng = gm[(size_t)cl(4) * ns], nv = p.V(3 <= L ? 3 : L), nl = lv[(size_t)cl(3) * ns];
    typename P::XR nx = p.xr(cl(3));

    // lvj = level_var(k+1) after the elimination sweep
    auto layer = [&](double lvj) {
        const int j = k + 1;
        const double bs = lvj - gC * bsC;  // back substitution (313-315)
        double l = bs;
        if (j == L) {  // 316-317
            l = dmin2(l, dmax2(vD, vC));
            l = dmax2(l, dmin2(vD, vC));
        }
        if (j == 2) {  // 318-319
            l = dmin2(l, dmax2(vD, vC));
            l = dmax2(l, dmin2(vD, vC));
        }
        if (j >= 3 && j <= L - 1) {  // 325-334 at index j-1: gam(j+1)*gam(j-1)
            const double lo = dmin2(vC, vD), hi = dmax2(vC, vD);
            if (gE * gC > 0) {
                l = dmin2(l, hi);
                l = dmax2(l, lo);
            } else if (gE > 0) {
                l = dmax2(l, lo);
            } else {
                l = dmin2(l, hi);
            }
        }
        bsN = bs;
        lfN = l;
        t = l;    // top_edge(k) = level_var(k+1)
        b = lfC;  // btm_edge(k) = level_var(k)
        if (k >= 3) {  // 340-352 at index k-1
            double pmp1 = vC - 2. * gB;
            double lac1 = pmp1 + 1.5 * gA;
            t = dmin2(dmax2(t, dmin3(vC, pmp1, lac1)), dmax3(vC, pmp1, lac1));
            pmp1 = vC + 2. * gC;
            lac1 = pmp1 - 1.5 * gB;
            b = dmin2(dmax2(b, dmin3(vC, pmp1, lac1)), dmax3(vC, pmp1, lac1));
        }
        cv = 6. * vC - 3. * (t + b);
    };
    auto advance = [&]() {
        k++;
        if (k > L) return;
        gA = gB;
        gB = gC;
        gC = gD;
        gD = gE;
        gE = k + 2 <= L ? ng : 0.;         // G(k + 2)
        vC = vD;
        vD = k + 1 <= L ? nv : 0.;         // V(k + 1)
        bsC = bsN;
        lfC = lfN;
        sk = sk1;
        sk1 = p.xv(nx);                    // X(k + 1)
        const double lvj = nl;             // level_var(k + 1)
        ng = gm[(size_t)cl(k + 3) * ns];
        nv = p.V(k + 2 <= L ? k + 2 : L);
        nl = lv[(size_t)cl(k + 2) * ns];
        nx = p.xr(cl(k + 2));
        layer(lvj);
    };

    layer(lv[(size_t)2 * ns]);
    double dm1 = p.D(1);
    typename P::XR nd = p.dr(2);  // D(m + 1), loaded one layer ahead
    for (int m = 1; m <= L; m++) {
        const double dm0 = dm1;
        dm1 = p.xv(nd);
        nd = p.dr(cl(m + 2));
        while (k <= L && sk1 >= dm0) advance();
        double acc = 0.;
        if (k <= L && sk > dm1) {
            for (;;) {
                const double dsrc = sk1 - sk;
                const double s1 = (dm0 - sk) / dsrc;
                const double s2 = (dm1 - sk) / dsrc;
                const double factor = dsrc / (dm1 - dm0);
                const double tmp3 = 2. * (s1 * s1);
                if (dm1 <= sk1 && sk1 < dm0 && dm0 < sk) {  // case 3
                    double tmp1 = s1 - 1., tmp2 = s1 + 1.;
                    acc = acc + tmp1 / 6. * (3. * (b * tmp1 - t * tmp2) + cv * (tmp3 - tmp2)) * factor;
                }
                if (sk1 < dm1 && dm1 < sk && sk <= dm0) {  // case 4
                    acc = acc + (s2 / 2.) * (b * (2. - s2) + s2 * (t + cv * (1. - ATM_W2_3 * s2))) * factor;
                }
                if (dm1 <= sk1 && sk <= dm0) {  // case 5
                    acc = acc + vC * factor;
                }
                if (sk1 < dm1 && dm0 < sk) {  // case 6
                    double tmp1 = s1 + s2;
                    acc = acc + (s1 - s2) / 6. *
                                    (3. * (b * (tmp1 - 2.) - t * tmp1) + cv * (tmp3 + tmp1 * (2. * s2 - 3.))) *
                                    factor;
                }
                if (k == L || sk1 <= dm1) break;
                advance();
            }
        }
        sink(m, acc, dm0, dm1);
    }
}
