// k_dyn_ref.hip -- reference-structured dynamics kernels (gfx950).
//
// One kernel per loop nest of src/mod_sw_dyn.f90, each launched over all
// levels at once (grid z = level), with the reference's exact index ranges
// and statement order, so every intermediate array matches the reference's
// module array bit for bit (given the same libm results).  The fused
// LDS-tiled path in k_dyn_fused.hip is checked against this one.
//
// The reference keeps its per-level work arrays (tmp, laplacian, cx, ...)
// as 2-D planes reused level by level; here each has one plane per level so
// that all levels run in one launch.  The A/B-grid winds ua, va, ub, vb are
// consumed by the kinetic-energy passes after the level loop that made
// them, so only their last-level (k = nlay) values are ever read; they are
// kept as single planes written at k = nlay.
#include "atm_common.hpp"

namespace {

const dim3 BLK(64, 4, 1);

struct Rng {
    int I0, I1, J0, J1;  // inclusive
};

#define RANGE_GUARD(r)                              \
    const int I = (r).I0 + blockIdx.x * blockDim.x + threadIdx.x; \
    const int J = (r).J0 + blockIdx.y * blockDim.y + threadIdx.y; \
    const int K = blockIdx.z;                       \
    if (I > (r).I1 || J > (r).J1) return;

template <typename F>
__global__ void dummy(F) {}

inline dim3 gridfor(const Rng &r, int nk) { return grid3(r.I1 - r.I0 + 1, r.J1 - r.J0 + 1, nk, BLK); }

// ---- column scans -----------------------------------------------------------
// plev(k) = plev(k+1) + src(k); pkap = plev**kappa, k = nlay..1, over a box of
// columns (cgrid 89-92, dgrid 433-437).
__global__ void k_scan_pkap(Geo g, Rng r, const double *__restrict__ src, double *__restrict__ plev,
                            double *__restrict__ pkap) {
    const int I = r.I0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = r.J0 + blockIdx.y * blockDim.y + threadIdx.y;
    if (I > r.I1 || J > r.J1) return;
    double p = plev[g.ix(I, J, g.L)];
    for (int K = g.L - 1; K >= 0; K--) {
        p = p + src[g.ix(I, J, K)];
        plev[g.ix(I, J, K)] = p;
        pkap[g.ix(I, J, K)] = pow_kappa(p);
    }
}

// plev/pkap from src top-down, then gz(k+1) = gz(k) + cp*th(k)*(pkap(k)-pkap(k+1))
// bottom-up (cgrid 276-292, dgrid 606-626).
__global__ void k_scan_pkap_gz(Geo g, Rng r, const double *__restrict__ src,
                               const double *__restrict__ th, double *__restrict__ plev,
                               double *__restrict__ pkap, double *__restrict__ gz) {
    const int I = r.I0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = r.J0 + blockIdx.y * blockDim.y + threadIdx.y;
    if (I > r.I1 || J > r.J1) return;
    double p = plev[g.ix(I, J, g.L)];
    for (int K = g.L - 1; K >= 0; K--) {
        p = p + src[g.ix(I, J, K)];
        plev[g.ix(I, J, K)] = p;
        pkap[g.ix(I, J, K)] = pow_kappa(p);
    }
    double z = gz[g.ix(I, J, 0)];
    double pk = pkap[g.ix(I, J, 0)];
    for (int K = 0; K < g.L; K++) {
        double pk1 = pkap[g.ix(I, J, K + 1)];
        z = z + ATM_CP_DRY * th[g.ix(I, J, K)] * (pk - pk1);
        gz[g.ix(I, J, K + 1)] = z;
        pk = pk1;
    }
}

// ---- C grid, first level loop (94-270) ---------------------------------------
// playkap and heating rate over is:ie (99-106)
__global__ void k_heat(Geo g, Rng r, DynCoef c, const double *__restrict__ pkap,
                       const double *__restrict__ nf, const double *__restrict__ dpl,
                       double *__restrict__ playkap, double *__restrict__ heat) {
    RANGE_GUARD(r);
    double p1 = pkap[g.ix(I, J, K + 1)], p0 = pkap[g.ix(I, J, K)];
    double pk = (p1 - p0) / log(p1 / p0);
    playkap[g.ix(I, J, K)] = pk;
    heat[g.ix(I, J, K)] = (nf[g.ix(I, J, K + 1)] - nf[g.ix(I, J, K)]) /
                          (ATM_CP_DRY * dpl[g.ix(I, J, K)] / c.gravity) / pk;
}

// B/C/A winds and Courant numbers (109-134)
__global__ void k_c_winds(Geo g, Rng r, DynCoef c, const double *__restrict__ ud,
                          const double *__restrict__ vd, double *__restrict__ uc,
                          double *__restrict__ vc, double *__restrict__ cx, double *__restrict__ cy,
                          double *__restrict__ ua, double *__restrict__ va) {
    RANGE_GUARD(r);
#define U(a, b) ud[g.ix(I + (a), J + (b), K)]
#define V(a, b) vd[g.ix(I + (a), J + (b), K)]
    double ub = .5 * (U(0, 0) + U(-1, 0));
    double vb = .5 * (V(0, 0) + V(0, -1));
    double u = .5 * (ub + .5 * (U(0, 1) + U(-1, 1)));
    double v = .5 * (vb + .5 * (V(1, 0) + V(1, -1)));
    uc[g.ix(I, J, K)] = u;
    vc[g.ix(I, J, K)] = v;
    cx[g.ix(I, J, K)] = c.dtdx * u;
    cy[g.ix(I, J, K)] = c.dtdx * v;  // Q2: y Courant number uses dt/dx
    if (K == g.L - 1) {
        ua[g.ix2(I, J)] = .5 * (u + .25 * (U(1, 0) + U(0, 0) + U(1, 1) + U(0, 1)));
        va[g.ix2(I, J)] = .5 * (v + .25 * (V(0, 1) + V(0, 0) + V(1, 1) + V(1, 0)));
    }
#undef U
#undef V
}

// denominators from C winds (136-139)
__global__ void k_denoms(Geo g, Rng r, DynCoef c, const double *__restrict__ uc,
                         const double *__restrict__ vc, double *__restrict__ denx,
                         double *__restrict__ deny) {
    RANGE_GUARD(r);
    denx[g.ix(I, J, K)] = 1. / (1. - c.dtdx * (uc[g.ix(I + 1, J, K)] - uc[g.ix(I, J, K)]));
    deny[g.ix(I, J, K)] = 1. / (1. - c.dtdy * (vc[g.ix(I, J + 1, K)] - vc[g.ix(I, J, K)]));
}

// inner y step: PCM y-fluxes of q with Courant cy, tmp and laplacian
// (142-155, 209-222 and the D-grid 470-482, 535-547)
__global__ void k_inner_y(Geo g, Rng r, DynCoef c, const double *__restrict__ q,
                          const double *__restrict__ cy, const double *__restrict__ deny,
                          double *__restrict__ tmp, double *__restrict__ lap) {
    RANGE_GUARD(r);
    double q0 = q[g.ix(I, J, K)];
    double f0 = pcm_face(q[g.ix(I, J - 1, K)], q0, cy[g.ix(I, J, K)]);
    double f1 = pcm_face(q0, q[g.ix(I, J + 1, K)], cy[g.ix(I, J + 1, K)]);
    tmp[g.ix(I, J, K)] = .5 * (q0 + deny[g.ix(I, J, K)] * (q0 - c.rA * (f1 - f0)));
    lap[g.ix(I, J, K)] = -2. * q0 * c.rdx2rdy2 + (q[g.ix(I + 1, J, K)] + q[g.ix(I - 1, J, K)]) * c.rdx2 +
                         (q[g.ix(I, J + 1, K)] + q[g.ix(I, J - 1, K)]) * c.rdy2;
}

// inner x step: PCM x-fluxes of q with Courant cx, tmp (176-185, 241-250, D 503-511, 566-574)
__global__ void k_inner_x(Geo g, Rng r, DynCoef c, const double *__restrict__ q,
                          const double *__restrict__ cx, const double *__restrict__ denx,
                          double *__restrict__ tmp) {
    RANGE_GUARD(r);
    double q0 = q[g.ix(I, J, K)];
    double f0 = pcm_face(q[g.ix(I - 1, J, K)], q0, cx[g.ix(I, J, K)]);
    double f1 = pcm_face(q0, q[g.ix(I + 1, J, K)], cx[g.ix(I + 1, J, K)]);
    tmp[g.ix(I, J, K)] = .5 * (q0 + denx[g.ix(I, J, K)] * (q0 - c.rA * (f1 - f0)));
}

// outer PPM face values along x or y over a range of faces; out = f*cf when
// mass_flux, else f.  DIR 0 = x, 1 = y.
template <int V, int DIR>
__global__ void k_outer(Geo g, Rng r, const double *__restrict__ tmp, const double *__restrict__ cc,
                        double *__restrict__ out, int mass_flux) {
    RANGE_GUARD(r);
    const int sI = DIR == 0 ? 1 : 0, sJ = DIR == 0 ? 0 : 1;
#define T(m) tmp[g.ix(I + (m) * sI, J + (m) * sJ, K)]
    double cf = cc[g.ix(I, J, K)];
    double f = ppm_face<V>(T(-3), T(-2), T(-1), T(0), T(1), T(2), cf);
#undef T
    out[g.ix(I, J, K)] = mass_flux ? f * cf : f;
}

// dpc x part (167-171 / D 494-498)
__global__ void k_dpc_x(Geo g, Rng r, DynCoef c, const double *__restrict__ dpl,
                        const double *__restrict__ dpfx, const double *__restrict__ lap,
                        double *__restrict__ dpc) {
    RANGE_GUARD(r);
    dpc[g.ix(I, J, K)] = dpl[g.ix(I, J, K)] -
                         (dpfx[g.ix(I + 1, J, K)] - dpfx[g.ix(I, J, K)] +
                          c.dtdx * c.vx * (lap[g.ix(I + 1, J, K)] - 2. * lap[g.ix(I, J, K)] + lap[g.ix(I - 1, J, K)]));
}

// dpc y part (197-201 / D 523-527)
__global__ void k_dpc_y(Geo g, Rng r, DynCoef c, const double *__restrict__ dpfy,
                        const double *__restrict__ lap, double *__restrict__ dpc) {
    RANGE_GUARD(r);
    dpc[g.ix(I, J, K)] = dpc[g.ix(I, J, K)] -
                         (dpfy[g.ix(I, J + 1, K)] - dpfy[g.ix(I, J, K)] +
                          c.dtdy * c.vy * (lap[g.ix(I, J + 1, K)] - 2. * lap[g.ix(I, J, K)] + lap[g.ix(I, J - 1, K)]));
}

// ptc x part (231-236 / D 556-561)
__global__ void k_ptc_x(Geo g, Rng r, DynCoef c, const double *__restrict__ dpl,
                        const double *__restrict__ ptl, const double *__restrict__ fx,
                        const double *__restrict__ dpfx, const double *__restrict__ lap,
                        double *__restrict__ ptc) {
    RANGE_GUARD(r);
    double d1 = dpfx[g.ix(I + 1, J, K)], d0 = dpfx[g.ix(I, J, K)];
    double l1 = lap[g.ix(I + 1, J, K)], l0 = lap[g.ix(I, J, K)], lm = lap[g.ix(I - 1, J, K)];
    ptc[g.ix(I, J, K)] = dpl[g.ix(I, J, K)] * ptl[g.ix(I, J, K)] -
                         (fx[g.ix(I + 1, J, K)] * d1 - fx[g.ix(I, J, K)] * d0 +
                          c.vx * ((l1 - l0) * d1 - (l0 - lm) * d0));
}

// ptc y part and division by dpc (259-266 / D 583-590)
__global__ void k_ptc_y(Geo g, Rng r, DynCoef c, const double *__restrict__ dpl,
                        const double *__restrict__ fy, const double *__restrict__ dpfy,
                        const double *__restrict__ lap, const double *__restrict__ heat,
                        const double *__restrict__ dpc, double *__restrict__ ptc) {
    RANGE_GUARD(r);
    double d1 = dpfy[g.ix(I, J + 1, K)], d0 = dpfy[g.ix(I, J, K)];
    double l1 = lap[g.ix(I, J + 1, K)], l0 = lap[g.ix(I, J, K)], lm = lap[g.ix(I, J - 1, K)];
    ptc[g.ix(I, J, K)] = (ptc[g.ix(I, J, K)] -
                          (fy[g.ix(I, J + 1, K)] * d1 - fy[g.ix(I, J, K)] * d0 +
                           c.vy * ((l1 - l0) * d1 - (l0 - lm) * d0)) +
                          c.dt * dpl[g.ix(I, J, K)] * heat[g.ix(I, J, K)]) /
                         dpc[g.ix(I, J, K)];
}

// ---- C grid, second level loop (294-414) ---------------------------------------
// kinetic energy (299-317): Q4 shifted pencils, last-level ua/va (see header)
__global__ void k_c_ke(Geo g, Rng r, DynCoef c, const double *__restrict__ uc,
                       const double *__restrict__ vc, const double *__restrict__ ua,
                       const double *__restrict__ va, double *__restrict__ ke) {
    RANGE_GUARD(r);
    double vaij = va[g.ix2(I, J)], uaij = ua[g.ix2(I, J)];
#define VQ(m) vc[g.ix(I, J + 1 + (m), K)]
    double fy = ppm_face<V_PPM_U>(VQ(-3), VQ(-2), VQ(-1), VQ(0), VQ(1), VQ(2), c.dtdy * vaij);
#undef VQ
#define UQ(m) uc[g.ix(I + 1 + (m), J, K)]
    double fx = ppm_face<V_PPM_U>(UQ(-3), UQ(-2), UQ(-1), UQ(0), UQ(1), UQ(2), c.dtdx * uaij);
#undef UQ
    double k = vaij * fy;
    ke[g.ix(I, J, K)] = .5 * (k + uaij * fx);
}

// C-grid vorticity (319-323)
__global__ void k_c_vort(Geo g, Rng r, DynCoef c, const double *__restrict__ uc,
                         const double *__restrict__ vc, double *__restrict__ vort) {
    RANGE_GUARD(r);
    vort[g.ix(I, J, K)] = c.coriolis - (uc[g.ix(I, J, K)] - uc[g.ix(I, J - 1, K)]) * c.rdy +
                          (vc[g.ix(I, J, K)] - vc[g.ix(I - 1, J, K)]) * c.rdx;
}

// C-grid pressure gradient force (325-339)
__global__ void k_c_pgf(Geo g, Rng r, const double *__restrict__ pkap, const double *__restrict__ gz,
                        double *__restrict__ pgfx, double *__restrict__ pgfy) {
    RANGE_GUARD(r);
    double a = pkap[g.ix(I, J, K)], d = pkap[g.ix(I, J, K + 1)];
    double b = pkap[g.ix(I - 1, J, K + 1)], cc = pkap[g.ix(I - 1, J, K)];
    pgfx[g.ix(I, J, K)] = ((a - b) * (gz[g.ix(I - 1, J, K)] - gz[g.ix(I, J, K + 1)]) +
                           (cc - d) * (gz[g.ix(I - 1, J, K + 1)] - gz[g.ix(I, J, K)])) /
                          (a - b + cc - d);
    b = pkap[g.ix(I, J - 1, K + 1)];
    cc = pkap[g.ix(I, J - 1, K)];
    pgfy[g.ix(I, J, K)] = ((a - b) * (gz[g.ix(I, J - 1, K)] - gz[g.ix(I, J, K + 1)]) +
                           (cc - d) * (gz[g.ix(I, J - 1, K + 1)] - gz[g.ix(I, J, K)])) /
                          (a - b + cc - d);
}

// denominators and Courant numbers from D winds (344-350)
__global__ void k_c_denoms2(Geo g, Rng r, DynCoef c, const double *__restrict__ ud,
                            const double *__restrict__ vd, double *__restrict__ denx,
                            double *__restrict__ deny, double *__restrict__ cx,
                            double *__restrict__ cy) {
    RANGE_GUARD(r);
    double u0 = ud[g.ix(I, J, K)], v0 = vd[g.ix(I, J, K)];
    denx[g.ix(I, J, K)] = 1. / (1. - c.dtdx * (u0 - ud[g.ix(I - 1, J, K)]));
    deny[g.ix(I, J, K)] = 1. / (1. - c.dtdy * (v0 - vd[g.ix(I, J - 1, K)]));
    cx[g.ix(I, J, K)] = c.dtdx * u0;
    cy[g.ix(I, J, K)] = c.dtdx * v0;  // Q2
}

// inner x PCM of vorticity with Courant cx shifted by sh (C grid: sh = -1, Q4;
// D grid: 0), tmp and laplacian (353-367 / D 708-722)
// kc >= 0 reads cx/denx of level kc instead of the current level (D grid:
// those 2-D arrays still hold the first level loop's last level, see
// atm_ref_dgrid)
__global__ void k_vort_inner_x(Geo g, Rng r, DynCoef c, int sh, int kc, const double *__restrict__ vort,
                               const double *__restrict__ cx, const double *__restrict__ denx,
                               double *__restrict__ tmp, double *__restrict__ lap) {
    RANGE_GUARD(r);
    const int Kc = kc >= 0 ? kc : K;
    double w0 = vort[g.ix(I, J, K)];
    double f0 = pcm_face(vort[g.ix(I - 1, J, K)], w0, cx[g.ix(I + sh, J, Kc)]);
    double f1 = pcm_face(w0, vort[g.ix(I + 1, J, K)], cx[g.ix(I + 1 + sh, J, Kc)]);
    tmp[g.ix(I, J, K)] = .5 * (w0 + denx[g.ix(I, J, Kc)] * (w0 - c.rA * (f1 - f0)));
    lap[g.ix(I, J, K)] = -2. * w0 * c.rdx2rdy2 + (vort[g.ix(I + 1, J, K)] + vort[g.ix(I - 1, J, K)]) * c.rdx2 +
                         (vort[g.ix(I, J + 1, K)] + vort[g.ix(I, J - 1, K)]) * c.rdy2;
}

// inner y PCM of vorticity with Courant cy shifted by sh (388-397 / D 744-753)
__global__ void k_vort_inner_y(Geo g, Rng r, DynCoef c, int sh, int kc, const double *__restrict__ vort,
                               const double *__restrict__ cy, const double *__restrict__ deny,
                               double *__restrict__ tmp) {
    RANGE_GUARD(r);
    const int Kc = kc >= 0 ? kc : K;
    double w0 = vort[g.ix(I, J, K)];
    double f0 = pcm_face(vort[g.ix(I, J - 1, K)], w0, cy[g.ix(I, J + sh, Kc)]);
    double f1 = pcm_face(w0, vort[g.ix(I, J + 1, K)], cy[g.ix(I, J + 1 + sh, Kc)]);
    tmp[g.ix(I, J, K)] = .5 * (w0 + deny[g.ix(I, J, Kc)] * (w0 - c.rA * (f1 - f0)));
}

// uc update (369-382): outer PPM_U along y on tmp shifted by +1 (Q4)
__global__ void k_c_uc(Geo g, Rng r, DynCoef c, const double *__restrict__ tmp,
                       const double *__restrict__ cy, const double *__restrict__ vd,
                       const double *__restrict__ lap, const double *__restrict__ ke,
                       const double *__restrict__ pgfx, double *__restrict__ uc) {
    RANGE_GUARD(r);
#define TQ(m) tmp[g.ix(I, J + 1 + (m), K)]
    double fy = ppm_face<V_PPM_U>(TQ(-3), TQ(-2), TQ(-1), TQ(0), TQ(1), TQ(2), cy[g.ix(I, J, K)]);
#undef TQ
    uc[g.ix(I, J, K)] = uc[g.ix(I, J, K)] +
                        c.dt * (vd[g.ix(I, J, K)] * fy + c.vy * (lap[g.ix(I, J + 1, K)] - lap[g.ix(I, J, K)])) -
                        c.dtdx * (ke[g.ix(I, J, K)] - ke[g.ix(I - 1, J, K)] - pgfx[g.ix(I, J, K)]);
}

// vc update (399-412): outer PPM_U along x on tmp shifted by +1 (Q4)
__global__ void k_c_vc(Geo g, Rng r, DynCoef c, const double *__restrict__ tmp,
                       const double *__restrict__ cx, const double *__restrict__ ud,
                       const double *__restrict__ lap, const double *__restrict__ ke,
                       const double *__restrict__ pgfy, double *__restrict__ vc) {
    RANGE_GUARD(r);
#define TQ(m) tmp[g.ix(I + 1 + (m), J, K)]
    double fx = ppm_face<V_PPM_U>(TQ(-3), TQ(-2), TQ(-1), TQ(0), TQ(1), TQ(2), cx[g.ix(I, J, K)]);
#undef TQ
    vc[g.ix(I, J, K)] = vc[g.ix(I, J, K)] -
                        c.dt * (ud[g.ix(I, J, K)] * fx + c.vx * (lap[g.ix(I + 1, J, K)] - lap[g.ix(I, J, K)])) -
                        c.dtdy * (ke[g.ix(I, J, K)] - ke[g.ix(I, J - 1, K)] - pgfy[g.ix(I, J, K)]);
}

// ---- D grid ---------------------------------------------------------------------
// B winds (last level only), Courant numbers and denominators (441-454)
__global__ void k_d_winds(Geo g, Rng r, DynCoef c, const double *__restrict__ uc,
                          const double *__restrict__ vc, double *__restrict__ cx,
                          double *__restrict__ cy, double *__restrict__ denx,
                          double *__restrict__ deny, double *__restrict__ ub,
                          double *__restrict__ vb) {
    RANGE_GUARD(r);  // range is:ie
    double u0 = uc[g.ix(I, J, K)], v0 = vc[g.ix(I, J, K)];
    cx[g.ix(I, J, K)] = c.dtdx * u0;
    cy[g.ix(I, J, K)] = c.dtdx * v0;  // Q2
    if (I >= 1 && J >= 1 && I <= g.NI - 2 && J <= g.NJ - 2) {
        denx[g.ix(I, J, K)] = 1. / (1. - c.dtdx * (uc[g.ix(I + 1, J, K)] - u0));
        deny[g.ix(I, J, K)] = 1. / (1. - c.dtdy * (vc[g.ix(I, J + 1, K)] - v0));
    }
    if (K == g.L - 1 && I >= 1 && J >= 1) {
        ub[g.ix2(I, J)] = .5 * (u0 + uc[g.ix(I, J - 1, K)]);
        vb[g.ix2(I, J)] = .5 * (v0 + vc[g.ix(I - 1, J, K)]);
    }
}

// copy back dp = dpc, pt = ptc on is+4..ie-4 (594-598)
__global__ void k_copy2(Geo g, Rng r, const double *__restrict__ a, const double *__restrict__ b,
                        double *__restrict__ oa, double *__restrict__ ob) {
    RANGE_GUARD(r);
    oa[g.ix(I, J, K)] = a[g.ix(I, J, K)];
    ob[g.ix(I, J, K)] = b[g.ix(I, J, K)];
}

// interpolate_to_corners (921-956) of q(level Kq) into out(level Ko) over the
// slab [s0, s1] in both directions; out range s0+2 .. s1-1.
__device__ __forceinline__ double corner_value(Geo g, const double *__restrict__ q, int I, int J, int Kq) {
#define Q(a, b) q[g.ix(I + (a), J + (b), Kq)]
#define TX(b) (ATM_W7_12 * (Q(0, b) + Q(-1, b)) - ATM_W1_12 * (Q(1, b) + Q(-2, b)))
#define TY(a) (ATM_W7_12 * (Q(a, 0) + Q(a, -1)) - ATM_W1_12 * (Q(a, 1) + Q(a, -2)))
    double o = (9. / 32.) * (TX(0) + TX(-1)) - (1. / 32.) * (TX(1) + TX(-2));
    o = o + (9. / 32.) * (TY(0) + TY(-1)) - (1. / 32.) * (TY(1) + TY(-2));
    return o;
#undef Q
#undef TX
#undef TY
}

__global__ void k_corners(Geo g, Rng r, int kofs, const double *__restrict__ q, double *__restrict__ out) {
    RANGE_GUARD(r);
    out[g.ix(I, J, K + kofs)] = corner_value(g, q, I, J, K + kofs);
}

// D-grid kinetic energy (639-657), PPM_CONSTRAINED, last-level ub/vb
__global__ void k_d_ke(Geo g, Rng r, DynCoef c, const double *__restrict__ ud,
                       const double *__restrict__ vd, const double *__restrict__ ub,
                       const double *__restrict__ vb, double *__restrict__ ke) {
    RANGE_GUARD(r);
    double vbij = vb[g.ix2(I, J)], ubij = ub[g.ix2(I, J)];
#define VQ(m) vd[g.ix(I, J + (m), K)]
    double fy = ppm_face<V_PPM_C>(VQ(-3), VQ(-2), VQ(-1), VQ(0), VQ(1), VQ(2), c.dtdy * vbij);
#undef VQ
#define UQ(m) ud[g.ix(I + (m), J, K)]
    double fx = ppm_face<V_PPM_C>(UQ(-3), UQ(-2), UQ(-1), UQ(0), UQ(1), UQ(2), c.dtdx * ubij);
#undef UQ
    double k = vbij * fy;
    ke[g.ix(I, J, K)] = .5 * (k + ubij * fx);
}

// D-grid vorticity (659-663)
__global__ void k_d_vort(Geo g, Rng r, DynCoef c, const double *__restrict__ ud,
                         const double *__restrict__ vd, double *__restrict__ vort) {
    RANGE_GUARD(r);
    vort[g.ix(I, J, K)] = c.coriolis - (ud[g.ix(I, J + 1, K)] - ud[g.ix(I, J, K)]) * c.rdy +
                          (vd[g.ix(I + 1, J, K)] - vd[g.ix(I, J, K)]) * c.rdx;
}

// D-grid pressure gradient force on the B grid (665-679)
__global__ void k_d_pgf(Geo g, Rng r, const double *__restrict__ pkb, const double *__restrict__ gzb,
                        double *__restrict__ pgfx, double *__restrict__ pgfy) {
    RANGE_GUARD(r);
    double a = pkb[g.ix(I + 1, J, K)], b = pkb[g.ix(I, J, K + 1)];
    double cc = pkb[g.ix(I, J, K)], d = pkb[g.ix(I + 1, J, K + 1)];
    pgfx[g.ix(I, J, K)] = ((a - b) * (gzb[g.ix(I, J, K)] - gzb[g.ix(I + 1, J, K + 1)]) +
                           (cc - d) * (gzb[g.ix(I, J, K + 1)] - gzb[g.ix(I + 1, J, K)])) /
                          (a - b + cc - d);
    a = pkb[g.ix(I, J + 1, K)];
    d = pkb[g.ix(I, J + 1, K + 1)];
    pgfy[g.ix(I, J, K)] = ((a - b) * (gzb[g.ix(I, J, K)] - gzb[g.ix(I, J + 1, K + 1)]) +
                           (cc - d) * (gzb[g.ix(I, J, K + 1)] - gzb[g.ix(I, J + 1, K)])) /
                          (a - b + cc - d);
}

// divergence (682-685)
__global__ void k_d_div(Geo g, Rng r, DynCoef c, const double *__restrict__ ud,
                        const double *__restrict__ vd, double *__restrict__ div) {
    RANGE_GUARD(r);
    div[g.ix(I, J, K)] = c.rdx * (ud[g.ix(I, J, K)] - ud[g.ix(I - 1, J, K)]) +
                         c.rdy * (vd[g.ix(I, J, K)] - vd[g.ix(I, J - 1, K)]);
}

// Laplacian of a (688-699); scale = 0 writes it as is, else multiplies (701-702)
__global__ void k_del2(Geo g, Rng r, DynCoef c, const double *__restrict__ a, double *__restrict__ out,
                       int scale) {
    RANGE_GUARD(r);
    double v = -2. * a[g.ix(I, J, K)] * c.rdx2rdy2 + (a[g.ix(I + 1, J, K)] + a[g.ix(I - 1, J, K)]) * c.rdx2 +
               (a[g.ix(I, J + 1, K)] + a[g.ix(I, J - 1, K)]) * c.rdy2;
    out[g.ix(I, J, K)] = scale ? c.divdamp * v : v;
}

// ud update (724-738); cy of the last level (see atm_ref_dgrid)
__global__ void k_d_ud(Geo g, Rng r, DynCoef c, const double *__restrict__ tmp,
                       const double *__restrict__ cy, const double *__restrict__ vc,
                       const double *__restrict__ lap, const double *__restrict__ ke,
                       const double *__restrict__ div, const double *__restrict__ pgfx,
                       double *__restrict__ ud) {
    RANGE_GUARD(r);
#define TQ(m) tmp[g.ix(I, J + (m), K)]
    double fy = ppm_face<V_PPM_C>(TQ(-3), TQ(-2), TQ(-1), TQ(0), TQ(1), TQ(2), cy[g.ix(I, J, g.L - 1)]);
#undef TQ
    ud[g.ix(I, J, K)] = ud[g.ix(I, J, K)] +
                        c.dt * (vc[g.ix(I, J, K)] * fy + c.vy * (lap[g.ix(I, J, K)] - lap[g.ix(I, J - 1, K)])) -
                        c.dtdx * (ke[g.ix(I + 1, J, K)] - ke[g.ix(I, J, K)] - div[g.ix(I + 1, J, K)] +
                                  div[g.ix(I, J, K)] - pgfx[g.ix(I, J, K)]);
}

// vd update (755-769); cx of the last level (see atm_ref_dgrid)
__global__ void k_d_vd(Geo g, Rng r, DynCoef c, const double *__restrict__ tmp,
                       const double *__restrict__ cx, const double *__restrict__ uc,
                       const double *__restrict__ lap, const double *__restrict__ ke,
                       const double *__restrict__ div, const double *__restrict__ pgfy,
                       double *__restrict__ vd) {
    RANGE_GUARD(r);
#define TQ(m) tmp[g.ix(I + (m), J, K)]
    double fx = ppm_face<V_PPM_C>(TQ(-3), TQ(-2), TQ(-1), TQ(0), TQ(1), TQ(2), cx[g.ix(I, J, g.L - 1)]);
#undef TQ
    vd[g.ix(I, J, K)] = vd[g.ix(I, J, K)] -
                        c.dt * (uc[g.ix(I, J, K)] * fx + c.vx * (lap[g.ix(I, J, K)] - lap[g.ix(I - 1, J, K)])) -
                        c.dtdy * (ke[g.ix(I, J + 1, K)] - ke[g.ix(I, J, K)] - div[g.ix(I, J + 1, K)] +
                                  div[g.ix(I, J, K)] - pgfy[g.ix(I, J, K)]);
}

// fill a plane of a level field with a constant (pkapb top, 991)
__global__ void k_fill(double *a, size_t n, double v) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

}  // namespace

// Range helper: reference bounds is+a .. ie-b (same in j)
static inline Rng R(const Geo &g, int a, int b) { return Rng{a, g.NI - 1 - b, a, g.NJ - 1 - b}; }
static inline Rng RX(const Geo &g, int ai, int bi, int aj, int bj) {
    return Rng{ai, g.NI - 1 - bi, aj, g.NJ - 1 - bj};
}

#define LAUNCH(kern, rng, nk, ...)                                                          \
    do {                                                                                    \
        Rng r_ = (rng);                                                                     \
        hipLaunchKernelGGL(kern, gridfor(r_, (nk)), BLK, 0, c->stream, c->g, r_, ##__VA_ARGS__); \
    } while (0)

int atm_launch_fill(atm_ctx *c, double *a, size_t n, double v) {
    if (n == 0) return ATM_OK;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a, n, v);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}

// init_sw_dyn (958-1023): coefficients, pkapb top, gzb(:,:,1)
int atm_ref_init_sw_dyn(atm_ctx *c) {
    const Geo &g = c->g;
    DynCoef &k = c->coef;
    k.rdx = 1. / c->dx;
    k.rdx2 = 1. / (c->dx * c->dx);
    k.rdy = 1. / c->dy;
    k.rdy2 = 1. / (c->dy * c->dy);
    k.rdx2rdy2 = k.rdx2 + k.rdy2;
    k.rA = 1. / (c->dx * c->dy);
    double x = c->prm.divergence_damping_coefficient * c->dx * c->dy;
    k.divdamp = x * x;
    double vort = c->prm.vorticity_damping_coefficient * c->dx * c->dy;
    k.vx = c->dy * k.rdx * vort;
    k.vy = c->dx * k.rdy * vort;
    k.coriolis = c->prm.coriolis_parameter;
    k.gravity = c->prm.gravity;
    int rc = atm_launch_fill(c, c->pkapb + g.plane() * g.L, g.plane(), pow(c->ak0, ATM_KAPPA));
    if (rc) return rc;
    // gzb(:,:,1) = corners(gz(:,:,1)) over the whole halo'd plane
    LAUNCH(k_corners, R(g, 2, 1), 1, 0, c->gz, c->gzb);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}

static void set_dt(atm_ctx *c, double dt) {
    c->coef.dt = dt;
    c->coef.dtdx = dt * c->coef.rdx;
    c->coef.dtdy = dt * c->coef.rdy;
}

// cgrid_dynamics_step (80-416)
int atm_ref_cgrid(atm_ctx *c, double dt) {
    const Geo &g = c->g;
    set_dt(c, dt);
    const DynCoef &k = c->coef;
    const int L = g.L;
    dim3 cb(64, 4);
    Rng all = R(g, 0, 0);
    hipLaunchKernelGGL(k_scan_pkap, grid3(g.NI, g.NJ, 1, cb), cb, 0, c->stream, g, all, c->dp, c->plev, c->pkap);
    LAUNCH(k_heat, all, L, k, c->pkap, c->net_flux, c->dp, c->playkap, c->heat);
    LAUNCH(k_c_winds, R(g, 1, 1), L, k, c->ud, c->vd, c->uc, c->vc, c->cx, c->cy, c->ua, c->va);
    LAUNCH(k_denoms, R(g, 2, 2), L, k, c->uc, c->vc, c->denx, c->deny);
    // dpc
    LAUNCH(k_inner_y, R(g, 2, 2), L, k, c->dp, c->cy, c->deny, c->tmp, c->lap);
    LAUNCH((k_outer<V_PPM_U, 0>), RX(g, 5, 4, 5, 5), L, c->tmp, c->cx, c->dpfx, 1);
    LAUNCH(k_dpc_x, R(g, 5, 5), L, k, c->dp, c->dpfx, c->lap, c->dpc);
    LAUNCH(k_inner_x, R(g, 2, 2), L, k, c->dp, c->cx, c->denx, c->tmp);
    LAUNCH((k_outer<V_PPM_U, 1>), RX(g, 5, 5, 5, 4), L, c->tmp, c->cy, c->dpfy, 1);
    LAUNCH(k_dpc_y, R(g, 5, 5), L, k, c->dpfy, c->lap, c->dpc);
    // ptc
    LAUNCH(k_inner_y, R(g, 2, 2), L, k, c->pt, c->cy, c->deny, c->tmp, c->lap);
    LAUNCH((k_outer<V_PPM_U, 0>), RX(g, 5, 4, 5, 5), L, c->tmp, c->cx, c->fxq, 0);
    LAUNCH(k_ptc_x, R(g, 5, 5), L, k, c->dp, c->pt, c->fxq, c->dpfx, c->lap, c->ptc);
    LAUNCH(k_inner_x, R(g, 2, 2), L, k, c->pt, c->cx, c->denx, c->tmp);
    LAUNCH((k_outer<V_PPM_U, 1>), RX(g, 5, 5, 5, 4), L, c->tmp, c->cy, c->fyq, 0);
    LAUNCH(k_ptc_y, R(g, 5, 5), L, k, c->dp, c->fyq, c->dpfy, c->lap, c->heat, c->dpc, c->ptc);
    // pressure and geopotential on is+5..ie-5
    {
        Rng r5 = R(g, 5, 5);
        hipLaunchKernelGGL(k_scan_pkap_gz, grid3(r5.I1 - r5.I0 + 1, r5.J1 - r5.J0 + 1, 1, cb), cb, 0,
                           c->stream, g, r5, c->dpc, c->ptc, c->plev, c->pkap, c->gz);
    }
    // second level loop
    LAUNCH(k_c_ke, R(g, 5, 5), L, k, c->uc, c->vc, c->ua, c->va, c->ke);
    LAUNCH(k_c_vort, R(g, 2, 2), L, k, c->uc, c->vc, c->vort);
    LAUNCH(k_c_pgf, R(g, 7, 7), L, c->pkap, c->gz, c->pgfx, c->pgfy);
    LAUNCH(k_c_denoms2, R(g, 1, 1), L, k, c->ud, c->vd, c->denx, c->deny, c->cx, c->cy);
    LAUNCH(k_vort_inner_x, R(g, 3, 3), L, k, -1, -1, c->vort, c->cx, c->denx, c->tmp, c->lap);
    LAUNCH(k_c_uc, R(g, 7, 7), L, k, c->tmp, c->cy, c->vd, c->lap, c->ke, c->pgfx, c->uc);
    LAUNCH(k_vort_inner_y, R(g, 3, 3), L, k, -1, -1, c->vort, c->cy, c->deny, c->tmp);
    LAUNCH(k_c_vc, R(g, 7, 7), L, k, c->tmp, c->cx, c->ud, c->lap, c->ke, c->pgfy, c->vc);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}

// dgrid_dynamics_step (424-773).  The second level loop (634-771) does not
// recompute the 2-D work arrays cx, cy, denom_x, denom_y, ub, vb set by the
// first loop (441-454), so every level's vorticity fluxes and kinetic energy
// use their values from the last level, k = nlay; the kernels below read
// level nlay of those arrays.
int atm_ref_dgrid(atm_ctx *c, double dt) {
    const Geo &g = c->g;
    set_dt(c, dt);
    const DynCoef &k = c->coef;
    const int L = g.L;
    dim3 cb(64, 4);
    Rng all = R(g, 0, 0);
    hipLaunchKernelGGL(k_scan_pkap, grid3(g.NI, g.NJ, 1, cb), cb, 0, c->stream, g, all, c->dpc, c->plev, c->pkap);
    LAUNCH(k_d_winds, all, L, k, c->uc, c->vc, c->cx, c->cy, c->denx, c->deny, c->ub, c->vb);
    LAUNCH(k_heat, all, L, k, c->pkap, c->net_flux, c->dpc, c->playkap, c->heat);
    // dp
    LAUNCH(k_inner_y, R(g, 1, 1), L, k, c->dp, c->cy, c->deny, c->tmp, c->lap);
    LAUNCH((k_outer<V_PPM_NN, 0>), RX(g, 4, 3, 4, 4), L, c->tmp, c->cx, c->dpfx, 1);
    LAUNCH(k_dpc_x, R(g, 4, 4), L, k, c->dp, c->dpfx, c->lap, c->dpc);
    LAUNCH(k_inner_x, R(g, 1, 1), L, k, c->dp, c->cx, c->denx, c->tmp);
    LAUNCH((k_outer<V_PPM_NN, 1>), RX(g, 4, 4, 4, 3), L, c->tmp, c->cy, c->dpfy, 1);
    LAUNCH(k_dpc_y, R(g, 4, 4), L, k, c->dpfy, c->lap, c->dpc);
    // pt
    LAUNCH(k_inner_y, R(g, 1, 1), L, k, c->pt, c->cy, c->deny, c->tmp, c->lap);
    LAUNCH((k_outer<V_PPM_NN, 0>), RX(g, 4, 3, 4, 4), L, c->tmp, c->cx, c->fxq, 0);
    LAUNCH(k_ptc_x, R(g, 4, 4), L, k, c->dp, c->pt, c->fxq, c->dpfx, c->lap, c->ptc);
    LAUNCH(k_inner_x, R(g, 1, 1), L, k, c->pt, c->cx, c->denx, c->tmp);
    LAUNCH((k_outer<V_PPM_NN, 1>), RX(g, 4, 4, 4, 3), L, c->tmp, c->cy, c->fyq, 0);
    LAUNCH(k_ptc_y, R(g, 4, 4), L, k, c->dp, c->fyq, c->dpfy, c->lap, c->heat, c->dpc, c->ptc);
    LAUNCH(k_copy2, R(g, 4, 4), L, c->dpc, c->ptc, c->dp, c->pt);
    // pressure, geopotential and their corner values on is+4..ie-4
    {
        Rng r4 = R(g, 4, 4);
        hipLaunchKernelGGL(k_scan_pkap_gz, grid3(r4.I1 - r4.I0 + 1, r4.J1 - r4.J0 + 1, 1, cb), cb, 0,
                           c->stream, g, r4, c->dp, c->pt, c->plev, c->pkap, c->gz);
    }
    LAUNCH(k_corners, R(g, 6, 5), L, 0, c->pkap, c->pkapb);
    LAUNCH(k_corners, R(g, 6, 5), L, 1, c->gz, c->gzb);
    // second level loop
    LAUNCH(k_d_ke, R(g, 4, 4), L, k, c->ud, c->vd, c->ub, c->vb, c->ke);
    LAUNCH(k_d_vort, R(g, 1, 1), L, k, c->ud, c->vd, c->vort);
    LAUNCH(k_d_pgf, R(g, 7, 7), L, c->pkapb, c->gzb, c->pgfx, c->pgfy);
    LAUNCH(k_d_div, R(g, 4, 4), L, k, c->ud, c->vd, c->divg);
    LAUNCH(k_del2, R(g, 5, 5), L, k, c->divg, c->tmp, 0);
    LAUNCH(k_del2, R(g, 6, 6), L, k, c->tmp, c->divg, 1);
    LAUNCH(k_vort_inner_x, R(g, 4, 4), L, k, 0, L - 1, c->vort, c->cx, c->denx, c->tmp, c->lap);
    LAUNCH(k_d_ud, R(g, 7, 7), L, k, c->tmp, c->cy, c->vc, c->lap, c->ke, c->divg, c->pgfx, c->ud);
    LAUNCH(k_vort_inner_y, R(g, 4, 4), L, k, 0, L - 1, c->vort, c->cy, c->deny, c->tmp);
    LAUNCH(k_d_vd, R(g, 7, 7), L, k, c->tmp, c->cx, c->uc, c->lap, c->ke, c->divg, c->pgfy, c->vd);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}
