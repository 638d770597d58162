// k_remap.hip -- vertical Lagrangian -> hybrid sigma-p remap (gfx950).
//
// src/mod_remap.f90:27-163.  One thread per column.  Per-column work arrays
// (face coordinates, tridiagonal coefficients, PPM edges) live in a device
// scratch laid out [slot][column] so that a wavefront's 64 columns read and
// write each slot contiguously.
//
// remap_column (165-285) tests every (source k, destination m) pair.  For
// monotone coordinates the pairs that contribute (the test at 203-204 fails)
// are, for each m, a contiguous run of k that only moves upward as m grows;
// the kernel walks exactly those pairs, for each m in ascending k, with the
// reference's formulas, so every destination value is accumulated from the
// same terms in the same order.
#include "atm_common.hpp"

struct RemapWork {
    double *scr = nullptr;      // [slots][plane]
    double *tplev = nullptr;    // target_plev  (NI x NJ x L+1), used on isd-3..ied+3
    double *tpkap = nullptr;    // target_pkap
    double *tplk = nullptr;     // target_playkap (L)
    double *uremap = nullptr, *vremap = nullptr, *ptremap = nullptr;  // NI x NJ x L
    int slots = 0;
};

namespace {

// scratch slot bases (multiples of the plane size); L+2 per array
struct Slots {
    int L;
    __device__ int src() const { return 0; }
    __device__ int dst() const { return (L + 2); }
    __device__ int gam() const { return 2 * (L + 2); }
    __device__ int lv() const { return 3 * (L + 2); }
    __device__ int top() const { return 4 * (L + 2); }
    __device__ int btm() const { return 5 * (L + 2); }
    __device__ int crv() const { return 6 * (L + 2); }
    __device__ int var() const { return 7 * (L + 2); }
    static int count(int L) { return 8 * (L + 2); }
};

struct Col {
    double *s;
    size_t plane, col;
    __device__ double &at(int slot, int k) const { return s[(size_t)(slot + k) * plane + col]; }
};

// vertical_ppm (287-356) with 1-based k as in the reference.
__device__ void vertical_ppm(const Col &c, const Slots &S, int L) {
#define V(k) c.at(S.var(), (k))
#define X(k) c.at(S.src(), (k))
#define G(k) c.at(S.gam(), (k))
#define LV(k) c.at(S.lv(), (k))
#define TOP(k) c.at(S.top(), (k))
#define BTM(k) c.at(S.btm(), (k))
    double alpha = (X(L) - X(L - 1)) / (X(L + 1) - X(L));
    double beta = alpha * (.5 + alpha);
    LV(L + 1) = (2. * alpha * (1. + alpha) * V(L) + V(L - 1)) / beta;
    G(L) = (1. + alpha * (1.5 + alpha)) / beta;
    for (int k = L; k >= 2; k--) {
        alpha = (X(k + 1) - X(k)) / (X(k) - X(k - 1));
        beta = 2. + 2. * alpha - G(k);
        LV(k) = (3. * (V(k) + alpha * V(k - 1)) - LV(k + 1)) / beta;
        G(k - 1) = alpha / beta;
    }
    beta = 1. + alpha * (1.5 + alpha);
    LV(1) = (2. * alpha * (1. + alpha) * V(1) + V(2) - beta * LV(2)) / (alpha * (.5 + alpha) - beta * G(1));
    {
        double prev = LV(1);
        for (int k = 2; k <= L + 1; k++) {
            double v = LV(k) - G(k - 1) * prev;
            LV(k) = v;
            prev = v;
        }
    }
    LV(L) = dmin2(LV(L), dmax2(V(L), V(L - 1)));
    LV(L) = dmax2(LV(L), dmin2(V(L), V(L - 1)));
    LV(2) = dmin2(LV(2), dmax2(V(2), V(1)));
    LV(2) = dmax2(LV(2), dmin2(V(2), V(1)));
    // monotonicity (325-334).  At k = L-1 the reference reads gam(L+1) out of
    // bounds (Q7); every branch is then a no-op on the already clamped
    // level_var(L), so that iteration is skipped.
    for (int k = 2; k <= L - 2; k++) {
        double gk2 = G(k + 2), gk = G(k);
        double lo = dmin2(V(k), V(k + 1)), hi = dmax2(V(k), V(k + 1));
        double x = LV(k + 1);
        if (gk2 * gk > 0) {
            x = dmin2(x, hi);
            x = dmax2(x, lo);
        } else if (gk2 > 0) {
            x = dmax2(x, lo);
        } else {
            x = dmin2(x, hi);
        }
        LV(k + 1) = x;
    }
    for (int k = 1; k <= L; k++) {
        TOP(k) = LV(k + 1);
        BTM(k) = LV(k);
    }
    // Huynh constraint (340-352), Q8
    for (int k = 2; k <= L - 1; k++) {
        double vk1 = V(k + 1);
        double pmp1 = vk1 - 2. * G(k);
        double lac1 = pmp1 + 1.5 * G(k - 1);
        TOP(k + 1) = dmin2(dmax2(TOP(k + 1), dmin3(vk1, pmp1, lac1)), dmax3(vk1, pmp1, lac1));
        pmp1 = vk1 + 2. * G(k + 1);
        lac1 = pmp1 - 1.5 * G(k);
        BTM(k + 1) = dmin2(dmax2(BTM(k + 1), dmin3(vk1, pmp1, lac1)), dmax3(vk1, pmp1, lac1));
    }
    for (int k = 1; k <= L; k++) c.at(S.crv(), k) = 6. * V(k) - 3. * (TOP(k) + BTM(k));
#undef V
#undef X
#undef G
#undef LV
#undef TOP
#undef BTM
}

// remap_column (165-285): the source variable is in slot var, coordinates in
// slots src/dst; result written to out(m) = out[(m-1)*ostride].
__device__ void remap_column(const Col &c, const Slots &S, int L, double *out, size_t ostride) {
    vertical_ppm(c, S, L);
#define SC(k) c.at(S.src(), (k))
#define DC(k) c.at(S.dst(), (k))
    int k0 = 1;
    for (int m = 1; m <= L; m++) {
        const double dm0 = DC(m), dm1 = DC(m + 1);
        while (k0 <= L && SC(k0 + 1) >= dm0) k0++;
        double acc = 0.;
        for (int k = k0; k <= L; k++) {
            const double sk = SC(k), sk1 = SC(k + 1);
            if (sk <= dm1) break;
            const double dsrc = sk1 - sk;
            const double s1 = (dm0 - sk) / dsrc;
            const double s2 = (dm1 - sk) / dsrc;
            const double factor = dsrc / (dm1 - dm0);
            const double tmp3 = 2. * (s1 * s1);
            const double b = c.at(S.btm(), k), t = c.at(S.top(), k), cv = c.at(S.crv(), k);
            if (dm1 <= sk1 && sk1 < dm0 && dm0 < sk) {  // case 3
                double tmp1 = s1 - 1., tmp2 = s1 + 1.;
                acc = acc + tmp1 / 6. * (3. * (b * tmp1 - t * tmp2) + cv * (tmp3 - tmp2)) * factor;
            }
            if (sk1 < dm1 && dm1 < sk && sk <= dm0) {  // case 4
                acc = acc + (s2 / 2.) * (b * (2. - s2) + s2 * (t + cv * (1. - ATM_W2_3 * s2))) * factor;
            }
            if (dm1 <= sk1 && sk <= dm0) {  // case 5
                acc = acc + c.at(S.var(), k) * factor;
            }
            if (sk1 < dm1 && dm0 < sk) {  // case 6
                double tmp1 = s1 + s2;
                acc = acc + (s1 - s2) / 6. * (3. * (b * (tmp1 - 2.) - t * tmp1) + cv * (tmp3 + tmp1 * (2. * s2 - 3.))) * factor;
            }
        }
        out[(size_t)(m - 1) * ostride] = acc;
    }
#undef SC
#undef DC
}

// The winds remap reads ud/vd on the interior and on the ring around it.
// The ring holds the halo values the previous step's halo exchange left
// there (the D-grid step updates the interior only, Q5); with separate
// input buffers for interior and ring, `pick` returns the value the
// reference's single array holds at that point.
__device__ __forceinline__ double pick(const Geo &g, const double *__restrict__ in,
                                       const double *__restrict__ ring, int I, int J, int K) {
    const bool inside = I >= ATM_HALO && I < ATM_HALO + g.nx && J >= ATM_HALO && J < ATM_HALO + g.ny;
    return inside ? in[g.ix(I, J, K)] : ring[g.ix(I, J, K)];
}

// current and target grids and gz over isd-3..ied+3 (42-78)
__global__ void k_remap_grid(Geo g, double top, const double *__restrict__ ak, const double *__restrict__ bk,
                             const double *__restrict__ dp, const double *__restrict__ pt,
                             double *__restrict__ plev, double *__restrict__ pkap, double *__restrict__ playkap,
                             double *__restrict__ gz, double *__restrict__ tplev, double *__restrict__ tpkap,
                             double *__restrict__ tplk) {
    const int I = 4 + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = 4 + blockIdx.y * blockDim.y + threadIdx.y;
    if (I > g.NI - 5 || J > g.NJ - 5) return;
    const int L = g.L;
    double p = plev[g.ix(I, J, L)];
    double pk1 = pkap[g.ix(I, J, L)];
    for (int K = L - 1; K >= 0; K--) {
        p = p + dp[g.ix(I, J, K)];
        plev[g.ix(I, J, K)] = p;
        double pk = pow(p, ATM_KAPPA);
        pkap[g.ix(I, J, K)] = pk;
        double plog = log(pk1 / pk);
        playkap[g.ix(I, J, K)] = (pk1 - pk) / plog;
        pk1 = pk;
    }
    const double ps = p;
    double tk1 = tpkap[g.ix(I, J, L)];
    for (int K = L - 1; K >= 0; K--) {
        double tp = ak[K] + bk[K] * (ps - top);
        tplev[g.ix(I, J, K)] = tp;
        double tk = pow(tp, ATM_KAPPA);
        tpkap[g.ix(I, J, K)] = tk;
        double plog = log(tk1 / tk);
        tplk[g.ix(I, J, K)] = (tk1 - tk) / plog;
        tk1 = tk;
    }
    double z = gz[g.ix(I, J, 0)];
    double pk = pkap[g.ix(I, J, 0)];
    for (int K = 0; K < L; K++) {
        double pkn = pkap[g.ix(I, J, K + 1)];
        z = z + ATM_CP_DRY * pt[g.ix(I, J, K)] * (pk - pkn);
        gz[g.ix(I, J, K + 1)] = z;
        pk = pkn;
    }
}

// winds on faces over isd-1..ied+1 (80-105)
__global__ void k_remap_winds(Geo g, double *__restrict__ scr, const double *__restrict__ plev,
                              const double *__restrict__ tplev, const double *__restrict__ ud,
                              const double *__restrict__ vd, const double *__restrict__ udr,
                              const double *__restrict__ vdr, double *__restrict__ ur, double *__restrict__ vr) {
    const int I = 6 + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = 6 + blockIdx.y * blockDim.y + threadIdx.y;
    if (I > g.NI - 7 || J > g.NJ - 7) return;
    const int L = g.L;
    Slots S{L};
    Col c{scr, g.plane(), g.ix2(I, J)};
    for (int K = 0; K <= L; K++) {
        c.at(S.src(), K + 1) = ATM_W7_12 * (plev[g.ix(I, J, K)] + plev[g.ix(I, J - 1, K)]) -
                               ATM_W1_12 * (plev[g.ix(I, J + 1, K)] + plev[g.ix(I, J - 2, K)]);
        c.at(S.dst(), K + 1) = ATM_W7_12 * (tplev[g.ix(I, J, K)] + tplev[g.ix(I, J - 1, K)]) -
                               ATM_W1_12 * (tplev[g.ix(I, J + 1, K)] + tplev[g.ix(I, J - 2, K)]);
    }
    for (int K = 0; K < L; K++) c.at(S.var(), K + 1) = pick(g, ud, udr, I, J, K);
    remap_column(c, S, L, ur + g.ix(I, J, 0), g.plane());
    for (int K = 0; K <= L; K++) {
        c.at(S.src(), K + 1) = ATM_W7_12 * (plev[g.ix(I, J, K)] + plev[g.ix(I - 1, J, K)]) -
                               ATM_W1_12 * (plev[g.ix(I + 1, J, K)] + plev[g.ix(I - 2, J, K)]);
        c.at(S.dst(), K + 1) = ATM_W7_12 * (tplev[g.ix(I, J, K)] + tplev[g.ix(I - 1, J, K)]) -
                               ATM_W1_12 * (tplev[g.ix(I + 1, J, K)] + tplev[g.ix(I - 2, J, K)]);
    }
    for (int K = 0; K < L; K++) c.at(S.var(), K + 1) = pick(g, vd, vdr, I, J, K);
    remap_column(c, S, L, vr + g.ix(I, J, 0), g.plane());
}

// total energy remap and pt reconstruction on the interior (111-155)
__global__ void k_remap_energy(Geo g, double *__restrict__ scr, const double *__restrict__ plev,
                               const double *__restrict__ tplev, const double *__restrict__ tpkap,
                               const double *__restrict__ tplk, const double *__restrict__ playkap,
                               const double *__restrict__ gz, const double *__restrict__ pt,
                               const double *__restrict__ ud, const double *__restrict__ vd,
                               const double *__restrict__ udr, const double *__restrict__ vdr,
                               const double *__restrict__ ur, const double *__restrict__ vr,
                               double *__restrict__ ptr) {
    const int I = 7 + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = 7 + blockIdx.y * blockDim.y + threadIdx.y;
    if (I > g.NI - 8 || J > g.NJ - 8) return;
    const int L = g.L;
    Slots S{L};
    Col c{scr, g.plane(), g.ix2(I, J)};
#define UD(a, b) pick(g, ud, udr, I + (a), J + (b), K)
#define VD(a, b) pick(g, vd, vdr, I + (a), J + (b), K)
    for (int K = 0; K < L; K++) {
        double ua = (.25 * (UD(0, 0) + UD(0, 1)) + .125 * (UD(-1, 0) + UD(-1, 1) + UD(1, 0) + UD(1, 1)));
        double va = (.25 * (VD(0, 0) + VD(1, 0)) + .125 * (VD(0, -1) + VD(1, -1) + VD(0, 1) + VD(1, 1)));
        c.at(S.var(), K + 1) = ATM_CV_DRY * pt[g.ix(I, J, K)] * playkap[g.ix(I, J, K)] +
                               .5 * (gz[g.ix(I, J, K + 1)] + gz[g.ix(I, J, K)] + ua * ua + va * va);
    }
#undef UD
#undef VD
    for (int K = 0; K <= L; K++) {
        c.at(S.src(), K + 1) = plev[g.ix(I, J, K)];
        c.at(S.dst(), K + 1) = tplev[g.ix(I, J, K)];
    }
    // energy_remap(m) goes to ptr (overwritten by pt_remap below, level by level)
    remap_column(c, S, L, ptr + g.ix(I, J, 0), g.plane());
#define UR(a, b) ur[g.ix(I + (a), J + (b), K)]
#define VR(a, b) vr[g.ix(I + (a), J + (b), K)]
    double gzr = gz[g.ix(I, J, 0)];
    for (int K = 0; K < L; K++) {
        double ua = (.25 * (UR(0, 0) + UR(0, 1)) + .125 * (UR(-1, 0) + UR(-1, 1) + UR(1, 0) + UR(1, 1)));
        double va = (.25 * (VR(0, 0) + VR(1, 0)) + .125 * (VR(0, -1) + VR(1, -1) + VR(0, 1) + VR(1, 1)));
        double tk = tpkap[g.ix(I, J, K)], tk1 = tpkap[g.ix(I, J, K + 1)];
        double e = ptr[g.ix(I, J, K)];
        double p = (e - gzr - .5 * (ua * ua + va * va)) /
                   (ATM_CV_DRY * tplk[g.ix(I, J, K)] + .5 * ATM_CP_DRY * (tk - tk1));
        ptr[g.ix(I, J, K)] = p;
        gzr = gzr + ATM_CP_DRY * p * (tk - tk1);
    }
#undef UR
#undef VR
}

// write back on the interior (157-161)
__global__ void k_remap_write(Geo g, const double *__restrict__ tplev, const double *__restrict__ ur,
                              const double *__restrict__ vr, const double *__restrict__ ptr,
                              double *__restrict__ dp, double *__restrict__ ud, double *__restrict__ vd,
                              double *__restrict__ pt) {
    const int I = 7 + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = 7 + blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I > g.NI - 8 || J > g.NJ - 8) return;
    size_t x = g.ix(I, J, K);
    dp[x] = tplev[x] - tplev[g.ix(I, J, K + 1)];
    ud[x] = ur[x];
    vd[x] = vr[x];
    pt[x] = ptr[x];
}

__global__ void k_fill_plane(double *a, size_t n, double v) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

}  // namespace

int atm_remap_alloc(atm_ctx *c) {
    const Geo &g = c->g;
    if (!c->remap) {
        RemapWork *w = new RemapWork();
        c->remap = w;
        w->slots = Slots::count(g.L);
        size_t p = g.plane();
        ATM_HIP(c, hipMalloc(&w->scr, sizeof(double) * p * w->slots));
        ATM_HIP(c, hipMalloc(&w->tplev, sizeof(double) * p * (g.L + 1)));
        ATM_HIP(c, hipMalloc(&w->tpkap, sizeof(double) * p * (g.L + 1)));
        ATM_HIP(c, hipMalloc(&w->tplk, sizeof(double) * p * g.L));
        ATM_HIP(c, hipMalloc(&w->uremap, sizeof(double) * p * g.L));
        ATM_HIP(c, hipMalloc(&w->vremap, sizeof(double) * p * g.L));
        ATM_HIP(c, hipMalloc(&w->ptremap, sizeof(double) * p * g.L));
    }
    // target_plev(:,:,nlay+1) = top; target_pkap(:,:,nlay+1) = top**kappa (381-382)
    RemapWork *w = c->remap;
    size_t p = g.plane();
    hipLaunchKernelGGL(k_fill_plane, dim3((unsigned)((p + 255) / 256)), dim3(256), 0, c->stream,
                       w->tplev + p * g.L, p, c->ak0);
    hipLaunchKernelGGL(k_fill_plane, dim3((unsigned)((p + 255) / 256)), dim3(256), 0, c->stream,
                       w->tpkap + p * g.L, p, pow(c->ak0, ATM_KAPPA));
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}

void atm_remap_free(atm_ctx *c) {
    if (!c->remap) return;
    RemapWork *w = c->remap;
    (void)hipFree(w->scr);
    (void)hipFree(w->tplev);
    (void)hipFree(w->tpkap);
    (void)hipFree(w->tplk);
    (void)hipFree(w->uremap);
    (void)hipFree(w->vremap);
    (void)hipFree(w->ptremap);
    delete w;
    c->remap = nullptr;
}

int atm_remap_run_split(atm_ctx *c, const double *dp, const double *pt, const double *ud_in,
                        const double *vd_in, const double *ud_ring, const double *vd_ring, double *dp_out,
                        double *pt_out, double *ud_out, double *vd_out) {
    const Geo &g = c->g;
    RemapWork *w = c->remap;
    dim3 b(64, 2);
    atm_timer_begin(c, ATM_T_REMAP);
    hipLaunchKernelGGL(k_remap_grid, grid3(g.nx + 6, g.ny + 6, 1, b), b, 0, c->stream, g, c->ak0, c->ak, c->bk,
                       dp, pt, c->plev, c->pkap, c->playkap, c->gz, w->tplev, w->tpkap, w->tplk);
    hipLaunchKernelGGL(k_remap_winds, grid3(g.nx + 2, g.ny + 2, 1, b), b, 0, c->stream, g, w->scr, c->plev,
                       w->tplev, ud_in, vd_in, ud_ring, vd_ring, w->uremap, w->vremap);
    hipLaunchKernelGGL(k_remap_energy, grid3(g.nx, g.ny, 1, b), b, 0, c->stream, g, w->scr, c->plev, w->tplev,
                       w->tpkap, w->tplk, c->playkap, c->gz, pt, ud_in, vd_in, ud_ring, vd_ring, w->uremap,
                       w->vremap, w->ptremap);
    dim3 bw(64, 4);
    hipLaunchKernelGGL(k_remap_write, grid3(g.nx, g.ny, g.L, bw), bw, 0, c->stream, g, w->tplev, w->uremap,
                       w->vremap, w->ptremap, dp_out, ud_out, vd_out, pt_out);
    atm_timer_end(c, ATM_T_REMAP);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}

// in-place form (reference-structured path)
int atm_remap_run(atm_ctx *c) {
    return atm_remap_run_split(c, c->dp, c->pt, c->ud, c->vd, c->ud, c->vd, c->dp, c->pt, c->ud, c->vd);
}
