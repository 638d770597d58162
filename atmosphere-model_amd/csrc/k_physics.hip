// k_physics.hip -- mod_physics on the device (gfx950): dry convective
// adjustment, surface heating and the application of the tendencies
// (src/mod_physics.f90:27-214).
//
// calculate_physics_tendencies is column-local: one thread per column runs
// the interface pressure / pkap scan down the column (49-55), geopotential
// and playkap up the column (57-70, Q17's physics form), and on the ring
// isd-2..ied+2 the mixing sweep of dry_convective_adjustment (80-150) as a
// single ascending pass: the sweep at interface k changes only layers k
// and k+1, so layer k is final once interface k is done and its wind and
// potential-temperature tendencies (with the gz_adj reconstruction chain)
// are emitted right there, from registers.
#include "atm_common.hpp"

struct PhysWork {
    double *u_tend = nullptr, *v_tend = nullptr, *pt_tend = nullptr;  // NI x NJ x L, valid on the ring
    double *ts_tend = nullptr;                                        // NI x NJ, valid on the interior
};

namespace {

__global__ void __launch_bounds__(64) k_phys_tend(Geo g, const double *__restrict__ dp,
                                                  const double *__restrict__ pt, const double *__restrict__ ud,
                                                  const double *__restrict__ vd, const double *__restrict__ nf,
                                                  double *__restrict__ plev, double *__restrict__ pkap,
                                                  double *__restrict__ gz, double *__restrict__ playkap,
                                                  double *__restrict__ u_tend, double *__restrict__ v_tend,
                                                  double *__restrict__ pt_tend, double *__restrict__ ts_tend) {
    // columns is+1..ie-1 x js+1..je-1
    const int I = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = 1 + blockIdx.y;
    if (I > g.NI - 2) return;
    const int L = g.L;
    const size_t pl = g.plane(), c0 = g.ix2(I, J);
    // interface pressure and pkap from the top (49-55)
    double p = plev[c0 + (size_t)L * pl];
    for (int K = L - 1; K >= 0; K--) {
        p = p + dp[c0 + (size_t)K * pl];
        plev[c0 + (size_t)K * pl] = p;
        pkap[c0 + (size_t)K * pl] = pow_kappa(p);
    }
    // geopotential and playkap from the surface (57-70)
    {
        double z = gz[c0], pk0 = pkap[c0], pl0 = plev[c0];
        for (int K = 0; K < L; K++) {
            const size_t x = c0 + (size_t)K * pl;
            const double pk1 = pkap[x + pl], pl1 = plev[x + pl];
            z = z + ATM_CP_DRY * pt[x] * (pk0 - pk1);
            gz[x + pl] = z;
            playkap[x] = (pk1 - pk0) / log(pl1 / pl0) / ATM_KAPPA;
            pk0 = pk1;
            pl0 = pl1;
        }
    }
    // surface heating on the interior (152-160): 5 m of water, 41680*5 J/m2/K
    const bool interior = I >= ATM_HALO && I < ATM_HALO + g.nx && J >= ATM_HALO && J < ATM_HALO + g.ny;
    if (interior) ts_tend[c0] = nf[c0] / (double)(41680 * 5);
    // dry convective adjustment on the ring isd-2..ied+2 (80-150)
    if (I < ATM_HALO - 2 || I > ATM_HALO + g.nx + 1 || J < ATM_HALO - 2 || J > ATM_HALO + g.ny + 1) return;
    const double min_v2 = (double)1e-4f;  // default-real literal
    const double crit = .25, timescale = 600;
    const size_t sJ = g.SI;
    // A-grid winds of the column (35-45)
    auto ua_at = [&](size_t x) {
        return (.25 * (ud[x] + ud[x + sJ]) + .125 * (ud[x - 1] + ud[x - 1 + sJ] + ud[x + 1] + ud[x + 1 + sJ]));
    };
    auto va_at = [&](size_t x) {
        return (.25 * (vd[x] + vd[x + 1]) + .125 * (vd[x - sJ] + vd[x + 1 - sJ] + vd[x + sJ] + vd[x + 1 + sJ]));
    };
    // layer k (0-based K) of the sweep: original values and adjusted ones
    size_t x0 = c0;
    double dp0 = dp[x0], pt0 = pt[x0], ua0 = ua_at(x0), va0 = va_at(x0);
    double gz0 = gz[x0], gz1 = gz[x0 + pl];
    double plk0 = playkap[x0];
    double uA = ua0, vA = va0;
    double eA = ATM_CV_DRY * pt0 * plk0 + .5 * (gz1 + gz0 + ua0 * ua0 + va0 * va0);
    double gz_adj = gz0;  // gz_adj(1) = gz(1)
    auto emit = [&](size_t x, double ua, double va, double ptv, double plk, double u, double v, double e) {
        u_tend[x] = 0. + (u - ua) / timescale;
        v_tend[x] = 0. + (v - va) / timescale;
        const double dpk = pkap[x] - pkap[x + pl];
        const double pta = (e - gz_adj - .5 * (u * u + v * v)) / (ATM_CV_DRY * plk + .5 * ATM_CP_DRY * dpk);
        gz_adj = gz_adj + ATM_CP_DRY * pta * dpk;
        pt_tend[x] = 0. + (pta - ptv) / timescale;
    };
    for (int K = 0; K < L - 1; K++) {
        const size_t x1 = x0 + pl;
        const double dp1 = dp[x1], pt1 = pt[x1], ua1 = ua_at(x1), va1 = va_at(x1);
        const double gz2 = gz[x1 + pl], plk1 = playkap[x1];
        double uB = ua1, vB = va1;
        double eB = ATM_CV_DRY * pt1 * plk1 + .5 * (gz2 + gz1 + ua1 * ua1 + va1 * va1);
        const double du = ua1 - ua0, dv = va1 - va0;
        const double ri = (pt1 - pt0) * (dp0 * (gz1 - gz0) + dp1 * (gz2 - gz1)) /
                          ((dp0 * pt0 + dp1 * pt1) * (du * du + dv * dv + min_v2));
        if (ri < crit) {
            const double m = 1. - dmax2(0.0, ri / crit);
            const double mf = m * m * dp0 * dp1 / (dp0 + dp1);
            double ex = mf * (uA - uB);
            uA = uA - ex / dp0;
            uB = uB + ex / dp1;
            ex = mf * (vA - vB);
            vA = vA - ex / dp0;
            vB = vB + ex / dp1;
            ex = mf * (eA - eB);
            eA = eA - ex / dp0;
            eB = eB + ex / dp1;
        }
        emit(x0, ua0, va0, pt0, plk0, uA, vA, eA);  // layer K is final
        x0 = x1;
        dp0 = dp1;
        pt0 = pt1;
        ua0 = ua1;
        va0 = va1;
        gz0 = gz1;
        gz1 = gz2;
        plk0 = plk1;
        uA = uB;
        vA = vB;
        eA = eB;
    }
    emit(x0, ua0, va0, pt0, plk0, uA, vA, eA);
}

// apply_physics_tendencies on the interior (162-185).  dp_tend is zero
// everywhere (28), so dp + dt * dp_tend is dp itself and is not formed.
__global__ void k_phys_apply(Geo g, double dt, const double *__restrict__ u_tend,
                             const double *__restrict__ v_tend, const double *__restrict__ pt_tend,
                             const double *__restrict__ ts_tend, double *__restrict__ pt, double *__restrict__ ud,
                             double *__restrict__ vd, double *__restrict__ ts, int *nfw) {
    const int I = ATM_HALO + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = ATM_HALO + blockIdx.y;
    const int K = blockIdx.z;
    if (I >= ATM_HALO + g.nx) return;
    const size_t x = g.ix(I, J, K), sJ = g.SI;
    const double p = pt[x] + dt * pt_tend[x];
    const double u = ud[x] + dt * (ATM_W7_12 * (u_tend[x] + u_tend[x - sJ]) - ATM_W1_12 * (u_tend[x + sJ] + u_tend[x - 2 * sJ]));
    const double v = vd[x] + dt * (ATM_W7_12 * (v_tend[x] + v_tend[x - 1]) - ATM_W1_12 * (v_tend[x + 1] + v_tend[x - 2]));
    pt[x] = p;
    ud[x] = u;
    vd[x] = v;
    atm_report_nonfinite(nfw, p);
    atm_report_nonfinite(nfw, u);
    atm_report_nonfinite(nfw, v);
    if (K == 0) {
        const size_t c = g.ix2(I, J);
        ts[c] = ts[c] + dt * ts_tend[c];
    }
}

}  // namespace

// Frees what a PhysWork holds; also for one whose allocation stopped part way.
static void phys_release(PhysWork *w) {
    (void)hipFree(w->u_tend);
    (void)hipFree(w->v_tend);
    (void)hipFree(w->pt_tend);
    (void)hipFree(w->ts_tend);
    delete w;
}

// The context's PhysWork is published only once every buffer exists and is
// cleared, so a failed initialisation leaves physics uninitialised and a
// retry allocates afresh.
int atm_phys_alloc(atm_ctx *c) {
    if (c->phys) return ATM_OK;
    const Geo &g = c->g;
    PhysWork *w = new PhysWork();
    const size_t p = g.plane();
    double **bufs[] = {&w->u_tend, &w->v_tend, &w->pt_tend, &w->ts_tend};
    const size_t n[] = {p * g.L, p * g.L, p * g.L, p};
    for (int i = 0; i < 4; i++) {
        hipError_t e = hipMalloc(bufs[i], sizeof(double) * n[i]);
        if (e == hipSuccess) e = hipMemsetAsync(*bufs[i], 0, sizeof(double) * n[i], c->stream);
        if (e != hipSuccess) {
            phys_release(w);
            return atm_fail(c, ATM_ERR_HIP, std::string("physics buffers: ") + hipGetErrorString(e));
        }
    }
    c->phys = w;
    return ATM_OK;
}

void atm_phys_free(atm_ctx *c) {
    if (!c->phys) return;
    phys_release(c->phys);
    c->phys = nullptr;
}

int atm_phys_tendencies(atm_ctx *c) {
    const Geo &g = c->g;
    PhysWork *w = c->phys;
    atm_timer_begin(c, ATM_T_PHYS);
    hipLaunchKernelGGL(k_phys_tend, dim3((g.NI - 2 + 63) / 64, g.NJ - 2), dim3(64), 0, c->stream, g, c->dp, c->pt,
                       c->ud, c->vd, c->net_flux, c->plev, c->pkap, c->gz, c->playkap, w->u_tend, w->v_tend,
                       w->pt_tend, w->ts_tend);
    atm_timer_end(c, ATM_T_PHYS);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}

int atm_phys_apply(atm_ctx *c, double dt) {
    const Geo &g = c->g;
    PhysWork *w = c->phys;
    atm_timer_begin(c, ATM_T_PHYS);
    hipLaunchKernelGGL(k_phys_apply, dim3((g.nx + 63) / 64, g.ny, g.L), dim3(64), 0, c->stream, g, dt, w->u_tend,
                       w->v_tend, w->pt_tend, w->ts_tend, c->pt, c->ud, c->vd, c->ts, c->nfw);
    atm_timer_end(c, ATM_T_PHYS);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}
