// k_rad.hip -- two-stream correlated-k radiation (gfx950).
//
// src/mod_radiation.f90:41-156 with the gas optics of
// src/mod_gas_optics.f90 and the solar source of src/mod_solar.f90.
// Layer/level profiles are built over the reference's ring (is+4..ie-4);
// columns are solved on the interior only: the ring results are always
// overwritten by the halo exchange that follows (mod_model.f90:122-124) and
// radius is undefined there (mod_input.f90:233-243).
//
// One thread per column, looping over the 16 g-points in the reference
// order so the flux sums accumulate in the same order.  Per-column work
// arrays live in a device scratch laid out [slot][column].
//
// Table lookups: the reference scans its tables linearly for the first
// node >= x (lookup_1d 4556-4579, lookup_pT_table 4581-4609).  Every table
// axis is strictly increasing, so a binary search returns the same node and
// the interpolation is then computed with the reference's formula.
#include <vector>

#include "atm_common.hpp"
#include "gas_optics_tables.inc"

struct Tables {
    const float *P, *T, *sw, *ray, *lw, *planckT, *planck, *frac;
};

struct RadWork {
    double *scr = nullptr;  // [slots][plane]
    double *layer_T = nullptr, *level_T = nullptr;
    float *tab = nullptr;
    Tables t;
    int slots = 0;
};

namespace {

// first index k in [0, n) with x <= xs[k]; caller guarantees xs[0] <= x <= xs[n-1]
__device__ __forceinline__ int first_ge(const float *xs, int n, double x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (x > (double)xs[mid]) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// lookup_1d (4556-4579); *err |= 1 outside the table
__device__ __forceinline__ double lookup_1d(double x, const float *xs, const float *ys, int n, int *err) {
    if (x < (double)xs[0] || x > (double)xs[n - 1]) { *err |= 1; return 0.; }
    int k = first_ge(xs, n, x);
    double xk = xs[k];
    if (x == xk) return (double)ys[k];
    double xm = xs[k - 1], yk = ys[k], ym = ys[k - 1];
    return (yk * (xm - x) + ym * (x - xk)) / (xm - xk);
}

// lookup_pT_table (4581-4609); table is the (6, 53) slab of one g-point
__device__ __forceinline__ double lookup_pT(const Tables &t, double p, double T, const float *table, int *err) {
    if (p < (double)t.P[0] || p > (double)t.P[52]) { *err |= 1; return 0.; }
    int k = first_ge(t.P, 53, p);
    if (p == (double)t.P[k]) return lookup_1d(T, t.T + 6 * k, table + 6 * k, 6, err);
    double t1 = lookup_1d(T, t.T + 6 * k, table + 6 * k, 6, err);
    double t2 = lookup_1d(T, t.T + 6 * (k - 1), table + 6 * (k - 1), 6, err);
    double pk = t.P[k], pm = t.P[k - 1];
    return (t1 * (pm - p) + t2 * (p - pk)) / (pm - pk);
}

struct RCol {
    double *s;
    size_t plane, col;
    __device__ double &at(int slot, int k) const { return s[(size_t)(slot + k) * plane + col]; }
};

// scratch slots, L+1 wide each
struct RSlots {
    int n;  // L+1
    __device__ int R() const { return 0; }
    __device__ int T() const { return n; }
    __device__ int Su() const { return 2 * n; }
    __device__ int Sd() const { return 3 * n; }
    __device__ int B() const { return 4 * n; }   // solar beam (SW) / level Planck (LW)
    __device__ int alb() const { return 5 * n; }
    __device__ int emi() const { return 6 * n; }
    __device__ int bet() const { return 7 * n; }
    __device__ int swu() const { return 8 * n; }
    __device__ int swd() const { return 9 * n; }
    __device__ int lwu() const { return 10 * n; }
    __device__ int lwd() const { return 11 * n; }
    static int count(int L) { return 12 * (L + 1); }
};

// calculate_fluxes (349-381); adds the result into (accU, accD) slots.
// For the shortwave the direct beam (slot B) is added to the downward flux.
__device__ void fluxes_accumulate(const RCol &c, const RSlots &S, int L, double toa, double sfc_up,
                                  double sfc_alb, int accU, int accD, bool add_beam) {
    c.at(S.alb(), 0) = sfc_alb;
    c.at(S.emi(), 0) = sfc_up;
    double a = sfc_alb, e = sfc_up;
    for (int k = 0; k < L; k++) {
        double R = c.at(S.R(), k), T = c.at(S.T(), k);
        double b = 1. / (1. - R * a);
        double an = R + (T * T) * b * a;
        double en = c.at(S.Su(), k) + T * b * (e + c.at(S.Sd(), k) * a);
        c.at(S.bet(), k) = b;
        c.at(S.alb(), k + 1) = an;
        c.at(S.emi(), k + 1) = en;
        a = an;
        e = en;
    }
    double fd = toa;
    double fu = a * toa + e;
    c.at(accU, L) = c.at(accU, L) + fu;
    c.at(accD, L) = c.at(accD, L) + (add_beam ? fd + c.at(S.B(), L) : fd);
    for (int k = L; k >= 1; k--) {
        double b = c.at(S.bet(), k - 1);
        double fdn = b * (c.at(S.Sd(), k - 1) + c.at(S.T(), k - 1) * fd + c.at(S.R(), k - 1) * c.at(S.emi(), k - 1));
        double fup = c.at(S.alb(), k - 1) * fdn + c.at(S.emi(), k - 1);
        c.at(accU, k - 1) = c.at(accU, k - 1) + fup;
        c.at(accD, k - 1) = c.at(accD, k - 1) + (add_beam ? fdn + c.at(S.B(), k - 1) : fdn);
        fd = fdn;
    }
}

// layer and level profiles over is+4..ie-4 (53-101)
__global__ void k_rad_profiles(Geo g, const double *__restrict__ dp, const double *__restrict__ pt,
                               double *__restrict__ plev, double *__restrict__ pkap, double *__restrict__ play,
                               double *__restrict__ playkap, double *__restrict__ lT, double *__restrict__ vT) {
    const int I = 4 + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = 4 + blockIdx.y * blockDim.y + threadIdx.y;
    if (I > g.NI - 5 || J > g.NJ - 5) return;
    const int L = g.L;
    double p1v = plev[g.ix(I, J, L)];
    double pk1 = pkap[g.ix(I, J, L)];
    double lay_next = 0., lT_next = 0.;
    for (int K = L - 1; K >= 0; K--) {
        double p = p1v + dp[g.ix(I, J, K)];
        plev[g.ix(I, J, K)] = p;
        double pk = pow(p, ATM_KAPPA);
        pkap[g.ix(I, J, K)] = pk;
        double plog = log(p / p1v);
        double pl = dp[g.ix(I, J, K)] / plog;
        play[g.ix(I, J, K)] = pl;
        double plk = (pk1 - pk) / (-plog * ATM_KAPPA);
        playkap[g.ix(I, J, K)] = plk;
        double t = pt[g.ix(I, J, K)] * plk;
        lT[g.ix(I, J, K)] = t;
        if (K < L - 1) {
            double a1 = log(pl), a2 = log(lay_next), pg = log(p1v);
            vT[g.ix(I, J, K + 1)] = (t * (a2 - pg) + lT_next * (pg - a1)) / (a2 - a1);
        }
        p1v = p;
        pk1 = pk;
        lay_next = pl;
        lT_next = t;
    }
    // top and bottom extrapolation (84-101)
    {
        double a1 = log(play[g.ix(I, J, L - 2)]), a2 = log(play[g.ix(I, J, L - 1)]);
        double pg = log(plev[g.ix(I, J, L)]);
        vT[g.ix(I, J, L)] = (lT[g.ix(I, J, L - 2)] * (a2 - pg) + lT[g.ix(I, J, L - 1)] * (pg - a1)) / (a2 - a1);
        a1 = log(play[g.ix(I, J, 0)]);
        a2 = log(play[g.ix(I, J, 1)]);
        pg = log(plev[g.ix(I, J, 0)]);
        vT[g.ix(I, J, 0)] = (lT[g.ix(I, J, 0)] * (a2 - pg) + lT[g.ix(I, J, 1)] * (pg - a1)) / (a2 - a1);
    }
}

// column solves (117-154)
__global__ void k_rad_columns(Geo g, Tables t, double gravity, double *__restrict__ scr,
                              const double *__restrict__ dp, const double *__restrict__ play,
                              const double *__restrict__ lT, const double *__restrict__ vT,
                              const double *__restrict__ ts, const double *__restrict__ radius,
                              double *__restrict__ net_flux, int *flag) {
    const int I = 7 + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = 7 + blockIdx.y * blockDim.y + threadIdx.y;
    if (I > g.NI - 8 || J > g.NJ - 8) return;
    const int L = g.L;
    RSlots S{L + 1};
    RCol c{scr, g.plane(), g.ix2(I, J)};
    int err = 0;
    for (int k = 0; k <= L; k++) {
        c.at(S.swu(), k) = 0.;
        c.at(S.swd(), k) = 0.;
        c.at(S.lwu(), k) = 0.;
        c.at(S.lwd(), k) = 0.;
    }
    // solar_irradiance (mod_solar.f90:26-35)
    double r = radius[g.ix2(I, J)] / (double)100000e3f;
    const double irr = 600. * (1. + (double)0.2f / (1. + r * r));
    const double mu0 = cos(0.);
    const double sfc_alb = (double)0.3f, emiss = (double)0.99f;
    const double tsur = ts[g.ix2(I, J)];
    for (int gp = 0; gp < 16; gp++) {
        // ---- shortwave (158-268) ----
        const float *swt = t.sw + 6 * 53 * gp;
        const double k_sca = (double)t.ray[gp];
        c.at(S.B(), L) = mu0 * ((double)t.frac[gp] * irr);
        for (int K = L - 1; K >= 0; K--) {
            double dm = dp[g.ix(I, J, K)] / gravity;
            double k_abs = lookup_pT(t, play[g.ix(I, J, K)], lT[g.ix(I, J, K)], swt, &err);
            double k_ext = k_abs + k_sca;
            double tau = k_ext * dm;
            double ssa = k_sca / k_ext;
            double asy = 0.;
            // delta_eddington_rescaling (383-398)
            {
                double f = asy * asy;
                double coef = 1. - ssa * f;
                tau = tau * coef;
                ssa = ssa * (1. - f) / coef;
                asy = asy / (1. + asy);
            }
            double asy2 = asy * asy;
            double back = (2. - 3. * asy * mu0) / 4.;
            double tmp1 = 4. * (1. - asy2 * (1. - mu0));
            double gamma3 = ssa * asy2;
            double gamma4 = 4. * back + 3. * asy;
            double tmp2 = 3. * asy;
            double gamma1 = (7. - 3. * asy2 - ssa * (4. + tmp2) + gamma3 * gamma4) / tmp1;
            double gamma2 = (-1. + asy2 + ssa * (4. - tmp2) + gamma3 * (gamma4 - 4.)) / tmp1;
            gamma3 = back;
            gamma4 = 1. - gamma3;
            double alpha1 = gamma1 * gamma4 + gamma2 * gamma3;
            double alpha2 = gamma2 * gamma3 + gamma2 * gamma4;
            double lambda = sqrt(gamma1 * gamma1 - gamma2 * gamma2);
            double lam_mu = lambda * mu0;
            double e_minus = exp(-lambda * tau);
            double e_2minus = e_minus * e_minus;
            double Tdir = exp(-tau / mu0);
            tmp1 = ssa / ((1. - lam_mu * lam_mu) * (lambda + gamma1 + (lambda - gamma1) * e_2minus));
            tmp2 = lambda * gamma3;
            double Rdif = tmp1 * ((1. - lam_mu) * (alpha2 + tmp2) - (1. + lam_mu) * (alpha2 - tmp2) * e_2minus -
                                  2. * lambda * (gamma3 - alpha2 * mu0) * e_minus * Tdir);
            tmp2 = lambda * gamma4;
            double Tdif = tmp1 * (2. * lambda * (gamma4 + alpha1 * mu0) * e_minus -
                                  Tdir * ((1. + lam_mu) * (alpha1 + tmp2) - (1. - lam_mu) * (alpha1 - tmp2) * e_2minus));
            double bk1 = c.at(S.B(), K + 1);
            c.at(S.B(), K) = Tdir * bk1;
            c.at(S.Su(), K) = Rdif * bk1;
            c.at(S.Sd(), K) = Tdif * bk1;
            back = (4. + asy) / (8. * (1. + asy));
            gamma1 = 2. * (1. - ssa * (1. - back));
            gamma2 = 2. * ssa * back;
            lambda = sqrt(gamma1 * gamma1 - gamma2 * gamma2);
            e_2minus = e_minus * e_minus;  // Q6: e_minus of the direct-beam lambda
            tmp1 = 1. / (lambda + gamma1 + (lambda - gamma1) * e_2minus);
            c.at(S.R(), K) = tmp1 * gamma2 * (1. - e_2minus);
            c.at(S.T(), K) = 2. * tmp1 * lambda * e_minus;
        }
        fluxes_accumulate(c, S, L, c.at(S.B(), L), 0., sfc_alb, S.swu(), S.swd(), true);
        // ---- longwave (270-347) ----
        const float *lwt = t.lw + 6 * 53 * gp;
        const float *pl = t.planck + 231 * gp;
        const double Bs = lookup_1d(tsur, t.planckT, pl, 231, &err);
        double Bk = lookup_1d(vT[g.ix(I, J, 0)], t.planckT, pl, 231, &err);
        for (int K = 0; K < L; K++) {
            double dm = dp[g.ix(I, J, K)] / gravity;
            double k_abs = lookup_pT(t, play[g.ix(I, J, K)], lT[g.ix(I, J, K)], lwt, &err);
            double k_sca = 0.;
            double k_ext = k_abs + k_sca;
            double tau = k_ext * dm;
            double ssa = k_sca / k_ext;
            double asy = 0.;
            {
                double f = asy * asy;
                double coef = 1. - ssa * f;
                tau = tau * coef;
                ssa = ssa * (1. - f) / coef;
                asy = asy / (1. + asy);
            }
            double back = (4. - 3. * asy) / 8.;
            double gamma1 = 2. * (1. - ssa * (1. - back));
            double gamma2 = 2. * ssa * back;
            double lambda = sqrt(gamma1 * gamma1 - gamma2 * gamma2);
            double e_minus = exp(-lambda * tau);
            double e_2minus = e_minus * e_minus;
            double tmp1 = 1. / (lambda + gamma1 + (lambda - gamma1) * e_2minus);
            double R = tmp1 * gamma2 * (1. - e_2minus);
            double T = 2. * tmp1 * lambda * e_minus;
            c.at(S.R(), K) = R;
            c.at(S.T(), K) = T;
            double Bk1 = lookup_1d(vT[g.ix(I, J, K + 1)], t.planckT, pl, 231, &err);
            double pdiff = (1. + R - T) * (Bk - Bk1) / (tau * (gamma1 + gamma2));
            tmp1 = 1. - R;
            c.at(S.Su(), K) = ATM_PI * (tmp1 * Bk1 - T * Bk + pdiff);
            c.at(S.Sd(), K) = ATM_PI * (tmp1 * Bk - T * Bk1 - pdiff);
            Bk = Bk1;
        }
        fluxes_accumulate(c, S, L, 0., emiss * ATM_PI * Bs, 1. - emiss, S.lwu(), S.lwd(), false);
    }
    for (int k = 0; k <= L; k++)
        net_flux[g.ix(I, J, k)] = (c.at(S.swd(), k) + c.at(S.lwd(), k)) - (c.at(S.swu(), k) + c.at(S.lwu(), k));
    if (err) atomicOr(flag, 2);
}

}  // namespace

int atm_rad_alloc(atm_ctx *c) {
    const Geo &g = c->g;
    if (c->rad) return ATM_OK;
    RadWork *w = new RadWork();
    c->rad = w;
    w->slots = RSlots::count(g.L);
    size_t p = g.plane();
    ATM_HIP(c, hipMalloc(&w->scr, sizeof(double) * p * w->slots));
    ATM_HIP(c, hipMalloc(&w->layer_T, sizeof(double) * p * g.L));
    ATM_HIP(c, hipMalloc(&w->level_T, sizeof(double) * p * (g.L + 1)));
    // gas-optics tables (binary32 values of the reference literals)
    std::vector<float> h;
    size_t off[8];
    auto add = [&](const float *a, size_t n, int i) { off[i] = h.size(); h.insert(h.end(), a, a + n); };
    add(GO_PRESSURES, GO_PRESSURES_N, 0);
    add(GO_TEMPERATURES, GO_TEMPERATURES_N, 1);
    add(GO_SW_ABS, GO_SW_ABS_N, 2);
    add(GO_RAYLEIGH, GO_RAYLEIGH_N, 3);
    add(GO_LW_ABS, GO_LW_ABS_N, 4);
    add(GO_PLANCK_T, GO_PLANCK_T_N, 5);
    add(GO_PLANCK, GO_PLANCK_N, 6);
    add(GO_SOLAR_FRACTION, GO_SOLAR_FRACTION_N, 7);
    ATM_HIP(c, hipMalloc(&w->tab, sizeof(float) * h.size()));
    ATM_HIP(c, hipMemcpy(w->tab, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    w->t = Tables{w->tab + off[0], w->tab + off[1], w->tab + off[2], w->tab + off[3],
                  w->tab + off[4], w->tab + off[5], w->tab + off[6], w->tab + off[7]};
    return ATM_OK;
}

void atm_rad_free(atm_ctx *c) {
    if (!c->rad) return;
    (void)hipFree(c->rad->scr);
    (void)hipFree(c->rad->layer_T);
    (void)hipFree(c->rad->level_T);
    (void)hipFree(c->rad->tab);
    delete c->rad;
    c->rad = nullptr;
}

// calculate_radiative_heating (41-156); the caller checks the flag
int atm_rad_run(atm_ctx *c, double time) {
    (void)time;  // unused by solar_irradiance (mod_solar.f90:26-35)
    const Geo &g = c->g;
    RadWork *w = c->rad;
    dim3 b(64, 2);
    atm_timer_begin(c, ATM_T_RAD);
    hipLaunchKernelGGL(k_rad_profiles, grid3(g.nx + 6, g.ny + 6, 1, b), b, 0, c->stream, g, c->dp, c->pt,
                       c->plev, c->pkap, c->play, c->playkap, w->layer_T, w->level_T);
    ATM_HIP(c, hipMemsetAsync(c->net_flux, 0, sizeof(double) * g.plane() * (g.L + 1), c->stream));
    hipLaunchKernelGGL(k_rad_columns, grid3(g.nx, g.ny, 1, b), b, 0, c->stream, g, w->t, c->prm.gravity, w->scr,
                       c->dp, c->play, w->layer_T, w->level_T, c->ts, c->radius, c->net_flux, c->flag);
    atm_timer_end(c, ATM_T_RAD);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}
