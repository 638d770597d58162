// k_remap.hip -- vertical Lagrangian -> hybrid sigma-p remap (gfx950).
//
// src/mod_remap.f90:27-163.  One thread per column (the u and v columns of
// a point on separate threads), columns flattened so every wavefront is
// full; the per-level energy inputs come from a kernel with one thread per
// cell.  Each remap_column (165-285) runs in two passes:
//
//   forward   the elimination sweep of vertical_ppm (287-356) from k = nlay
//             down to 2, the only part that runs against the final order;
//             it leaves the eliminated edge values and gam in a scratch
//             laid out [slot][column] (two slots per level, coalesced);
//   merge     one ascending pass that back-substitutes, clamps (316-319),
//             applies the monotonicity test (325-334) and the Huynh
//             constraint (340-352), forms each source layer's PPM
//             coefficients in registers as the layer comes up, and walks
//             the source/destination overlaps (remap_column 193-283).
//
// For monotone coordinates the (source k, destination m) pairs that pass
// the overlap tests are, for each m, a contiguous run of k that only moves
// upward as m grows; the merge visits exactly those pairs, for each m in
// ascending k, with the reference's formulas, so every destination value
// is accumulated from the same terms in the same order.  Layers are
// produced in ascending order only, so the per-layer limiter state is a
// window of a few registers (gam(k-2..k+2), var(k..k+1), edges(k..k+1)).
#include "atm_common.hpp"
#include "atm_fused.hpp"
#include "remap_column.hpp"

struct RemapWork {
    double *scr = nullptr;      // [4 * (L+2)][ncol]: four column slots (see Scr)
    double *tplev = nullptr;    // target_plev  (NI x NJ x L+1), used on isd-3..ied+3
    double *tpkap = nullptr;    // target_pkap
    double *tplk = nullptr;     // target_playkap (L)
    double *uremap = nullptr, *vremap = nullptr;  // NI x NJ x L
    size_t ncol = 0;            // columns of the widest remap domain ((nx+2) x (ny+2))
};

namespace {

// Column scratch: slot s, 1-based level k -> scr[(s*(L+2) + k) * ncol + t].
// The wind remap uses slots 0-1 (u: eliminated edges, gam) and 2-3 (v); the
// energy remap slots 0-1 (edges, gam), 2 (energy) and 3 (kinetic energy of
// the remapped winds).
struct Scr {
    double *s;
    size_t ncol, t;
    int L;
    __device__ double *slot(int q) const { return s + (size_t)q * (L + 2) * ncol + t; }
};

// Column sources.  X(k): source face coordinate, D(k): destination face
// coordinate (k = 1..L+1), V(k): layer value (k = 1..L).
//
// Winds (80-105): face coordinates are the 4-point averages of plev /
// target_plev along J (u) or I (v); the layer values are the D-grid winds.
struct WindCol {
    const double *plev, *tplev, *v;
    size_t base, s, plane;
    struct XR {  // the four values a face coordinate averages
        double a0, am, ap, am2;
    };
    __device__ XR raw(const double *a, int k) const {
        const size_t o = base + (size_t)(k - 1) * plane;
        return {a[o], a[o - s], a[o + s], a[o - 2 * s]};
    }
    __device__ double xv(const XR &r) const { return ATM_W7_12 * (r.a0 + r.am) - ATM_W1_12 * (r.ap + r.am2); }
    __device__ XR xr(int k) const { return raw(plev, k); }
    __device__ XR dr(int k) const { return raw(tplev, k); }
    __device__ double X(int k) const { return xv(xr(k)); }
    __device__ double D(int k) const { return xv(dr(k)); }
    __device__ double V(int k) const { return v[base + (size_t)(k - 1) * plane]; }
};

// Total energy (111-150): coordinates are plev / target_plev of the column,
// the layer values are those k_remap_energy_in left in the scratch's slot 2.
struct EnergyCol {
    const double *plev, *tplev, *var;
    size_t base, plane, ncol;
    using XR = double;
    __device__ double X(int k) const { return plev[base + (size_t)(k - 1) * plane]; }
    __device__ double D(int k) const { return tplev[base + (size_t)(k - 1) * plane]; }
    __device__ double V(int k) const { return var[(size_t)k * ncol]; }
    __device__ XR xr(int k) const { return X(k); }
    __device__ XR dr(int k) const { return D(k); }
    __device__ double xv(XR r) const { return r; }
};

// Logical block of a 1-D launch, consecutive logical blocks kept on one XCD
// (hardware block b runs on XCD b % 8) so that neighbouring rows of columns
// share that XCD's L2.  The grid is padded to a multiple of 8 blocks.
__device__ __forceinline__ unsigned xcd_block1d() {
    const unsigned n = gridDim.x, b = blockIdx.x;
    return (b & 7u) * (n >> 3) + (b >> 3);
}

// current and target grids and gz over isd-3..ied+3 (42-78).  Levels go in
// chunks of 4 whose pow / log evaluations are independent.
// The kernels' bodies take the thread's column (or cell) from the caller:
// each member of a batched ensemble launch (atm_ens_launch_remap) runs
// the body its own launch runs.
__device__ __forceinline__ void remap_grid_body(int I, int J, Geo g, double top, const double *__restrict__ ak,
                                                const double *__restrict__ bk, const double *__restrict__ dp,
                                                const double *__restrict__ pt, double *__restrict__ plev,
                                                double *__restrict__ pkap, double *__restrict__ playkap,
                                                double *__restrict__ gz, double *__restrict__ tplev,
                                                double *__restrict__ tpkap, double *__restrict__ tplk) {
    constexpr int CH = 4;
    if (I > g.NI - 5 || J > g.NJ - 5) return;
    const int L = g.L;
    const size_t pl = g.plane(), c0 = g.ix2(I, J);
    double p = plev[c0 + (size_t)L * pl];
    double pk1 = pkap[c0 + (size_t)L * pl];
    // plev, pkap, playkap of the current grid from the top down (42-58)
    {
        auto level = [&](int K, double pv, double pk) {
            const size_t x = c0 + (size_t)K * pl;
            plev[x] = pv;
            pkap[x] = pk;
            double plog = log(pk1 / pk);
            playkap[x] = (pk1 - pk) / plog;
            pk1 = pk;
        };
        int K = L - 1;
        for (; K >= CH - 1; K -= CH) {
            double pv[CH], pk[CH];
#pragma unroll
            for (int u = 0; u < CH; u++) {
                p = p + dp[c0 + (size_t)(K - u) * pl];
                pv[u] = p;
            }
#pragma unroll
            for (int u = 0; u < CH; u++) pk[u] = pow_kappa(pv[u]);
#pragma unroll
            for (int u = 0; u < CH; u++) level(K - u, pv[u], pk[u]);
        }
        for (; K >= 0; K--) {
            p = p + dp[c0 + (size_t)K * pl];
            level(K, p, pow_kappa(p));
        }
    }
    // target grid from the surface pressure (60-70)
    const double ps = p;
    double tk1 = tpkap[c0 + (size_t)L * pl];
    {
        auto level = [&](int K, double tp, double tk) {
            const size_t x = c0 + (size_t)K * pl;
            tplev[x] = tp;
            tpkap[x] = tk;
            double plog = log(tk1 / tk);
            tplk[x] = (tk1 - tk) / plog;
            tk1 = tk;
        };
        int K = L - 1;
        for (; K >= CH - 1; K -= CH) {
            double tp[CH], tk[CH];
#pragma unroll
            for (int u = 0; u < CH; u++) tp[u] = ak[K - u] + bk[K - u] * (ps - top);
#pragma unroll
            for (int u = 0; u < CH; u++) tk[u] = pow_kappa(tp[u]);
#pragma unroll
            for (int u = 0; u < CH; u++) level(K - u, tp[u], tk[u]);
        }
        for (; K >= 0; K--) {
            const double tp = ak[K] + bk[K] * (ps - top);
            level(K, tp, pow_kappa(tp));
        }
    }
    // geopotential (72-78)
    double z = gz[c0];
    double pk = pkap[c0];
    for (int K = 0; K < L; K++) {
        double pkn = pkap[c0 + (size_t)(K + 1) * pl];
        z = z + ATM_CP_DRY * pt[c0 + (size_t)K * pl] * (pk - pkn);
        gz[c0 + (size_t)(K + 1) * pl] = z;
        pk = pkn;
    }
}

__global__ void k_remap_grid(Geo g, double top, const double *__restrict__ ak, const double *__restrict__ bk,
                             const double *__restrict__ dp, const double *__restrict__ pt,
                             double *__restrict__ plev, double *__restrict__ pkap, double *__restrict__ playkap,
                             double *__restrict__ gz, double *__restrict__ tplev, double *__restrict__ tpkap,
                             double *__restrict__ tplk) {
    remap_grid_body(4 + blockIdx.x * blockDim.x + threadIdx.x, 4 + blockIdx.y * blockDim.y + threadIdx.y, g, top, ak,
                    bk, dp, pt, plev, pkap, playkap, gz, tplev, tpkap, tplk);
}

// winds on faces over isd-1..ied+1 (80-105): u for blockIdx.y = 0, v for
// blockIdx.y = 1, so the two columns of a point run on separate threads.
// The ring of the D-grid winds holds the halo values the previous step's
// halo exchange left there (the D-grid step updates the interior only): a
// column outside the interior reads its winds from `ring`, an interior
// column from `in`.
__device__ __forceinline__ void remap_winds_body(size_t t, bool is_v, Geo g, double *__restrict__ scr, size_t ncol,
                                                 const double *__restrict__ plev, const double *__restrict__ tplev,
                                                 const double *__restrict__ ud, const double *__restrict__ vd,
                                                 const double *__restrict__ udr, const double *__restrict__ vdr,
                                                 double *__restrict__ ur, double *__restrict__ vr) {
    const int W = g.nx + 2;
    if (t >= ncol) return;
    const int I = 6 + (int)(t % W), J = 6 + (int)(t / W);
    const int L = g.L;
    const bool inside = I >= ATM_HALO && I < ATM_HALO + g.nx && J >= ATM_HALO && J < ATM_HALO + g.ny;
    const size_t base = g.ix2(I, J), plane = g.plane();
    const Scr s{scr, ncol, t, L};
    // u: faces averaged along J (stride SI); v: along I (stride 1)
    const WindCol c{plev, tplev, is_v ? (inside ? vd : vdr) : (inside ? ud : udr), base,
                    is_v ? (size_t)1 : (size_t)g.SI, plane};
    double *lv = s.slot(is_v ? 2 : 0), *gm = s.slot(is_v ? 3 : 1);
    double *o = is_v ? vr : ur;
    ppm_forward(c, L, lv, gm, ncol);
    auto out = [&](int m, double acc, double, double) { o[base + (size_t)(m - 1) * plane] = acc; };
    ppm_merge(c, L, lv, gm, ncol, out);
}

__global__ void __launch_bounds__(256) k_remap_winds(Geo g, double *__restrict__ scr, size_t ncol,
                                                     const double *__restrict__ plev,
                                                     const double *__restrict__ tplev,
                                                     const double *__restrict__ ud, const double *__restrict__ vd,
                                                     const double *__restrict__ udr,
                                                     const double *__restrict__ vdr, double *__restrict__ ur,
                                                     double *__restrict__ vr) {
    remap_winds_body((size_t)xcd_block1d() * blockDim.x + threadIdx.x, blockIdx.y == 1, g, scr, ncol, plev, tplev, ud,
                     vd, udr, vdr, ur, vr);
}

// The per-level inputs of the energy remap, over the interior columns and
// all levels at once (one thread per cell; column t = (J-7) nx + (I-7)):
// the total energy of each layer (111-128) into slot 2 and the kinetic
// energy of the remapped winds (the term of 137-142) into slot 3, each the
// reference's expression, so the column pass reads one value per level for
// each.  The winds of a point outside the interior come from the ring.
__device__ __forceinline__ void remap_energy_in_body(int i, int j, int K, Geo g, double *__restrict__ scr,
                                                    size_t ncol, const double *__restrict__ playkap,
                                                    const double *__restrict__ gz, const double *__restrict__ pt,
                                                    const double *__restrict__ ud, const double *__restrict__ vd,
                                                    const double *__restrict__ udr, const double *__restrict__ vdr,
                                                    const double *__restrict__ ur, const double *__restrict__ vr) {
    if (i >= g.nx || j >= g.ny) return;
    const int I = ATM_HALO + i, J = ATM_HALO + j, L = g.L;
    const size_t t = (size_t)j * g.nx + i, plane = g.plane(), x = g.ix(I, J, K);
    const Scr s{scr, ncol, t, L};
    auto w = [&](const double *in, const double *ring, int a, int b) {
        const int ii = I + a, jj = J + b;
        const bool inside = ii >= ATM_HALO && ii < ATM_HALO + g.nx && jj >= ATM_HALO && jj < ATM_HALO + g.ny;
        return (inside ? in : ring)[g.ix(ii, jj, K)];
    };
    {
        const double ua = (.25 * (w(ud, udr, 0, 0) + w(ud, udr, 0, 1)) +
                           .125 * (w(ud, udr, -1, 0) + w(ud, udr, -1, 1) + w(ud, udr, 1, 0) + w(ud, udr, 1, 1)));
        const double va = (.25 * (w(vd, vdr, 0, 0) + w(vd, vdr, 1, 0)) +
                           .125 * (w(vd, vdr, 0, -1) + w(vd, vdr, 1, -1) + w(vd, vdr, 0, 1) + w(vd, vdr, 1, 1)));
        s.slot(2)[(size_t)(K + 1) * ncol] =
            ATM_CV_DRY * pt[x] * playkap[x] + .5 * (gz[x + plane] + gz[x] + ua * ua + va * va);
    }
    {
        const size_t sN = g.SI;  // +1 in J
        const double ua =
            (.25 * (ur[x] + ur[x + sN]) + .125 * (ur[x - 1] + ur[x - 1 + sN] + ur[x + 1] + ur[x + 1 + sN]));
        const double va =
            (.25 * (vr[x] + vr[x + 1]) + .125 * (vr[x - sN] + vr[x + 1 - sN] + vr[x + sN] + vr[x + 1 + sN]));
        s.slot(3)[(size_t)(K + 1) * ncol] = .5 * (ua * ua + va * va);
    }
}

__global__ void k_remap_energy_in(Geo g, double *__restrict__ scr, size_t ncol, const double *__restrict__ playkap,
                                  const double *__restrict__ gz, const double *__restrict__ pt,
                                  const double *__restrict__ ud, const double *__restrict__ vd,
                                  const double *__restrict__ udr, const double *__restrict__ vdr,
                                  const double *__restrict__ ur, const double *__restrict__ vr) {
    remap_energy_in_body(blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y, blockIdx.z, g,
                         scr, ncol, playkap, gz, pt, ud, vd, udr, vdr, ur, vr);
}

// Total energy remap, pt reconstruction (111-155) and the write-back of
// the interior (157-161), from the energies k_remap_energy_in left in the
// scratch.  With write_winds the remapped winds are copied into ud_out /
// vd_out here (only when those are not the buffers the energy of
// neighbouring columns is formed from).
__device__ __forceinline__ void remap_energy_body(
    size_t t, Geo g, double *__restrict__ scr, size_t ncol, const double *__restrict__ plev,
    const double *__restrict__ tplev, const double *__restrict__ tpkap, const double *__restrict__ tplk,
    const double *__restrict__ gz, const double *__restrict__ ur, const double *__restrict__ vr, double *pt_out,
    double *__restrict__ dp_out, double *__restrict__ ud_out, double *__restrict__ vd_out, int write_winds,
    int *nfw) {
    const int W = g.nx;
    if (t >= ncol) return;
    const int I = 7 + (int)(t % W), J = 7 + (int)(t / W);
    const int L = g.L;
    const size_t plane = g.plane(), base = g.ix2(I, J);
    const Scr s{scr, ncol, t, L};
    const double *ke = s.slot(3);
    const EnergyCol c{plev, tplev, s.slot(2), base, plane, ncol};
    ppm_forward(c, L, s.slot(0), s.slot(1), ncol);
    double gzr = gz[base];
    auto out = [&](int m, double e, double dm0, double dm1) {
        const size_t x = base + (size_t)(m - 1) * plane;
        const double tk = tpkap[x], tk1 = tpkap[x + plane];
        const double p = (e - gzr - ke[(size_t)m * ncol]) / (ATM_CV_DRY * tplk[x] + .5 * ATM_CP_DRY * (tk - tk1));
        pt_out[x] = p;
        gzr = gzr + ATM_CP_DRY * p * (tk - tk1);
        const double d = dm0 - dm1;  // target_plev(k) - target_plev(k+1)
        dp_out[x] = d;
        atm_report_nonfinite(nfw, p);
        atm_report_nonfinite(nfw, d);
        if (write_winds) {
            ud_out[x] = ur[x];
            vd_out[x] = vr[x];
            atm_report_nonfinite(nfw, ur[x]);
            atm_report_nonfinite(nfw, vr[x]);
        }
    };
    ppm_merge(c, L, s.slot(0), s.slot(1), ncol, out);
}

__global__ void __launch_bounds__(256) k_remap_energy(
    Geo g, double *__restrict__ scr, size_t ncol, const double *__restrict__ plev,
    const double *__restrict__ tplev, const double *__restrict__ tpkap, const double *__restrict__ tplk,
    const double *__restrict__ gz, const double *__restrict__ ur, const double *__restrict__ vr, double *pt_out,
    double *__restrict__ dp_out, double *__restrict__ ud_out, double *__restrict__ vd_out, int write_winds,
    int *nfw) {
    remap_energy_body((size_t)xcd_block1d() * blockDim.x + threadIdx.x, g, scr, ncol, plev, tplev, tpkap, tplk, gz, ur,
                      vr, pt_out, dp_out, ud_out, vd_out, write_winds, nfw);
}

// interior wind write-back when the output winds are the input buffers
__global__ void k_remap_write_winds(Geo g, const double *__restrict__ ur, const double *__restrict__ vr,
                                    double *__restrict__ ud, double *__restrict__ vd, int *nfw) {
    const int I = 7 + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = 7 + blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I > g.NI - 8 || J > g.NJ - 8) return;
    size_t x = g.ix(I, J, K);
    const double u = ur[x], v = vr[x];
    ud[x] = u;
    vd[x] = v;
    atm_report_nonfinite(nfw, u);
    atm_report_nonfinite(nfw, v);
}

// Batched remap of every member of an ensemble (atm_ensemble_remap): a
// member's span of workgroups runs, stage by stage, the bodies its own
// launches run (same columns, cells and operations, so the same values),
// reading the copy of the fields its parity bit names and writing the
// other, as atm_remap does after a D step.
constexpr int ENS_RM_BLOCK = 256;
__device__ __forceinline__ unsigned ens_rm_blocks(size_t n) { return (unsigned)((n + ENS_RM_BLOCK - 1) / ENS_RM_BLOCK); }

__global__ void __launch_bounds__(128) k_remap_grid_ens(const EnsMember *__restrict__ tab, EnsIndex ix, int nm,
                                                        unsigned long long par) {
    int m;
    unsigned l;
    if (!ens_find(ix, nm, m, l)) return;
    const EnsMember &e = tab[m];
    const int q = (int)((par >> m) & 1ull);
    const unsigned gx = (unsigned)(e.g.nx + 6 + 63) / 64;  // the 64 x 2 blocks of k_remap_grid
    const int bx = (int)(l % gx), by = (int)(l / gx);
    remap_grid_body(4 + bx * 64 + (int)threadIdx.x % 64, 4 + by * 2 + (int)threadIdx.x / 64, e.g, e.ak0, e.ak, e.bk,
                    e.p[0][q], e.p[1][q], e.plev, e.pkap, e.playkap, e.gz, e.tplev, e.tpkap, e.tplk);
}

__global__ void __launch_bounds__(ENS_RM_BLOCK) k_remap_winds_ens(const EnsMember *__restrict__ tab, EnsIndex ix,
                                                                  int nm, unsigned long long par) {
    int m;
    unsigned l;
    if (!ens_find(ix, nm, m, l)) return;
    const EnsMember &e = tab[m];
    const int q = (int)((par >> m) & 1ull);
    const size_t nw = (size_t)(e.g.nx + 2) * (e.g.ny + 2);
    const unsigned nb = ens_rm_blocks(nw);  // u columns, then v columns
    const bool is_v = l >= nb;
    remap_winds_body((size_t)(is_v ? l - nb : l) * ENS_RM_BLOCK + threadIdx.x, is_v, e.g, e.rscr, nw, e.plev, e.tplev,
                     e.p[2][q], e.p[3][q], e.p[2][1 - q], e.p[3][1 - q], e.ur, e.vr);
}

__global__ void __launch_bounds__(ENS_RM_BLOCK) k_remap_energy_in_ens(const EnsMember *__restrict__ tab, EnsIndex ix,
                                                                      int nm, unsigned long long par) {
    int m;
    unsigned l;
    if (!ens_find(ix, nm, m, l)) return;
    const EnsMember &e = tab[m];
    const int q = (int)((par >> m) & 1ull);
    const unsigned gx = (unsigned)(e.g.nx + 63) / 64, gy = (unsigned)(e.g.ny + 3) / 4;  // the 64 x 4 blocks
    const int bx = (int)(l % gx), by = (int)(l / gx % gy), K = (int)(l / gx / gy);
    remap_energy_in_body(bx * 64 + (int)threadIdx.x % 64, by * 4 + (int)threadIdx.x / 64, K, e.g, e.rscr,
                         (size_t)e.g.nx * e.g.ny, e.playkap, e.gz, e.p[1][q], e.p[2][q], e.p[3][q], e.p[2][1 - q],
                         e.p[3][1 - q], e.ur, e.vr);
}

__global__ void __launch_bounds__(ENS_RM_BLOCK) k_remap_energy_ens(const EnsMember *__restrict__ tab, EnsIndex ix,
                                                                   int nm, unsigned long long par) {
    int m;
    unsigned l;
    if (!ens_find(ix, nm, m, l)) return;
    const EnsMember &e = tab[m];
    const int o = 1 - (int)((par >> m) & 1ull);
    remap_energy_body((size_t)l * ENS_RM_BLOCK + threadIdx.x, e.g, e.rscr, (size_t)e.g.nx * e.g.ny, e.plev, e.tplev,
                      e.tpkap, e.tplk, e.gz, e.ur, e.vr, e.p[1][o], e.p[0][o], e.p[2][o], e.p[3][o], 1, e.nfw);
}

__global__ void k_fill_plane(double *a, size_t n, double v) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

unsigned blocks8(size_t n, unsigned bs) {
    size_t b = (n + bs - 1) / bs;
    return (unsigned)((b + 7) / 8 * 8);
}

}  // namespace

bool atm_remap_member(const atm_ctx *c, EnsMember &m) {
    const RemapWork *w = c->remap;
    if (!w) return false;
    m.ak = c->ak;
    m.bk = c->bk;
    m.playkap = c->playkap;
    m.rscr = w->scr;
    m.tplev = w->tplev;
    m.tpkap = w->tpkap;
    m.tplk = w->tplk;
    m.ur = w->uremap;
    m.vr = w->vremap;
    return true;
}

// stages: 0 grid (128-thread blocks), 1 winds, 2 energy inputs, 3 energy
unsigned atm_ens_blocks_remap(const Geo &g, int stage) {
    auto blocks = [](size_t n, size_t b) { return (unsigned)((n + b - 1) / b); };
    switch (stage) {
        case 0: return blocks(g.nx + 6, 64) * blocks(g.ny + 6, 2);
        case 1: return 2u * blocks((size_t)(g.nx + 2) * (g.ny + 2), ENS_RM_BLOCK);
        case 2: return blocks(g.nx, 64) * blocks(g.ny, 4) * (unsigned)g.L;
        default: return blocks((size_t)g.nx * g.ny, ENS_RM_BLOCK);
    }
}

void atm_ens_launch_remap(hipStream_t s, const EnsMember *tab, const EnsIndex *ix, int nm, unsigned long long par,
                          const unsigned *n) {
    hipLaunchKernelGGL(k_remap_grid_ens, dim3(n[0]), dim3(128), 0, s, tab, ix[0], nm, par);
    hipLaunchKernelGGL(k_remap_winds_ens, dim3(n[1]), dim3(ENS_RM_BLOCK), 0, s, tab, ix[1], nm, par);
    hipLaunchKernelGGL(k_remap_energy_in_ens, dim3(n[2]), dim3(ENS_RM_BLOCK), 0, s, tab, ix[2], nm, par);
    hipLaunchKernelGGL(k_remap_energy_ens, dim3(n[3]), dim3(ENS_RM_BLOCK), 0, s, tab, ix[3], nm, par);
}

// Frees what a RemapWork holds; also for one whose allocation stopped part way.
static void remap_release(RemapWork *w) {
    (void)hipFree(w->scr);
    (void)hipFree(w->tplev);
    (void)hipFree(w->tpkap);
    (void)hipFree(w->tplk);
    (void)hipFree(w->uremap);
    (void)hipFree(w->vremap);
    delete w;
}

// The context's RemapWork is published only once every buffer exists, so a
// failed initialisation leaves the remap uninitialised and a retry
// allocates afresh.
int atm_remap_alloc(atm_ctx *c) {
    const Geo &g = c->g;
    if (!c->remap) {
        RemapWork *w = new RemapWork();
        w->ncol = (size_t)(g.nx + 2) * (g.ny + 2);
        size_t p = g.plane();
        double **bufs[] = {&w->scr, &w->tplev, &w->tpkap, &w->tplk, &w->uremap, &w->vremap};
        const size_t n[] = {w->ncol * 4 * (g.L + 2), p * (g.L + 1), p * (g.L + 1), p * g.L, p * g.L, p * g.L};
        for (int i = 0; i < 6; i++) {
            const hipError_t e = hipMalloc(bufs[i], sizeof(double) * n[i]);
            if (e != hipSuccess) {
                remap_release(w);
                return atm_fail(c, ATM_ERR_HIP, std::string("hipMalloc (remap): ") + hipGetErrorString(e));
            }
        }
        c->remap = w;
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
    remap_release(c->remap);
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
    const size_t nw = (size_t)(g.nx + 2) * (g.ny + 2), ne = (size_t)g.nx * g.ny;
    const unsigned bs = 256;
    hipLaunchKernelGGL(k_remap_winds, dim3(blocks8(nw, bs), 2), dim3(bs), 0, c->stream, g, w->scr, nw, c->plev,
                       w->tplev, ud_in, vd_in, ud_ring, vd_ring, w->uremap, w->vremap);
    {
        const dim3 be(64, 4);
        hipLaunchKernelGGL(k_remap_energy_in, grid3(g.nx, g.ny, g.L, be), be, 0, c->stream, g, w->scr, ne,
                           c->playkap, c->gz, pt, ud_in, vd_in, ud_ring, vd_ring, w->uremap, w->vremap);
    }
    // the remapped winds go to the output buffers from the energy kernel
    // unless those are the input winds, whose interior neighbouring columns
    // still read there (a ring-only source buffer is read outside the
    // interior only)
    const int write_winds = (ud_out != ud_in && vd_out != vd_in) ? 1 : 0;
    hipLaunchKernelGGL(k_remap_energy, dim3(blocks8(ne, bs)), dim3(bs), 0, c->stream, g, w->scr, ne, c->plev,
                       w->tplev, w->tpkap, w->tplk, c->gz, w->uremap, w->vremap, pt_out, dp_out, ud_out, vd_out,
                       write_winds, c->nfw);
    if (!write_winds) {
        dim3 bw(64, 4);
        hipLaunchKernelGGL(k_remap_write_winds, grid3(g.nx, g.ny, g.L, bw), bw, 0, c->stream, g, w->uremap,
                           w->vremap, ud_out, vd_out, c->nfw);
    }
    atm_timer_end(c, ATM_T_REMAP);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}

// in-place form (reference-structured path)
int atm_remap_run(atm_ctx *c) {
    return atm_remap_run_split(c, c->dp, c->pt, c->ud, c->vd, c->ud, c->vd, c->dp, c->pt, c->ud, c->vd);
}
