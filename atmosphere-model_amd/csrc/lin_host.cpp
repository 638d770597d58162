// lin_host.cpp -- the linearised step of lin_step.hpp run by host loops,
// for CPU checks of the tangent-linear and adjoint models (finite
// differences, the dot-product test) on small grids.  It runs the same
// per-point bodies and the same step program as the HIP kernels of
// k_lin.hip; the library never calls it.
//
// Host arrays use the reference's unpadded halo'd layout (row stride nx+14).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

// Frozen-branch mode: a tape of the program's branch decisions (LIN_DECIDE
// in lin_types.hpp).  Recording appends each decision of a run; replaying
// returns the recorded decisions in the same order, so a run from a
// perturbed state takes the branches of the recorded one, and counts the
// decisions the perturbed state alone would have taken otherwise.  The host
// program runs its loops in a fixed order, so the k-th decision of a replay
// is the k-th of the recording.
namespace lintape {
struct Tape {
    int mode = 0;  // 0 off, 1 record, 2 replay
    std::vector<bool> bits;
    size_t pos = 0;
    long long flips = 0, overrun = 0;
};
inline Tape g;
inline bool decide(bool b) {
    if (g.mode == 1) {
        g.bits.push_back(b);
        return b;
    }
    if (g.mode == 2) {
        if (g.pos >= g.bits.size()) {
            g.overrun++;
            return b;
        }
        const bool r = g.bits[g.pos++];
        if (r != b) g.flips++;
        return r;
    }
    return b;
}
}  // namespace lintape
#define LIN_DECIDE(b) (lintape::decide(b))

#include "lin_step.hpp"

namespace {

struct HostAdd {
    static void add(double *p, double v) { *p += v; }
};

struct HostBackend {
    double *alloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
    void release(double *p) { free(p); }
    void copy(double *d, const double *s, size_t n) { memmove(d, s, sizeof(double) * n); }
    void fill(double *d, size_t n, double v) {
        for (size_t i = 0; i < n; i++) d[i] = v;
    }
    template <class F>
    void fwd(const LinArgs &x, Rng r, int k0, int nk) {
        for (int K = k0; K < k0 + nk; K++)
            for (int J = r.J0; J <= r.J1; J++)
                for (int I = r.I0; I <= r.I1; I++) {
                    AccFwd a{x};
                    F::run(a, x, I, J, K);
                }
    }
    template <class F>
    void tan(const LinArgs &x, Rng r, int k0, int nk) {
        for (int K = k0; K < k0 + nk; K++)
            for (int J = r.J0; J <= r.J1; J++)
                for (int I = r.I0; I <= r.I1; I++) {
                    AccTan a{x};
                    F::run(a, x, I, J, K);
                }
    }
    template <class F>
    void adj(const LinArgs &x, Rng r, int k0, int nk) {
        for (int K = k0; K < k0 + nk; K++)
            for (int J = r.J0; J <= r.J1; J++)
                for (int I = r.I0; I <= r.I1; I++) {
                    AccAdj<F::NR, HostAdd> a(x);
                    F::run(a, x, I, J, K);
                    a.flush();
                }
    }
    void scan(int mode, bool gz, const Geo &g, Rng r, const double *sv, double *sd, const double *tv, double *td,
              double *pv, double *pd, double *kv, double *kd, double *gv, double *gd) {
        for (int J = r.J0; J <= r.J1; J++)
            for (int I = r.I0; I <= r.I1; I++) {
                const size_t c0 = g.ix2(I, J);
                if (mode == LIN_FWD) {
                    if (gz)
                        scan_column<double, true>(g, c0, sv, nullptr, tv, nullptr, pv, nullptr, kv, nullptr, gv, nullptr);
                    else
                        scan_column<double, false>(g, c0, sv, nullptr, tv, nullptr, pv, nullptr, kv, nullptr, gv, nullptr);
                } else if (mode == LIN_TAN) {
                    if (gz)
                        scan_column<Jac<1>, true>(g, c0, sv, sd, tv, td, pv, pd, kv, kd, gv, gd);
                    else
                        scan_column<Jac<1>, false>(g, c0, sv, sd, tv, td, pv, pd, kv, kd, gv, gd);
                } else {
                    if (gz)
                        scan_column_adj<true>(g, c0, sd, tv, td, pv, pd, kv, kd, gd);
                    else
                        scan_column_adj<false>(g, c0, sd, tv, td, pv, pd, kv, kd, gd);
                }
            }
    }
    void columns(int mode, int kind, const LinColOp *ops, int n) {
        for (int i = 0; i < n; i++) column(mode, kind, ops[i].x, ops[i].r, ops[i].w);
    }
    void column(int mode, int kind, const LinArgs &x, Rng r, double *ws) {
        const Geo &g = x.g;
        const size_t W = (size_t)(r.I1 - r.I0 + 1), ncol = W * (size_t)(r.J1 - r.J0 + 1);
        for (int J = r.J0; J <= r.J1; J++)
            for (int I = r.I0; I <= r.I1; I++) {
                const RmWS w{ws, ncol, (size_t)(J - r.J0) * W + (size_t)(I - r.I0), g.L};
                const size_t c0 = g.ix2(I, J);
                if (kind == RM_COL_REMAP) {
                    if (mode == LIN_FWD) rm_column<double>(x, c0, w);
                    else if (mode == LIN_TAN) rm_column<Jac<1>>(x, c0, w);
                    else rm_column_adj(x, c0, w);
                } else {
                    if (mode == LIN_FWD) rm_pt_column<double>(x, c0, w);
                    else if (mode == LIN_TAN) rm_pt_column<Jac<1>>(x, c0, w);
                    else rm_pt_column_adj(x, c0, w);
                }
            }
    }
    void halo_copy(int mode, const Geo &g, int nk, const double *sv, double *sd, double *dv, double *dd) {
        for (int K = 0; K < nk; K++)
            for (int J = 0; J < g.NJ; J++)
                for (int I = 0; I < g.NI; I++) {
                    const size_t d = g.ix(I, J, K), s = g.ix(wrap_index(I, g.nx), wrap_index(J, g.ny), K);
                    if (mode == LIN_ADJ) {
                        const double a = dd[d];
                        dd[d] = 0.;
                        sd[s] += a;
                    } else {
                        dv[d] = sv[s];
                        if (mode == LIN_TAN) dd[d] = sd[s];
                    }
                }
    }
};

struct HostLin {
    HostBackend be;
    LinCore<HostBackend> core;
    double *gz_sfc = nullptr, *gzb_sfc = nullptr, *nf = nullptr;
    double *ak = nullptr, *bk = nullptr;
    double *tl[4] = {}, *ad[4] = {};
    int last = -1;
    explicit HostLin(const Geo &g) : core(be, g) {}
};

Geo host_geo(int nx, int ny, int nlay) {
    Geo g{};
    g.nx = nx;
    g.ny = ny;
    g.L = nlay;
    g.NI = nx + 2 * ATM_HALO;
    g.NJ = ny + 2 * ATM_HALO;
    g.SI = g.NI;
    return g;
}

void prepare(HostLin *h, int kind) {
    LinCore<HostBackend> &k = h->core;
    if (h->last != kind) {
        for (int i = 0; i < A_COUNT; i++)
            if (k.drv[i]) h->be.fill(k.drv[i], k.count(i), 0.);
        h->last = kind;
    }
    // corner values of the surface geopotential (init_sw_dyn 995-996) over
    // the halo'd plane, then the constants of the step
    LinArgs x{};
    x.g = k.g;
    x.iv[0] = h->gz_sfc;
    x.ov[0] = h->gzb_sfc;
    h->be.fwd<OpCorners>(x, lin_box(k.g, 2, 1), 0, 1);
    k.set_constants(h->gz_sfc, h->gzb_sfc, h->nf);
}

}  // namespace

extern "C" {

// params: gravity, coriolis, divergence damping, vorticity damping
void *lin_host_create(int nx, int ny, int nlay, double dx, double dy, const double *params, const double *ak,
                      const double *bk) {
    const Geo g = host_geo(nx, ny, nlay);
    HostLin *h = new HostLin(g);
    LinCore<HostBackend> &k = h->core;
    if (!k.allocate()) {
        k.release();
        delete h;
        return nullptr;
    }
    const size_t p = g.plane(), n = p * nlay;
    h->gz_sfc = h->be.alloc(p);
    h->gzb_sfc = h->be.alloc(p);
    h->nf = h->be.alloc(p * (nlay + 1));
    h->ak = h->be.alloc(nlay + 1);
    h->bk = h->be.alloc(nlay + 1);
    memcpy(h->ak, ak, sizeof(double) * (nlay + 1));
    memcpy(h->bk, bk, sizeof(double) * (nlay + 1));
    for (int i = 0; i < 4; i++) {
        h->tl[i] = h->be.alloc(n);
        h->ad[i] = h->be.alloc(n);
    }
    // init_sw_dyn (mod_sw_dyn.f90:958-1023)
    DynCoef &c = k.c;
    c.rdx = 1. / dx;
    c.rdx2 = 1. / (dx * dx);
    c.rdy = 1. / dy;
    c.rdy2 = 1. / (dy * dy);
    c.rdx2rdy2 = c.rdx2 + c.rdy2;
    c.rA = 1. / (dx * dy);
    const double dd = params[2] * dx * dy;
    c.divdamp = dd * dd;
    const double vort = params[3] * dx * dy;
    c.vx = dy * c.rdx * vort;
    c.vy = dx * c.rdy * vort;
    c.coriolis = params[1];
    c.gravity = params[0];
    k.ak0 = ak[0];
    k.pktop = pow(ak[0], ATM_KAPPA);
    k.ak = h->ak;
    k.bk = h->bk;
    return h;
}

// tape mode: 0 off, 1 record (clears the tape), 2 replay (from its start)
void lin_host_tape(int mode) {
    lintape::Tape &t = lintape::g;
    t.mode = mode;
    if (mode == 1) t.bits.clear();
    t.pos = 0;
    t.flips = t.overrun = 0;
}

// 0: decisions recorded, 1: decisions replayed, 2: replayed decisions the
// live comparison would have taken the other way, 3: decisions past the
// tape's end
long long lin_host_tape_stat(int which) {
    const lintape::Tape &t = lintape::g;
    switch (which) {
    case 0: return (long long)t.bits.size();
    case 1: return (long long)t.pos;
    case 2: return t.flips;
    default: return t.overrun;
    }
}

void lin_host_destroy(void *p) {
    HostLin *h = (HostLin *)p;
    if (!h) return;
    h->core.release();
    for (double *a : {h->gz_sfc, h->gzb_sfc, h->nf, h->ak, h->bk}) free(a);
    for (int i = 0; i < 4; i++) {
        free(h->tl[i]);
        free(h->ad[i]);
    }
    delete h;
}

// which: 0-3 state dp, pt, ud, vd; 4-7 their tangents; 8-11 their adjoints
// (halo'd, nlay levels); 12 surface geopotential (a plane); 13 net flux
// (nlay + 1 levels)
static double *host_array(HostLin *h, int which, size_t *n) {
    const Geo &g = h->core.g;
    const size_t p = g.plane();
    *n = p * g.L;
    if (which >= 0 && which < 4) return h->core.val[A_XDP + which];
    if (which >= 4 && which < 8) return h->tl[which - 4];
    if (which >= 8 && which < 12) return h->ad[which - 8];
    if (which == 12) {
        *n = p;
        return h->gz_sfc;
    }
    if (which == 13) {
        *n = p * (g.L + 1);
        return h->nf;
    }
    return nullptr;
}

int lin_host_put(void *p, int which, const double *a) {
    size_t n = 0;
    double *d = host_array((HostLin *)p, which, &n);
    if (!d) return 1;
    memcpy(d, a, sizeof(double) * n);
    return 0;
}

int lin_host_get(void *p, int which, double *a) {
    size_t n = 0;
    double *d = host_array((HostLin *)p, which, &n);
    if (!d) return 1;
    memcpy(a, d, sizeof(double) * n);
    return 0;
}

// value of one single-assignment array (LinArr id) after the last sweep
int lin_host_array(void *p, int id, double *a) {
    HostLin *h = (HostLin *)p;
    if (id < 0 || id >= A_COUNT) return 1;
    memcpy(a, h->core.val[id], sizeof(double) * h->core.count(id));
    return 0;
}

// one step forward and tangent-linear: the state and the tangents advance
int lin_host_tl_step(void *p, double dt, int remap) {
    HostLin *h = (HostLin *)p;
    LinCore<HostBackend> &k = h->core;
    prepare(h, LIN_TAN);
    const size_t n = k.g.plane() * k.g.L;
    for (int i = 0; i < 4; i++) memcpy(k.drv[A_XDP + i], h->tl[i], sizeof(double) * n);
    k.sweep(LIN_TAN, dt, remap != 0);
    int end[4];
    lin_end_arrays(remap != 0, end);
    for (int i = 0; i < 4; i++) {
        // the state keeps the interior of the end arrays (its halo is the
        // next step's fill)
        double *sv = k.val[A_XDP + i];
        const Geo &g = k.g;
        for (int K = 0; K < g.L; K++)
            for (int J = ATM_HALO; J < ATM_HALO + g.ny; J++)
                for (int I = ATM_HALO; I < ATM_HALO + g.nx; I++) {
                    const size_t x = g.ix(I, J, K);
                    sv[x] = k.val[end[i]][x];
                    h->tl[i][x] = k.drv[end[i]][x];
                }
    }
    return 0;
}

// adjoint of one step about the current state (unchanged)
int lin_host_ad_step(void *p, double dt, int remap) {
    HostLin *h = (HostLin *)p;
    LinCore<HostBackend> &k = h->core;
    const Geo &g = k.g;
    const size_t n = g.plane() * g.L;
    prepare(h, LIN_ADJ);
    int end[4];
    lin_end_arrays(remap != 0, end);
    for (int i = 0; i < 4; i++)
        for (int K = 0; K < g.L; K++)
            for (int J = ATM_HALO; J < ATM_HALO + g.ny; J++)
                for (int I = ATM_HALO; I < ATM_HALO + g.nx; I++) {
                    const size_t x = g.ix(I, J, K);
                    k.drv[end[i]][x] = h->ad[i][x];
                }
    k.adjoint(dt, remap != 0);
    for (int i = 0; i < 4; i++) {
        memcpy(h->ad[i], k.drv[A_XDP + i], sizeof(double) * n);
        memset(k.drv[A_XDP + i], 0, sizeof(double) * n);
    }
    return 0;
}

}  // extern "C"
