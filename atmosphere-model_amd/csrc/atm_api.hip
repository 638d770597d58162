// atm_api.hip -- C ABI entry points (include/atm.h) of the gfx950 library.
//
// Each entry point replaces one Fortran module procedure of the reference
// (cited per function); device buffers mirror the reference's module
// arrays in its halo'd Fortran layout.
#include <cstdio>
#include <cstring>

#include <cstdlib>
#include <algorithm>
#include <vector>

#include "atm_common.hpp"
#include "atm_fused.hpp"

int atm_fail(atm_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define CHECK_CTX(c) \
    if (!(c)) return ATM_ERR_ARGUMENT

// entry points that wait on the device, copy between host and device or
// allocate cannot be recorded into a graph
#define CHECK_NOT_CAPTURING(c) \
    if ((c)->capturing) return atm_fail(c, ATM_ERR_STATE, "not allowed while recording a graph")

static double *field_ptr(atm_ctx *c, int id, int *nk) {
    const int L = c->g.L;
    switch (id) {
    case ATM_DP: *nk = L; return c->dp;
    case ATM_PT: *nk = L; return c->pt;
    case ATM_UD: *nk = L; return c->ud;
    case ATM_VD: *nk = L; return c->vd;
    case ATM_TS: *nk = 1; return c->ts;
    case ATM_GZ: *nk = L + 1; return c->gz;
    case ATM_NET_FLUX: *nk = L + 1; return c->net_flux;
    case ATM_PLEV: *nk = L + 1; return c->plev;
    case ATM_PKAP: *nk = L + 1; return c->pkap;
    case ATM_RADIUS: *nk = 1; return c->radius;
    case ATM_DPC: *nk = L; return c->dpc;
    case ATM_PTC: *nk = L; return c->ptc;
    case ATM_UC: *nk = L; return c->fused ? c->uc2 : c->uc;
    case ATM_VC: *nk = L; return c->fused ? c->vc2 : c->vc;
    case ATM_TL_DP: case ATM_TL_PT: case ATM_TL_UD: case ATM_TL_VD:
    case ATM_AD_DP: case ATM_AD_PT: case ATM_AD_UD: case ATM_AD_VD:
        *nk = L;
        return atm_lin_field(c, id);
    default: *nk = 0; return nullptr;
    }
}

// n doubles plus the tail that overhanging tiles read (zero-filled)
static int dalloc(atm_ctx *c, double **p, size_t n) {
    const size_t m = n + atm_tile_slack(c->g);
    ATM_HIP(c, hipMalloc(p, sizeof(double) * m));
    c->field_bufs.push_back({*p, n});
    ATM_HIP(c, hipMemsetAsync(*p, 0, sizeof(double) * m, c->stream));
    return ATM_OK;
}

static void release(atm_ctx *c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->dp = c->pt = c->ud = c->vd = nullptr;  // owned through pbuf
    double **all[] = {&c->ak, &c->bk, &c->ts, &c->gz, &c->net_flux,
                      &c->plev, &c->pkap, &c->play, &c->playkap, &c->radius, &c->dpc, &c->ptc,
                      &c->uc, &c->vc, &c->heat, &c->pkapb, &c->gzb, &c->tmp, &c->lap, &c->cx,
                      &c->cy, &c->denx, &c->deny, &c->dpfx, &c->dpfy, &c->fxq, &c->fyq, &c->vort,
                      &c->ke, &c->pgfx, &c->pgfy, &c->divg, &c->ua, &c->va, &c->ub, &c->vb};
    for (auto p : all)
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    for (int f = 0; f < 4; f++)
        for (int b = 0; b < 2; b++)
            if (c->pbuf[f][b]) { (void)hipFree(c->pbuf[f][b]); c->pbuf[f][b] = nullptr; }
    c->field_bufs.clear();
    double **fz[] = {&c->uc2, &c->vc2, &c->heatd, &c->plk};
    for (auto p : fz)
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    for (auto &t : c->timers) {
        for (int i = 0; i < t.cap; i++) (void)hipEventDestroy(t.ev[i]);
        delete[] t.ev;
        t.ev = nullptr;
        t.cap = t.used = 0;
    }
    if (c->flag) (void)hipFree(c->flag);
    if (c->flag_host) (void)hipHostFree(c->flag_host);
    if (c->chk_host) (void)hipHostFree(c->chk_host);
    c->chk_host = nullptr;
    if (c->chk_ev) (void)hipEventDestroy(c->chk_ev);
    c->chk_ev = nullptr;
    c->chk_pending = false;
    if (c->nfw) (void)hipFree(c->nfw);
    atm_remap_free(c);
    atm_rad_free(c);
    atm_phys_free(c);
    atm_lin_free(c);
    if (c->outf) { (void)hipFree(c->outf); c->outf = nullptr; }
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
}

// dp/pt/ud/vd point at the current copy of the prognostic fields
void atm_point_current(atm_ctx *c) {
    c->dp = c->pbuf[0][c->cur];
    c->pt = c->pbuf[1][c->cur];
    c->ud = c->pbuf[2][c->cur];
    c->vd = c->pbuf[3][c->cur];
}

static FusedBufs bufs_of(atm_ctx *c, int which) {
    return FusedBufs{c->pbuf[0][which], c->pbuf[1][which], c->pbuf[2][which], c->pbuf[3][which]};
}

void atm_timer_begin(atm_ctx *c, int id) {
    if (!c->timing || !((c->timer_mask >> id) & 1u)) return;
    AtmTimer &t = c->timers[id];
    if (t.used + 2 > t.cap) {
        int ncap = t.cap ? 2 * t.cap : 256;
        hipEvent_t *ev = new hipEvent_t[ncap];
        for (int i = 0; i < t.cap; i++) ev[i] = t.ev[i];
        for (int i = t.cap; i < ncap; i++) (void)hipEventCreate(&ev[i]);
        delete[] t.ev;
        t.ev = ev;
        t.cap = ncap;
    }
    (void)hipEventRecord(t.ev[t.used], c->stream);
}

void atm_timer_end(atm_ctx *c, int id) {
    if (!c->timing || !((c->timer_mask >> id) & 1u)) return;
    AtmTimer &t = c->timers[id];
    (void)hipEventRecord(t.ev[t.used + 1], c->stream);
    t.used += 2;
}

extern "C" {

int atm_create(const atm_grid *grid, const atm_params *params, int precision, atm_ctx **out) {
    if (!grid || !params || !out) return ATM_ERR_ARGUMENT;
    *out = nullptr;
    if (precision != ATM_PRECISION_FP64 && precision != ATM_PRECISION_MIXED) return ATM_ERR_ARGUMENT;
    // the halo is 7 wide and vertical_ppm needs at least 4 layers
    if (grid->nx < ATM_HALO || grid->ny < ATM_HALO || grid->nlay < 4) return ATM_ERR_ARGUMENT;
    if (!(grid->dx > 0) || !(grid->dy > 0)) return ATM_ERR_ARGUMENT;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return ATM_ERR_HIP;
    atm_ctx *c = new atm_ctx();
    c->g.nx = grid->nx;
    c->g.ny = grid->ny;
    c->g.L = grid->nlay;
    c->g.NI = grid->nx + 2 * ATM_HALO;
    c->g.NJ = grid->ny + 2 * ATM_HALO;
    c->g.SI = (c->g.NI + 15) / 16 * 16;
    c->dx = grid->dx;
    c->dy = grid->dy;
    c->prm = *params;
    int rc = ATM_OK;
    auto fail = [&](int code) {
        release(c);
        delete c;
        return code;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return fail(ATM_ERR_HIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(ATM_ERR_HIP);
    c->own_stream = true;
    const size_t p = c->g.plane(), L = c->g.L;
    const char *path = getenv("ATM_PATH");
    c->fused = !(path && std::string(path) == "reference");
    c->mixed = precision == ATM_PRECISION_MIXED;
    if (c->mixed && !c->fused) return fail(ATM_ERR_ARGUMENT);  // the mixed form is on the fused path only
    struct { double **ptr; size_t n; } bufs[] = {
        {&c->pbuf[0][0], p * L}, {&c->pbuf[1][0], p * L}, {&c->pbuf[2][0], p * L}, {&c->pbuf[3][0], p * L},
        {&c->ak, L + 1}, {&c->bk, L + 1}, {&c->ts, p}, {&c->gz, p * (L + 1)}, {&c->net_flux, p * (L + 1)},
        {&c->plev, p * (L + 1)}, {&c->pkap, p * (L + 1)}, {&c->play, p * L}, {&c->playkap, p * L},
        {&c->radius, p}, {&c->dpc, p * L}, {&c->ptc, p * L}, {&c->uc, p * L}, {&c->vc, p * L},
        {&c->heat, p * L}, {&c->pkapb, p * (L + 1)}, {&c->gzb, p * (L + 1)}, {&c->tmp, p * L},
        {&c->lap, p * L}, {&c->cx, p * L}, {&c->cy, p * L}, {&c->denx, p * L}, {&c->deny, p * L},
        {&c->dpfx, p * L}, {&c->dpfy, p * L}, {&c->fxq, p * L}, {&c->fyq, p * L}, {&c->vort, p * L},
        {&c->ke, p * L}, {&c->pgfx, p * L}, {&c->pgfy, p * L}, {&c->divg, p * L}, {&c->ua, p},
        {&c->va, p}, {&c->ub, p}, {&c->vb, p}};
    for (auto &b : bufs) {
        rc = dalloc(c, b.ptr, b.n);
        if (rc) return fail(rc);
    }
    if (c->fused) {
        double **fz[] = {&c->pbuf[0][1], &c->pbuf[1][1], &c->pbuf[2][1], &c->pbuf[3][1], &c->uc2, &c->vc2,
                         &c->heatd, &c->plk};
        for (auto q : fz) {
            rc = dalloc(c, q, p * L);
            if (rc) return fail(rc);
        }
    }
    atm_point_current(c);
    if (hipMalloc(&c->flag, sizeof(int)) != hipSuccess) return fail(ATM_ERR_HIP);
    if (hipMalloc(&c->nfw, sizeof(int)) != hipSuccess) return fail(ATM_ERR_HIP);
    if (hipMemsetAsync(c->nfw, 0, sizeof(int), c->stream) != hipSuccess) return fail(ATM_ERR_HIP);
    if (hipHostMalloc((void **)&c->flag_host, sizeof(int), hipHostMallocDefault) != hipSuccess)
        return fail(ATM_ERR_HIP);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(ATM_ERR_HIP);
    *out = c;
    return ATM_OK;
}

int atm_destroy(atm_ctx *c) {
    CHECK_CTX(c);
    if (c->capturing) {  // discard an unfinished recording
        hipGraph_t gr = nullptr;
        if (hipStreamEndCapture(c->stream, &gr) == hipSuccess && gr) (void)hipGraphDestroy(gr);
        c->capturing = false;
    }
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    release(c);
    delete c;
    return ATM_OK;
}

const char *atm_last_error(const atm_ctx *c) { return c ? c->err.c_str() : "null context"; }

int atm_stream_create(void **s) {
    if (!s) return ATM_ERR_ARGUMENT;
    hipStream_t h = nullptr;
    if (hipStreamCreateWithFlags(&h, hipStreamNonBlocking) != hipSuccess) return ATM_ERR_HIP;
    *s = (void *)h;
    return ATM_OK;
}

int atm_stream_destroy(void *s) {
    if (!s) return ATM_ERR_ARGUMENT;
    return hipStreamDestroy((hipStream_t)s) == hipSuccess ? ATM_OK : ATM_ERR_HIP;
}

int atm_stream_synchronize(void *s) {
    return hipStreamSynchronize((hipStream_t)s) == hipSuccess ? ATM_OK : ATM_ERR_HIP;
}

namespace {
struct StreamTimer {
    hipEvent_t a = nullptr, b = nullptr;
    bool begun = false;
    ~StreamTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
thread_local StreamTimer g_stream_timer;
}  // namespace

int atm_stream_timer_begin(void *s) {
    StreamTimer &t = g_stream_timer;
    if (!t.a && (hipEventCreate(&t.a) != hipSuccess || hipEventCreate(&t.b) != hipSuccess)) return ATM_ERR_HIP;
    if (hipEventRecord(t.a, (hipStream_t)s) != hipSuccess) return ATM_ERR_HIP;
    t.begun = true;
    return ATM_OK;
}

int atm_stream_timer_end(void *s, double *ms) {
    StreamTimer &t = g_stream_timer;
    if (!ms || !t.begun) return ATM_ERR_ARGUMENT;
    t.begun = false;
    float f = 0.f;
    if (hipEventRecord(t.b, (hipStream_t)s) != hipSuccess || hipEventSynchronize(t.b) != hipSuccess ||
        hipEventElapsedTime(&f, t.a, t.b) != hipSuccess)
        return ATM_ERR_HIP;
    *ms = f;
    return ATM_OK;
}

int atm_set_stream(atm_ctx *c, void *s) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    ATM_HIP(c, hipStreamSynchronize(c->stream));
    if (c->own_stream) {
        if (s == nullptr) return ATM_OK;
        (void)hipStreamDestroy(c->stream);
        c->own_stream = false;
        c->stream = (hipStream_t)s;
        return ATM_OK;
    }
    if (s == nullptr) {
        ATM_HIP(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
    } else {
        c->stream = (hipStream_t)s;
    }
    return ATM_OK;
}

int atm_set_coordinate(atm_ctx *c, const double *ak, const double *bk) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (!ak || !bk) return atm_fail(c, ATM_ERR_ARGUMENT, "null coordinate array");
    const size_t n = c->g.L + 1;
    ATM_HIP(c, hipMemcpyAsync(c->ak, ak, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    ATM_HIP(c, hipMemcpyAsync(c->bk, bk, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    ATM_HIP(c, hipStreamSynchronize(c->stream));
    c->ak0 = ak[0];
    c->pktop = pow(ak[0], ATM_KAPPA);  // host libm, as the reference's top_pressure**kappa
    c->have_coord = true;
    return ATM_OK;
}

// Between atm_dgrid_step and atm_dgrid_halo the double-buffered path keeps
// the new D-grid winds in the current copy and the wind ring the reference's
// arrays still hold there (the previous step's halo, read by the remap) in
// the previous copy.  Transfers of UD / VD in that window see the
// reference's array: reads compose ring and interior, writes go to both.
static bool split_winds(const atm_ctx *c, int id) {
    return c->fused && c->after_d && (id == ATM_UD || id == ATM_VD);
}

static int xfer(atm_ctx *c, int id, void *dst, const void *src, bool to_dev, hipMemcpyKind kind, bool block) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    int nk = 0;
    double *d = field_ptr(c, id, &nk);
    if (!d && id >= ATM_TL_DP && id < ATM_NFIELDS) return atm_fail(c, ATM_ERR_STATE, "atm_lin_init first");
    if (!d) return atm_fail(c, ATM_ERR_ARGUMENT, "unknown field id " + std::to_string(id));
    if (!(to_dev ? src : dst)) return atm_fail(c, ATM_ERR_ARGUMENT, "null array");
    const Geo &g = c->g;
    // the caller's array has rows of NI values back to back (the reference's
    // layout); the device's rows have the stride SI: one 2-D copy of the
    // NJ * nk rows converts
    const size_t hp = sizeof(double) * g.NI, dpch = sizeof(double) * g.SI, w = sizeof(double) * g.NI;
    const size_t rows = (size_t)g.NJ * nk;
    auto to_device = [&](double *dev, const void *from) {
        return hipMemcpy2DAsync(dev, dpch, from, hp, w, rows, kind, c->stream);
    };
    auto from_device = [&](void *to, const double *dev) {
        return hipMemcpy2DAsync(to, hp, dev, dpch, w, rows, kind, c->stream);
    };
    if (to_dev && id == ATM_DP) c->plk_valid = false;
    if (to_dev && (id == ATM_DP || id == ATM_PT || id == ATM_UD || id == ATM_VD)) atm_untrack(c);
    if (split_winds(c, id)) {
        double *ring = c->pbuf[id == ATM_UD ? 2 : 3][1 - c->cur];
        if (to_dev) {
            ATM_HIP(c, to_device(d, src));
            ATM_HIP(c, to_device(ring, src));
        } else {
            ATM_HIP(c, from_device(dst, ring));
            for (int K = 0; K < nk; K++) {
                const size_t ho = ((size_t)K * g.NJ + ATM_HALO) * g.NI + ATM_HALO;
                ATM_HIP(c, hipMemcpy2DAsync((double *)dst + ho, hp, d + g.ix(ATM_HALO, ATM_HALO, K), dpch,
                                            sizeof(double) * g.nx, g.ny, kind, c->stream));
            }
        }
    } else if (to_dev) {
        ATM_HIP(c, to_device(d, src));
    } else {
        ATM_HIP(c, from_device(dst, d));
    }
    if (block) ATM_HIP(c, hipStreamSynchronize(c->stream));
    return ATM_OK;
}

int atm_put_field(atm_ctx *c, int id, const double *h) {
    return xfer(c, id, nullptr, h, true, hipMemcpyHostToDevice, true);
}
int atm_get_field(atm_ctx *c, int id, double *h) {
    return xfer(c, id, h, nullptr, false, hipMemcpyDeviceToHost, true);
}
int atm_put_field_device(atm_ctx *c, int id, const double *d) {
    return xfer(c, id, nullptr, d, true, hipMemcpyDeviceToDevice, false);
}
int atm_get_field_device(atm_ctx *c, int id, double *d) {
    return xfer(c, id, d, nullptr, false, hipMemcpyDeviceToDevice, false);
}

// init_prognostic_fields (mod_input.f90:229-231)
int atm_init_prognostic_fields(atm_ctx *c) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (!c->have_coord) return atm_fail(c, ATM_ERR_STATE, "atm_set_coordinate first");
    const size_t p = c->g.plane();
    int rc = atm_launch_fill(c, c->plev + p * c->g.L, p, c->ak0);
    if (rc) return rc;
    return atm_launch_fill(c, c->pkap + p * c->g.L, p, c->pktop);
}

// initial_halo_exchange (mod_input.f90:247-252)
int atm_initial_halo_exchange(atm_ctx *c) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    const int L = c->g.L;
    double *f[] = {c->dp, c->pt, c->gz, c->ud, c->vd};
    for (double *a : f) {
        int rc = atm_launch_halo(c, a, L);
        if (rc) return rc;
    }
    return atm_launch_halo(c, c->ts, 1);
}

int atm_init_sw_dyn(atm_ctx *c) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (!c->have_coord) return atm_fail(c, ATM_ERR_STATE, "atm_set_coordinate first");
    int rc = atm_ref_init_sw_dyn(c);
    if (rc) return rc;
    c->sw_init = true;
    return ATM_OK;
}

int atm_cgrid_step(atm_ctx *c, double dt) {
    CHECK_CTX(c);
    if (!c->sw_init) return atm_fail(c, ATM_ERR_STATE, "atm_init_sw_dyn first");
    if (c->after_d) return atm_fail(c, ATM_ERR_STATE, "atm_dgrid_halo must follow atm_dgrid_step");
    if (c->fused) {
        FusedBufs b = bufs_of(c, c->cur);
        return atm_fused_cgrid(c, b, dt);
    }
    return atm_ref_cgrid(c, dt);
}

int atm_cgrid_halo(atm_ctx *c) {
    CHECK_CTX(c);
    if (c->fused) return atm_fused_cgrid_halo(c);
    double *f[] = {c->dpc, c->ptc, c->uc, c->vc};
    for (double *a : f) {
        int rc = atm_launch_halo(c, a, c->g.L);
        if (rc) return rc;
    }
    return ATM_OK;
}

// The fused D step writes the other copy of dp, pt, ud, vd and makes it
// current; the previous copy keeps the halo the wind remap reads (Q5) until
// atm_dgrid_halo.  Between atm_dgrid_step and atm_dgrid_halo the interior
// of dp/pt/ud/vd is the reference's; their halo is refreshed by the fill.
int atm_dgrid_step(atm_ctx *c, double dt) {
    CHECK_CTX(c);
    if (!c->sw_init) return atm_fail(c, ATM_ERR_STATE, "atm_init_sw_dyn first");
    if (c->after_d) return atm_fail(c, ATM_ERR_STATE, "atm_dgrid_halo must follow atm_dgrid_step");
    if (c->fused) {
        FusedBufs cur = bufs_of(c, c->cur), nxt = bufs_of(c, 1 - c->cur);
        int rc = atm_fused_dgrid(c, cur, nxt, dt);
        if (rc) return rc;
        c->plk_valid = true;
        c->cur = 1 - c->cur;
        atm_point_current(c);
        c->after_d = true;
        return ATM_OK;
    }
    c->plk_valid = false;
    atm_untrack(c);  // the reference-structured kernels do not report
    return atm_ref_dgrid(c, dt);
}

int atm_dgrid_halo(atm_ctx *c) {
    CHECK_CTX(c);
    double *f[] = {c->dp, c->pt, c->ud, c->vd};
    const int nk[] = {c->g.L, c->g.L, c->g.L, c->g.L};
    atm_timer_begin(c, ATM_T_HALO);
    int rc = atm_launch_halo_multi(c, f, nk, 4);
    atm_timer_end(c, ATM_T_HALO);
    if (rc) return rc;
    c->after_d = false;
    return ATM_OK;
}

int atm_remap_init(atm_ctx *c) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (!c->have_coord) return atm_fail(c, ATM_ERR_STATE, "atm_set_coordinate first");
    int rc = atm_remap_alloc(c);
    if (rc) return rc;
    c->remap_init = true;
    return ATM_OK;
}

int atm_remap(atm_ctx *c) {
    CHECK_CTX(c);
    if (!c->remap_init) return atm_fail(c, ATM_ERR_STATE, "atm_remap_init first");
    c->plk_valid = false;
    if (c->fused && c->after_d) {
        // read the D step's state (current copy) with the wind ring of the
        // previous copy; write the remapped interior into the previous copy
        const int o = 1 - c->cur;
        FusedBufs cur = bufs_of(c, c->cur), prv = bufs_of(c, o);
        int rc = atm_remap_run_split(c, cur.dp, cur.pt, cur.ud, cur.vd, prv.ud, prv.vd, prv.dp, prv.pt, prv.ud,
                                     prv.vd);
        if (rc) return rc;
        c->cur = o;
        atm_point_current(c);
        return ATM_OK;
    }
    return atm_remap_run(c);
}

int atm_radiation_init(atm_ctx *c) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    int rc = atm_rad_alloc(c);
    if (rc) return rc;
    c->rad_init = true;
    return ATM_OK;
}

int atm_radiation(atm_ctx *c, double time) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (!c->rad_init) return atm_fail(c, ATM_ERR_STATE, "atm_radiation_init first");
    ATM_HIP(c, hipMemsetAsync(c->flag, 0, sizeof(int), c->stream));
    int rc = atm_rad_run(c, time);
    if (rc) return rc;
    ATM_HIP(c, hipMemcpyAsync(c->flag_host, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    ATM_HIP(c, hipStreamSynchronize(c->stream));
    if (*c->flag_host & 2)
        return atm_fail(c, ATM_ERR_TABLE_RANGE,
                        "gas-optics lookup outside the tabulated pressure/temperature range");
    return ATM_OK;
}

int atm_radiation_halo(atm_ctx *c) {
    CHECK_CTX(c);
    return atm_launch_halo(c, c->net_flux, c->g.L + 1);
}

int atm_physics_init(atm_ctx *c) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    int rc = atm_phys_alloc(c);
    if (rc) return rc;
    c->phys_init = true;
    return ATM_OK;
}

int atm_physics_tendencies(atm_ctx *c) {
    CHECK_CTX(c);
    if (!c->phys_init) return atm_fail(c, ATM_ERR_STATE, "atm_physics_init first");
    if (c->after_d) return atm_fail(c, ATM_ERR_STATE, "physics tendencies come before the C-grid step");
    return atm_phys_tendencies(c);
}

int atm_apply_physics(atm_ctx *c, double dt) {
    CHECK_CTX(c);
    if (!c->phys_init) return atm_fail(c, ATM_ERR_STATE, "atm_physics_init first");
    return atm_phys_apply(c, dt);
}

int atm_physics_halo(atm_ctx *c) {
    CHECK_CTX(c);
    return atm_launch_halo(c, c->ts, 1);
}

int atm_output_init(atm_ctx *c) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    int rc = atm_output_alloc(c);
    if (rc) return rc;
    c->out_init = true;
    return ATM_OK;
}

int atm_accumulate_output(atm_ctx *c, double dt) {
    CHECK_CTX(c);
    if (!c->out_init) return atm_fail(c, ATM_ERR_STATE, "atm_output_init first");
    if (c->after_d) return atm_fail(c, ATM_ERR_STATE, "accumulate after atm_dgrid_halo");
    return atm_output_accumulate(c, dt);
}

int atm_output_mean(atm_ctx *c, double *host) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (!c->out_init) return atm_fail(c, ATM_ERR_STATE, "atm_output_init first");
    if (!host) return atm_fail(c, ATM_ERR_ARGUMENT, "null output array");
    if (!(c->acc_time > 0.)) return atm_fail(c, ATM_ERR_STATE, "nothing accumulated");
    return atm_output_take_mean(c, host);
}

int atm_halo_exchange(atm_ctx *c, int id) {
    CHECK_CTX(c);
    int nk = 0;
    double *d = field_ptr(c, id, &nk);
    if (!d) return atm_fail(c, ATM_ERR_ARGUMENT, "unknown field id " + std::to_string(id));
    if (id == ATM_GZ) nk = c->g.L;  // the 5-argument specific sees gz(:,:,1:nlay)
    return atm_launch_halo(c, d, nk);
}

// check_stability (mod_fields.f90:69-106): all of interior dp, pt, ud, vd
// finite.  While the state is tracked (atm_ctx::nfw) a clear tracking word
// answers; otherwise, or when a reported value may since have been
// overwritten, the four fields are scanned.
int atm_check_stability(atm_ctx *c, int *stable) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (!stable) return atm_fail(c, ATM_ERR_ARGUMENT, "null result pointer");
    if (c->chk_pending) return atm_fail(c, ATM_ERR_STATE, "a deferred check is pending: atm_check_stability_end first");
    if (c->nfw_tracked) {
        ATM_HIP(c, hipMemcpyAsync(c->flag_host, c->nfw, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        ATM_HIP(c, hipStreamSynchronize(c->stream));
        if (!(*c->flag_host & 1)) {
            *stable = 1;
            return ATM_OK;
        }
    }
    int rc = atm_launch_stability(c);
    if (rc) return rc;
    ATM_HIP(c, hipMemcpyAsync(c->flag_host, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    ATM_HIP(c, hipMemsetAsync(c->nfw, 0, sizeof(int), c->stream));
    ATM_HIP(c, hipStreamSynchronize(c->stream));
    *stable = (*c->flag_host & 1) ? 0 : 1;
    c->nfw_tracked = *stable != 0;
    return ATM_OK;
}

// The deferred form of the check: _begin queues it on the stream behind
// the work issued so far and returns; _end waits for it alone (not for
// work issued after _begin) and reports.  A non-finite value written into
// a prognostic field by a reporting kernel reads as unstable here, without
// the confirming scan atm_check_stability makes.
int atm_check_stability_begin(atm_ctx *c) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (c->chk_pending) return atm_fail(c, ATM_ERR_STATE, "a deferred check is pending: atm_check_stability_end first");
    if (!c->chk_ev) ATM_HIP(c, hipEventCreateWithFlags(&c->chk_ev, hipEventDisableTiming));
    if (!c->chk_host) ATM_HIP(c, hipHostMalloc((void **)&c->chk_host, sizeof(int), hipHostMallocDefault));
    c->chk_scan = !c->nfw_tracked;
    if (c->chk_scan) {
        int rc = atm_launch_stability(c);
        if (rc) return rc;
        ATM_HIP(c, hipMemcpyAsync(c->chk_host, c->flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        ATM_HIP(c, hipMemsetAsync(c->nfw, 0, sizeof(int), c->stream));
    } else {
        ATM_HIP(c, hipMemcpyAsync(c->chk_host, c->nfw, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    }
    ATM_HIP(c, hipEventRecord(c->chk_ev, c->stream));
    c->chk_epoch = c->untrack_epoch;
    c->chk_pending = true;
    return ATM_OK;
}

int atm_check_stability_end(atm_ctx *c, int *stable) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (!stable) return atm_fail(c, ATM_ERR_ARGUMENT, "null result pointer");
    if (!c->chk_pending) return atm_fail(c, ATM_ERR_STATE, "no deferred check: atm_check_stability_begin first");
    c->chk_pending = false;
    ATM_HIP(c, hipEventSynchronize(c->chk_ev));
    const bool bad = (*c->chk_host & 1) != 0;
    *stable = bad ? 0 : 1;
    if (bad) {
        atm_untrack(c);
    } else if (c->chk_scan && c->untrack_epoch == c->chk_epoch) {
        // the scan found the fields finite and every write since has
        // reported into the word it cleared
        c->nfw_tracked = true;
    }
    return ATM_OK;
}

int atm_set_path(atm_ctx *c, int path) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (path != ATM_PATH_FUSED && path != ATM_PATH_REFERENCE) return atm_fail(c, ATM_ERR_ARGUMENT, "bad path");
    if (c->after_d) return atm_fail(c, ATM_ERR_STATE, "finish the step (atm_dgrid_halo) first");
    c->plk_valid = false;
    atm_untrack(c);
    if (path == ATM_PATH_FUSED) {
        if (!c->uc2)
            return atm_fail(c, ATM_ERR_ARGUMENT,
                            "the context was created on the reference path (ATM_PATH=reference): no fused buffers");
        c->fused = true;
    } else {
        if (c->mixed) return atm_fail(c, ATM_ERR_ARGUMENT, "the mixed-precision form runs on the fused path only");
        c->fused = false;
    }
    return ATM_OK;
}

int atm_get_path(atm_ctx *c, int *path) {
    CHECK_CTX(c);
    if (!path) return ATM_ERR_ARGUMENT;
    *path = c->fused ? ATM_PATH_FUSED : ATM_PATH_REFERENCE;
    return ATM_OK;
}

// One fused step (C step, its halo fill, D step) from the current state
// into the other copy of the fields, with the host-side state left as it
// was: the next real step reads only the current copy and rewrites every
// array the trial wrote before reading it (playkap, which the trial's D
// step leaves for the other copy, is marked stale by the caller).
static int trial_step(atm_ctx *c, double dt) {
    FusedBufs cur = bufs_of(c, c->cur), nxt = bufs_of(c, 1 - c->cur);
    int rc = atm_fused_cgrid(c, cur, .5 * dt);
    if (!rc) rc = atm_fused_cgrid_halo(c);
    if (!rc) rc = atm_fused_dgrid(c, cur, nxt, dt);
    return rc;
}

// The tuning of the fused path: each candidate launch shape is timed
// inside whole trial steps on this context's grid, state and stream (one
// warm-up step, then `reps` steps between two events), so that a shape's
// effect on its neighbours in the step (the corner planes D2 reads back)
// counts; the fastest is kept when it beats the current choice by more
// than 1% (three interleaved timings of each candidate, after a warm-up
// until the clocks settle).  Only shapes that give the same values are
// candidates: the rows of the dp/pt and of the wind tiles (8, 16) and the
// corner kernel's levels per block (1, 2, 4, 8).
int atm_autotune(atm_ctx *c, int reps) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (reps < 1) return atm_fail(c, ATM_ERR_ARGUMENT, "reps must be >= 1");
    if (!c->fused) return ATM_OK;
    if (!c->sw_init) return atm_fail(c, ATM_ERR_STATE, "atm_init_sw_dyn first");
    if (c->after_d) return atm_fail(c, ATM_ERR_STATE, "the autotuner runs at a step boundary");
    // the last step's dt, or a Courant number of about 0.2 at 100 m/s
    const DynCoef coef = c->coef;
    const double dt = coef.dt > 0. ? coef.dt : .2 * std::min(c->dx, c->dy) / 100.;
    const bool timing = c->timing;
    const int before[3] = {c->tile_pair, c->tile_wind, c->bc_levels};
    c->timing = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    int rc = ATM_OK;
    // the trial steps' time with the current choice
    auto timed = [&](float &ms) {
        ms = 0.f;
        if (e != hipSuccess || rc) return false;
        rc = trial_step(c, dt);
        if (!rc) e = hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps && !rc && e == hipSuccess; r++) rc = trial_step(c, dt);
        if (!rc && e == hipSuccess) e = hipEventRecord(e1, c->stream);
        if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
        if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        return !rc && e == hipSuccess;
    };
    // The GPU's clocks rise under load for tens of milliseconds: trial
    // steps run first until two consecutive timings agree to 1%.
    {
        float prev = 0.f, ms = 0.f;
        for (int w = 0; w < 40; w++) {
            if (!timed(ms)) break;
            if (w > 0 && std::fabs(ms - prev) < .01f * prev) break;
            prev = ms;
        }
    }
    // One knob at a time, the others at their current choice: the dp/pt
    // tile rows, the wind tile rows, then the corner depth.  Each knob's
    // candidates are timed in three rounds, in alternating order, and each
    // keeps its fastest time; a candidate replaces the current choice only
    // when it is the fastest and faster by more than 1% (the shipped shapes
    // are the measured best at the benchmark grid).
    int *knob[] = {&c->tile_pair, &c->tile_wind, &c->bc_levels};
    const std::vector<int> cand[] = {{8, 16}, {16, 8}, {1, 2, 4, 8}};
    for (int q = 0; q < 3; q++) {
        const int start = *knob[q];
        const size_t nc = cand[q].size();
        std::vector<float> best(nc, 0.f);
        bool ok = true;
        for (int round = 0; round < 3 && ok; round++)
            for (size_t j = 0; j < nc && ok; j++) {
                const size_t i = round % 2 ? nc - 1 - j : j;
                *knob[q] = cand[q][i];
                float ms;
                ok = timed(ms);
                if (ok && (round == 0 || ms < best[i])) best[i] = ms;
            }
        *knob[q] = start;
        if (!ok) break;
        size_t pick = 0, cur = 0;
        for (size_t i = 0; i < nc; i++) {
            if (best[i] < best[pick]) pick = i;
            if (cand[q][i] == start) cur = i;
        }
        if (best[pick] < .99f * best[cur]) *knob[q] = cand[q][pick];
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    c->coef = coef;
    c->timing = timing;
    c->plk_valid = false;
    if (rc || e != hipSuccess) {
        c->tile_pair = before[0];
        c->tile_wind = before[1];
        c->bc_levels = before[2];
        if (rc) return rc;
        return atm_fail(c, ATM_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
    }
    return ATM_OK;
}

int atm_get_tuning(atm_ctx *c, int *corner_levels) {
    CHECK_CTX(c);
    if (!corner_levels) return atm_fail(c, ATM_ERR_ARGUMENT, "null output");
    *corner_levels = c->bc_levels;
    return ATM_OK;
}

int atm_set_tuning(atm_ctx *c, int corner_levels) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (corner_levels != 1 && corner_levels != 2 && corner_levels != 4 && corner_levels != 8)
        return atm_fail(c, ATM_ERR_ARGUMENT, "corner levels per workgroup: 1, 2, 4 or 8");
    c->bc_levels = corner_levels;
    return ATM_OK;
}

int atm_get_tiles(atm_ctx *c, int *pair_rows, int *wind_rows) {
    CHECK_CTX(c);
    if (!pair_rows || !wind_rows) return atm_fail(c, ATM_ERR_ARGUMENT, "null output");
    *pair_rows = c->tile_pair;
    *wind_rows = c->tile_wind;
    return ATM_OK;
}

int atm_set_tiles(atm_ctx *c, int pair_rows, int wind_rows) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if ((pair_rows != 8 && pair_rows != 16) || (wind_rows != 8 && wind_rows != 16))
        return atm_fail(c, ATM_ERR_ARGUMENT, "tile rows: 8 or 16");
    c->tile_pair = pair_rows;
    c->tile_wind = wind_rows;
    return ATM_OK;
}

int atm_poison(atm_ctx *c, int what, double value) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (what & ~(ATM_POISON_PADDING | ATM_POISON_OUTER_HALO)) return atm_fail(c, ATM_ERR_ARGUMENT, "poison: bad class");
    for (const auto &b : c->field_bufs) {
        const int rc = atm_launch_poison(c, b.first, b.second, what, value);
        if (rc) return rc;
    }
    return ATM_OK;
}

int atm_timer_enable(atm_ctx *c, int on) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    ATM_HIP(c, hipStreamSynchronize(c->stream));
    for (auto &t : c->timers) t.used = 0;
    c->timing = on != 0;
    return ATM_OK;
}

int atm_timer_select(atm_ctx *c, unsigned groups) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (groups >> ATM_T_COUNT) return atm_fail(c, ATM_ERR_ARGUMENT, "timer groups: bits 0-8");
    ATM_HIP(c, hipStreamSynchronize(c->stream));
    for (auto &t : c->timers) t.used = 0;
    c->timer_mask = groups;
    return ATM_OK;
}

int atm_timer_read(atm_ctx *c, int kernel, int64_t *launches, double *total_ms) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (kernel < 0 || kernel >= ATM_T_COUNT || !launches || !total_ms) return ATM_ERR_ARGUMENT;
    ATM_HIP(c, hipStreamSynchronize(c->stream));
    AtmTimer &t = c->timers[kernel];
    double sum = 0;
    for (int i = 0; i + 1 < t.used; i += 2) {
        float ms = 0;
        ATM_HIP(c, hipEventElapsedTime(&ms, t.ev[i], t.ev[i + 1]));
        sum += ms;
    }
    *launches = t.used / 2;
    *total_ms = sum;
    return ATM_OK;
}

int atm_sync(atm_ctx *c) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    ATM_HIP(c, hipStreamSynchronize(c->stream));
    return ATM_OK;
}

// Graphs of whole steps.  Between atm_graph_begin and atm_graph_end the
// step entry points (C and D steps, their halo fills, remap, physics
// tendencies and their application, halo fills, output accumulation) are
// recorded on the context's stream instead of run; atm_graph_launch runs
// the recording.  The host-side state those calls advance (which copy of
// the double-buffered fields is current, the pending D-step halo) must be
// back where it started when the recording ends, so that every launch
// finds the buffers the recording used: whole steps, and on the fused path
// an even number of buffer switches (each D step and each remap switch).
// whether a recording may start on this context: at a step boundary, with
// no deferred check pending
static int cap_ready(atm_ctx *c) {
    if (!c->sw_init) return atm_fail(c, ATM_ERR_STATE, "atm_init_sw_dyn first");
    if (c->after_d) return atm_fail(c, ATM_ERR_STATE, "a graph starts at a step boundary");
    if (c->chk_pending) return atm_fail(c, ATM_ERR_STATE, "a deferred check is pending: atm_check_stability_end first");
    return ATM_OK;
}

// the host-side state of a context when its calls start being recorded
static void cap_enter(atm_ctx *c) {
    c->cap_cur = c->cur;
    c->cap_timing = c->timing;
    c->timing = false;  // timer events are not recorded into the graph
    // A launch may follow any eager work (a D step, a remap, a put of dp),
    // so the recording cannot assume playkap valid: its first C step forms
    // it afresh.
    c->cap_plk_valid = c->plk_valid;
    c->plk_valid = false;
    // set until a recorded call writes dp, pt, ud or vd without reporting
    // into the stability word (a reference-path D step)
    c->cap_nfw_tracked = c->nfw_tracked;
    c->nfw_tracked = true;
    c->cap_dts.clear();
    c->capturing = true;
}

// What the recorded calls changed on the host side, for each launch to
// apply; the context goes back to its state at cap_enter (nothing recorded
// has run).  Returns whether the recording holds whole steps with an even
// number of buffer switches.
static bool cap_leave(atm_ctx *c, GraphEffects &fx) {
    c->capturing = false;
    c->timing = c->cap_timing;
    const bool whole = !c->after_d && c->cur == c->cap_cur;
    fx.cur = c->cap_cur;
    fx.plk_end = c->plk_valid;
    fx.tracks = c->nfw_tracked;
    fx.dts.clear();
    fx.dts.swap(c->cap_dts);
    c->after_d = false;
    if (c->cur != c->cap_cur) {
        c->cur = c->cap_cur;
        atm_point_current(c);
    }
    c->plk_valid = c->cap_plk_valid;
    c->nfw_tracked = c->cap_nfw_tracked;
    return whole;
}

// whether a recording made from this context's state may launch now
static int fx_ready(atm_ctx *c, const GraphEffects &fx) {
    if (c->after_d) return atm_fail(c, ATM_ERR_STATE, "a graph launches at a step boundary");
    if (c->cur != fx.cur)
        return atm_fail(c, ATM_ERR_STATE,
                        "the current copy of the fields is not the one the graph was recorded on "
                        "(an odd number of eager D steps + remaps since)");
    return ATM_OK;
}

// the host-side effects of one launch of the recorded steps, as if run
// eagerly: accumulation time in the recorded order, playkap's validity, and
// the stability tracking (lost when a recorded write does not report)
static void fx_apply(atm_ctx *c, const GraphEffects &fx) {
    for (double dt : fx.dts) c->acc_time = c->acc_time + dt;
    c->plk_valid = fx.plk_end;
    if (!fx.tracks) atm_untrack(c);
}

static const char *const ATM_GRAPH_WHOLE =
    "a graph must hold whole steps with an even number of buffer switches (D steps + remaps)";

int atm_graph_begin(atm_ctx *c) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    int rc = cap_ready(c);
    if (rc) return rc;
    ATM_HIP(c, hipStreamSynchronize(c->stream));
    if (c->gexec) {
        ATM_HIP(c, hipGraphExecDestroy(c->gexec));
        c->gexec = nullptr;
    }
    ATM_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    cap_enter(c);
    c->gfx.dts.clear();
    return ATM_OK;
}

int atm_graph_end(atm_ctx *c) {
    CHECK_CTX(c);
    if (!c->capturing) return atm_fail(c, ATM_ERR_STATE, "atm_graph_begin first");
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &gr);
    GraphEffects fx;
    const bool whole = cap_leave(c, fx);
    if (e != hipSuccess || !gr) return atm_fail(c, ATM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    if (!whole) {
        (void)hipGraphDestroy(gr);
        return atm_fail(c, ATM_ERR_STATE, ATM_GRAPH_WHOLE);
    }
    const hipError_t ei = hipGraphInstantiate(&c->gexec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (ei != hipSuccess) {
        c->gexec = nullptr;
        return atm_fail(c, ATM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    c->gfx = std::move(fx);
    return ATM_OK;
}

int atm_graph_launch(atm_ctx *c, int times) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (!c->gexec) return atm_fail(c, ATM_ERR_STATE, "no recorded graph");
    if (times < 0) return atm_fail(c, ATM_ERR_ARGUMENT, "times must be >= 0");
    // the recording reads and writes the buffers that were current when it
    // began, at a step boundary
    int rc = fx_ready(c, c->gfx);
    if (rc) return rc;
    for (int i = 0; i < times; i++) {
        ATM_HIP(c, hipGraphLaunch(c->gexec, c->stream));
        fx_apply(c, c->gfx);
    }
    return ATM_OK;
}

int atm_graph_free(atm_ctx *c) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (c->gexec) {
        ATM_HIP(c, hipStreamSynchronize(c->stream));
        ATM_HIP(c, hipGraphExecDestroy(c->gexec));
        c->gexec = nullptr;
    }
    return ATM_OK;
}

// ---------------------------------------------------------------------------
// Batched ensemble steps: the dynamics stages of every member in one launch
// each (atm_fused.hpp).  The calls mirror atm_cgrid_step, atm_cgrid_halo,
// atm_dgrid_step and atm_dgrid_halo member by member, with the same checks
// and the same host-side state changes; errors are reported on the member
// concerned, or on the first.
// ---------------------------------------------------------------------------
struct atm_ensemble {
    std::vector<atm_ctx *> m;
    EnsMember *tab = nullptr;  // device
    // per launch kind: the members' workgroup spans and the launch size (a
    // multiple of 8)
    EnsIndex ix_pair{}, ix_wind{}, ix_corner{}, ix_scan{}, ix_halo[4]{}, ix_remap[4]{}, ix_rad[3]{};
    unsigned n_pair = 0, n_wind = 0, n_corner = 0, n_scan = 0, n_halo[4] = {0, 0, 0, 0}, n_remap[4] = {0, 0, 0, 0},
             n_rad[3] = {0, 0, 0};
    // every member's remap / radiation initialised when the table was made
    bool remap_ready = false, rad_ready = false;
    int lmax = 0;  // the members' largest nlay
    // The batched launches go to the first member's stream.  Members may run
    // on other streams (their remap and radiation overlap there): a batched
    // call first makes the launch stream wait for every other member stream
    // and afterwards makes those wait for it (events, one per stream).
    std::vector<hipEvent_t> ev;          // n + 1
    std::vector<hipStream_t> others;     // the other member streams of the current call
    // a recording of batched and member calls (atm_ensemble_graph_*): one
    // graph on the launch stream, with every member's host-side effects
    bool capturing = false;
    hipGraphExec_t gexec = nullptr;
    std::vector<GraphEffects> fx;
};

// the launch stream waits for the work already issued on every other member stream
static int ens_join(atm_ensemble *e) {
    atm_ctx *c0 = e->m[0];
    const hipStream_t s0 = c0->stream;
    e->others.clear();
    for (atm_ctx *c : e->m) {
        if (c->stream == s0 || std::find(e->others.begin(), e->others.end(), c->stream) != e->others.end()) continue;
        hipEvent_t ev = e->ev[e->others.size()];
        e->others.push_back(c->stream);
        ATM_HIP(c0, hipEventRecord(ev, c->stream));
        ATM_HIP(c0, hipStreamWaitEvent(s0, ev, 0));
    }
    return ATM_OK;
}

// every other member stream waits for the batched launches
static int ens_fork(atm_ensemble *e) {
    atm_ctx *c0 = e->m[0];
    if (e->others.empty()) return ATM_OK;
    hipEvent_t ev = e->ev.back();
    ATM_HIP(c0, hipEventRecord(ev, c0->stream));
    for (hipStream_t s : e->others) ATM_HIP(c0, hipStreamWaitEvent(s, ev, 0));
    return ATM_OK;
}

static unsigned round8(unsigned long long n) { return (unsigned)((n + 7ull) / 8ull * 8ull); }

// every member on the fused path, in one precision, on one device, with
// its dynamics initialised, not recording a graph of its own
static int ens_usable(atm_ensemble *e) {
    atm_ctx *c0 = e->m[0];
    for (atm_ctx *c : e->m) {
        if (!c->sw_init) return atm_fail(c, ATM_ERR_STATE, "ensemble: atm_init_sw_dyn first");
        if (!c->fused) return atm_fail(c, ATM_ERR_STATE, "ensemble: every member on the fused path");
        if (c->mixed != c0->mixed) return atm_fail(c, ATM_ERR_STATE, "ensemble: every member in one precision");
        if (c->device != c0->device) return atm_fail(c, ATM_ERR_STATE, "ensemble: every member on one device");
        if (c->capturing != e->capturing) return atm_fail(c, ATM_ERR_STATE, "ensemble: not while recording a graph");
    }
    return ATM_OK;
}

// the distinct member streams other than the launch stream
static std::vector<hipStream_t> ens_streams(atm_ensemble *e) {
    std::vector<hipStream_t> v;
    for (atm_ctx *c : e->m)
        if (c->stream != e->m[0]->stream && std::find(v.begin(), v.end(), c->stream) == v.end()) v.push_back(c->stream);
    return v;
}

// ends a recording that will not be used; the members go back to their
// state at its start
static void ens_discard(atm_ensemble *e) {
    hipGraph_t gr = nullptr;
    if (hipStreamEndCapture(e->m[0]->stream, &gr) == hipSuccess && gr) (void)hipGraphDestroy(gr);
    GraphEffects fx;
    for (atm_ctx *c : e->m)
        if (c->capturing) cap_leave(c, fx);
    e->capturing = false;
}

// bit m: member m's current copy of dp, pt, ud, vd
static unsigned long long ens_parity(atm_ensemble *e) {
    unsigned long long par = 0;
    for (size_t i = 0; i < e->m.size(); i++)
        if (e->m[i]->cur) par |= 1ull << i;
    return par;
}

int atm_ensemble_create(atm_ctx *const *members, int n, atm_ensemble **out) {
    if (!out || !members || n < 1) return ATM_ERR_ARGUMENT;
    *out = nullptr;
    for (int i = 0; i < n; i++)
        if (!members[i]) return ATM_ERR_ARGUMENT;
    if (n > ATM_ENS_MAX) return atm_fail(members[0], ATM_ERR_ARGUMENT, "ensemble: at most 64 members");
    atm_ensemble *e = new atm_ensemble();
    e->m.assign(members, members + n);
    int rc = ens_usable(e);
    std::vector<EnsMember> h(n);
    unsigned long long np = 0, nw = 0, nc = 0, ns = 0, nh[4] = {0, 0, 0, 0}, nr[4] = {0, 0, 0, 0}, nd[3] = {0, 0, 0};
    e->remap_ready = true;
    e->rad_ready = true;
    for (int i = 0; rc == ATM_OK && i < n; i++) {
        e->ix_pair.first[i] = (unsigned)np;
        e->ix_wind.first[i] = (unsigned)nw;
        e->ix_corner.first[i] = (unsigned)nc;
        e->ix_scan.first[i] = (unsigned)ns;
        for (int k = 0; k < 4; k++) e->ix_halo[k].first[i] = (unsigned)nh[k];
        for (int k = 0; k < 4; k++) e->ix_remap[k].first[i] = (unsigned)nr[k];
        for (int k = 0; k < 3; k++) e->ix_rad[k].first[i] = (unsigned)nd[k];
        atm_ctx *c = members[i];
        if (!atm_ens_scan_fits(c->g)) {
            rc = atm_fail(c, ATM_ERR_ARGUMENT, "ensemble: nlay <= 64 for the batched column scans");
            break;
        }
        EnsMember &m = h[i];
        m.g = c->g;
        m.coef = c->coef;
        m.ak0 = c->ak0;
        m.pktop = c->pktop;
        for (int f = 0; f < 4; f++)
            for (int w = 0; w < 2; w++) m.p[f][w] = c->pbuf[f][w];
        m.dpc = c->dpc; m.ptc = c->ptc; m.uc = c->uc; m.vc = c->vc; m.ua = c->ua; m.va = c->va;
        m.uc2 = c->uc2; m.vc2 = c->vc2; m.plk = c->plk; m.net_flux = c->net_flux; m.heatd = c->heatd;
        m.plev = c->plev; m.pkap = c->pkap; m.gz = c->gz; m.pkapb = c->pkapb; m.gzb = c->gzb; m.nfw = c->nfw;
        if (!atm_remap_member(c, m)) e->remap_ready = false;
        if (!atm_rad_member(c, m)) e->rad_ready = false;
        e->lmax = std::max(e->lmax, c->g.L);
        np += atm_ens_blocks_pair(c->g);
        nw += atm_ens_blocks_wind(c->g);
        nc += atm_ens_blocks_corner(c->g);
        ns += atm_ens_blocks_scan(c->g);
        for (int k = 0; k < 4; k++) nh[k] += atm_ens_blocks_halo(c->g, k);
        for (int k = 0; k < 4; k++) nr[k] += atm_ens_blocks_remap(c->g, k);
        for (int k = 0; k < 3; k++) nd[k] += atm_ens_blocks_rad(c->g, k);
    }
    if (rc == ATM_OK && std::max(std::max(std::max(np, nw), std::max(nc, ns)),
                                 std::max(std::max(nr[1], nr[2]), std::max(nd[1], nd[2]))) > 0x7fffffffull)
        rc = atm_fail(members[0], ATM_ERR_ARGUMENT, "ensemble: too many workgroups for one launch");
    if (rc == ATM_OK) {
        e->ix_pair.first[n] = (unsigned)np;
        e->ix_wind.first[n] = (unsigned)nw;
        e->ix_corner.first[n] = (unsigned)nc;
        e->ix_scan.first[n] = (unsigned)ns;
        for (int k = 0; k < 4; k++) e->ix_halo[k].first[n] = (unsigned)nh[k];
        for (int k = 0; k < 4; k++) e->ix_remap[k].first[n] = (unsigned)nr[k];
        for (int k = 0; k < 3; k++) e->ix_rad[k].first[n] = (unsigned)nd[k];
    }
    if (rc == ATM_OK) {
        hipError_t err = hipMalloc(&e->tab, sizeof(EnsMember) * n);
        if (err == hipSuccess) err = hipMemcpy(e->tab, h.data(), sizeof(EnsMember) * n, hipMemcpyHostToDevice);
        for (int i = 0; err == hipSuccess && i <= n; i++) {
            hipEvent_t ev = nullptr;
            err = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            if (err == hipSuccess) e->ev.push_back(ev);
        }
        if (err != hipSuccess) rc = atm_fail(members[0], ATM_ERR_HIP, std::string("ensemble: ") + hipGetErrorString(err));
    }
    if (rc != ATM_OK) {
        for (hipEvent_t ev : e->ev) (void)hipEventDestroy(ev);
        if (e->tab) (void)hipFree(e->tab);
        delete e;
        return rc;
    }
    e->n_pair = round8(np);
    e->n_wind = round8(nw);
    e->n_corner = round8(nc);
    e->n_scan = round8(ns);
    for (int k = 0; k < 4; k++) e->n_halo[k] = round8(nh[k]);
    for (int k = 0; k < 4; k++) e->n_remap[k] = round8(nr[k]);
    for (int k = 0; k < 3; k++) e->n_rad[k] = round8(nd[k]);
    *out = e;
    return ATM_OK;
}

int atm_ensemble_destroy(atm_ensemble *e) {
    if (!e) return ATM_OK;
    if (e->capturing) ens_discard(e);
    for (atm_ctx *c : e->m) (void)hipStreamSynchronize(c->stream);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    for (hipEvent_t ev : e->ev) (void)hipEventDestroy(ev);
    if (e->tab) (void)hipFree(e->tab);
    delete e;
    return ATM_OK;
}

int atm_ensemble_cgrid_step(atm_ensemble *e, double dt) {
    if (!e) return ATM_ERR_ARGUMENT;
    int rc = ens_usable(e);
    if (rc) return rc;
    for (atm_ctx *c : e->m)
        if (c->after_d) return atm_fail(c, ATM_ERR_STATE, "atm_dgrid_halo must follow atm_dgrid_step");
    atm_ctx *c0 = e->m[0];
    const hipStream_t s = c0->stream;
    const int nm = (int)e->m.size();
    for (atm_ctx *c : e->m) {
        c->coef.dt = dt;
        c->coef.dtdx = dt * c->coef.rdx;
        c->coef.dtdy = dt * c->coef.rdy;
        if (!c->plk_valid) {  // dp changed since the last D step's scan (a remap, a put)
            rc = atm_fused_plk(c);
            if (rc) return rc;
        }
    }
    const unsigned long long par = ens_parity(e);
    rc = ens_join(e);
    if (rc) return rc;
    atm_timer_begin(c0, ATM_T_C1);
    atm_ens_launch_pair(s, e->tab, e->ix_pair, nm, par, dt, true, c0->mixed, e->n_pair);
    atm_timer_end(c0, ATM_T_C1);
    atm_timer_begin(c0, ATM_T_HALO);
    atm_ens_launch_halo(s, e->tab, e->ix_halo[ENS_HALO_C1], nm, par, ENS_HALO_C1, e->n_halo[ENS_HALO_C1]);
    atm_timer_end(c0, ATM_T_HALO);
    atm_timer_begin(c0, ATM_T_SCAN);
    atm_ens_launch_scan(s, e->tab, e->ix_scan, nm, par, false, c0->mixed, e->n_scan);
    atm_timer_end(c0, ATM_T_SCAN);
    atm_timer_begin(c0, ATM_T_C2);
    atm_ens_launch_c2(s, e->tab, e->ix_wind, nm, par, dt, c0->mixed, e->n_wind);
    atm_timer_end(c0, ATM_T_C2);
    ATM_HIP(c0, hipGetLastError());
    return ens_fork(e);
}

int atm_ensemble_cgrid_halo(atm_ensemble *e) {
    if (!e) return ATM_ERR_ARGUMENT;
    int rc = ens_usable(e);
    if (rc) return rc;
    atm_ctx *c0 = e->m[0];
    rc = ens_join(e);
    if (rc) return rc;
    atm_timer_begin(c0, ATM_T_HALO);
    atm_ens_launch_halo(c0->stream, e->tab, e->ix_halo[ENS_HALO_CH], (int)e->m.size(), ens_parity(e), ENS_HALO_CH, e->n_halo[ENS_HALO_CH]);
    atm_timer_end(c0, ATM_T_HALO);
    ATM_HIP(c0, hipGetLastError());
    return ens_fork(e);
}

int atm_ensemble_dgrid_step(atm_ensemble *e, double dt) {
    if (!e) return ATM_ERR_ARGUMENT;
    int rc = ens_usable(e);
    if (rc) return rc;
    for (atm_ctx *c : e->m)
        if (c->after_d) return atm_fail(c, ATM_ERR_STATE, "atm_dgrid_halo must follow atm_dgrid_step");
    atm_ctx *c0 = e->m[0];
    const hipStream_t s = c0->stream;
    const int nm = (int)e->m.size();
    for (atm_ctx *c : e->m) {
        c->coef.dt = dt;
        c->coef.dtdx = dt * c->coef.rdx;
        c->coef.dtdy = dt * c->coef.rdy;
    }
    const unsigned long long par = ens_parity(e);
    rc = ens_join(e);
    if (rc) return rc;
    atm_timer_begin(c0, ATM_T_D1);
    atm_ens_launch_pair(s, e->tab, e->ix_pair, nm, par, dt, false, c0->mixed, e->n_pair);
    atm_timer_end(c0, ATM_T_D1);
    atm_timer_begin(c0, ATM_T_HALO);
    atm_ens_launch_halo(s, e->tab, e->ix_halo[ENS_HALO_D1], nm, par, ENS_HALO_D1, e->n_halo[ENS_HALO_D1]);
    atm_timer_end(c0, ATM_T_HALO);
    atm_timer_begin(c0, ATM_T_SCAN);
    atm_ens_launch_scan(s, e->tab, e->ix_scan, nm, par, true, c0->mixed, e->n_scan);
    atm_ens_launch_corners(s, e->tab, e->ix_corner, nm, e->n_corner);
    atm_timer_end(c0, ATM_T_SCAN);
    atm_timer_begin(c0, ATM_T_D2);
    atm_ens_launch_d2(s, e->tab, e->ix_wind, nm, par, dt, c0->mixed, e->n_wind);
    atm_timer_end(c0, ATM_T_D2);
    ATM_HIP(c0, hipGetLastError());
    rc = ens_fork(e);
    if (rc) return rc;
    for (atm_ctx *c : e->m) {  // as atm_dgrid_step on the fused path
        c->plk_valid = true;
        c->cur = 1 - c->cur;
        atm_point_current(c);
        c->after_d = true;
    }
    return ATM_OK;
}

int atm_ensemble_dgrid_halo(atm_ensemble *e) {
    if (!e) return ATM_ERR_ARGUMENT;
    int rc = ens_usable(e);
    if (rc) return rc;
    atm_ctx *c0 = e->m[0];
    rc = ens_join(e);
    if (rc) return rc;
    atm_timer_begin(c0, ATM_T_HALO);
    atm_ens_launch_halo(c0->stream, e->tab, e->ix_halo[ENS_HALO_DH], (int)e->m.size(), ens_parity(e), ENS_HALO_DH, e->n_halo[ENS_HALO_DH]);
    atm_timer_end(c0, ATM_T_HALO);
    ATM_HIP(c0, hipGetLastError());
    for (atm_ctx *c : e->m) c->after_d = false;
    return ens_fork(e);
}

// perform_vertical_remapping of every member after the batched D step:
// the four remap stages, each one launch for all members, each member's
// values those of its own atm_remap.
int atm_ensemble_remap(atm_ensemble *e) {
    if (!e) return ATM_ERR_ARGUMENT;
    int rc = ens_usable(e);
    if (rc) return rc;
    atm_ctx *c0 = e->m[0];
    for (atm_ctx *c : e->m) {
        if (!c->remap_init) return atm_fail(c, ATM_ERR_STATE, "atm_remap_init first");
        if (!c->after_d) return atm_fail(c, ATM_ERR_STATE, "ensemble: the remap follows atm_ensemble_dgrid_step");
    }
    if (!e->remap_ready)
        return atm_fail(c0, ATM_ERR_STATE, "ensemble: every member's atm_remap_init before atm_ensemble_create");
    const unsigned long long par = ens_parity(e);
    rc = ens_join(e);
    if (rc) return rc;
    atm_timer_begin(c0, ATM_T_REMAP);
    atm_ens_launch_remap(c0->stream, e->tab, e->ix_remap, (int)e->m.size(), par, e->n_remap);
    atm_timer_end(c0, ATM_T_REMAP);
    ATM_HIP(c0, hipGetLastError());
    rc = ens_fork(e);
    if (rc) return rc;
    for (atm_ctx *c : e->m) {  // as atm_remap after a fused D step
        c->plk_valid = false;
        c->cur = 1 - c->cur;
        atm_point_current(c);
    }
    return ATM_OK;
}

// calculate_radiative_heating of every member (atm_radiation for each):
// the profiles, the net flux's zero fill and the column solves, each one
// launch for all members; one wait for every member's range flag, an error
// reported on the first member whose lookup left the tables.
int atm_ensemble_radiation(atm_ensemble *e, double time) {
    (void)time;  // as atm_radiation: unused by solar_irradiance (mod_solar.f90:26-35)
    if (!e) return ATM_ERR_ARGUMENT;
    int rc = ens_usable(e);
    if (rc) return rc;
    atm_ctx *c0 = e->m[0];
    for (atm_ctx *c : e->m) {
        CHECK_NOT_CAPTURING(c);
        if (!c->rad_init) return atm_fail(c, ATM_ERR_STATE, "atm_radiation_init first");
    }
    if (!e->rad_ready)
        return atm_fail(c0, ATM_ERR_STATE, "ensemble: every member's atm_radiation_init before atm_ensemble_create");
    const unsigned long long par = ens_parity(e);
    rc = ens_join(e);
    if (rc) return rc;
    const hipStream_t s = c0->stream;
    for (atm_ctx *c : e->m) ATM_HIP(c0, hipMemsetAsync(c->flag, 0, sizeof(int), s));
    atm_timer_begin(c0, ATM_T_RAD);
    atm_ens_launch_rad(s, e->tab, e->ix_rad, (int)e->m.size(), par, e->n_rad, e->lmax, c0->mixed);
    atm_timer_end(c0, ATM_T_RAD);
    ATM_HIP(c0, hipGetLastError());
    for (atm_ctx *c : e->m)
        ATM_HIP(c0, hipMemcpyAsync(c->flag_host, c->flag, sizeof(int), hipMemcpyDeviceToHost, s));
    ATM_HIP(c0, hipStreamSynchronize(s));
    for (atm_ctx *c : e->m)
        if (*c->flag_host & 2)
            return atm_fail(c, ATM_ERR_TABLE_RANGE, "gas-optics lookup outside the tabulated pressure/temperature range");
    return ens_fork(e);
}

// Recording of batched steps: the launch stream starts a capture and the
// other member streams join it (they wait for an event recorded in it), so
// the batched launches and any member's own calls (its playkap scan, an
// atm_remap of its own) land in one graph; at the end every member stream
// is joined back.  Each member's
// host-side state is handled as atm_graph_begin / end handle one context's.
int atm_ensemble_graph_begin(atm_ensemble *e) {
    if (!e) return ATM_ERR_ARGUMENT;
    if (e->capturing) return atm_fail(e->m[0], ATM_ERR_STATE, "ensemble: already recording");
    int rc = ens_usable(e);
    if (rc) return rc;
    for (atm_ctx *c : e->m) {
        rc = cap_ready(c);
        if (rc) return rc;
    }
    atm_ctx *c0 = e->m[0];
    for (atm_ctx *c : e->m) ATM_HIP(c, hipStreamSynchronize(c->stream));
    if (e->gexec) {
        ATM_HIP(c0, hipGraphExecDestroy(e->gexec));
        e->gexec = nullptr;
    }
    ATM_HIP(c0, hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
    const std::vector<hipStream_t> others = ens_streams(e);
    hipError_t err = others.empty() ? hipSuccess : hipEventRecord(e->ev.back(), c0->stream);
    for (size_t i = 0; err == hipSuccess && i < others.size(); i++) err = hipStreamWaitEvent(others[i], e->ev.back(), 0);
    for (atm_ctx *c : e->m) cap_enter(c);
    e->capturing = true;
    if (err != hipSuccess) {
        ens_discard(e);
        return atm_fail(c0, ATM_ERR_HIP, std::string("ensemble graph: ") + hipGetErrorString(err));
    }
    return ATM_OK;
}

int atm_ensemble_graph_end(atm_ensemble *e) {
    if (!e) return ATM_ERR_ARGUMENT;
    atm_ctx *c0 = e->m[0];
    if (!e->capturing) return atm_fail(c0, ATM_ERR_STATE, "atm_ensemble_graph_begin first");
    const std::vector<hipStream_t> others = ens_streams(e);
    hipError_t err = hipSuccess;
    for (size_t i = 0; err == hipSuccess && i < others.size(); i++) {
        err = hipEventRecord(e->ev[i], others[i]);
        if (err == hipSuccess) err = hipStreamWaitEvent(c0->stream, e->ev[i], 0);
    }
    hipGraph_t gr = nullptr;
    const hipError_t ee = hipStreamEndCapture(c0->stream, &gr);
    e->capturing = false;
    std::vector<GraphEffects> fx(e->m.size());
    atm_ctx *broken = nullptr;
    for (size_t i = 0; i < e->m.size(); i++)
        if (!cap_leave(e->m[i], fx[i]) && !broken) broken = e->m[i];
    if (err == hipSuccess) err = ee;
    if (err != hipSuccess || !gr) {
        if (gr) (void)hipGraphDestroy(gr);
        return atm_fail(c0, ATM_ERR_HIP, std::string("ensemble graph: ") + hipGetErrorString(err));
    }
    if (broken) {
        (void)hipGraphDestroy(gr);
        return atm_fail(broken, ATM_ERR_STATE, ATM_GRAPH_WHOLE);
    }
    const hipError_t ei = hipGraphInstantiate(&e->gexec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (ei != hipSuccess) {
        e->gexec = nullptr;
        return atm_fail(c0, ATM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    e->fx.swap(fx);
    return ATM_OK;
}

// Each launch follows the work already issued on every member stream and
// precedes what is issued there afterwards.
int atm_ensemble_graph_launch(atm_ensemble *e, int times) {
    if (!e) return ATM_ERR_ARGUMENT;
    atm_ctx *c0 = e->m[0];
    if (!e->gexec) return atm_fail(c0, ATM_ERR_STATE, "no recorded ensemble graph");
    if (times < 0) return atm_fail(c0, ATM_ERR_ARGUMENT, "times must be >= 0");
    int rc = ens_usable(e);
    if (rc) return rc;
    for (size_t i = 0; i < e->m.size(); i++) {
        rc = fx_ready(e->m[i], e->fx[i]);
        if (rc) return rc;
    }
    rc = ens_join(e);
    if (rc) return rc;
    for (int t = 0; t < times; t++) {
        ATM_HIP(c0, hipGraphLaunch(e->gexec, c0->stream));
        for (size_t i = 0; i < e->m.size(); i++) fx_apply(e->m[i], e->fx[i]);
    }
    return ens_fork(e);
}

int atm_lin_init(atm_ctx *c) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (!c->sw_init) return atm_fail(c, ATM_ERR_STATE, "atm_init_sw_dyn first");
    return atm_lin_alloc(c);
}

static int lin_check(atm_ctx *c, double dt, int remap) {
    CHECK_CTX(c);
    CHECK_NOT_CAPTURING(c);
    if (!c->lin) return atm_fail(c, ATM_ERR_STATE, "atm_lin_init first");
    if (c->after_d) return atm_fail(c, ATM_ERR_STATE, "a linearised step starts at a step boundary");
    if (!(dt > 0.)) return atm_fail(c, ATM_ERR_ARGUMENT, "dt must be positive");
    if (remap && !c->remap_init) return atm_fail(c, ATM_ERR_STATE, "atm_remap_init first");
    return ATM_OK;
}

int atm_tl_step(atm_ctx *c, double dt, int remap) {
    int rc = lin_check(c, dt, remap);
    if (rc) return rc;
    rc = atm_lin_tl_step(c, dt, remap);
    // the state changed outside the tracked kernels
    c->plk_valid = false;
    atm_untrack(c);
    return rc;
}

int atm_ad_step(atm_ctx *c, double dt, int remap) {
    int rc = lin_check(c, dt, remap);
    if (rc) return rc;
    return atm_lin_ad_step(c, dt, remap);
}

const char *atm_build_info(void) {
    return "atmosphere-model_amd gfx950 FP64: fused LDS-tiled dynamics (32x16 tiles) with the "
           "reference-structured kernels as the alternative path; -O3 -ffp-contract=off";
}

}  // extern "C"
