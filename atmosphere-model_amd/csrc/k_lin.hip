// k_lin.hip -- tangent-linear and adjoint of one model step on the GPU
// (gfx950).  The step program and its per-point bodies are in
// lin_step.hpp / lin_ops.hpp; this file runs them with HIP kernels:
//   forward   one thread per point of an operation's box, the body in double;
//   tangent   the same body on Jac<1> (value and tangent read, both written);
//   adjoint   the same body on Jac<NR>: each output's adjoint times each
//             partial, summed over the body's outputs, is what the point
//             owes each value it read; a workgroup sums these in LDS planes
//             and adds each touched value to global memory once per plane
//             with an FP64 atomic (neighbouring workgroups touch the same
//             values); outputs whose adjoint is zero add nothing, so an
//             adjoint confined to a small region costs little of the grid.
// Column scans run one thread per column, the periodic halo fills one
// thread per point (the adjoint folds each halo point's adjoint onto its
// interior source with an atomic add).
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "atm_common.hpp"
#include "lin_step.hpp"

namespace {

#define LIN_POINT                                                     \
    const int I = r.I0 + blockIdx.x * blockDim.x + threadIdx.x;       \
    const int J = r.J0 + blockIdx.y * blockDim.y + threadIdx.y;       \
    const int K = k0 + blockIdx.z;                                    \
    if (I > r.I1 || J > r.J1) return;

template <class F>
__global__ void __launch_bounds__(256) k_lin_fwd(LinArgs x, Rng r, int k0) {
    LIN_POINT
    AccFwd a{x};
    F::run(a, x, I, J, K);
}

template <class F>
__global__ void __launch_bounds__(256) k_lin_tan(LinArgs x, Rng r, int k0) {
    LIN_POINT
    AccTan a{x};
    F::run(a, x, I, J, K);
}

// Adjoint body with one global atomic add per read and point.
struct DevAdd {
    __device__ static void add(double *p, double v) { unsafeAtomicAdd(p, v); }
};

template <class F>
__global__ void __launch_bounds__(256) k_lin_adj_direct(LinArgs x, Rng r, int k0) {
    LIN_POINT
    AccAdj<F::NR, DevAdd> a(x);
    F::run(a, x, I, J, K);
    a.flush();
}
#undef LIN_POINT

// ---------------------------------------------------------------------------
// Adjoint bodies accumulated in LDS.  A workgroup takes ADJ_X x ADJ_Y points
// of one level.  Each point evaluates its body on Jac<NR>; what it owes a
// value it read goes into an LDS plane kept for that operand and level
// offset, which covers the workgroup's points plus ADJ_H on every side.
// Neighbouring points read the same values (a stencil reads one array at up
// to 16 offsets), so their contributions sum in LDS, and every value that
// received any gets one global atomic add per plane, instead of one per
// read per point.  A read beyond the planes' reach, or of an operand past
// the first ADJ_PMAX planes, adds to global memory directly.
//
// Every read of a body names a compile-time slot, operand and offset from
// the point (lin_ops.hpp), so each thread derives the same plane table from
// its own slots.  Threads past the box run the body on the box's last point
// (valid memory) with their writes and additions switched off, so they keep
// the table and every barrier.
// ---------------------------------------------------------------------------
constexpr int ADJ_X = 32, ADJ_Y = 8, ADJ_H = 3, ADJ_PMAX = 6;

// An operation whose reads share few values between neighbouring points
// gains less from the planes than the planes cost; it declares
// ADJ_DIRECT = true and takes the direct kernel above.
template <class F, class = void>
struct AdjDirect : std::false_type {};
template <class F>
struct AdjDirect<F, std::void_t<decltype(F::ADJ_DIRECT)>> : std::integral_constant<bool, F::ADJ_DIRECT> {};
template <class F>
constexpr bool adj_in_lds() { return !AdjDirect<F>::value; }
constexpr int ADJ_W = ADJ_X + 2 * ADJ_H, ADJ_PL = ADJ_W * (ADJ_Y + 2 * ADJ_H);

template <int NR>
struct AccAdjTile {
    using S = Jac<NR>;
    const LinArgs &x;
    const int I, J, K;
    const bool active;
    int sarr[NR], sdi[NR], sdj[NR], sdk[NR];
    double acc[NR];
    __device__ AccAdjTile(const LinArgs &a, int I_, int J_, int K_, bool act)
        : x(a), I(I_), J(J_), K(K_), active(act) {
        for (int s = 0; s < NR; s++) acc[s] = 0.;
    }
    template <int s>
    __device__ S rd(int arr, int Ir, int Jr, int Kr) {
        static_assert(s < NR, "read slot");
        sarr[s] = arr;
        sdi[s] = Ir - I;
        sdj[s] = Jr - J;
        sdk[s] = Kr - K;
        return S::seed(x.iv[arr][x.g.ix(Ir, Jr, Kr)], s);
    }
    __device__ void wr(int o, int Iw, int Jw, int Kw, const S &v) {
        if (!active) return;
        const size_t i = x.g.ix(Iw, Jw, Kw);
        const double ab = x.ot[o][i];
        if (ab == 0.) return;
        x.ot[o][i] = 0.;
#pragma unroll
        for (int s = 0; s < NR; s++) acc[s] += ab * v.d[s];
    }
};

template <class F>
__global__ void __launch_bounds__(ADJ_X * ADJ_Y) k_lin_adj(LinArgs x, Rng r, int k0) {
    constexpr int NR = F::NR, NT = ADJ_X * ADJ_Y;
    __shared__ double pl[ADJ_PMAX * ADJ_PL];
    const int tid = threadIdx.y * ADJ_X + threadIdx.x;
    const int Ib = r.I0 + blockIdx.x * ADJ_X, Jb = r.J0 + blockIdx.y * ADJ_Y, K = k0 + blockIdx.z;
    const int Iq = Ib + threadIdx.x, Jq = Jb + threadIdx.y;
    const bool active = Iq <= r.I1 && Jq <= r.J1;
    AccAdjTile<NR> a(x, min(Iq, r.I1), min(Jq, r.J1), K, active);
    F::run(a, x, a.I, a.J, K);
    // plane of each slot: one per (operand, level offset) in slot order
    // (register arrays indexed only by unrolled loop counters)
    int np = 0, parr[ADJ_PMAX], pdk[ADJ_PMAX], sp[NR];
#pragma unroll
    for (int q = 0; q < ADJ_PMAX; q++) parr[q] = pdk[q] = 0;
#pragma unroll
    for (int s = 0; s < NR; s++) {
        int p = -1;
#pragma unroll
        for (int q = 0; q < ADJ_PMAX; q++)
            if (p < 0 && q < np && parr[q] == a.sarr[s] && pdk[q] == a.sdk[s]) p = q;
        if (p < 0 && np < ADJ_PMAX && x.it[a.sarr[s]]) {
#pragma unroll
            for (int q = 0; q < ADJ_PMAX; q++)
                if (q == np) {
                    parr[q] = a.sarr[s];
                    pdk[q] = a.sdk[s];
                }
            p = np++;
        }
        sp[s] = p;
    }
    for (int c = tid; c < np * ADJ_PL; c += NT) pl[c] = 0.;
    __syncthreads();  // planes cleared
#pragma unroll
    for (int s = 0; s < NR; s++) {
        const double v = a.acc[s];
        double *t = x.it[a.sarr[s]];
        if (!t || v == 0.) continue;
        const int px = (int)threadIdx.x + ADJ_H + a.sdi[s], py = (int)threadIdx.y + ADJ_H + a.sdj[s];
        if (sp[s] >= 0 && px >= 0 && px < ADJ_W && py >= 0 && py < ADJ_Y + 2 * ADJ_H)
            atomicAdd(&pl[sp[s] * ADJ_PL + py * ADJ_W + px], v);
        else
            unsafeAtomicAdd(t + x.g.ix(a.I + a.sdi[s], a.J + a.sdj[s], K + a.sdk[s]), v);
    }
    __syncthreads();  // contributions summed
#pragma unroll
    for (int q = 0; q < ADJ_PMAX; q++) {
        if (q >= np) break;
        double *t = x.it[parr[q]];
        for (int c = tid; c < ADJ_PL; c += NT) {
            const double v = pl[q * ADJ_PL + c];
            if (v != 0.)
                unsafeAtomicAdd(t + x.g.ix(Ib - ADJ_H + c % ADJ_W, Jb - ADJ_H + c / ADJ_W, K + pdk[q]), v);
        }
    }
}

template <bool GZ>
__global__ void k_lin_scan(int mode, Geo g, Rng r, const double *sv, double *sd, const double *tv, double *td,
                           double *pv, double *pd, double *kv, double *kd, double *gv, double *gd) {
    const int I = r.I0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = r.J0 + blockIdx.y * blockDim.y + threadIdx.y;
    if (I > r.I1 || J > r.J1) return;
    const size_t c0 = g.ix2(I, J);
    if (mode == LIN_FWD)
        scan_column<double, GZ>(g, c0, sv, nullptr, tv, nullptr, pv, nullptr, kv, nullptr, gv, nullptr);
    else if (mode == LIN_TAN)
        scan_column<Jac<1>, GZ>(g, c0, sv, sd, tv, td, pv, pd, kv, kd, gv, gd);
    else
        scan_column_adj<GZ>(g, c0, sd, tv, td, pv, pd, kv, kd, gd);
}

// column operations of the remap, one thread per column of each box;
// blockIdx.y picks the operation (independent operations run together)
struct LinColOps {
    LinColOp op[LIN_MAXCOL];
};

template <int KIND>
__global__ void __launch_bounds__(256) k_lin_column(int mode, LinColOps ops) {
    const LinColOp &o = ops.op[blockIdx.y];
    const Rng r = o.r;
    const LinArgs &x = o.x;
    const size_t W = (size_t)(r.I1 - r.I0 + 1), ncol = W * (size_t)(r.J1 - r.J0 + 1);
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncol) return;
    const int I = r.I0 + (int)(t % W), J = r.J0 + (int)(t / W);
    const RmWS w{o.w, ncol, t, x.g.L};
    const size_t c0 = x.g.ix2(I, J);
    if (KIND == RM_COL_REMAP) {
        if (mode == LIN_FWD) rm_column<double>(x, c0, w);
        else if (mode == LIN_TAN) rm_column<Jac<1>>(x, c0, w);
        else rm_column_adj(x, c0, w);
    } else {
        if (mode == LIN_FWD) rm_pt_column<double>(x, c0, w);
        else if (mode == LIN_TAN) rm_pt_column<Jac<1>>(x, c0, w);
        else rm_pt_column_adj(x, c0, w);
    }
}

// dst = periodic fill of src's interior over the whole halo'd plane
__global__ void k_lin_halo(Geo g, int nk, const double *__restrict__ sv, const double *__restrict__ st,
                           double *__restrict__ dv, double *__restrict__ dt) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I >= g.NI || J >= g.NJ || K >= nk) return;
    const size_t d = g.ix(I, J, K), s = g.ix(wrap_index(I, g.nx), wrap_index(J, g.ny), K);
    dv[d] = sv[s];
    if (dt) dt[d] = st[s];
}

// adjoint of the fill: each point's adjoint onto its interior source
__global__ void k_lin_halo_adj(Geo g, int nk, double *__restrict__ sd, double *__restrict__ dd) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y * blockDim.y + threadIdx.y;
    const int K = blockIdx.z;
    if (I >= g.NI || J >= g.NJ || K >= nk) return;
    const size_t d = g.ix(I, J, K);
    const double a = dd[d];
    if (a == 0.) return;
    dd[d] = 0.;
    unsafeAtomicAdd(sd + g.ix(wrap_index(I, g.nx), wrap_index(J, g.ny), K), a);
}

__global__ void k_lin_fill(double *a, size_t n, double v) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

// interior of src (nk levels) into dst
__global__ void k_lin_copy_interior(Geo g, int nk, const double *__restrict__ src, double *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, K = blockIdx.z;
    if (i >= g.nx) return;
    const size_t x = g.ix(ATM_HALO + i, ATM_HALO + j, K);
    dst[x] = src[x];
}

const dim3 LBLK(64, 4, 1);

inline dim3 lin_grid(const Rng &r, int nk) {
    return dim3((unsigned)((r.I1 - r.I0 + LBLK.x) / LBLK.x), (unsigned)((r.J1 - r.J0 + LBLK.y) / LBLK.y),
                (unsigned)nk);
}

}  // namespace

struct GpuBackend {
    hipStream_t s = nullptr;
    bool failed = false;  // an allocation failed

    double *alloc(size_t n) {
        double *p = nullptr;
        if (hipMalloc(&p, sizeof(double) * (n ? n : 1)) != hipSuccess) {
            failed = true;
            return nullptr;
        }
        fill(p, n, 0.);
        return p;
    }
    void release(double *p) {
        if (p) (void)hipFree(p);
    }
    void copy(double *d, const double *src, size_t n) {
        (void)hipMemcpyAsync(d, src, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
    }
    void fill(double *d, size_t n, double v) {
        if (n) hipLaunchKernelGGL(k_lin_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, n, v);
    }
    template <class F>
    void fwd(const LinArgs &x, Rng r, int k0, int nk) {
        hipLaunchKernelGGL(k_lin_fwd<F>, lin_grid(r, nk), LBLK, 0, s, x, r, k0);
    }
    template <class F>
    void tan(const LinArgs &x, Rng r, int k0, int nk) {
        hipLaunchKernelGGL(k_lin_tan<F>, lin_grid(r, nk), LBLK, 0, s, x, r, k0);
    }
    template <class F>
    void adj(const LinArgs &x, Rng r, int k0, int nk) {
        if (!adj_in_lds<F>()) {
            hipLaunchKernelGGL(k_lin_adj_direct<F>, lin_grid(r, nk), LBLK, 0, s, x, r, k0);
            return;
        }
        const dim3 grid((unsigned)((r.I1 - r.I0 + ADJ_X) / ADJ_X), (unsigned)((r.J1 - r.J0 + ADJ_Y) / ADJ_Y),
                        (unsigned)nk);
        hipLaunchKernelGGL(k_lin_adj<F>, grid, dim3(ADJ_X, ADJ_Y), 0, s, x, r, k0);
    }
    void scan(int mode, bool gz, const Geo &g, Rng r, const double *sv, double *sd, const double *tv, double *td,
              double *pv, double *pd, double *kv, double *kd, double *gv, double *gd) {
        const dim3 grid = lin_grid(r, 1);
        if (gz)
            hipLaunchKernelGGL(k_lin_scan<true>, grid, LBLK, 0, s, mode, g, r, sv, sd, tv, td, pv, pd, kv, kd, gv, gd);
        else
            hipLaunchKernelGGL(k_lin_scan<false>, grid, LBLK, 0, s, mode, g, r, sv, sd, tv, td, pv, pd, kv, kd, gv,
                               gd);
    }
    void columns(int mode, int kind, const LinColOp *ops, int n) {
        LinColOps a{};
        size_t ncol = 0;
        for (int i = 0; i < n; i++) {
            a.op[i] = ops[i];
            const Rng &r = ops[i].r;
            ncol = std::max(ncol, (size_t)(r.I1 - r.I0 + 1) * (size_t)(r.J1 - r.J0 + 1));
        }
        const dim3 grid((unsigned)((ncol + 255) / 256), (unsigned)n);
        if (kind == RM_COL_REMAP)
            hipLaunchKernelGGL(k_lin_column<RM_COL_REMAP>, grid, dim3(256), 0, s, mode, a);
        else
            hipLaunchKernelGGL(k_lin_column<RM_COL_PT>, grid, dim3(256), 0, s, mode, a);
    }
    void halo_copy(int mode, const Geo &g, int nk, const double *sv, double *sd, double *dv, double *dd) {
        const dim3 grid((unsigned)((g.NI + LBLK.x - 1) / LBLK.x), (unsigned)((g.NJ + LBLK.y - 1) / LBLK.y),
                        (unsigned)nk);
        if (mode == LIN_ADJ)
            hipLaunchKernelGGL(k_lin_halo_adj, grid, LBLK, 0, s, g, nk, sd, dd);
        else
            hipLaunchKernelGGL(k_lin_halo, grid, LBLK, 0, s, g, nk, sv, mode == LIN_TAN ? sd : nullptr, dv,
                               mode == LIN_TAN ? dd : nullptr);
    }
    void copy_interior(const Geo &g, int nk, const double *src, double *dst) {
        hipLaunchKernelGGL(k_lin_copy_interior, dim3((unsigned)((g.nx + 63) / 64), (unsigned)g.ny, (unsigned)nk),
                           dim3(64), 0, s, g, nk, src, dst);
    }
};

struct LinWork {
    GpuBackend be;
    LinCore<GpuBackend> core;
    double *tl[4] = {}, *ad[4] = {};  // tangent and adjoint of dp, pt, ud, vd (halo'd layout)
    int last = -1;                    // sweep kind the derivative arrays last held
    explicit LinWork(const Geo &g) : core(be, g) {}
};

double *atm_lin_field(atm_ctx *c, int id) {
    if (!c->lin) return nullptr;
    if (id >= ATM_TL_DP && id <= ATM_TL_VD) return c->lin->tl[id - ATM_TL_DP];
    if (id >= ATM_AD_DP && id <= ATM_AD_VD) return c->lin->ad[id - ATM_AD_DP];
    return nullptr;
}

void atm_lin_free(atm_ctx *c) {
    LinWork *w = c->lin;
    if (!w) return;
    w->core.release();
    for (int i = 0; i < 4; i++) {
        w->be.release(w->tl[i]);
        w->be.release(w->ad[i]);
    }
    delete w;
    c->lin = nullptr;
}

int atm_lin_alloc(atm_ctx *c) {
    if (c->lin) return ATM_OK;
    LinWork *w = new LinWork(c->g);
    w->be.s = c->stream;
    const size_t n = c->g.plane() * c->g.L;
    bool ok = w->core.allocate();
    for (int i = 0; ok && i < 4; i++) {
        w->tl[i] = w->be.alloc(n);
        w->ad[i] = w->be.alloc(n);
        ok = w->tl[i] && w->ad[i];
    }
    if (!ok) {
        c->lin = w;
        atm_lin_free(c);
        return atm_fail(c, ATM_ERR_HIP, "hipMalloc failed for the linearised step's arrays");
    }
    ATM_HIP(c, hipGetLastError());
    c->lin = w;
    return ATM_OK;
}

// Derivative arrays hold tangents in a tangent sweep and adjoints in an
// adjoint sweep; switching clears them (constant levels and the adjoint
// bookkeeping rely on zeros).
static void lin_prepare(atm_ctx *c, int kind) {
    LinWork *w = c->lin;
    LinCore<GpuBackend> &k = w->core;
    w->be.s = c->stream;
    k.c = c->coef;
    k.ak0 = c->ak0;
    k.pktop = c->pktop;
    k.ak = c->ak;
    k.bk = c->bk;
    if (w->last != kind) {
        for (int i = 0; i < A_COUNT; i++)
            if (k.drv[i]) w->be.fill(k.drv[i], k.count(i), 0.);
        w->last = kind;
    }
    // constants of this step: surface gz and its corner values, the net flux
    k.set_constants(c->gz, c->gzb, c->net_flux);
    // the start state
    const size_t n = c->g.plane() * c->g.L;
    const double *st[4] = {c->dp, c->pt, c->ud, c->vd};
    for (int i = 0; i < 4; i++) w->be.copy(k.val[A_XDP + i], st[i], n);
}

int atm_lin_tl_step(atm_ctx *c, double dt, int remap) {
    LinWork *w = c->lin;
    LinCore<GpuBackend> &k = w->core;
    lin_prepare(c, LIN_TAN);
    const size_t n = c->g.plane() * c->g.L;
    for (int i = 0; i < 4; i++) w->be.copy(k.drv[A_XDP + i], w->tl[i], n);
    k.sweep(LIN_TAN, dt, remap != 0);
    int end[4];
    lin_end_arrays(remap != 0, end);
    double *st[4] = {c->dp, c->pt, c->ud, c->vd};
    for (int i = 0; i < 4; i++) {
        w->be.copy(st[i], k.val[end[i]], n);
        w->be.copy(w->tl[i], k.drv[end[i]], n);
    }
    const int nk[] = {c->g.L, c->g.L, c->g.L, c->g.L};
    int rc = atm_launch_halo_multi(c, st, nk, 4);
    if (rc) return rc;
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}

int atm_lin_ad_step(atm_ctx *c, double dt, int remap) {
    LinWork *w = c->lin;
    LinCore<GpuBackend> &k = w->core;
    lin_prepare(c, LIN_ADJ);
    int end[4];
    lin_end_arrays(remap != 0, end);
    for (int i = 0; i < 4; i++) w->be.copy_interior(c->g, c->g.L, w->ad[i], k.drv[end[i]]);
    k.adjoint(dt, remap != 0);
    const size_t n = c->g.plane() * c->g.L;
    for (int i = 0; i < 4; i++) {
        w->be.copy(w->ad[i], k.drv[A_XDP + i], n);
        w->be.fill(k.drv[A_XDP + i], n, 0.);
    }
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}
