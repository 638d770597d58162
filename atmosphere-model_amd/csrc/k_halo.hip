// k_halo.hip -- periodic halo fill and the stability check (gfx950).
//
// halo_exchange on one image (src/mod_sync.f90:41-217) copies interior
// strips into the 7-wide halo of the same tile: every halo cell (i, j)
// receives q(wrap(i), wrap(j)), edges and corners alike.  One thread per
// halo cell; the halo cells of a level are enumerated as two full-height
// side strips (7 x NJ each, corners included) and two interior-width
// bottom/top strips (nx x 7 each).  Reads touch the interior only and
// writes the halo only, so the fill is race-free in one launch.
#include "atm_common.hpp"
#include "atm_fused.hpp"

namespace {

// up to ATM_HALO_MAXF fields in one launch (grid z = field)
struct HaloSet {
    double *a[ATM_HALO_MAXF];
    int nk[ATM_HALO_MAXF];
};

// halo cell t of level K of a (nk levels)
__device__ __forceinline__ void halo_cell(const Geo &g, double *__restrict__ a, int nk, int t, int K) {
    const int H = ATM_HALO;
    const int side = H * g.NJ;       // per side strip
    const int tb = g.nx * H;         // per top/bottom strip
    const int per = 2 * side + 2 * tb;
    if (t >= per || K >= nk) return;
    int I, J;
    if (t < 2 * side) {
        int s = t / side, r = t - s * side;
        J = r / H;
        int m = r - J * H;
        I = s == 0 ? m : H + g.nx + m;
    } else {
        int r = t - 2 * side;
        int s = r / tb;
        r -= s * tb;
        int m = r / g.nx;
        I = H + (r - m * g.nx);
        J = s == 0 ? m : H + g.ny + m;
    }
    int Is = I < H ? I + g.nx : (I >= H + g.nx ? I - g.nx : I);
    int Js = J < H ? J + g.ny : (J >= H + g.ny ? J - g.ny : J);
    a[g.ix(I, J, K)] = a[g.ix(Is, Js, K)];
}

__global__ void k_halo_fill(Geo g, HaloSet set) {
    const int f = blockIdx.z;
    halo_cell(g, set.a[f], set.nk[f], blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// The halo fills of a batched ensemble stage (atm_fused.hpp): the fields of
// `kind` for every member, each as the member's own fill launch does them.
__device__ __forceinline__ unsigned ens_halo_strip_blocks(const Geo &g) {
    return (unsigned)((2 * ATM_HALO * g.NJ + 2 * g.nx * ATM_HALO + 255) / 256);
}
__global__ void __launch_bounds__(256) k_halo_ens(const EnsMember *__restrict__ tab, EnsIndex ix, int nm,
                                                 unsigned long long par, int kind) {
    int m;
    unsigned l;
    if (!ens_find(ix, nm, m, l)) return;
    const EnsMember &e = tab[m];
    const unsigned nb = ens_halo_strip_blocks(e.g);
    const int bx = (int)(l % nb), K = (int)((l / nb) % (unsigned)e.g.L), f = (int)(l / nb / (unsigned)e.g.L);
    const int p = (int)((par >> m) & 1ull), L = e.g.L;
    double *a;
    int nk = L;
    if (kind == ENS_HALO_C1) {  // after C1: dpc, ptc, uc, vc and the last level's A winds
        a = f == 0 ? e.dpc : f == 1 ? e.ptc : f == 2 ? e.uc : f == 3 ? e.vc : f == 4 ? e.ua : e.va;
        if (f >= 4) nk = 1;
    } else if (kind == ENS_HALO_CH) {  // cgrid_halo_exchange: the updated C winds
        a = f == 0 ? e.uc2 : e.vc2;
    } else if (kind == ENS_HALO_D1) {  // after D1: the copy of dp, pt it wrote
        a = e.p[f][1 - p];
    } else {  // dgrid_halo_exchange: the current dp, pt, ud, vd
        a = e.p[f][p];
    }
    halo_cell(e.g, a, nk, bx * 256 + (int)threadIdx.x, K);
}

// check_stability (mod_fields.f90:69-106): flag bit set when any interior
// value of dp, pt, ud, vd is not finite.
__global__ void k_finite(Geo g, const double *__restrict__ a, const double *__restrict__ b,
                         const double *__restrict__ c, const double *__restrict__ d, int *flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int K = blockIdx.z;
    if (i >= g.nx) return;
    size_t x = g.ix(i + ATM_HALO, j + ATM_HALO, K);
    bool bad = !isfinite(a[x]) || !isfinite(b[x]) || !isfinite(c[x]) || !isfinite(d[x]);
    if (bad) atomicOr(flag, 1);
}

// atm_poison over the planes of one buffer: the row padding (i >= NI)
// and/or the halo cells farther than 4 from the interior
__global__ void k_poison_planes(Geo g, double *__restrict__ a, int what, double value) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y;
    const size_t K = blockIdx.z;
    if (i >= g.SI) return;
    const int far = ATM_HALO - 4;  // halo cells [0, far) and past the interior + 4
    bool hit;
    if (i >= g.NI)
        hit = (what & ATM_POISON_PADDING) != 0;
    else
        hit = (what & ATM_POISON_OUTER_HALO) != 0 &&
              (i < far || i >= ATM_HALO + g.nx + 4 || J < far || J >= ATM_HALO + g.ny + 4);
    if (hit) a[K * g.plane() + (size_t)J * g.SI + i] = value;
}

__global__ void k_poison_span(double *__restrict__ a, size_t n, double value) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = value;
}

}  // namespace

int atm_launch_poison(atm_ctx *c, double *buf, size_t n, int what, double value) {
    const Geo &g = c->g;
    const size_t pl = g.plane();
    if (n >= pl && n % pl == 0) {
        const dim3 b(256), gr((g.SI + 255) / 256, g.NJ, (unsigned)(n / pl));
        hipLaunchKernelGGL(k_poison_planes, gr, b, 0, c->stream, g, buf, what, value);
    }
    if (what & ATM_POISON_PADDING) {  // the tail past the last plane
        const size_t m = atm_tile_slack(g);
        hipLaunchKernelGGL(k_poison_span, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->stream, buf + n, m,
                           value);
    }
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}

int atm_launch_halo_multi(atm_ctx *c, double *const *a, const int *nk, int n) {
    const Geo &g = c->g;
    if (n < 1 || n > ATM_HALO_MAXF) return atm_fail(c, ATM_ERR_ARGUMENT, "halo fill: bad field count");
    HaloSet set{};
    int kmax = 0;
    for (int i = 0; i < n; i++) {
        set.a[i] = a[i];
        set.nk[i] = nk[i];
        kmax = nk[i] > kmax ? nk[i] : kmax;
    }
    int per = 2 * ATM_HALO * g.NJ + 2 * g.nx * ATM_HALO;
    dim3 b(256), gr((per + 255) / 256, kmax, n);
    hipLaunchKernelGGL(k_halo_fill, gr, b, 0, c->stream, g, set);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}

unsigned atm_ens_blocks_halo(const Geo &g, int kind) {
    const int nf = kind == ENS_HALO_C1 ? 6 : kind == ENS_HALO_DH ? 4 : 2;
    return (unsigned)((2 * ATM_HALO * g.NJ + 2 * g.nx * ATM_HALO + 255) / 256) * (unsigned)g.L * (unsigned)nf;
}

void atm_ens_launch_halo(hipStream_t s, const EnsMember *tab, const EnsIndex &ix, int nm, unsigned long long par,
                         int kind, unsigned n) {
    hipLaunchKernelGGL(k_halo_ens, dim3(n), dim3(256), 0, s, tab, ix, nm, par, kind);
}

int atm_launch_halo(atm_ctx *c, double *a, int nk) {
    return atm_launch_halo_multi(c, &a, &nk, 1);
}

int atm_launch_stability(atm_ctx *c) {
    const Geo &g = c->g;
    ATM_HIP(c, hipMemsetAsync(c->flag, 0, sizeof(int), c->stream));
    dim3 b(256), gr((g.nx + 255) / 256, g.ny, g.L);
    hipLaunchKernelGGL(k_finite, gr, b, 0, c->stream, g, c->dp, c->pt, c->ud, c->vd, c->flag);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}
