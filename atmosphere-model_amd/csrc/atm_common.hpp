// atm_common.hpp -- shared definitions of the gfx950 HIP library.
//
// Layout: every field is FP64 over the reference's halo'd index space,
// index (I, J, K) with I = i - is in [0, NI), J = j - js in [0, NJ),
// K = k - 1, NI = nx + 14, NJ = ny + 14.  On the device a row of NI values
// is stored with the stride SI = NI rounded up to 16 values (128 bytes), so
// that every row and plane starts on a cache line: element
// ((K*NJ)+J)*SI + I.  Host arrays at the boundary use SI = NI (the
// reference's own layout); the transfers convert.  Reference index "is+a"
// is I = a; "ie-b" is I = NI-1-b; the interior isd..ied is I = 7 .. 7+nx-1.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/atm.h"

#define ATM_HALO 7

// Physical constants of src/mod_constants.f90:8-14.  Default-real literals
// are binary32-rounded then widened (pi, R_dry); cp_dry is an integer.
#define ATM_PI ((double)3.1415926535897932384626433f)
#define ATM_R_DRY ((double)287.052874f)
#define ATM_CP_DRY 1005.0
#define ATM_CV_DRY (ATM_CP_DRY - ATM_R_DRY)
#define ATM_KAPPA (ATM_R_DRY / ATM_CP_DRY)
// single-precision constant expressions (7./12., 1./12., 2./3.) widened
#define ATM_W7_12 ((double)(7.0f / 12.0f))
#define ATM_W1_12 ((double)(1.0f / 12.0f))
#define ATM_W2_3 ((double)(2.0f / 3.0f))

enum { V_PCM = 0, V_PPM_U = 1, V_PPM_C = 2, V_PPM_NN = 3 };  // mod_sw_dyn.f90:73-76

struct Geo {
    int NI, NJ, L, nx, ny;
    int SI;  // row stride on the device, NI rounded up to a multiple of 16
    __host__ __device__ size_t plane() const { return (size_t)SI * NJ; }
    __host__ __device__ size_t ix(int I, int J, int K) const {
        return ((size_t)K * NJ + (size_t)J) * SI + (size_t)I;
    }
    __host__ __device__ size_t ix2(int I, int J) const { return (size_t)J * SI + (size_t)I; }
};

// Per-step scalars of mod_sw_dyn (init_sw_dyn 958-1023 and dt-derived).
struct DynCoef {
    double dt, dtdx, dtdy;
    double rdx, rdy, rdx2, rdy2, rdx2rdy2, rA;
    double vx, vy, divdamp, coriolis, gravity;
};

// ---------------------------------------------------------------------------
// Device numerics shared by all dynamics kernels (mod_sw_dyn.f90:781-919).
// ---------------------------------------------------------------------------
// Fortran MAX/MIN on finite values: one v_max / v_min each (the
// compare-and-select form costs three instructions).  They differ from a
// compare-and-select only for NaN operands and in the sign of an exact zero.
__host__ __device__ __forceinline__ double dmax2(double a, double b) { return __builtin_fmax(a, b); }
__host__ __device__ __forceinline__ double dmin2(double a, double b) { return __builtin_fmin(a, b); }
__host__ __device__ __forceinline__ float dmax2(float a, float b) { return __builtin_fmaxf(a, b); }
__host__ __device__ __forceinline__ float dmin2(float a, float b) { return __builtin_fminf(a, b); }
template <class T>
__host__ __device__ __forceinline__ T dmax3(T a, T b, T c) { return dmax2(dmax2(a, b), c); }
template <class T>
__host__ __device__ __forceinline__ T dmin3(T a, T b, T c) { return dmin2(dmin2(a, b), c); }
// Fortran SIGN(a, b)
__host__ __device__ __forceinline__ double fsign(double a, double b) { return copysign(fabs(a), b); }
__host__ __device__ __forceinline__ float fsign(float a, float b) { return copysignf(fabsf(a), b); }

// x / 3 rounded to nearest, without the division sequence: the product
// with RN(1/3) is corrected by one FMA step using the exact remainder
// x - 3q.  For every finite x away from the subnormal and overflow ranges
// this equals the IEEE quotient (checked on 1.6e9 random doubles spanning
// 2^-300..2^300 against x / 3.; in binary32 the same argument holds: x / 3
// is never a rounding midpoint); a zero product keeps the sign of x.
__host__ __device__ __forceinline__ double div3(double x) {
    const double r3 = 1.0 / 3.0;
    const double q = x * r3;
    if (q == 0.) return q;
    return fma(fma(-q, 3.0, x), r3, q);
}
__host__ __device__ __forceinline__ float div3(float x) {
    const float r3 = 1.0f / 3.0f;
    const float q = x * r3;
    if (q == 0.f) return q;
    return fmaf(fmaf(-q, 3.0f, x), r3, q);
}

// x**kappa, the reference's pressure-to-the-kappa (a libm pow of the
// double kappa = R_dry / cp_dry), for the column scans.  Written for this
// exponent from the textbook reduction (e.g. Muller, "Elementary
// Functions", ch. 11 and 13):
//   ln x = e ln2 + ln m,  x = m 2^e with m in [1/sqrt2, sqrt2),
//   ln m = 2 atanh(f) = 2f + f^3 (2/3 + 2/5 f^2 + ... + 2/23 f^20),
//          f = (m-1)/(m+1), |f| < 0.1716 (truncation below 2^-60),
//   y = kappa ln x as a double-double (ln2 in two parts, Fast2Sum),
//   e^y = 2^n (1 + p), n = rint(y / ln2), p = e^r - 1 by the Taylor
//          series to degree 13 in r = y - n ln2 (|r| <= ln2/2, Cody-Waite).
// The result is within one ulp of the exact power (checked on the host
// against long-double powl over the pressures' range, where it also
// agrees with the libm pow in all but a small fraction of cases, each one
// ulp apart), in about a quarter of the general pow's instructions.
// x = 0 gives 0, x = +inf gives +inf, x < 0 and NaN give NaN.
// N values at once, each step over all N before the next: a coefficient
// is brought into registers once per step instead of once per value.
// Each result is pow_kappa's, operation for operation.
template <int N>
__host__ __device__ inline void pow_kappa_n(double (&x)[N]) {
    constexpr double kappa = ATM_KAPPA;
    constexpr double ln2_hi = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000: e ln2_hi exact for |e| < 2^11
    constexpr double ln2_lo = 1.90821492927058770002e-10;
    constexpr double inv_ln2 = 1.44269504088896338700e+00;
    constexpr double cl[10] = {2. / 21., 2. / 19., 2. / 17., 2. / 15., 2. / 13., 2. / 11.,
                               2. / 9.,  2. / 7.,  2. / 5.,  2. / 3.};
    constexpr double ce[11] = {1. / 479001600., 1. / 39916800., 1. / 3628800., 1. / 362880.,
                               1. / 40320.,     1. / 5040.,     1. / 720.,     1. / 120.,
                               1. / 24.,        1. / 6.,        .5};
    double f[N], f2[N], q[N], v[N];
    int e[N];
    for (int i = 0; i < N; i++) {
        double m = frexp(x[i], &e[i]);  // [0.5, 1)
        if (m < 0.70710678118654752440) {
            m = m + m;
            e[i] = e[i] - 1;
        }
        f[i] = (m - 1.) / (m + 1.);  // m - 1 exact
        f2[i] = f[i] * f[i];
        q[i] = 2. / 23.;
    }
    for (int c = 0; c < 10; c++)
        for (int i = 0; i < N; i++) q[i] = fma(q[i], f2[i], cl[c]);
    double yh[N], yl[N];
    for (int i = 0; i < N; i++) {
        const double lnm = fma(f[i] * f2[i], q[i], f[i] + f[i]);
        // ln x = s + lo
        const double ed = (double)e[i];
        const double a = ed * ln2_hi;
        const double s = a + lnm;
        const double lo = (lnm - (s - a)) + ed * ln2_lo;  // Fast2Sum error (|a| >= |lnm| or a = 0)
        yh[i] = kappa * s;
        yl[i] = fma(kappa, s, -yh[i]) + kappa * lo;
    }
    // e^(yh + yl) = 2^n e^(rh + rl): rh exact (Cody-Waite), 1 + rh exact
    // as h + l (Fast2Sum), the rest added to l, one rounding at the end
    double n[N], rh[N];
    for (int i = 0; i < N; i++) {
        n[i] = rint(yh[i] * inv_ln2);
        rh[i] = fma(-n[i], ln2_hi, yh[i]);
        q[i] = 1. / 6227020800.;  // 1/13!
    }
    for (int c = 0; c < 11; c++)
        for (int i = 0; i < N; i++) q[i] = fma(q[i], rh[i], ce[c]);
    for (int i = 0; i < N; i++) {
        const double rl = yl[i] - n[i] * ln2_lo;
        const double h = 1. + rh[i];
        const double l = (1. - h) + rh[i];
        const double t = rh[i] * rh[i] * q[i];  // e^rh = h + l + t
        // n is clamped before its conversion to int, which would be undefined
        // for the NaN of x < 0 or x = NaN (whose result is replaced below);
        // |n| < 310 for every finite positive x, so no value changes
        v[i] = ldexp(h + (t + fma(rl, h + t, l)), (int)dmin2(dmax2(n[i], -2000.), 2000.));
        const double xi = x[i];
        x[i] = (xi > 0. && xi < __builtin_inf()) ? v[i] : xi == 0. ? 0. : xi == __builtin_inf() ? xi : __builtin_nan("");
    }
}

__host__ __device__ inline double pow_kappa(double x) {
    double a[1] = {x};
    pow_kappa_n<1>(a);
    return a[0];
}

// report a non-finite result value into the stability word (see atm_ctx::nfw)
__device__ __forceinline__ void atm_report_nonfinite(int *nfw, double v) {
    if (!isfinite(v)) atomicOr(nfw, 1);
}

// The PPM operators below are templates on the arithmetic type T: double
// for the FP64 path (the reference's operations and rounding), float for
// the mixed-precision path's reconstruction and fluxes.

// dqi_mono of the middle cell from (q(m-1), q(m), q(m+1)): ppm 870-875
template <class T>
__host__ __device__ __forceinline__ T dq_mono(T qm, T q0, T qp) {
    T d = (qp - qm) / T(4);
    T dmax = dmax3(qp, q0, qm) - q0;
    T dmin = q0 - dmin3(qp, q0, qm);
    return fsign(dmin3(T(fabs(d)), dmin, dmax), d);
}

// Edge anomalies of the middle cell from q(m-1..m+1) and the dqi_mono of
// cells m-1, m, m+1 (ppm 877-911).  Each cell's dqi_mono depends only on
// its own three values, so a tile can form it once per cell and share it
// between the faces that need it; ppm_edges below forms it in place.
template <int V, class T>
__host__ __device__ __forceinline__ void ppm_edges_dq(T qm1, T q0, T qp1, T dmm, T dm0, T dmp, T &eL, T &eR) {
    eL = (qm1 - q0) / T(2) + div3(dmm - dm0);
    eR = (qp1 - q0) / T(2) + div3(dm0 - dmp);
    if (V == V_PPM_C || V == V_PPM_NN) {
        T d = (qp1 - qm1) / T(4);
        eL = -fsign(dmin2(T(2) * T(fabs(d)), T(fabs(eL))), d);
        eR = +fsign(dmin2(T(2) * T(fabs(d)), T(fabs(eR))), d);
        if (V == V_PPM_NN) {
            eL = dmax2(eL, -q0);
            eR = dmax2(eR, -q0);
            T a = eR - eL;
            T b = T(-3) * (eR + eL);
            // The test below holds only where q0 <= |b| / 3: |a| < -b makes
            // a^2 / (4b) > b / 4.  Where q0 > 0.34 |b| the computed sum is
            // >= 0 in either precision (0.34 - 1/4 - 1/12 leaves far more
            // than the rounding of three operations), so those cells (almost
            // all of a smooth dp or pt field) skip the division; a NaN fails
            // this pre-test and takes the full test as before.
            // b / 12 as div3(b / 4): the quarter is exact, so this is the
            // correctly rounded b / 12 without a division sequence
            if (!(q0 > T(0.34) * T(fabs(b))) && (T(fabs(a)) < -b) &&
                (q0 + a * a / (T(4) * b) + div3(b * T(0.25)) < T(0))) {
                if (eL * eR > T(0)) {
                    eL = T(0);
                    eR = T(0);
                } else if (eL > eR) {
                    eL = T(2) * T(fabs(eR));
                } else {
                    eR = T(2) * T(fabs(eL));
                }
            }
        }
    }
}

// Edge anomalies of the middle cell from q(m-2..m+2): ppm 877-911.
template <int V, class T>
__host__ __device__ __forceinline__ void ppm_edges(T qm2, T qm1, T q0, T qp1, T qp2, T &eL, T &eR) {
    ppm_edges_dq<V>(qm1, q0, qp1, dq_mono(qm2, qm1, q0), dq_mono(qm1, q0, qp1), dq_mono(q0, qp1, qp2), eL, eR);
}

// Upwind PPM value at the face between cells i-1 and i (flux 802-829) from
// q(i-2..i+1), the dqi_mono of cells i-2..i+1 (dm2, dm1, d0, dp1) and the
// Courant number c(i).  q(i-3) and q(i+2) enter the reference's flux only
// through dqi_mono(i-2) and dqi_mono(i+1).
template <int V, class T>
__host__ __device__ __forceinline__ T ppm_face_dq(T qm2, T qm1, T q0, T qp1, T dm2, T dm1, T d0, T dp1, T c) {
    T s = T(fabs(c));
    T k1 = (T(1) - T(2) * s) * (T(1) - s);
    T k2 = s * (s - T(1));
    if (V == V_PPM_C) {
        T eLm, eRm, eL0, eR0;
        ppm_edges_dq<V>(qm2, qm1, q0, dm2, dm1, d0, eLm, eRm);
        ppm_edges_dq<V>(qm1, q0, qp1, dm1, d0, dp1, eL0, eR0);
        if ((T(3) * T(fabs(eL0 + eR0)) > T(fabs(eL0 - eR0))) && (T(3) * T(fabs(eLm + eRm)) > T(fabs(eLm - eRm))))
            return c > T(0) ? qm1 : q0;
        if (c > T(0)) return qm1 + k1 * eRm + k2 * eLm;
        return q0 + k1 * eL0 + k2 * eR0;
    } else {
        // only the upwind cell's edges are needed
        bool up = c > T(0);
        T a1 = up ? qm2 : qm1, a0 = up ? qm1 : q0, b1 = up ? q0 : qp1;
        T da = up ? dm2 : dm1, db = up ? dm1 : d0, dc = up ? d0 : dp1;
        T eL, eR;
        ppm_edges_dq<V>(a1, a0, b1, da, db, dc, eL, eR);
        if (up) return a0 + k1 * eR + k2 * eL;
        return a0 + k1 * eL + k2 * eR;
    }
}

// The same from q(i-3..i+2), each dqi_mono formed here.
template <int V, class T>
__host__ __device__ __forceinline__ T ppm_face(T qm3, T qm2, T qm1, T q0, T qp1, T qp2, T c) {
    if (V == V_PPM_C)
        return ppm_face_dq<V>(qm2, qm1, q0, qp1, dq_mono(qm3, qm2, qm1), dq_mono(qm2, qm1, q0),
                              dq_mono(qm1, q0, qp1), dq_mono(q0, qp1, qp2), c);
    // only the upwind cell's edges are needed
    T s = T(fabs(c));
    T k1 = (T(1) - T(2) * s) * (T(1) - s);
    T k2 = s * (s - T(1));
    bool up = c > T(0);
    T a2 = up ? qm3 : qm2, a1 = up ? qm2 : qm1, a0 = up ? qm1 : q0;
    T b1 = up ? q0 : qp1, b2 = up ? qp1 : qp2;
    T eL, eR;
    ppm_edges<V>(a2, a1, a0, b1, b2, eL, eR);
    if (up) return a0 + k1 * eR + k2 * eL;
    return a0 + k1 * eL + k2 * eR;
}

// Piecewise-constant upwind value at the face between cells i-1 and i: flux 832-838.
template <class T>
__host__ __device__ __forceinline__ T pcm_face(T qm1, T q0, T c) { return c > T(0) ? qm1 : q0; }

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct RemapWork;  // k_remap.hip
struct RadWork;    // k_rad.hip
struct PhysWork;   // k_physics.hip
struct LinWork;    // k_lin.hip

// kernels timed with HIP events when timing is enabled (atm_timer_*)
enum {
    ATM_T_C1 = 0, ATM_T_C2, ATM_T_D1, ATM_T_D2, ATM_T_SCAN, ATM_T_HALO, ATM_T_REMAP,
    ATM_T_RAD, ATM_T_PHYS, ATM_T_COUNT
};
struct AtmTimer {
    hipEvent_t *ev = nullptr;  // pairs (start, stop)
    int cap = 0, used = 0;
};

// What one launch of a recorded graph changes on a context's host side:
// the buffer parity it was recorded on (and leaves), the accumulation time
// steps in the recorded order, playkap's validity after the recorded steps
// and whether every recorded write of dp, pt, ud, vd reports into the
// stability word.
struct GraphEffects {
    int cur = 0;
    std::vector<double> dts;
    bool plk_end = false, tracks = false;
};

struct atm_ctx {
    Geo g;
    double dx, dy;
    atm_params prm;
    DynCoef coef;  // dt-independent part filled by init_sw_dyn
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    bool have_coord = false, sw_init = false, remap_init = false, rad_init = false, phys_init = false;
    double ak0 = 0;    // Ak(1) = top pressure
    double pktop = 0;  // Ak(1)**kappa, evaluated on the host as the reference does
    // execution path: fused LDS-tiled kernels (the default) or the
    // reference-structured kernels of k_dyn_ref.hip
    bool fused = false;
    // mixed precision (fused path only): the tile kernels reconstruct and
    // form fluxes in float, update the fields in double
    bool mixed = false;
    // prognostic fields: two copies each for the fused path; dp/pt/ud/vd
    // below point at the current copy, prv at the previous one
    double *pbuf[4][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    int cur = 0;
    bool after_d = false;  // a fused D step ran and its halo fill is pending
    double *uc2 = nullptr, *vc2 = nullptr, *heatd = nullptr;  // fused-path C-step outputs
    int bc_levels = 4;  // fused path: levels per block of the corner kernel (atm_autotune)
    // fused path: rows of the dp/pt tiles (8 or 16) and of the wind tiles
    // (16 or 8), each 32 cells wide (atm_set_tiles, atm_autotune)
    int tile_pair = 8, tile_wind = 16;
    // fused path: playkap of the current dp on the interior (the C step's
    // heating rate needs it); left by the D step's scan, recomputed by the
    // C step when dp changed since (remap, put)
    double *plk = nullptr;
    bool plk_valid = false;
    bool timing = false;
    unsigned timer_mask = ~0u;  // the kernel groups timed while timing (bit per ATM_T_*)
    AtmTimer timers[ATM_T_COUNT];
    double *ak = nullptr, *bk = nullptr;  // device, nlay+1
    // mod_fields
    double *dp = nullptr, *pt = nullptr, *ud = nullptr, *vd = nullptr, *ts = nullptr;
    double *gz = nullptr, *net_flux = nullptr, *plev = nullptr, *pkap = nullptr;
    double *play = nullptr, *playkap = nullptr, *radius = nullptr;
    // mod_sw_dyn module arrays
    double *dpc = nullptr, *ptc = nullptr, *uc = nullptr, *vc = nullptr, *heat = nullptr;
    double *pkapb = nullptr, *gzb = nullptr;
    // per-level 2-D work arrays, one plane per level
    double *tmp = nullptr, *lap = nullptr, *cx = nullptr, *cy = nullptr, *denx = nullptr,
           *deny = nullptr, *dpfx = nullptr, *dpfy = nullptr, *fxq = nullptr, *fyq = nullptr,
           *vort = nullptr, *ke = nullptr, *pgfx = nullptr, *pgfy = nullptr, *divg = nullptr;
    double *ua = nullptr, *va = nullptr, *ub = nullptr, *vb = nullptr;  // 2-D (last level)
    // every field buffer of the context with its size in doubles (without
    // the tail), for atm_poison
    std::vector<std::pair<double *, size_t>> field_bufs;
    int *flag = nullptr;        // device status word
    int *flag_host = nullptr;   // pinned mirror
    // Stability tracking.  Every kernel that writes interior dp, pt, ud or
    // vd as a step's result (D step of the fused path, remap, physics)
    // sets bit 0 of nfw when a value it writes is not finite.  While
    // nfw_tracked holds, the last check found the state finite and every
    // write since went through such a kernel, so nfw == 0 means the state
    // is finite; otherwise the check scans the four fields.
    int *nfw = nullptr;
    bool nfw_tracked = false;
    // bumped whenever nfw_tracked is cleared (atm_untrack), so that a
    // deferred check started before knows tracking was lost since
    unsigned long long untrack_epoch = 0;
    // the deferred check (atm_check_stability_begin / _end)
    int *chk_host = nullptr;  // pinned: the word the check copied
    hipEvent_t chk_ev = nullptr;
    bool chk_pending = false, chk_scan = false;
    unsigned long long chk_epoch = 0;
    RemapWork *remap = nullptr;
    RadWork *rad = nullptr;
    PhysWork *phys = nullptr;
    LinWork *lin = nullptr;  // tangent-linear / adjoint step (atm_lin_init)
    // HIP graph of a recorded step sequence (atm_graph_begin / end, or a
    // batched ensemble's recording): while capturing, the entry points that
    // wait on the device or write memory from the host refuse; cap_cur is
    // the buffer parity at the start.  Host-side state that recorded calls
    // would advance is kept aside while recording (cap_dts collects the
    // accumulation time steps) and applied by each launch instead (gfx).
    bool capturing = false;
    int cap_cur = 0;
    bool cap_timing = false, cap_plk_valid = false, cap_nfw_tracked = false;
    std::vector<double> cap_dts;
    GraphEffects gfx;
    hipGraphExec_t gexec = nullptr;
    // time-mean output accumulators (nx, ny, nlay+1, 6) and their time
    double *outf = nullptr;
    double acc_time = 0.;
    bool out_init = false;
};

// error helpers
int atm_fail(atm_ctx *c, int code, const std::string &msg);
#define ATM_HIP(c, call)                                                                 \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return atm_fail((c), ATM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Workgroups are dispatched round-robin over the 8 XCDs of the part, each
// with its own L2.  For a tiled stencil launch this maps the hardware block
// index to a logical tile so that each XCD receives one contiguous range of
// tiles (neighbouring tiles, whose halos overlap, then share an L2).
__device__ __forceinline__ void xcd_tile(int gx, int gy, int gz, int &bx, int &by, int &bz) {
    const int n = gx * gy * gz;
    int b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if ((n & 7) == 0) b = (b & 7) * (n >> 3) + (b >> 3);
    bx = b % gx;
    b /= gx;
    by = b % gy;
    bz = b / gy;
    // uniform over the workgroup: scalar registers for the tile origin and
    // the level offset derived from them (the divisions above run on the
    // vector unit, whose results the compiler cannot otherwise prove uniform)
    bx = __builtin_amdgcn_readfirstlane(bx);
    by = __builtin_amdgcn_readfirstlane(by);
    bz = __builtin_amdgcn_readfirstlane(bz);
}

// launch helpers: 2-D tiles over I x J, z = level
// Workgroup index -> work item so that each XCD gets a contiguous range:
// the dispatcher hands workgroups to the 8 XCDs in turn, so the m-th
// workgroup an XCD x receives is workgroup 8 m + x; with n = 8 q + r work
// items, XCDs 0..r-1 take q + 1 consecutive items and the others q.  A
// bijection of [0, n); neighbouring items then share their XCD's L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned n) {
    const unsigned x = b % 8u, m = b / 8u, q = n / 8u, r = n % 8u;
    return (x < r ? x * (q + 1u) : r * (q + 1u) + (x - r) * q) + m;
}

static inline dim3 grid3(int nI, int nJ, int nK, dim3 b) {
    return dim3((nI + b.x - 1) / b.x, (nJ + b.y - 1) / b.y, nK);
}

// kernels implemented in the .hip files (host launchers)
int atm_launch_halo(atm_ctx *c, double *a, int nk);
#define ATM_HALO_MAXF 6
// periodic halo fill of n <= ATM_HALO_MAXF fields (nk[i] levels each) in one launch
// the state changed outside the kernels that report into nfw: the next
// stability check scans the fields
inline void atm_untrack(atm_ctx *c) {
    c->nfw_tracked = false;
    c->untrack_epoch++;
}
int atm_launch_halo_multi(atm_ctx *c, double *const *a, const int *nk, int n);
// fills the poisoned classes (ATM_POISON_*) of a field buffer of n doubles
// (whole planes, or fewer than one plane: then only its tail)
int atm_launch_poison(atm_ctx *c, double *buf, size_t n, int what, double value);
int atm_ref_cgrid(atm_ctx *c, double dt);
int atm_ref_dgrid(atm_ctx *c, double dt);
int atm_ref_init_sw_dyn(atm_ctx *c);
int atm_remap_alloc(atm_ctx *c);
void atm_remap_free(atm_ctx *c);
int atm_remap_run(atm_ctx *c);
int atm_phys_alloc(atm_ctx *c);
void atm_phys_free(atm_ctx *c);
int atm_phys_tendencies(atm_ctx *c);
int atm_phys_apply(atm_ctx *c, double dt);
int atm_output_alloc(atm_ctx *c);
int atm_output_accumulate(atm_ctx *c, double dt);
int atm_output_take_mean(atm_ctx *c, double *host);
int atm_lin_alloc(atm_ctx *c);
void atm_lin_free(atm_ctx *c);
double *atm_lin_field(atm_ctx *c, int id);
int atm_lin_tl_step(atm_ctx *c, double dt, int remap);
int atm_lin_ad_step(atm_ctx *c, double dt, int remap);
int atm_rad_alloc(atm_ctx *c);
void atm_rad_free(atm_ctx *c);
int atm_rad_run(atm_ctx *c, double time);
int atm_launch_stability(atm_ctx *c);
int atm_launch_fill(atm_ctx *c, double *a, size_t n, double v);
void atm_timer_begin(atm_ctx *c, int id);
void atm_timer_end(atm_ctx *c, int id);
int atm_remap_run_split(atm_ctx *c, const double *dp, const double *pt, const double *ud_in,
                        const double *vd_in, const double *ud_ring, const double *vd_ring, double *dp_out,
                        double *pt_out, double *ud_out, double *vd_out);
