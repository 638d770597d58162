// k_dyn_fused.hip -- fused, LDS-tiled dynamics kernels (gfx950, FP64).
//
// The C-grid and D-grid steps of src/mod_sw_dyn.f90 as four tile kernels
// per step (C1, C2, D1, D2) plus column scans, instead of one kernel per
// reference loop nest.  A workgroup of TX x TY threads owns a TX x TY tile of
// one level; every intermediate of the reference's level loop (C winds,
// denominators, inner PCM steps, Laplacians, outer PPM fluxes, vorticity,
// kinetic energy, divergence damping, B-grid pressure gradient) lives in LDS
// over the tile plus the halo it needs, so HBM sees only each kernel's
// inputs and outputs.
//
// Every value is computed with the same expression, operand order and
// rounding as the reference (and as k_dyn_ref.hip), so the two paths agree
// bit for bit.  Tile-local recomputation of an intermediate in the halo of
// neighbouring tiles repeats identical arithmetic on identical inputs.
//
// The reference evaluates phase 1 of each step on the ring around the
// interior as well; those ring values are periodic images of interior values
// because every input is periodic there.  Here each kernel writes the
// interior only and a periodic halo fill provides the ring, which is the
// same data.
//
// Tile coordinates: x = I - I0, y = J - J0 with (I0, J0) the tile origin in
// the halo'd index space; LDS arrays cover [x0, x1) x [y0, y1) regions.
//
// Grids of any nx, ny: the last tile of a row or column may overhang the
// domain.  Its threads past the domain run the same stages on whatever the
// memory holds there (the halo, the row padding, the next row, or the tail
// every buffer carries, atm_tile_slack), which only ever feeds cells past
// the domain, and store nothing: every store, and every report to the
// stability word, is for a cell inside the domain.  The stencils of cells
// inside reach at most 4 cells out, all inside the halo of 7.
#include <type_traits>

#include "atm_common.hpp"
#include "atm_fused.hpp"

namespace {

// Arithmetic type R of the tile kernels: double on the FP64 path; float on
// the mixed-precision path for the reconstruction, the fluxes and the other
// intermediates in LDS.  The final updates of dp, pt, uc, vc, ud, vd, the
// heating rate and the pressure gradient stay in double (each face flux is
// one value used by both cells it separates, so dp changes by a telescoping
// sum: mass is conserved to FP64 rounding).  Every R expression below is
// the FP64 path's expression when R is double.
template <class R>
struct CoefR {
    R dt, dtdx, dtdy, rdx, rdy, rdx2, rdy2, rdx2rdy2, rA, vx, vy, divdamp, coriolis;
    __device__ explicit CoefR(const DynCoef &c)
        : dt(R(c.dt)), dtdx(R(c.dtdx)), dtdy(R(c.dtdy)), rdx(R(c.rdx)), rdy(R(c.rdy)), rdx2(R(c.rdx2)),
          rdy2(R(c.rdy2)), rdx2rdy2(R(c.rdx2rdy2)), rA(R(c.rA)), vx(R(c.vx)), vy(R(c.vy)), divdamp(R(c.divdamp)),
          coriolis(R(c.coriolis)) {}
};

// 1 / x: the IEEE quotient in double; v_rcp_f32 (1 ulp) in float
__device__ __forceinline__ double rcp(double x) { return 1. / x; }
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

#define D(x) static_cast<double>(x)

constexpr int TX = ATM_TILE_X;
// Tile rows per kernel: the dp/pt kernels (the heaviest in registers) run
// 32 x 8 tiles, the wind kernels and the corner kernel 32 x 16 (less halo
// recomputation per cell).  Each kernel declares its own TY and NT (one
// thread per output cell).
constexpr int TY_PAIR = ATM_TILE_Y_PAIR;
constexpr int TY_WIND = ATM_TILE_Y;
constexpr int NT_PAIR = TX * TY_PAIR, NT_WIND = TX * TY_WIND;
static_assert((NT_PAIR & (NT_PAIR - 1)) == 0 && NT_PAIR % 64 == 0 && (NT_WIND & (NT_WIND - 1)) == 0 &&
                  NT_WIND % 64 == 0,
              "rounds rotate by whole wavefronts");

template <int X0, int X1, int Y0, int Y1>
struct Reg {
    static constexpr int x0 = X0, x1 = X1, y0 = Y0, y1 = Y1;
    static constexpr int w = X1 - X0, h = Y1 - Y0, n = w * h;
    __device__ static int at(int x, int y) { return (y - Y0) * w + (x - X0); }
};

// The items of a region are dealt to the block's threads in rounds of NT;
// the last round of a region is usually partial.  Each region starts its
// rounds at a different wavefront (__COUNTER__ rotation), so the partial
// rounds are spread over the block's wavefronts (and SIMDs) instead
// of all landing on wavefront 0.
#define FOR_REG(R)                                                                         \
    for (int t_ = (threadIdx.x + 64 * (__COUNTER__ % (NT / 64))) & (NT - 1); t_ < R::n; t_ += NT) { \
        const int x = R::x0 + t_ % R::w, y = R::y0 + t_ / R::w;
#define END_FOR }

// Tile-relative global addressing: a base that is uniform over the
// workgroup (level k, tile origin moved back by 4 rows and columns) plus a
// non-negative 32-bit per-lane offset, which the loads and stores take as
// a scalar base and a 32-bit vector offset.  Offsets x, y >= -4.
#define TBASE(k) ((size_t)(k) * g.plane() + (size_t)(J0 - 4) * g.SI + (size_t)(I0 - 4))
#define TOFF(x, y) ((unsigned)(((y) + 4) * g.SI + (x) + 4))

// This thread's offset TOFF(tx, ty), formed afresh from an opaque copy of
// the thread index: the compiler would otherwise hold the 64-bit offset of
// the first use in registers through every later stencil stage, and the C
// step's dp/pt kernel (the tightest in registers) spilled it.
__device__ __forceinline__ unsigned own_off(const Geo &g, int tx_n) {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const int tx = t % tx_n, ty = t / tx_n;
    return (unsigned)((ty + 4) * g.SI + tx + 4);
}

// Whether this thread's cell lies inside the domain (a tile may overhang
// it), formed afresh like own_off so that no register holds it between the
// stores.
__device__ __forceinline__ bool own_inside(const Geo &g, int I0, int J0, int tx_n) {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return I0 + t % tx_n < ATM_HALO + g.nx && J0 + t / tx_n < ATM_HALO + g.ny;
}

// ---------------------------------------------------------------------------
// C1 / D1: the dp and pt updates of one level (cgrid 94-268, dgrid 439-598).
// CGRID: the C winds come from the D winds (109-126) and are also written
// out; otherwise uc/vc of the C step are read (the D step).
// ---------------------------------------------------------------------------
// waves per SIMD of the dp/pt kernels (measured): 6 in double (5 and 7
// are slower); in float 7 for the C step (one spilled register) and 8 for
// the D step, whose registers fit without spills
template <class R, bool CGRID>
constexpr int pair_waves() { return std::is_same<R, double>::value ? 6 : CGRID ? 7 : 8; }

template <int VAR, bool CGRID, class R, int TY = TY_PAIR>
__device__ __forceinline__ void pair_body(int bx, int by, int K, Geo g, DynCoef c, const double *__restrict__ dp,
                                             const double *__restrict__ pt, const double *__restrict__ uw,
                                             const double *__restrict__ vw, const double *__restrict__ heat,
                                             const double *__restrict__ plk, const double *__restrict__ nf,
                                             double *__restrict__ dpo, double *__restrict__ pto,
                                             double *__restrict__ uco, double *__restrict__ vco,
                                             double *__restrict__ ua, double *__restrict__ va, int *nfw) {
    constexpr int NT = TX * TY;
    constexpr bool FP64 = std::is_same<R, double>::value;
    const CoefR<R> r(c);
    using RUC = Reg<0, TX + 1, -3, TY + 3>;
    using RVC = Reg<-3, TX + 3, 0, TY + 1>;
    using RDX = Reg<0, TX, -3, TY + 3>;  // tmp of the inner x step (denominators recomputed per pass)
    using RDY = Reg<-3, TX + 3, 0, TY>;  // tmp of the inner y step
    using RQ = Reg<-3, TX + 3, -3, TY + 3>;
    using RL = Reg<-1, TX + 1, -1, TY + 1>;
    using RFX = Reg<0, TX + 1, 0, TY>;
    using RFY = Reg<0, TX, 0, TY + 1>;
    __shared__ R sUC[RUC::n], sVC[RVC::n], sQ[RQ::n];
    __shared__ R sTX[RDY::n], sTY[RDX::n], sL[RL::n];
    // face values: the dp mass fluxes in the dp pass (each thread keeps the
    // four around its cell in registers), the pt face values in the pt pass
    __shared__ R sFX[RFX::n], sFY[RFY::n];

    const int I0 = ATM_HALO + bx * TX, J0 = ATM_HALO + by * TY;
    const size_t kb0 = TBASE(K);
#define G(a, x, y) (a + kb0)[TOFF(x, y)]
#define UC(x, y) sUC[RUC::at(x, y)]
#define VC(x, y) sVC[RVC::at(x, y)]
#define Q(x, y) sQ[RQ::at(x, y)]
#define LP(x, y) sL[RL::at(x, y)]

    // C winds (cgrid 109-117) or the C step's winds (dgrid)
    FOR_REG(RUC)
    if (CGRID) {
        double ub = .5 * (G(uw, x, y) + G(uw, x - 1, y));
        UC(x, y) = .5 * (ub + .5 * (G(uw, x, y + 1) + G(uw, x - 1, y + 1)));
    } else {
        UC(x, y) = G(uw, x, y);
    }
    END_FOR
    FOR_REG(RVC)
    if (CGRID) {
        double vb = .5 * (G(vw, x, y) + G(vw, x, y - 1));
        VC(x, y) = .5 * (vb + .5 * (G(vw, x + 1, y) + G(vw, x + 1, y - 1)));
    } else {
        VC(x, y) = G(vw, x, y);
    }
    END_FOR
    FOR_REG(RQ) Q(x, y) = G(dp, x, y); END_FOR
    __syncthreads();

    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    double dpv = 0., dpcv = 0., pt_new = 0.;
    R fx0 = 0., fx1 = 0., fy0 = 0., fy1 = 0.;  // dp mass fluxes on the cell's faces
    for (int pass = 0; pass < 2; pass++) {
        // inner y step (tmp of the outer x step), with deny; inner x step with denx; Laplacian
        FOR_REG(RDY)
        const R v0 = VC(x, y), v1 = VC(x, y + 1);
        const R dny = rcp(R(1) - r.dtdy * (v1 - v0));
        const R q0 = Q(x, y);
        const R f0 = pcm_face(Q(x, y - 1), q0, r.dtdx * v0);  // Q2: cy = dt/dx * vc
        const R f1 = pcm_face(q0, Q(x, y + 1), r.dtdx * v1);
        sTX[RDY::at(x, y)] = R(.5) * (q0 + dny * (q0 - r.rA * (f1 - f0)));
        END_FOR
        FOR_REG(RDX)
        const R u0 = UC(x, y), u1 = UC(x + 1, y);
        const R dnx = rcp(R(1) - r.dtdx * (u1 - u0));
        const R q0 = Q(x, y);
        const R f0 = pcm_face(Q(x - 1, y), q0, r.dtdx * u0);
        const R f1 = pcm_face(q0, Q(x + 1, y), r.dtdx * u1);
        sTY[RDX::at(x, y)] = R(.5) * (q0 + dnx * (q0 - r.rA * (f1 - f0)));
        END_FOR
        FOR_REG(RL)
        LP(x, y) = R(-2) * Q(x, y) * r.rdx2rdy2 + (Q(x + 1, y) + Q(x - 1, y)) * r.rdx2 +
                   (Q(x, y + 1) + Q(x, y - 1)) * r.rdy2;
        END_FOR
        __syncthreads();
        // outer steps: PPM fluxes on faces
        FOR_REG(RFX)
#define T(m) sTX[RDY::at(x + (m), y)]
        const R cx = r.dtdx * UC(x, y);
        const R f = ppm_face<VAR>(T(-3), T(-2), T(-1), T(0), T(1), T(2), cx);
#undef T
        sFX[RFX::at(x, y)] = pass == 0 ? f * cx : f;
        END_FOR
        FOR_REG(RFY)
#define T(m) sTY[RDX::at(x, y + (m))]
        const R cy = r.dtdx * VC(x, y);  // Q2
        const R f = ppm_face<VAR>(T(-3), T(-2), T(-1), T(0), T(1), T(2), cy);
#undef T
        sFY[RFY::at(x, y)] = pass == 0 ? f * cy : f;
        END_FOR
        __syncthreads();
        const R l0 = LP(tx, ty);
        const R lxp = LP(tx + 1, ty), lxm = LP(tx - 1, ty), lyp = LP(tx, ty + 1), lym = LP(tx, ty - 1);
        if (pass == 0) {
            fx0 = sFX[RFX::at(tx, ty)];
            fx1 = sFX[RFX::at(tx + 1, ty)];
            fy0 = sFY[RFY::at(tx, ty)];
            fy1 = sFY[RFY::at(tx, ty + 1)];
            if constexpr (FP64)
                dpv = Q(tx, ty);
            else
                dpv = G(dp, tx, ty);
            double d = dpv - (D(fx1) - D(fx0) + c.dtdx * c.vx * (D(lxp) - 2. * D(l0) + D(lxm)));
            d = d - (D(fy1) - D(fy0) + c.dtdy * c.vy * (D(lyp) - 2. * D(l0) + D(lym)));
            dpcv = d;
            if (own_inside(g, I0, J0, TX)) {
                (dpo + kb0)[CGRID ? own_off(g, TX) : TOFF(tx, ty)] = d;
                if (!CGRID) atm_report_nonfinite(nfw, d);
            }
            __syncthreads();
            FOR_REG(RQ) Q(x, y) = G(pt, x, y); END_FOR
            __syncthreads();
        } else {
            double ptv;
            if constexpr (FP64)
                ptv = Q(tx, ty);
            else
                ptv = G(pt, tx, ty);
            // heating rate: the C step forms it from dp and playkap(dp)
            // (cgrid 99-106); the D step reads the C step's (dgrid 457-465)
            double hr;
            if (CGRID) {
                const unsigned o = TOFF(tx, ty);
                if constexpr (FP64)
                    hr = ((nf + TBASE(K + 1))[o] - (nf + kb0)[o]) / (ATM_CP_DRY * dpv / c.gravity) / (plk + kb0)[o];
                else  // the mixed form's heating rate, a tendency, in float
                    hr = D(R((nf + TBASE(K + 1))[o] - (nf + kb0)[o]) *
                           rcp(R(ATM_CP_DRY / c.gravity) * R(dpv) * R((plk + kb0)[o])));
            } else {
                hr = G(heat, tx, ty);
            }
            double p = dpv * ptv - (D(sFX[RFX::at(tx + 1, ty)] * fx1) - D(sFX[RFX::at(tx, ty)] * fx0) +
                                    c.vx * (D((lxp - l0) * fx1) - D((l0 - lxm) * fx0)));
            p = (p - (D(sFY[RFY::at(tx, ty + 1)] * fy1) - D(sFY[RFY::at(tx, ty)] * fy0) +
                      c.vy * (D((lyp - l0) * fy1) - D((l0 - lym) * fy0))) +
                 c.dt * dpv * hr) /
                dpcv;
            // the float C step stores pt with its winds below: one masked
            // branch instead of two keeps its registers from spilling
            if (CGRID && !FP64) {
                pt_new = p;
            } else if (own_inside(g, I0, J0, TX)) {
                (pto + kb0)[CGRID ? own_off(g, TX) : TOFF(tx, ty)] = p;
                if (!CGRID) atm_report_nonfinite(nfw, p);
            }
        }
    }
    if (CGRID && own_inside(g, I0, J0, TX)) {
        double u, v;
        if constexpr (FP64) {
            u = UC(tx, ty);
            v = VC(tx, ty);
        } else {  // the C winds of the cell in double (109-117)
            const double ub = .5 * (G(uw, tx, ty) + G(uw, tx - 1, ty));
            u = .5 * (ub + .5 * (G(uw, tx, ty + 1) + G(uw, tx - 1, ty + 1)));
            const double vb = .5 * (G(vw, tx, ty) + G(vw, tx, ty - 1));
            v = .5 * (vb + .5 * (G(vw, tx + 1, ty) + G(vw, tx + 1, ty - 1)));
        }
        const unsigned o = own_off(g, TX);
        if (!FP64) (pto + kb0)[o] = pt_new;
        (uco + kb0)[o] = u;
        (vco + kb0)[o] = v;
        if (K == g.L - 1) {  // A winds of the last level (119-125), read by the KE pass
            const size_t p2 = g.ix2(I0 + tx, J0 + ty);
            ua[p2] = .5 * (u + .25 * (G(uw, tx + 1, ty) + G(uw, tx, ty) + G(uw, tx + 1, ty + 1) + G(uw, tx, ty + 1)));
            va[p2] = .5 * (v + .25 * (G(vw, tx, ty + 1) + G(vw, tx, ty) + G(vw, tx + 1, ty + 1) + G(vw, tx + 1, ty)));
        }
    }
#undef G
#undef UC
#undef VC
#undef Q
#undef LP
}

template <int VAR, bool CGRID, class R, int TY>
__global__ __launch_bounds__(TX * TY, (pair_waves<R, CGRID>())) void k_pair(Geo g, DynCoef c, const double *__restrict__ dp,
                                             const double *__restrict__ pt, const double *__restrict__ uw,
                                             const double *__restrict__ vw, const double *__restrict__ heat,
                                             const double *__restrict__ plk, const double *__restrict__ nf,
                                             double *__restrict__ dpo, double *__restrict__ pto,
                                             double *__restrict__ uco, double *__restrict__ vco,
                                             double *__restrict__ ua, double *__restrict__ va, int *nfw) {
    int bx, by, K;
    xcd_tile(gridDim.x, gridDim.y, gridDim.z, bx, by, K);
    pair_body<VAR, CGRID, R, TY>(bx, by, K, g, c, dp, pt, uw, vw, heat, plk, nf, dpo, pto, uco, vco, ua, va, nfw);
}

// ---------------------------------------------------------------------------
// C2: kinetic energy, vorticity, pressure gradient and the uc/vc update of
// one level (cgrid 294-414).
// ---------------------------------------------------------------------------
template <class R, int TY = TY_WIND>
__device__ __forceinline__ void c2_body(int bx, int by, int K, Geo g, DynCoef c, const double *__restrict__ uc1,
                                           const double *__restrict__ vc1, const double *__restrict__ ud,
                                           const double *__restrict__ vd, const double *__restrict__ pkap,
                                           const double *__restrict__ gz, const double *__restrict__ ua,
                                           const double *__restrict__ va, double *__restrict__ uco,
                                           double *__restrict__ vco) {
    constexpr int NT = TX * TY;
    constexpr bool FP64 = std::is_same<R, double>::value;
    const CoefR<R> r(c);
    using RB = Reg<-3, TX + 3, -3, TY + 3>;
    using RW = Reg<-2, TX + 3, -2, TY + 3>;
    using RK = Reg<-1, TX, -1, TY>;
    using RTV = Reg<0, TX, -2, TY + 3>;
    using RTW = Reg<-2, TX + 3, 0, TY>;
    using RLV = Reg<0, TX + 1, 0, TY + 1>;
    __shared__ R sU[RB::n], sV[RB::n], sW[RW::n], sK[RK::n], sTV[RTV::n], sTW[RTW::n], sL[RLV::n];
    const int I0 = ATM_HALO + bx * TX, J0 = ATM_HALO + by * TY;
    const size_t kb0 = TBASE(K);
#define G(a, x, y) (a + kb0)[TOFF(x, y)]
#define U(x, y) sU[RB::at(x, y)]
#define V(x, y) sV[RB::at(x, y)]
#define W(x, y) sW[RW::at(x, y)]
    FOR_REG(RB)
    U(x, y) = G(uc1, x, y);
    V(x, y) = G(vc1, x, y);
    END_FOR
    // pressure gradient force (325-339): global values only, formed before
    // the LDS stages so its loads' latency overlaps them (a thread past the
    // domain reads the halo or the tile slack and stores nothing)
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    double pgfx, pgfy;
    {
        const double *pk0 = pkap + kb0, *pk1 = pkap + TBASE(K + 1), *gz0 = gz + kb0, *gz1 = gz + TBASE(K + 1);
        double a = pk0[TOFF(tx, ty)], d = pk1[TOFF(tx, ty)];
        double b = pk1[TOFF(tx - 1, ty)], cc = pk0[TOFF(tx - 1, ty)];
        pgfx = ((a - b) * (gz0[TOFF(tx - 1, ty)] - gz1[TOFF(tx, ty)]) +
                (cc - d) * (gz1[TOFF(tx - 1, ty)] - gz0[TOFF(tx, ty)])) /
               (a - b + cc - d);
        b = pk1[TOFF(tx, ty - 1)];
        cc = pk0[TOFF(tx, ty - 1)];
        pgfy = ((a - b) * (gz0[TOFF(tx, ty - 1)] - gz1[TOFF(tx, ty)]) +
                (cc - d) * (gz1[TOFF(tx, ty - 1)] - gz0[TOFF(tx, ty)])) /
               (a - b + cc - d);
    }
    __syncthreads();
    FOR_REG(RW)  // vorticity (319-323)
    W(x, y) = r.coriolis - (U(x, y) - U(x, y - 1)) * r.rdy + (V(x, y) - V(x - 1, y)) * r.rdx;
    END_FOR
    FOR_REG(RK)  // kinetic energy (299-317), last-level A winds, Q4 shifts
    const size_t p2 = g.ix2(I0 + x, J0 + y);
    const R vaij = va[p2], uaij = ua[p2];
    const R fy = ppm_face<V_PPM_U>(V(x, y - 2), V(x, y - 1), V(x, y), V(x, y + 1), V(x, y + 2), V(x, y + 3),
                                   r.dtdy * vaij);
    const R fx = ppm_face<V_PPM_U>(U(x - 2, y), U(x - 1, y), U(x, y), U(x + 1, y), U(x + 2, y), U(x + 3, y),
                                   r.dtdx * uaij);
    const R k = vaij * fy;
    sK[RK::at(x, y)] = R(.5) * (k + uaij * fx);
    END_FOR
    __syncthreads();
    FOR_REG(RTV)  // inner x step of vorticity, Courant cx shifted by one (353-367, Q4)
    const R w0 = W(x, y);
    const R um = G(ud, x - 1, y), u0 = G(ud, x, y);
    const R dnx = rcp(R(1) - r.dtdx * (u0 - um));
    const R f0 = pcm_face(W(x - 1, y), w0, r.dtdx * um);
    const R f1 = pcm_face(w0, W(x + 1, y), r.dtdx * u0);
    sTV[RTV::at(x, y)] = R(.5) * (w0 + dnx * (w0 - r.rA * (f1 - f0)));
    END_FOR
    FOR_REG(RTW)  // inner y step of vorticity, Courant cy shifted by one (388-397, Q4)
    const R w0 = W(x, y);
    const R vm = G(vd, x, y - 1), v0 = G(vd, x, y);
    const R dny = rcp(R(1) - r.dtdy * (v0 - vm));
    const R f0 = pcm_face(W(x, y - 1), w0, r.dtdx * vm);
    const R f1 = pcm_face(w0, W(x, y + 1), r.dtdx * v0);
    sTW[RTW::at(x, y)] = R(.5) * (w0 + dny * (w0 - r.rA * (f1 - f0)));
    END_FOR
    FOR_REG(RLV)
    sL[RLV::at(x, y)] = R(-2) * W(x, y) * r.rdx2rdy2 + (W(x + 1, y) + W(x - 1, y)) * r.rdx2 +
                        (W(x, y + 1) + W(x, y - 1)) * r.rdy2;
    END_FOR
    __syncthreads();
    if (!(I0 + tx < ATM_HALO + g.nx && J0 + ty < ATM_HALO + g.ny)) return;  // no barrier follows
    const double udv = G(ud, tx, ty), vdv = G(vd, tx, ty);
    double ucv, vcv;  // this cell's C winds in double
    if constexpr (FP64) {
        ucv = U(tx, ty);
        vcv = V(tx, ty);
    } else {
        ucv = G(uc1, tx, ty);
        vcv = G(vc1, tx, ty);
    }
    // uc (369-382)
    {
#define T(m) sTV[RTV::at(tx, ty + 1 + (m))]
        const R fy = ppm_face<V_PPM_U>(T(-3), T(-2), T(-1), T(0), T(1), T(2), r.dtdx * R(vdv));
#undef T
        G(uco, tx, ty) = ucv + c.dt * (vdv * D(fy) + c.vy * (D(sL[RLV::at(tx, ty + 1)]) - D(sL[RLV::at(tx, ty)]))) -
                         c.dtdx * (D(sK[RK::at(tx, ty)]) - D(sK[RK::at(tx - 1, ty)]) - pgfx);
    }
    // vc (399-412)
    {
#define T(m) sTW[RTW::at(tx + 1 + (m), ty)]
        const R fx = ppm_face<V_PPM_U>(T(-3), T(-2), T(-1), T(0), T(1), T(2), r.dtdx * R(udv));
#undef T
        G(vco, tx, ty) = vcv - c.dt * (udv * D(fx) + c.vx * (D(sL[RLV::at(tx + 1, ty)]) - D(sL[RLV::at(tx, ty)]))) -
                         c.dtdy * (D(sK[RK::at(tx, ty)]) - D(sK[RK::at(tx, ty - 1)]) - pgfy);
    }
#undef G
#undef U
#undef V
#undef W
}

// waves per SIMD of the C-step wind kernel: 8 in 32 x 16 tiles; the LDS of
// 32 x 8 tiles holds 6
template <class R, int TY>
__global__ __launch_bounds__(TX * TY, TY == 16 ? 8 : 6) void k_c2(Geo g, DynCoef c, const double *__restrict__ uc1,
                                           const double *__restrict__ vc1, const double *__restrict__ ud,
                                           const double *__restrict__ vd, const double *__restrict__ pkap,
                                           const double *__restrict__ gz, const double *__restrict__ ua,
                                           const double *__restrict__ va, double *__restrict__ uco,
                                           double *__restrict__ vco) {
    int bx, by, K;
    xcd_tile(gridDim.x, gridDim.y, gridDim.z, bx, by, K);
    c2_body<R, TY>(bx, by, K, g, c, uc1, vc1, ud, vd, pkap, gz, ua, va, uco, vco);
}

// ---------------------------------------------------------------------------
// D2: kinetic energy, vorticity, divergence damping, B-grid pressure
// gradient and the ud/vd update of one level (dgrid 606-771).
//
// Its PPM_CONSTRAINED faces (the kinetic energy's and the vorticity
// fluxes') each need the dqi_mono of four cells; every cell's dqi_mono is
// formed once per tile, into LDS, before the faces that read it (storage
// of arrays that are dead or not yet written at that stage).
// ---------------------------------------------------------------------------
template <class R, int TY = TY_WIND>
__device__ __forceinline__ void d2_body(int bx, int by, int K, Geo g, DynCoef c, const double *__restrict__ ud,
                                           const double *__restrict__ vd, const double *__restrict__ uc,
                                           const double *__restrict__ vc, const double *__restrict__ pkapb,
                                           const double *__restrict__ gzb, double *__restrict__ udo,
                                           double *__restrict__ vdo, int *nfw) {
    constexpr int NT = TX * TY;
    const CoefR<R> r(c);
    using RU = Reg<-3, TX + 3, -3, TY + 4>;
    using RV = Reg<-3, TX + 4, -3, TY + 3>;
    using RKE = Reg<0, TX + 1, 0, TY + 1>;
    using RW = Reg<-3, TX + 3, -3, TY + 3>;
    using RDV = Reg<-2, TX + 3, -2, TY + 3>;
    using RD2 = Reg<-1, TX + 2, -1, TY + 2>;
    using RTU = Reg<0, TX, -3, TY + 3>;
    using RTW = Reg<-3, TX + 3, 0, TY>;
    using RLP = Reg<-1, TX, -1, TY>;
    // dqi_mono of the kinetic energy's faces: U along x, V along y
    using RQU = Reg<-2, TX + 2, 0, TY + 1>;
    using RQV = Reg<0, TX + 1, -2, TY + 2>;
    // ... and of the vorticity fluxes': TU along y, TW along x
    using RQTU = Reg<0, TX, -2, TY + 1>;
    using RQTW = Reg<-2, TX + 1, 0, TY>;
    __shared__ R sU[RU::n], sV[RV::n], sKE[RKE::n], sW[RW::n], sDV[RDV::n], sD2[RD2::n];
    __shared__ R sT[RTU::n + RTW::n], sL[RLP::n];
    R *const sTU = sT, *const sTW = sT + RTU::n;
    // del^4 divergence reuses the divergence's storage (dead once del^2 is formed)
    static_assert(RKE::n <= RDV::n, "sD4 fits in sDV");
    R *const sD4 = sDV;
    // the kinetic energy's dqi_mono live where TU and TW are formed later;
    // the vorticity fluxes' where U and V were (dead after the kinetic energy)
    static_assert(RQU::n + RQV::n <= RTU::n + RTW::n && RQTU::n <= RU::n && RQTW::n <= RV::n, "dq storage");
    R *const sQU = sT, *const sQV = sT + RQU::n, *const sQTU = sU, *const sQTW = sV;
    const int I0 = ATM_HALO + bx * TX, J0 = ATM_HALO + by * TY;
    const size_t kb0 = TBASE(K);
#define G(a, x, y) (a + kb0)[TOFF(x, y)]
#define U(x, y) sU[RU::at(x, y)]
#define V(x, y) sV[RV::at(x, y)]
#define W(x, y) sW[RW::at(x, y)]
#define DV(x, y) sDV[RDV::at(x, y)]
#define D2(x, y) sD2[RD2::at(x, y)]
    FOR_REG(RU) U(x, y) = G(ud, x, y); END_FOR
    FOR_REG(RV) V(x, y) = G(vd, x, y); END_FOR
    // The B-grid pressure gradient (665-679) needs global values only: it is
    // formed before the LDS stages, so its loads' latency overlaps them.  The
    // corner values of pkap and gz at levels k and k+1 come from k_bcorners
    // (pkapb(nlay+1) and gzb(1) from init_sw_dyn 991, 995).  A thread past
    // the domain reads the halo or the tile slack and stores nothing.
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int Lm = g.L - 1;
    const size_t kbL = TBASE(Lm);
#define GL(a, x, y) (a + kbL)[TOFF(x, y)]
    double pgfx, pgfy;
    {
        const size_t kb1 = TBASE(K + 1);
#define CB(a, kq, x, y) (a + ((kq) ? kb1 : kb0))[TOFF(x, y)]
        double a = CB(pkapb, 0, tx + 1, ty), b = CB(pkapb, 1, tx, ty), cc = CB(pkapb, 0, tx, ty),
               d = CB(pkapb, 1, tx + 1, ty);
        const double g00 = CB(gzb, 0, tx, ty), g01 = CB(gzb, 1, tx, ty);
        pgfx = ((a - b) * (g00 - CB(gzb, 1, tx + 1, ty)) + (cc - d) * (g01 - CB(gzb, 0, tx + 1, ty))) /
               (a - b + cc - d);
        a = CB(pkapb, 0, tx, ty + 1);
        d = CB(pkapb, 1, tx, ty + 1);
        pgfy = ((a - b) * (g00 - CB(gzb, 1, tx, ty + 1)) + (cc - d) * (g01 - CB(gzb, 0, tx, ty + 1))) /
               (a - b + cc - d);
#undef CB
    }
    __syncthreads();
    FOR_REG(RQU) sQU[RQU::at(x, y)] = dq_mono(U(x - 1, y), U(x, y), U(x + 1, y)); END_FOR
    FOR_REG(RQV) sQV[RQV::at(x, y)] = dq_mono(V(x, y - 1), V(x, y), V(x, y + 1)); END_FOR
    FOR_REG(RW)  // vorticity (659-663)
    W(x, y) = r.coriolis - (U(x, y + 1) - U(x, y)) * r.rdy + (V(x + 1, y) - V(x, y)) * r.rdx;
    END_FOR
    FOR_REG(RDV)  // divergence (682-685)
    DV(x, y) = r.rdx * (U(x, y) - U(x - 1, y)) + r.rdy * (V(x, y) - V(x, y - 1));
    END_FOR
    __syncthreads();
    FOR_REG(RKE)  // kinetic energy (639-657), PPM_CONSTRAINED, last-level B winds
    const R ubij = R(.5 * (GL(uc, x, y) + GL(uc, x, y - 1)));
    const R vbij = R(.5 * (GL(vc, x, y) + GL(vc, x - 1, y)));
#define QV(m) sQV[RQV::at(x, y + (m))]
    const R fy = ppm_face_dq<V_PPM_C>(V(x, y - 2), V(x, y - 1), V(x, y), V(x, y + 1), QV(-2), QV(-1), QV(0), QV(1),
                                      r.dtdy * vbij);
#undef QV
#define QU(m) sQU[RQU::at(x + (m), y)]
    const R fx = ppm_face_dq<V_PPM_C>(U(x - 2, y), U(x - 1, y), U(x, y), U(x + 1, y), QU(-2), QU(-1), QU(0), QU(1),
                                      r.dtdx * ubij);
#undef QU
    const R k = vbij * fy;
    sKE[RKE::at(x, y)] = R(.5) * (k + ubij * fx);
    END_FOR
    FOR_REG(RD2)  // del^2 divergence (688-692)
    D2(x, y) = R(-2) * DV(x, y) * r.rdx2rdy2 + (DV(x + 1, y) + DV(x - 1, y)) * r.rdx2 +
               (DV(x, y + 1) + DV(x, y - 1)) * r.rdy2;
    END_FOR
    FOR_REG(RLP)  // Laplacian of vorticity (717-720)
    sL[RLP::at(x, y)] = R(-2) * W(x, y) * r.rdx2rdy2 + (W(x + 1, y) + W(x - 1, y)) * r.rdx2 +
                        (W(x, y + 1) + W(x, y - 1)) * r.rdy2;
    END_FOR
    __syncthreads();
    // The reference's second level loop (634-771) reuses the 2-D arrays cx,
    // cy, denom_x, denom_y (and ub, vb) that its first loop (441-454) left
    // at the last level, k = nlay: the vorticity fluxes of every level use
    // the C winds of level nlay for their Courant numbers and denominators.
    FOR_REG(RTU)  // inner x step of vorticity for ud (708-716)
    const R w0 = W(x, y);
    const R u0 = GL(uc, x, y), u1 = GL(uc, x + 1, y);
    const R dnx = rcp(R(1) - r.dtdx * (u1 - u0));
    const R f0 = pcm_face(W(x - 1, y), w0, r.dtdx * u0);
    const R f1 = pcm_face(w0, W(x + 1, y), r.dtdx * u1);
    sTU[RTU::at(x, y)] = R(.5) * (w0 + dnx * (w0 - r.rA * (f1 - f0)));
    END_FOR
    FOR_REG(RTW)  // inner y step of vorticity for vd (744-752)
    const R w0 = W(x, y);
    const R v0 = GL(vc, x, y), v1 = GL(vc, x, y + 1);
    const R dny = rcp(R(1) - r.dtdy * (v1 - v0));
    const R f0 = pcm_face(W(x, y - 1), w0, r.dtdx * v0);  // Q2
    const R f1 = pcm_face(w0, W(x, y + 1), r.dtdx * v1);
    sTW[RTW::at(x, y)] = R(.5) * (w0 + dny * (w0 - r.rA * (f1 - f0)));
    END_FOR
    FOR_REG(RKE)  // del^4 divergence times the damping coefficient (695-702)
    sD4[RKE::at(x, y)] = r.divdamp * (R(-2) * D2(x, y) * r.rdx2rdy2 + (D2(x + 1, y) + D2(x - 1, y)) * r.rdx2 +
                                      (D2(x, y + 1) + D2(x, y - 1)) * r.rdy2);
    END_FOR
    __syncthreads();
    FOR_REG(RQTU) sQTU[RQTU::at(x, y)] = dq_mono(sTU[RTU::at(x, y - 1)], sTU[RTU::at(x, y)], sTU[RTU::at(x, y + 1)]); END_FOR
    FOR_REG(RQTW) sQTW[RQTW::at(x, y)] = dq_mono(sTW[RTW::at(x - 1, y)], sTW[RTW::at(x, y)], sTW[RTW::at(x + 1, y)]); END_FOR
    __syncthreads();
    if (!(I0 + tx < ATM_HALO + g.nx && J0 + ty < ATM_HALO + g.ny)) return;  // no barrier follows
    const double ucv = G(uc, tx, ty), vcv = G(vc, tx, ty);
    const double ucl = GL(uc, tx, ty), vcl = GL(vc, tx, ty);
    // this cell's D winds in double (their LDS tiles now hold dqi_mono)
    const double udv = G(ud, tx, ty), vdv = G(vd, tx, ty);
    {  // ud (724-738)
#define T(m) sTU[RTU::at(tx, ty + (m))]
#define Q(m) sQTU[RQTU::at(tx, ty + (m))]
        const R fy = ppm_face_dq<V_PPM_C>(T(-2), T(-1), T(0), T(1), Q(-2), Q(-1), Q(0), Q(1), r.dtdx * R(vcl));
#undef T
#undef Q
        const double un = udv +
                          c.dt * (vcv * D(fy) + c.vy * (D(sL[RLP::at(tx, ty)]) - D(sL[RLP::at(tx, ty - 1)]))) -
                          c.dtdx * (D(sKE[RKE::at(tx + 1, ty)]) - D(sKE[RKE::at(tx, ty)]) -
                                    D(sD4[RKE::at(tx + 1, ty)]) + D(sD4[RKE::at(tx, ty)]) - pgfx);
        G(udo, tx, ty) = un;
        atm_report_nonfinite(nfw, un);
    }
    {  // vd (755-769)
#define T(m) sTW[RTW::at(tx + (m), ty)]
#define Q(m) sQTW[RQTW::at(tx + (m), ty)]
        const R fx = ppm_face_dq<V_PPM_C>(T(-2), T(-1), T(0), T(1), Q(-2), Q(-1), Q(0), Q(1), r.dtdx * R(ucl));
#undef T
#undef Q
        const double vn = vdv -
                          c.dt * (ucv * D(fx) + c.vx * (D(sL[RLP::at(tx, ty)]) - D(sL[RLP::at(tx - 1, ty)]))) -
                          c.dtdy * (D(sKE[RKE::at(tx, ty + 1)]) - D(sKE[RKE::at(tx, ty)]) -
                                    D(sD4[RKE::at(tx, ty + 1)]) + D(sD4[RKE::at(tx, ty)]) - pgfy);
        G(vdo, tx, ty) = vn;
        atm_report_nonfinite(nfw, vn);
    }
#undef G
#undef GL
#undef U
#undef V
#undef W
#undef DV
#undef D2
}

template <class R, int TY>
__global__ __launch_bounds__(TX * TY) void k_d2(Geo g, DynCoef c, const double *__restrict__ ud,
                                           const double *__restrict__ vd, const double *__restrict__ uc,
                                           const double *__restrict__ vc, const double *__restrict__ pkapb,
                                           const double *__restrict__ gzb, double *__restrict__ udo,
                                           double *__restrict__ vdo, int *nfw) {
    int bx, by, K;
    xcd_tile(gridDim.x, gridDim.y, gridDim.z, bx, by, K);
    d2_body<R, TY>(bx, by, K, g, c, ud, vd, uc, vc, pkapb, gzb, udo, vdo, nfw);
}

// B-grid corner values of pkap and gz for the D step (dgrid 606-632): for
// every level k, pkapb(k) = corners(pkap(k)) and gzb(k+1) = corners(gz(k+1))
// on the interior plus one row and column (what the pressure gradient of
// the tile reads).  pkapb(nlay+1) and gzb(1) keep their init_sw_dyn values.
// interpolate_to_corners (921-956) is separable: each corner combines four
// x-direction 4-point averages (rows j-2..j+1) and four y-direction ones
// (columns i-2..i+1).  A TX x TY tile stages the field, forms each average
// once in LDS and combines them in the reference's order.  A block walks
// NLEV levels and loads the next level's tile into registers while the
// current one is combined, so each CU keeps loads in flight through the
// LDS stages; fewer levels per block give small grids more blocks.  NLEV
// is the context's tuning (atm_autotune); every choice gives the same values.
template <int NLEV>
__device__ __forceinline__ void bcorners_body(int bxi, int byi, int bzi, Geo g, const double *__restrict__ pkap,
                                                 const double *__restrict__ gz, double *__restrict__ pkapb,
                                                 double *__restrict__ gzb) {
    constexpr int TY = TY_WIND, NT = NT_WIND;
    using RQ = Reg<-2, TX + 1, -2, TY + 1>;  // field
    using RAX = Reg<0, TX, -2, TY + 1>;      // x averages
    using RAY = Reg<-2, TX + 1, 0, TY>;      // y averages
    constexpr int NR = (RQ::n + NT - 1) / NT;  // staged points per thread and field
    __shared__ double sQ[2][RQ::n], sAX[2][RAX::n], sAY[2][RAY::n];
    const int I0 = ATM_HALO + bxi * TX, J0 = ATM_HALO + byi * TY;
    const int K0 = bzi * NLEV, K1 = min(K0 + NLEV, g.L);
    // this thread's staged points: plane offsets, or -1 past the halo'd
    // plane (the last tile row / column holds one valid output row / column
    // and does not use them)
    long off[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int t = threadIdx.x + r * NT;
        const int I = I0 + RQ::x0 + t % RQ::w, J = J0 + RQ::y0 + t / RQ::w;
        off[r] = (t < RQ::n && I < g.NI && J < g.NJ) ? (long)g.ix2(I, J) : -1;
    }
    double q[2][NR];
    auto load = [&](int K) {
        const double *src[2] = {pkap + g.ix(0, 0, K), gz + g.ix(0, 0, K + 1)};
#pragma unroll
        for (int f = 0; f < 2; f++)
#pragma unroll
            for (int r = 0; r < NR; r++) q[f][r] = off[r] >= 0 ? src[f][off[r]] : 0.;
    };
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int I = I0 + tx, J = J0 + ty;
    const bool writes = I <= ATM_HALO + g.nx && J <= ATM_HALO + g.ny;
    load(K0);
    for (int K = K0; K < K1; K++) {
#pragma unroll
        for (int f = 0; f < 2; f++)
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int t = threadIdx.x + r * NT;
                if (t < RQ::n) sQ[f][t] = q[f][r];
            }
        __syncthreads();
        if (K + 1 < K1) load(K + 1);
#define Q(f, x, y) sQ[f][RQ::at(x, y)]
        for (int f = 0; f < 2; f++) {
            FOR_REG(RAX)
            sAX[f][RAX::at(x, y)] = ATM_W7_12 * (Q(f, x, y) + Q(f, x - 1, y)) - ATM_W1_12 * (Q(f, x + 1, y) + Q(f, x - 2, y));
            END_FOR
            FOR_REG(RAY)
            sAY[f][RAY::at(x, y)] = ATM_W7_12 * (Q(f, x, y) + Q(f, x, y - 1)) - ATM_W1_12 * (Q(f, x, y + 1) + Q(f, x, y - 2));
            END_FOR
        }
#undef Q
        __syncthreads();
        if (writes) {
            double out[2];
            for (int f = 0; f < 2; f++) {
#define AX(b) sAX[f][RAX::at(tx, ty + (b))]
#define AY(a) sAY[f][RAY::at(tx + (a), ty)]
                double o = (9. / 32.) * (AX(0) + AX(-1)) - (1. / 32.) * (AX(1) + AX(-2));
                o = o + (9. / 32.) * (AY(0) + AY(-1)) - (1. / 32.) * (AY(1) + AY(-2));
                out[f] = o;
#undef AX
#undef AY
            }
            pkapb[g.ix(I, J, K)] = out[0];
            gzb[g.ix(I, J, K + 1)] = out[1];
        }
    }
}

template <int NLEV>
__global__ void __launch_bounds__(NT_WIND) k_bcorners(Geo g, const double *__restrict__ pkap,
                                                 const double *__restrict__ gz, double *__restrict__ pkapb,
                                                 double *__restrict__ gzb) {
    bcorners_body<NLEV>((int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, g, pkap, gz, pkapb, gzb);
}

// k_bcorners with n levels per block (1, 2, 4 or 8)
void launch_bcorners(atm_ctx *c, int n) {
    const Geo &g = c->g;
    const dim3 blk(NT_WIND), grid((g.nx + TX) / TX, (g.ny + TY_WIND) / TY_WIND, (g.L + n - 1) / n);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, grid, blk, 0, c->stream, g, c->pkap, c->gz, c->pkapb, c->gzb);
    };
    switch (n) {
    case 1: go(k_bcorners<1>); break;
    case 2: go(k_bcorners<2>); break;
    case 8: go(k_bcorners<8>); break;
    default: go(k_bcorners<4>); break;
    }
}

// ---------------------------------------------------------------------------
// Column scans.  One thread per column, i fastest across the wavefront.  The
// running sum of dp is the only serial dependence: levels go in chunks of
// SCAN_CH whose pow / log evaluations are independent of each other, so a
// wavefront keeps SCAN_CH of them in flight.
// ---------------------------------------------------------------------------
#define SCAN_CH 4

// playkap of the C step from dp (cgrid 89-101) on interior columns; C1
// forms the heating rate from it (99-106).  Needed only when dp changed
// after the last D step's scan (which leaves the same values, see below).
__global__ void __launch_bounds__(64) k_scan_plk(Geo g, double ptop, double pktop,
                                                 const double *__restrict__ src, double *__restrict__ plk) {
    const int I = ATM_HALO + blockIdx.x * blockDim.x + threadIdx.x;
    const int J = ATM_HALO + blockIdx.y;
    if (I >= ATM_HALO + g.nx) return;
    const size_t pl = g.plane(), c0 = g.ix2(I, J);
    double p = ptop, pk1 = pktop;
    auto level = [&](int K, double pk) {
        plk[c0 + (size_t)K * pl] = (pk1 - pk) / log(pk1 / pk);
        pk1 = pk;
    };
    int K = g.L - 1;
    for (; K >= SCAN_CH - 1; K -= SCAN_CH) {
        double pv[SCAN_CH], pk[SCAN_CH];
#pragma unroll
        for (int u = 0; u < SCAN_CH; u++) {
            p = p + src[c0 + (size_t)(K - u) * pl];
            pv[u] = p;
        }
#pragma unroll
        for (int u = 0; u < SCAN_CH; u++) pk[u] = pow_kappa(pv[u]);
#pragma unroll
        for (int u = 0; u < SCAN_CH; u++) level(K - u, pk[u]);
    }
    for (; K >= 0; K--) {
        p = p + src[c0 + (size_t)K * pl];
        level(K, pow_kappa(p));
    }
}

// pkap and gz from (src, th) over all columns (cgrid 276-292, dgrid 433-437 +
// 606-626); optional plev output, optional heating rate of the D step and
// optional playkap on the interior.  After the D step the latter is what
// the next C step's scan would compute from the same dp, level by level
// with the same operations (cgrid 89-101).
__global__ void __launch_bounds__(64) k_scan_pg(Geo g, double ptop, double pktop, double gravity,
                                                const double *__restrict__ src, const double *__restrict__ th,
                                                const double *__restrict__ nf, double *__restrict__ plev,
                                                double *__restrict__ pkap, double *__restrict__ gz,
                                                double *__restrict__ heat, double *__restrict__ plk_out) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y;
    if (I >= g.NI) return;
    const size_t pl = g.plane(), c0 = g.ix2(I, J);
    const bool interior = I >= ATM_HALO && I < ATM_HALO + g.nx && J >= ATM_HALO && J < ATM_HALO + g.ny;
    const bool want_heat = heat && interior;
    const bool want_plk = plk_out && interior;
    double p = ptop, pk1 = pktop;
    double nf1 = want_heat ? nf[c0 + (size_t)g.L * pl] : 0.;
    auto level = [&](int K, double s, double pv, double pk) {
        const size_t x = c0 + (size_t)K * pl;
        if (plev) plev[x] = pv;
        pkap[x] = pk;
        if (want_heat) {
            const double plk = (pk1 - pk) / log(pk1 / pk);
            const double nf0 = nf[x];
            heat[x] = (nf1 - nf0) / (ATM_CP_DRY * s / gravity) / plk;
            nf1 = nf0;
        }
        if (want_plk) plk_out[x] = (pk1 - pk) / log(pk1 / pk);
        pk1 = pk;
    };
    int K = g.L - 1;
    for (; K >= SCAN_CH - 1; K -= SCAN_CH) {
        double s[SCAN_CH], pv[SCAN_CH], pk[SCAN_CH];
#pragma unroll
        for (int u = 0; u < SCAN_CH; u++) s[u] = src[c0 + (size_t)(K - u) * pl];
#pragma unroll
        for (int u = 0; u < SCAN_CH; u++) {
            p = p + s[u];
            pv[u] = p;
        }
#pragma unroll
        for (int u = 0; u < SCAN_CH; u++) pk[u] = pow_kappa(pv[u]);
#pragma unroll
        for (int u = 0; u < SCAN_CH; u++) level(K - u, s[u], pv[u], pk[u]);
    }
    for (; K >= 0; K--) {
        const double s = src[c0 + (size_t)K * pl];
        p = p + s;
        level(K, s, p, pow_kappa(p));
    }
    double z = gz[c0];
    double pk = pkap[c0];
    for (K = 0; K < g.L; K++) {
        const double pkn = K + 1 == g.L ? pktop : pkap[c0 + (size_t)(K + 1) * pl];
        z = z + ATM_CP_DRY * th[c0 + (size_t)K * pl] * (pk - pkn);
        gz[c0 + (size_t)(K + 1) * pl] = z;
        pk = pkn;
    }
}

// ---------------------------------------------------------------------------
// The same scans with a wavefront over 16 columns and 4 chunks of levels
// (nlay <= 4 CMAX): lane (c, q) = (lane % 16, lane / 16) holds the levels
// [q C, q C + n) of column c, C = ceil(nlay / 4).  The two serial sums (dp
// down from the top, gz up from the surface) pass from chunk to chunk in
// the order of the loops above, so every partial sum is the sequential one
// and every output the same double; the pow, log and divisions of the
// levels run on all 64 lanes, four times the columns' parallelism.  Outputs
// are optional (nullptr: not written); th == nullptr skips gz.
// ---------------------------------------------------------------------------
constexpr int SC_COLS = 16, SC_PARTS = 4;

// F32: the mixed form, whose heating rate and the playkap it divides by
// (tendencies of pt) are formed in float: playkap = pk d / log1p(d) with
// d = (pk1 - pk) / pk in double, so the logarithm of a ratio near 1 keeps
// float's relative precision; pkap, plev and gz stay double.
template <int CMAX, bool F32 = false>
__device__ __forceinline__ void scan_cols_body(int bxs, int bys, Geo g, int i0, int j0, int ni, double ptop, double pktop,
                                                  double gravity, const double *__restrict__ src,
                                                  const double *__restrict__ th, const double *__restrict__ nf,
                                                  double *__restrict__ plev, double *__restrict__ pkap,
                                                  double *__restrict__ gz, double *__restrict__ heat,
                                                  double *__restrict__ plk_out) {
    static_assert(SC_COLS * SC_PARTS == 64, "one wavefront");
    const int lane = threadIdx.x, q = lane / SC_COLS;
    const int I = i0 + bxs * SC_COLS + lane % SC_COLS, J = j0 + bys;
    const bool col = I < i0 + ni;  // lanes past the last column still pass the sums on
    const int C = (g.L + SC_PARTS - 1) / SC_PARTS;
    const int k0 = min(q * C, g.L), n = min(C, g.L - k0);
    const size_t pl = g.plane(), c0 = g.ix2(col ? I : i0, J);
    // the lane's first level; level k0 + u is u planes on (u pl is uniform)
    const size_t xb = c0 + (size_t)k0 * pl;
    const bool interior = col && I >= ATM_HALO && I < ATM_HALO + g.nx && J >= ATM_HALO && J < ATM_HALO + g.ny;
    const bool want_heat = heat && interior, want_plk = plk_out && interior;
    double s[CMAX], pk[CMAX];
#pragma unroll
    for (int u = 0; u < CMAX; u++) {
        s[u] = 0.;
        pk[u] = 1.;
        if (u < n && col) s[u] = src[xb + (size_t)u * pl];
    }
    // p = ptop + the dp above and at each level, from the top chunk down
    double p = ptop;
#pragma unroll
    for (int ph = SC_PARTS - 1; ph >= 0; ph--) {
        if (q == ph) {
#pragma unroll
            for (int u = CMAX - 1; u >= 0; u--)
                if (u < n) {
                    p = p + s[u];
                    pk[u] = p;
                }
        }
        if (ph > 0) {
            const double above = __shfl_down(p, SC_COLS);
            if (q == ph - 1) p = above;
        }
    }
#pragma unroll
    for (int u = 0; u < CMAX; u++)
        if (u < n && plev && col) plev[xb + (size_t)u * pl] = pk[u];
    static_assert(CMAX % 4 == 0, "pow_kappa in groups of 4");
#pragma unroll
    for (int u0 = 0; u0 < CMAX; u0 += 4) {  // levels past n hold 1 and stay unused
        double v[4] = {pk[u0], pk[u0 + 1], pk[u0 + 2], pk[u0 + 3]};
        pow_kappa_n<4>(v);
#pragma unroll
        for (int i = 0; i < 4; i++) pk[u0 + i] = v[i];
    }
#pragma unroll
    for (int u = 0; u < CMAX; u++)
        if (u < n && pkap && col) pkap[xb + (size_t)u * pl] = pk[u];
    // pkap of the level above the chunk: the next chunk's first, or the top
    double pk_up = __shfl_down(pk[0], SC_COLS);
    if (k0 + n >= g.L) pk_up = pktop;
    if (want_heat || want_plk) {
        double nfv[CMAX + 1];
#pragma unroll
        for (int u = 0; u <= CMAX; u++) nfv[u] = (want_heat && u <= n) ? nf[xb + (size_t)u * pl] : 0.;
#pragma unroll
        for (int u = 0; u < CMAX; u++)
            if (u < n) {
                const size_t x = xb + (size_t)u * pl;
                const double pk1 = u + 1 < n ? pk[u + 1] : pk_up;
                if constexpr (F32) {
                    const float d = (float)((pk1 - pk[u]) / pk[u]);
                    const float plk = (float)pk[u] * d * __builtin_amdgcn_rcpf(log1pf(d));
                    if (want_heat)
                        heat[x] = (double)((float)(nfv[u + 1] - nfv[u]) *
                                           __builtin_amdgcn_rcpf((float)(ATM_CP_DRY / gravity) * (float)s[u] * plk));
                    if (want_plk) plk_out[x] = (double)plk;
                } else {
                    const double plk = (pk1 - pk[u]) / log(pk1 / pk[u]);
                    if (want_heat) heat[x] = (nfv[u + 1] - nfv[u]) / (ATM_CP_DRY * s[u] / gravity) / plk;
                    if (want_plk) plk_out[x] = plk;
                }
            }
    }
    if (!th) return;
    // gz: the terms of the levels, then their sum from the surface chunk up
    double tz[CMAX];
#pragma unroll
    for (int u = 0; u < CMAX; u++) {
        tz[u] = 0.;
        if (u < n && col) {
            const double pkn = u + 1 < n ? pk[u + 1] : pk_up;
            tz[u] = ATM_CP_DRY * th[xb + (size_t)u * pl] * (pk[u] - pkn);
        }
    }
    double z = col ? gz[c0] : 0.;
#pragma unroll
    for (int ph = 0; ph < SC_PARTS; ph++) {
        if (q == ph) {
#pragma unroll
            for (int u = 0; u < CMAX; u++)
                if (u < n) {
                    z = z + tz[u];
                    tz[u] = z;
                }
        }
        if (ph + 1 < SC_PARTS) {
            const double below = __shfl_up(z, SC_COLS);
            if (q == ph + 1) z = below;
        }
    }
#pragma unroll
    for (int u = 0; u < CMAX; u++)
        if (u < n && col) gz[xb + (size_t)(u + 1) * pl] = tz[u];
}

template <int CMAX, bool F32>
__global__ void __launch_bounds__(64, 4) k_scan_cols(Geo g, int i0, int j0, int ni, double ptop, double pktop,
                                                  double gravity, const double *__restrict__ src,
                                                  const double *__restrict__ th, const double *__restrict__ nf,
                                                  double *__restrict__ plev, double *__restrict__ pkap,
                                                  double *__restrict__ gz, double *__restrict__ heat,
                                                  double *__restrict__ plk_out) {
    scan_cols_body<CMAX, F32>((int)blockIdx.x, (int)blockIdx.y, g, i0, j0, ni, ptop, pktop, gravity, src, th, nf, plev,
                              pkap, gz, heat, plk_out);
}

constexpr int SC_CMAX = 16;  // k_scan_cols for nlay <= 64

// the scans of the columns (i0.., j0..) x ni x nj: k_scan_cols when nlay
// fits, else one thread per column (k_scan_pg / k_scan_plk)
void launch_scans(atm_ctx *c, int i0, int j0, int ni, int nj, double gravity, const double *src, const double *th,
                  const double *nf, double *plev, double *pkap, double *gz, double *heat, double *plk_out) {
    const Geo &g = c->g;
    if (g.L <= SC_PARTS * SC_CMAX) {
        if (c->mixed)
            hipLaunchKernelGGL((k_scan_cols<SC_CMAX, true>), dim3((ni + SC_COLS - 1) / SC_COLS, nj), dim3(64), 0,
                               c->stream, g, i0, j0, ni, c->ak0, c->pktop, gravity, src, th, nf, plev, pkap, gz, heat,
                               plk_out);
        else
            hipLaunchKernelGGL((k_scan_cols<SC_CMAX, false>), dim3((ni + SC_COLS - 1) / SC_COLS, nj), dim3(64), 0,
                               c->stream, g, i0, j0, ni, c->ak0, c->pktop, gravity, src, th, nf, plev, pkap, gz, heat,
                               plk_out);
    } else if (th) {
        hipLaunchKernelGGL(k_scan_pg, dim3((g.NI + 63) / 64, g.NJ), dim3(64), 0, c->stream, g, c->ak0, c->pktop,
                           gravity, src, th, nf, plev, pkap, gz, heat, plk_out);
    } else {
        hipLaunchKernelGGL(k_scan_plk, dim3((g.nx + 63) / 64, g.ny), dim3(64), 0, c->stream, g, c->ak0, c->pktop, src,
                           plk_out);
    }
}

// ---------------------------------------------------------------------------
// Batched ensemble stages: each kernel runs one stage for every member of an
// ensemble in one launch (the member table, ens_find in atm_fused.hpp); a
// workgroup does what the member's own launch does for the same tile, with
// the same body, so every member's values are those of its own launches.
// par holds each member's current copy of the double-buffered fields (bit m).
// ---------------------------------------------------------------------------
constexpr int ENS_CORNER_LEVELS = 4;  // corner kernel levels per workgroup (every choice gives the same values)

__device__ __forceinline__ unsigned tiles_d(int n, int t) { return (unsigned)((n + t - 1) / t); }

// the member's coefficients for a step of dt (as atm_fused_cgrid/dgrid set them)
__device__ __forceinline__ DynCoef ens_coef(const EnsMember &e, double dt) {
    DynCoef c = e.coef;
    c.dt = dt;
    c.dtdx = dt * c.rdx;
    c.dtdy = dt * c.rdy;
    return c;
}

// C1 (cgrid) or D1 of every member
template <int VAR, bool CGRID, class R>
__global__ __launch_bounds__(NT_PAIR, (pair_waves<R, CGRID>())) void k_pair_ens(const EnsMember *__restrict__ tab,
                                                                              EnsIndex ix, int nm,
                                                                              unsigned long long par, double dt) {
    int m;
    unsigned l;
    if (!ens_find(ix, nm, m, l)) return;
    const EnsMember &e = tab[m];
    const int tx = (int)tiles_d(e.g.nx, TX), ty = (int)tiles_d(e.g.ny, TY_PAIR);
    const int bx = (int)(l % tx), by = (int)((l / tx) % ty), K = (int)(l / tx / ty);
    const int p = (int)((par >> m) & 1ull);
    if constexpr (CGRID)
        pair_body<VAR, true, R>(bx, by, K, e.g, ens_coef(e, dt), e.p[0][p], e.p[1][p], e.p[2][p], e.p[3][p],
                                nullptr, e.plk, e.net_flux, e.dpc, e.ptc, e.uc, e.vc, e.ua, e.va, nullptr);
    else
        pair_body<VAR, false, R>(bx, by, K, e.g, ens_coef(e, dt), e.p[0][p], e.p[1][p], e.uc2, e.vc2, e.heatd,
                                 nullptr, nullptr, e.p[0][1 - p], e.p[1][1 - p], nullptr, nullptr, nullptr, nullptr,
                                 e.nfw);
}

template <class R>
__global__ __launch_bounds__(NT_WIND, 8) void k_c2_ens(const EnsMember *__restrict__ tab, EnsIndex ix, int nm,
                                                      unsigned long long par, double dt) {
    int m;
    unsigned l;
    if (!ens_find(ix, nm, m, l)) return;
    const EnsMember &e = tab[m];
    const int tx = (int)tiles_d(e.g.nx, TX), ty = (int)tiles_d(e.g.ny, TY_WIND);
    const int bx = (int)(l % tx), by = (int)((l / tx) % ty), K = (int)(l / tx / ty);
    const int p = (int)((par >> m) & 1ull);
    c2_body<R>(bx, by, K, e.g, ens_coef(e, dt), e.uc, e.vc, e.p[2][p], e.p[3][p], e.pkap, e.gz, e.ua, e.va, e.uc2,
               e.vc2);
}

template <class R>
__global__ __launch_bounds__(NT_WIND) void k_d2_ens(const EnsMember *__restrict__ tab, EnsIndex ix, int nm,
                                                   unsigned long long par, double dt) {
    int m;
    unsigned l;
    if (!ens_find(ix, nm, m, l)) return;
    const EnsMember &e = tab[m];
    const int tx = (int)tiles_d(e.g.nx, TX), ty = (int)tiles_d(e.g.ny, TY_WIND);
    const int bx = (int)(l % tx), by = (int)((l / tx) % ty), K = (int)(l / tx / ty);
    const int p = (int)((par >> m) & 1ull);
    d2_body<R>(bx, by, K, e.g, ens_coef(e, dt), e.p[2][p], e.p[3][p], e.uc2, e.vc2, e.pkapb, e.gzb, e.p[2][1 - p],
               e.p[3][1 - p], e.nfw);
}

__global__ void __launch_bounds__(NT_WIND) k_bcorners_ens(const EnsMember *__restrict__ tab, EnsIndex ix, int nm) {
    int m;
    unsigned l;
    if (!ens_find(ix, nm, m, l)) return;
    const EnsMember &e = tab[m];
    const int tx = (e.g.nx + TX) / TX, ty = (e.g.ny + TY_WIND) / TY_WIND;
    bcorners_body<ENS_CORNER_LEVELS>((int)(l % tx), (int)((l / tx) % ty), (int)(l / tx / ty), e.g, e.pkap, e.gz,
                                     e.pkapb, e.gzb);
}

// the column scans after C1 (pkap, gz, heating rate) or after D1 (plev, pkap,
// gz and the next C step's playkap), over all columns of every member
template <bool DSTEP, bool F32>
__global__ void __launch_bounds__(64, 4) k_scan_ens(const EnsMember *__restrict__ tab, EnsIndex ix, int nm,
                                                   unsigned long long par) {
    int m;
    unsigned l;
    if (!ens_find(ix, nm, m, l)) return;
    const EnsMember &e = tab[m];
    const int tx = (int)tiles_d(e.g.NI, SC_COLS);
    const int bx = (int)(l % tx), by = (int)(l / tx);
    if constexpr (DSTEP) {
        const int q = 1 - (int)((par >> m) & 1ull);  // the copy D1 wrote
        scan_cols_body<SC_CMAX, F32>(bx, by, e.g, 0, 0, e.g.NI, e.ak0, e.pktop, e.coef.gravity, e.p[0][q], e.p[1][q],
                                e.net_flux, e.plev, e.pkap, e.gz, nullptr, e.plk);
    } else {
        scan_cols_body<SC_CMAX, F32>(bx, by, e.g, 0, 0, e.g.NI, e.ak0, e.pktop, e.coef.gravity, e.dpc, e.ptc,
                                e.net_flux, nullptr, e.pkap, e.gz, e.heatd, nullptr);
    }
}

#undef D

}  // namespace

static unsigned tiles_h(int n, int t) { return (unsigned)((n + t - 1) / t); }
unsigned atm_ens_blocks_pair(const Geo &g) { return tiles_h(g.nx, TX) * tiles_h(g.ny, TY_PAIR) * (unsigned)g.L; }
unsigned atm_ens_blocks_wind(const Geo &g) { return tiles_h(g.nx, TX) * tiles_h(g.ny, TY_WIND) * (unsigned)g.L; }
unsigned atm_ens_blocks_corner(const Geo &g) {
    return (unsigned)((g.nx + TX) / TX) * (unsigned)((g.ny + TY_WIND) / TY_WIND) * tiles_h(g.L, ENS_CORNER_LEVELS);
}
unsigned atm_ens_blocks_scan(const Geo &g) { return tiles_h(g.NI, SC_COLS) * (unsigned)g.NJ; }
bool atm_ens_scan_fits(const Geo &g) { return g.L <= SC_PARTS * SC_CMAX; }

void atm_ens_launch_pair(hipStream_t s, const EnsMember *tab, const EnsIndex &ix, int nm, unsigned long long par,
                         double dt, bool cgrid, bool mixed, unsigned n) {
    const dim3 grid(n), blk(NT_PAIR);
    if (cgrid) {
        if (mixed)
            hipLaunchKernelGGL((k_pair_ens<V_PPM_U, true, float>), grid, blk, 0, s, tab, ix, nm, par, dt);
        else
            hipLaunchKernelGGL((k_pair_ens<V_PPM_U, true, double>), grid, blk, 0, s, tab, ix, nm, par, dt);
    } else {
        if (mixed)
            hipLaunchKernelGGL((k_pair_ens<V_PPM_NN, false, float>), grid, blk, 0, s, tab, ix, nm, par, dt);
        else
            hipLaunchKernelGGL((k_pair_ens<V_PPM_NN, false, double>), grid, blk, 0, s, tab, ix, nm, par, dt);
    }
}
void atm_ens_launch_c2(hipStream_t s, const EnsMember *tab, const EnsIndex &ix, int nm, unsigned long long par,
                       double dt, bool mixed, unsigned n) {
    if (mixed)
        hipLaunchKernelGGL(k_c2_ens<float>, dim3(n), dim3(NT_WIND), 0, s, tab, ix, nm, par, dt);
    else
        hipLaunchKernelGGL(k_c2_ens<double>, dim3(n), dim3(NT_WIND), 0, s, tab, ix, nm, par, dt);
}
void atm_ens_launch_d2(hipStream_t s, const EnsMember *tab, const EnsIndex &ix, int nm, unsigned long long par,
                       double dt, bool mixed, unsigned n) {
    if (mixed)
        hipLaunchKernelGGL(k_d2_ens<float>, dim3(n), dim3(NT_WIND), 0, s, tab, ix, nm, par, dt);
    else
        hipLaunchKernelGGL(k_d2_ens<double>, dim3(n), dim3(NT_WIND), 0, s, tab, ix, nm, par, dt);
}
void atm_ens_launch_scan(hipStream_t s, const EnsMember *tab, const EnsIndex &ix, int nm, unsigned long long par,
                         bool dstep, bool mixed, unsigned n) {
    if (dstep && mixed)
        hipLaunchKernelGGL((k_scan_ens<true, true>), dim3(n), dim3(64), 0, s, tab, ix, nm, par);
    else if (dstep)
        hipLaunchKernelGGL((k_scan_ens<true, false>), dim3(n), dim3(64), 0, s, tab, ix, nm, par);
    else if (mixed)
        hipLaunchKernelGGL((k_scan_ens<false, true>), dim3(n), dim3(64), 0, s, tab, ix, nm, par);
    else
        hipLaunchKernelGGL((k_scan_ens<false, false>), dim3(n), dim3(64), 0, s, tab, ix, nm, par);
}
void atm_ens_launch_corners(hipStream_t s, const EnsMember *tab, const EnsIndex &ix, int nm, unsigned n) {
    hipLaunchKernelGGL(k_bcorners_ens, dim3(n), dim3(NT_WIND), 0, s, tab, ix, nm);
}

// ---------------------------------------------------------------------------
// Step orchestration of the fused path.  Buffers: prognostic fields have two
// copies (cur, nxt); the C step writes its own arrays; the D step writes nxt.
// ---------------------------------------------------------------------------
// tiles of t cells over n cells (the last may overhang)
static unsigned tiles(int n, int t) { return (unsigned)((n + t - 1) / t); }

// playkap of the current dp for the C step's heating rate, when dp changed
// since the last D step's scan left it (a remap, a put)
int atm_fused_plk(atm_ctx *c) {
    const Geo &g = c->g;
    atm_timer_begin(c, ATM_T_SCAN);
    launch_scans(c, ATM_HALO, ATM_HALO, g.nx, g.ny, c->coef.gravity, c->pbuf[0][c->cur], nullptr, nullptr, nullptr,
                 nullptr, nullptr, nullptr, c->plk);
    atm_timer_end(c, ATM_T_SCAN);
    c->plk_valid = true;
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}

// The stencil kernels at the context's tile heights (atm_set_tiles): dp/pt
// tiles of 32 x 8 or 32 x 16 cells, wind tiles of 32 x 16 or 32 x 8, one
// thread per cell.  Every height gives the same values.
template <int VAR, bool CGRID, class R, class... A>
static void launch_pair(atm_ctx *c, A... a) {
    const Geo &g = c->g;
    if (c->tile_pair == 16)
        hipLaunchKernelGGL((k_pair<VAR, CGRID, R, 16>), dim3(tiles(g.nx, TX), tiles(g.ny, 16), g.L), dim3(TX * 16), 0,
                           c->stream, a...);
    else
        hipLaunchKernelGGL((k_pair<VAR, CGRID, R, 8>), dim3(tiles(g.nx, TX), tiles(g.ny, 8), g.L), dim3(TX * 8), 0,
                           c->stream, a...);
}
template <class R, class... A>
static void launch_c2(atm_ctx *c, A... a) {
    const Geo &g = c->g;
    if (c->tile_wind == 8)
        hipLaunchKernelGGL((k_c2<R, 8>), dim3(tiles(g.nx, TX), tiles(g.ny, 8), g.L), dim3(TX * 8), 0, c->stream, a...);
    else
        hipLaunchKernelGGL((k_c2<R, 16>), dim3(tiles(g.nx, TX), tiles(g.ny, 16), g.L), dim3(TX * 16), 0, c->stream,
                           a...);
}
template <class R, class... A>
static void launch_d2(atm_ctx *c, A... a) {
    const Geo &g = c->g;
    if (c->tile_wind == 8)
        hipLaunchKernelGGL((k_d2<R, 8>), dim3(tiles(g.nx, TX), tiles(g.ny, 8), g.L), dim3(TX * 8), 0, c->stream, a...);
    else
        hipLaunchKernelGGL((k_d2<R, 16>), dim3(tiles(g.nx, TX), tiles(g.ny, 16), g.L), dim3(TX * 16), 0, c->stream,
                           a...);
}

int atm_fused_cgrid(atm_ctx *c, FusedBufs &b, double dt) {
    const Geo &g = c->g;
    c->coef.dt = dt;
    c->coef.dtdx = dt * c->coef.rdx;
    c->coef.dtdy = dt * c->coef.rdy;
    const DynCoef &k = c->coef;
    if (!c->plk_valid) {
        const int rc = atm_fused_plk(c);
        if (rc) return rc;
    }
    atm_timer_begin(c, ATM_T_C1);
    if (c->mixed)
        launch_pair<V_PPM_U, true, float>(c, g, k, b.dp, b.pt, b.ud, b.vd, (const double *)nullptr, c->plk,
                                          c->net_flux, c->dpc, c->ptc, c->uc, c->vc, c->ua, c->va, (int *)nullptr);
    else
        launch_pair<V_PPM_U, true, double>(c, g, k, b.dp, b.pt, b.ud, b.vd, (const double *)nullptr, c->plk,
                                           c->net_flux, c->dpc, c->ptc, c->uc, c->vc, c->ua, c->va, (int *)nullptr);
    atm_timer_end(c, ATM_T_C1);
    {
        double *fill[] = {c->dpc, c->ptc, c->uc, c->vc, c->ua, c->va};
        const int nk[] = {g.L, g.L, g.L, g.L, 1, 1};
        atm_timer_begin(c, ATM_T_HALO);
        int rc = atm_launch_halo_multi(c, fill, nk, 6);
        atm_timer_end(c, ATM_T_HALO);
        if (rc) return rc;
    }
    atm_timer_begin(c, ATM_T_SCAN);
    launch_scans(c, 0, 0, g.NI, g.NJ, k.gravity, c->dpc, c->ptc, c->net_flux, nullptr, c->pkap, c->gz, c->heatd,
                 nullptr);
    atm_timer_end(c, ATM_T_SCAN);
    atm_timer_begin(c, ATM_T_C2);
    if (c->mixed)
        launch_c2<float>(c, g, k, c->uc, c->vc, b.ud, b.vd, c->pkap, c->gz, c->ua, c->va, c->uc2, c->vc2);
    else
        launch_c2<double>(c, g, k, c->uc, c->vc, b.ud, b.vd, c->pkap, c->gz, c->ua, c->va, c->uc2, c->vc2);
    atm_timer_end(c, ATM_T_C2);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}

// the halo fill of cgrid_halo_exchange for the fused path: dpc/ptc are
// already periodic, uc/vc were updated on the interior
int atm_fused_cgrid_halo(atm_ctx *c) {
    double *fill[] = {c->uc2, c->vc2};
    const int nk[] = {c->g.L, c->g.L};
    atm_timer_begin(c, ATM_T_HALO);
    int rc = atm_launch_halo_multi(c, fill, nk, 2);
    atm_timer_end(c, ATM_T_HALO);
    return rc;
}

int atm_fused_dgrid(atm_ctx *c, FusedBufs &cur, FusedBufs &nxt, double dt) {
    const Geo &g = c->g;
    c->coef.dt = dt;
    c->coef.dtdx = dt * c->coef.rdx;
    c->coef.dtdy = dt * c->coef.rdy;
    const DynCoef &k = c->coef;
    atm_timer_begin(c, ATM_T_D1);
    if (c->mixed)
        launch_pair<V_PPM_NN, false, float>(c, g, k, cur.dp, cur.pt, c->uc2, c->vc2, c->heatd, (const double *)nullptr,
                                            (const double *)nullptr, nxt.dp, nxt.pt, (double *)nullptr,
                                            (double *)nullptr, (double *)nullptr, (double *)nullptr, c->nfw);
    else
        launch_pair<V_PPM_NN, false, double>(c, g, k, cur.dp, cur.pt, c->uc2, c->vc2, c->heatd,
                                             (const double *)nullptr, (const double *)nullptr, nxt.dp, nxt.pt,
                                             (double *)nullptr, (double *)nullptr, (double *)nullptr,
                                             (double *)nullptr, c->nfw);
    atm_timer_end(c, ATM_T_D1);
    {
        double *fill[] = {nxt.dp, nxt.pt};
        const int nk[] = {g.L, g.L};
        atm_timer_begin(c, ATM_T_HALO);
        int rc = atm_launch_halo_multi(c, fill, nk, 2);
        atm_timer_end(c, ATM_T_HALO);
        if (rc) return rc;
    }
    atm_timer_begin(c, ATM_T_SCAN);
    launch_scans(c, 0, 0, g.NI, g.NJ, k.gravity, nxt.dp, nxt.pt, c->net_flux, c->plev, c->pkap, c->gz, nullptr,
                 c->plk);
    launch_bcorners(c, c->bc_levels);
    atm_timer_end(c, ATM_T_SCAN);
    atm_timer_begin(c, ATM_T_D2);
    if (c->mixed)
        launch_d2<float>(c, g, k, cur.ud, cur.vd, c->uc2, c->vc2, c->pkapb, c->gzb, nxt.ud, nxt.vd, c->nfw);
    else
        launch_d2<double>(c, g, k, cur.ud, cur.vd, c->uc2, c->vc2, c->pkapb, c->gzb, nxt.ud, nxt.vd, c->nfw);
    atm_timer_end(c, ATM_T_D2);
    ATM_HIP(c, hipGetLastError());
    return ATM_OK;
}
