// lin_step.hpp -- one model step (C-grid step, halo fill, D-grid step,
// optional remap) as a single-assignment program, run forward, tangent-
// linear or adjoint by a backend (the HIP kernels of k_lin.hip, or the
// host loops of lin_host.cpp that the CPU tests use).
//
// The program follows the reference's call order (src/mod_model.f90:144-173)
// and each statement follows the loop nest it cites.  Its state is the
// interior of dp, pt, ud, vd at the start of a step; the halos are the
// periodic fill of the interior (mod_sync), formed by the program's first
// operations, so the wind ring the remap reads (Q5) is the start state's.
//
// Backend B provides, each over a box of (I, J) and levels [k0, k0 + nk):
//   template <class F> fwd / tan / adj (const LinArgs &, Rng, k0, nk)
//   scan(mode, gz, Rng, ...)            the column scans (lin_ops.hpp)
//   halo_copy(mode, nk, sv, st, dv, dt) dst = periodic fill of src's interior
//   columns(mode, kind, ops, n)         column operations of the remap
//   copy, fill, alloc, release          plain memory operations
#pragma once

#include <array>
#include <functional>
#include <initializer_list>
#include <type_traits>
#include <vector>

#include "lin_ops.hpp"

struct Rng {
    int I0, I1, J0, J1;  // inclusive
};

// One column operation of the remap: its box of columns, its operands and
// its workspace.  A backend's columns(mode, kind, ops, n) runs n of them,
// independent of each other, together.
struct LinColOp {
    Rng r;
    LinArgs x;
    double *w;
};
constexpr int LIN_MAXCOL = 3;

enum LinMode { LIN_FWD = 0, LIN_TAN = 1, LIN_ADJ = 2 };

// single-assignment arrays of the step
enum LinArr {
    // start state: X_* hold the interior (the step's input), *0 the filled arrays
    A_XDP, A_XPT, A_XUD, A_XVD, A_DP0, A_PT0, A_UD0, A_VD0,
    A_NF,  // net flux: a fixed forcing, not differentiated
    // C grid (mod_sw_dyn.f90:80-416)
    A_PLEV_C, A_PKAP_C, A_HEAT_C, A_UC0, A_VC0, A_CX0, A_CY0, A_UA, A_VA, A_DENX0, A_DENY0,
    A_TMP1, A_LAP1, A_DPFX, A_DPCX, A_TMP2, A_DPFY, A_DPC, A_TMP3, A_LAP2, A_FXQ, A_PTCX, A_TMP4, A_FYQ, A_PTC,
    A_PLEV_C2, A_PKAP_C2, A_GZ_C2, A_KE_C, A_VORT_C, A_PGFX_C, A_PGFY_C, A_DENX1, A_DENY1, A_CX1, A_CY1,
    A_TMP5, A_LAP3, A_UC1, A_TMP6, A_VC1,
    // after cgrid_halo_exchange
    A_DPCH, A_PTCH, A_UCH, A_VCH,
    // D grid (424-773)
    A_PLEV_D, A_PKAP_D, A_CXD, A_CYD, A_DENXD, A_DENYD, A_UB, A_VB, A_HEAT_D,
    A_TMP7, A_LAP4, A_DPFXD, A_DPCDX, A_TMP8, A_DPFYD, A_DP1, A_TMP9, A_LAP5, A_FXQD, A_PTCDX, A_TMP10, A_FYQD,
    A_PT1, A_PLEV_D2, A_PKAP_D2, A_GZ_D2, A_PKB, A_GZB, A_KE_D, A_VORT_D, A_PGFX_D, A_PGFY_D, A_DIV, A_TMP11,
    A_DIV4, A_TMP12, A_LAP6, A_UD1, A_TMP13, A_VD1,
    // remap (mod_remap.f90:27-163)
    A_RPLEV, A_RPKAP, A_RPLK, A_RGZ, A_TPLEV, A_TPKAP, A_TPLK, A_UFX, A_UFD, A_VFX, A_VFD, A_UDM, A_VDM,
    A_UR, A_VR, A_ENER, A_KER, A_EREM, A_DP2, A_PT2,
    A_COUNT
};

// levels of each array: L, L + 1 (levels), 1 (a plane)
inline int lin_levels(int id, int L) {
    switch (id) {
    case A_NF: case A_PLEV_C: case A_PKAP_C: case A_PLEV_C2: case A_PKAP_C2: case A_GZ_C2:
    case A_PLEV_D: case A_PKAP_D: case A_PLEV_D2: case A_PKAP_D2: case A_GZ_D2: case A_PKB: case A_GZB:
    case A_RPLEV: case A_RPKAP: case A_RGZ: case A_TPLEV: case A_TPKAP:
    case A_UFX: case A_UFD: case A_VFX: case A_VFD:
        return L + 1;
    case A_UA: case A_VA: case A_UB: case A_VB:
        return 1;
    default:
        return L;
    }
}

inline Rng lin_box(const Geo &g, int a, int b) { return Rng{a, g.NI - 1 - b, a, g.NJ - 1 - b}; }
inline Rng lin_box4(const Geo &g, int ai, int bi, int aj, int bj) { return Rng{ai, g.NI - 1 - bi, aj, g.NJ - 1 - bj}; }

#include "lin_remap.hpp"

// The step program.  E: the runner (LinCore below) with op<F>, scan,
// halo_copy and remap.
template <class E>
void lin_program(E &e, bool remap) {
    const Geo &g = e.g;
    const int L = g.L;
    auto R = [&](int a, int b) { return lin_box(g, a, b); };
    auto RX = [&](int ai, int bi, int aj, int bj) { return lin_box4(g, ai, bi, aj, bj); };
    const Rng all = R(0, 0);
    // the start state with its periodic halo
    e.halo_copy(A_XDP, A_DP0);
    e.halo_copy(A_XPT, A_PT0);
    e.halo_copy(A_XUD, A_UD0);
    e.halo_copy(A_XVD, A_VD0);
    // ---- cgrid_dynamics_step(dt / 2) (Q18) ----
    e.set_dt(e.step_dt / 2.0);
    e.scan(false, all, A_DP0, -1, A_PLEV_C, A_PKAP_C, -1);
    e.template op<OpHeat>(all, 0, L, {A_PKAP_C, A_NF, A_DP0}, {A_HEAT_C});
    e.template op<OpCWinds>(R(1, 1), 0, L, {A_UD0, A_VD0}, {A_UC0, A_VC0, A_CX0, A_CY0});
    e.template op<OpCUaVa>(R(1, 1), L - 1, 1, {A_UD0, A_VD0}, {A_UA, A_VA});
    e.template op<OpDenoms>(R(2, 2), 0, L, {A_UC0, A_VC0}, {A_DENX0, A_DENY0});
    e.template op<OpInnerY>(R(2, 2), 0, L, {A_DP0, A_CY0, A_DENY0}, {A_TMP1, A_LAP1});
    e.template op<OpOuter<V_PPM_U, 0, true>>(RX(5, 4, 5, 5), 0, L, {A_TMP1, A_CX0}, {A_DPFX});
    e.template op<OpDpcX>(R(5, 5), 0, L, {A_DP0, A_DPFX, A_LAP1}, {A_DPCX});
    e.template op<OpInnerX>(R(2, 2), 0, L, {A_DP0, A_CX0, A_DENX0}, {A_TMP2});
    e.template op<OpOuter<V_PPM_U, 1, true>>(RX(5, 5, 5, 4), 0, L, {A_TMP2, A_CY0}, {A_DPFY});
    e.template op<OpDpcY>(R(5, 5), 0, L, {A_DPCX, A_DPFY, A_LAP1}, {A_DPC});
    e.template op<OpInnerY>(R(2, 2), 0, L, {A_PT0, A_CY0, A_DENY0}, {A_TMP3, A_LAP2});
    e.template op<OpOuter<V_PPM_U, 0, false>>(RX(5, 4, 5, 5), 0, L, {A_TMP3, A_CX0}, {A_FXQ});
    e.template op<OpPtcX>(R(5, 5), 0, L, {A_DP0, A_PT0, A_FXQ, A_DPFX, A_LAP2}, {A_PTCX});
    e.template op<OpInnerX>(R(2, 2), 0, L, {A_PT0, A_CX0, A_DENX0}, {A_TMP4});
    e.template op<OpOuter<V_PPM_U, 1, false>>(RX(5, 5, 5, 4), 0, L, {A_TMP4, A_CY0}, {A_FYQ});
    e.template op<OpPtcY>(R(5, 5), 0, L, {A_DP0, A_FYQ, A_DPFY, A_LAP2, A_HEAT_C, A_DPC, A_PTCX}, {A_PTC});
    e.scan(true, R(5, 5), A_DPC, A_PTC, A_PLEV_C2, A_PKAP_C2, A_GZ_C2);
    e.template op<OpCKe>(R(5, 5), 0, L, {A_UC0, A_VC0, A_UA, A_VA}, {A_KE_C});
    e.template op<OpCVort>(R(2, 2), 0, L, {A_UC0, A_VC0}, {A_VORT_C});
    e.template op<OpCPgf>(R(7, 7), 0, L, {A_PKAP_C2, A_GZ_C2}, {A_PGFX_C, A_PGFY_C});
    e.template op<OpCDenoms2>(R(1, 1), 0, L, {A_UD0, A_VD0}, {A_DENX1, A_DENY1, A_CX1, A_CY1});
    e.template op<OpVortInnerX>(R(3, 3), 0, L, {A_VORT_C, A_CX1, A_DENX1}, {A_TMP5, A_LAP3}, -1, -1);
    e.template op<OpCUc>(R(7, 7), 0, L, {A_TMP5, A_CY1, A_VD0, A_LAP3, A_KE_C, A_PGFX_C, A_UC0}, {A_UC1});
    e.template op<OpVortInnerY>(R(3, 3), 0, L, {A_VORT_C, A_CY1, A_DENY1}, {A_TMP6}, -1, -1);
    e.template op<OpCVc>(R(7, 7), 0, L, {A_TMP6, A_CX1, A_UD0, A_LAP3, A_KE_C, A_PGFY_C, A_VC0}, {A_VC1});
    // ---- cgrid_halo_exchange ----
    e.halo_copy(A_DPC, A_DPCH);
    e.halo_copy(A_PTC, A_PTCH);
    e.halo_copy(A_UC1, A_UCH);
    e.halo_copy(A_VC1, A_VCH);
    // ---- dgrid_dynamics_step(dt) ----
    e.set_dt(e.step_dt);
    e.scan(false, all, A_DPCH, -1, A_PLEV_D, A_PKAP_D, -1);
    e.template op<OpDCourant>(all, 0, L, {A_UCH, A_VCH}, {A_CXD, A_CYD});
    e.template op<OpDenoms>(R(1, 1), 0, L, {A_UCH, A_VCH}, {A_DENXD, A_DENYD});
    e.template op<OpDUbVb>(Rng{1, g.NI - 1, 1, g.NJ - 1}, L - 1, 1, {A_UCH, A_VCH}, {A_UB, A_VB});
    e.template op<OpHeat>(all, 0, L, {A_PKAP_D, A_NF, A_DPCH}, {A_HEAT_D});
    e.template op<OpInnerY>(R(1, 1), 0, L, {A_DP0, A_CYD, A_DENYD}, {A_TMP7, A_LAP4});
    e.template op<OpOuter<V_PPM_NN, 0, true>>(RX(4, 3, 4, 4), 0, L, {A_TMP7, A_CXD}, {A_DPFXD});
    e.template op<OpDpcX>(R(4, 4), 0, L, {A_DP0, A_DPFXD, A_LAP4}, {A_DPCDX});
    e.template op<OpInnerX>(R(1, 1), 0, L, {A_DP0, A_CXD, A_DENXD}, {A_TMP8});
    e.template op<OpOuter<V_PPM_NN, 1, true>>(RX(4, 4, 4, 3), 0, L, {A_TMP8, A_CYD}, {A_DPFYD});
    e.template op<OpDpcY>(R(4, 4), 0, L, {A_DPCDX, A_DPFYD, A_LAP4}, {A_DP1});
    e.template op<OpInnerY>(R(1, 1), 0, L, {A_PT0, A_CYD, A_DENYD}, {A_TMP9, A_LAP5});
    e.template op<OpOuter<V_PPM_NN, 0, false>>(RX(4, 3, 4, 4), 0, L, {A_TMP9, A_CXD}, {A_FXQD});
    e.template op<OpPtcX>(R(4, 4), 0, L, {A_DP0, A_PT0, A_FXQD, A_DPFXD, A_LAP5}, {A_PTCDX});
    e.template op<OpInnerX>(R(1, 1), 0, L, {A_PT0, A_CXD, A_DENXD}, {A_TMP10});
    e.template op<OpOuter<V_PPM_NN, 1, false>>(RX(4, 4, 4, 3), 0, L, {A_TMP10, A_CYD}, {A_FYQD});
    e.template op<OpPtcY>(R(4, 4), 0, L, {A_DP0, A_FYQD, A_DPFYD, A_LAP5, A_HEAT_D, A_DP1, A_PTCDX}, {A_PT1});
    // dp = dpc, pt = ptc on is+4..ie-4 (594-598): DP1 and PT1 are those arrays
    e.scan(true, R(4, 4), A_DP1, A_PT1, A_PLEV_D2, A_PKAP_D2, A_GZ_D2);
    e.template op<OpCorners>(R(6, 5), 0, L, {A_PKAP_D2}, {A_PKB}, 0);
    e.template op<OpCorners>(R(6, 5), 0, L, {A_GZ_D2}, {A_GZB}, 1);
    e.template op<OpDKe>(R(4, 4), 0, L, {A_UD0, A_VD0, A_UB, A_VB}, {A_KE_D});
    e.template op<OpDVort>(R(1, 1), 0, L, {A_UD0, A_VD0}, {A_VORT_D});
    e.template op<OpDPgf>(R(7, 7), 0, L, {A_PKB, A_GZB}, {A_PGFX_D, A_PGFY_D});
    e.template op<OpDDiv>(R(4, 4), 0, L, {A_UD0, A_VD0}, {A_DIV});
    e.template op<OpDel2>(R(5, 5), 0, L, {A_DIV}, {A_TMP11}, 0);
    e.template op<OpDel2>(R(6, 6), 0, L, {A_TMP11}, {A_DIV4}, 1);
    e.template op<OpVortInnerX>(R(4, 4), 0, L, {A_VORT_D, A_CXD, A_DENXD}, {A_TMP12, A_LAP6}, 0, L - 1);
    e.template op<OpDUd>(R(7, 7), 0, L, {A_TMP12, A_CYD, A_VCH, A_LAP6, A_KE_D, A_DIV4, A_PGFX_D, A_UD0}, {A_UD1});
    e.template op<OpVortInnerY>(R(4, 4), 0, L, {A_VORT_D, A_CYD, A_DENYD}, {A_TMP13}, 0, L - 1);
    e.template op<OpDVd>(R(7, 7), 0, L, {A_TMP13, A_CXD, A_UCH, A_LAP6, A_KE_D, A_DIV4, A_PGFY_D, A_VD0}, {A_VD1});
    // ---- perform_vertical_remapping (the D winds' ring: the start state's halo, Q5) ----
    if (remap) lin_remap_program(e);
}

// the arrays holding the end state's interior
inline void lin_end_arrays(bool remap, int out[4]) {
    if (remap) {
        out[0] = A_DP2; out[1] = A_PT2; out[2] = A_UR; out[3] = A_VR;
    } else {
        out[0] = A_DP1; out[1] = A_PT1; out[2] = A_UD1; out[3] = A_VD1;
    }
}

// Arrays, constant levels and the three sweeps over the program.
template <class B>
struct LinCore {
    B &b;
    Geo g;
    DynCoef c;
    double ak0 = 0., pktop = 0.;
    double *val[A_COUNT] = {};
    double *drv[A_COUNT] = {};  // tangents (tangent sweep) or adjoints (adjoint sweep)
    LinMode mode = LIN_FWD;
    double step_dt = 0.;  // the model step; the C-grid half step uses step_dt / 2
    std::vector<std::function<void()>> rev;
    const double *ak = nullptr, *bk = nullptr;  // coordinates (device or host, as the backend)
    double *ws[RM_NWS] = {};                    // column workspaces of the remap

    LinCore(B &be, const Geo &geo) : b(be), g(geo) {}

    size_t count(int id) const { return g.plane() * (size_t)lin_levels(id, g.L); }
    // the remap's widest column box is isd-1..ied+1
    size_t ws_count() const { return (size_t)RM_SLOTS * (g.L + 2) * (size_t)(g.nx + 2) * (g.ny + 2); }

    bool allocate() {
        for (int i = 0; i < A_COUNT; i++) {
            val[i] = b.alloc(count(i));
            if (!val[i]) return false;
            if (i == A_NF) continue;  // not differentiated
            drv[i] = b.alloc(count(i));
            if (!drv[i]) return false;
        }
        for (int i = 0; i < RM_NWS; i++) {
            ws[i] = b.alloc(ws_count());
            if (!ws[i]) return false;
        }
        return true;
    }
    void release() {
        for (int i = 0; i < RM_NWS; i++) {
            b.release(ws[i]);
            ws[i] = nullptr;
        }
        for (int i = 0; i < A_COUNT; i++) {
            b.release(val[i]);
            b.release(drv[i]);
            val[i] = drv[i] = nullptr;
        }
    }

    // Constant levels: the top of plev / pkap (Ak(1), Ak(1)**kappa), the
    // surface of gz and of its corner values, the top of pkapb; gz_sfc and
    // gzb_sfc are planes, nf the net flux (nlay + 1 levels).
    void set_constants(const double *gz_sfc, const double *gzb_sfc, const double *nf) {
        const size_t p = g.plane(), L = g.L;
        for (int id : {A_PLEV_C, A_PLEV_C2, A_PLEV_D, A_PLEV_D2, A_RPLEV, A_TPLEV}) b.fill(val[id] + p * L, p, ak0);
        for (int id : {A_PKAP_C, A_PKAP_C2, A_PKAP_D, A_PKAP_D2, A_RPKAP, A_TPKAP}) b.fill(val[id] + p * L, p, pktop);
        b.fill(val[A_PKB] + p * L, p, pktop);
        for (int id : {A_GZ_C2, A_GZ_D2, A_RGZ}) b.copy(val[id], gz_sfc, p);
        b.copy(val[A_GZB], gzb_sfc, p);
        b.copy(val[A_NF], nf, p * (L + 1));
    }

    // ---- the runner interface of lin_program ----
    LinArgs args(std::initializer_list<int> in, std::initializer_list<int> out, int e0, int e1) const {
        LinArgs x{};
        x.g = g;
        x.c = c;
        x.e0 = e0;
        x.e1 = e1;
        x.ak = ak;
        x.bk = bk;
        x.top = ak0;
        int k = 0;
        for (int id : in) {
            x.iv[k] = val[id];
            x.it[k] = drv[id];
            k++;
        }
        k = 0;
        for (int id : out) {
            x.ov[k] = val[id];
            x.ot[k] = drv[id];
            k++;
        }
        return x;
    }

    template <class F>
    void op(Rng r, int k0, int nk, std::initializer_list<int> in, std::initializer_list<int> out, int e0 = 0,
            int e1 = 0) {
        const LinArgs x = args(in, out, e0, e1);
        if (mode == LIN_FWD)
            b.template fwd<F>(x, r, k0, nk);
        else if (mode == LIN_TAN)
            b.template tan<F>(x, r, k0, nk);
        else
            rev.push_back([this, x, r, k0, nk]() { b.template adj<F>(x, r, k0, nk); });
    }

    void scan(bool gz, Rng r, int src, int th, int plev, int pkap, int gzo) {
        const double *sv = val[src], *tv = th >= 0 ? val[th] : nullptr;
        double *sd = drv[src], *td = th >= 0 ? drv[th] : nullptr;
        double *pv = val[plev], *pd = drv[plev], *kv = val[pkap], *kd = drv[pkap];
        double *gv = gzo >= 0 ? val[gzo] : nullptr, *gd = gzo >= 0 ? drv[gzo] : nullptr;
        if (mode == LIN_ADJ)
            rev.push_back([=]() { b.scan(LIN_ADJ, gz, g, r, sv, sd, tv, td, pv, pd, kv, kd, gv, gd); });
        else
            b.scan(mode, gz, g, r, sv, sd, tv, td, pv, pd, kv, kd, gv, gd);
    }

    void halo_copy(int src, int dst) {
        const int nk = lin_levels(src, g.L);
        const double *sv = val[src];
        double *sd = drv[src], *dv = val[dst], *dd = drv[dst];
        if (mode == LIN_ADJ)
            rev.push_back([=]() { b.halo_copy(LIN_ADJ, g, nk, sv, sd, dv, dd); });
        else
            b.halo_copy(mode, g, nk, sv, sd, dv, dd);
    }

    // column operations of the remap (lin_remap.hpp), each over a box of
    // columns with its own workspace, independent of each other
    LinColOp col(Rng r, std::initializer_list<int> in, std::initializer_list<int> out, int wsi) const {
        return LinColOp{r, args(in, out, 0, 0), ws[wsi]};
    }
    void columns(int kind, std::initializer_list<LinColOp> ops) {
        std::array<LinColOp, LIN_MAXCOL> a{};
        int n = 0;
        for (const LinColOp &o : ops) a[n++] = o;
        if (mode == LIN_ADJ)
            rev.push_back([=]() { b.columns(LIN_ADJ, kind, a.data(), n); });
        else
            b.columns(mode, kind, a.data(), n);
    }

    void set_dt(double dt) {
        c.dt = dt;
        c.dtdx = dt * c.rdx;
        c.dtdy = dt * c.rdy;
    }

    // Forward (values) or forward + tangent (LIN_TAN) through the program.
    void sweep(LinMode m, double dt, bool remap) {
        step_dt = dt;
        mode = m;
        lin_program(*this, remap);
        mode = LIN_FWD;
    }

    // The adjoint: the forward values of the step first, then the
    // operations' adjoints in reverse order.  On entry the adjoints of the
    // end arrays hold the end state's adjoint (interior), every other
    // adjoint array is zero where it will be read; on return the adjoint
    // of X_* holds the start state's.
    void adjoint(double dt, bool remap) {
        sweep(LIN_FWD, dt, remap);
        step_dt = dt;
        mode = LIN_ADJ;
        rev.clear();
        lin_program(*this, remap);
        mode = LIN_FWD;
        for (auto it = rev.rbegin(); it != rev.rend(); ++it) (*it)();
        rev.clear();
    }
};
