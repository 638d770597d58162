// atm_fused.hpp -- declarations shared by the fused dynamics path and the API.
#pragma once

#include "atm_common.hpp"

// output tiles of one workgroup (one cell per thread): 32 x 8 for the dp/pt
// kernels, 32 x 16 for the wind and corner kernels.  Any nx, ny: the last
// tile of a row or column may reach past the domain; its cells there are
// computed from whatever the memory holds and never stored.
#define ATM_TILE_X 32
#define ATM_TILE_Y_PAIR 8
#define ATM_TILE_Y 16

// Doubles allocated past the end of every field buffer.  A tile reads up to
// 4 rows and columns around its 32 x TY cells (TY <= ATM_TILE_Y); a last
// tile that overhangs the domain reads at most ATM_TILE_Y - 4 rows past the
// halo'd plane (ATM_TILE_Y - 3 with the columns past the row stride, which
// run on into the next row), which on the last plane of a buffer lie in this
// tail.
inline size_t atm_tile_slack(const Geo &g) { return (size_t)(ATM_TILE_Y + 1) * (size_t)g.SI; }

struct FusedBufs {
    double *dp, *pt, *ud, *vd;
};

int atm_fused_cgrid(atm_ctx *c, FusedBufs &b, double dt);
int atm_fused_cgrid_halo(atm_ctx *c);
int atm_fused_dgrid(atm_ctx *c, FusedBufs &cur, FusedBufs &nxt, double dt);
int atm_fused_plk(atm_ctx *c);  // playkap of the current dp (the C step needs it valid)

// ---------------------------------------------------------------------------
// Batched ensemble steps (atm_ensemble_*): one launch per stage over every
// member of an ensemble.  The member table, in device memory, holds each
// member's grid, coefficients and arrays (both copies of the double-buffered
// prognostic fields); a workgroup finds its member and tile from its index
// in the launch (ens_find), then does exactly what the member's own launch
// would do for that tile.
// ---------------------------------------------------------------------------
struct EnsMember {
    Geo g;
    DynCoef coef;      // the dt-independent part (init_sw_dyn)
    double ak0, pktop;
    double *p[4][2];   // dp, pt, ud, vd: copy 0 and copy 1
    double *dpc, *ptc, *uc, *vc, *ua, *va, *uc2, *vc2;
    double *plk, *net_flux, *heatd, *plev, *pkap, *gz, *pkapb, *gzb;
    int *nfw;
    // the remap's (atm_remap_member; null before atm_remap_init)
    const double *ak, *bk;
    double *playkap, *rscr, *tplev, *tpkap, *tplk, *ur, *vr;
    // the radiation's (atm_rad_member; null before atm_radiation_init)
    const float *rtab[8];
    double rgrav;
    double *play, *lT, *vT;
    const double *ts, *radius;
    int *rflag;
};
#define ATM_ENS_MAX 64  // members per ensemble (one bit each in a parity mask)

// The workgroup spans of a batched launch: member m owns the (XCD-permuted)
// workgroups [first[m], first[m+1]); first[nm] is the members' total.
// Passed by value, so a workgroup finds its member by a binary search in
// the kernel arguments.
struct EnsIndex {
    unsigned first[ATM_ENS_MAX + 1];
};

// Workgroup b of a launch of n workgroups (n a multiple of 8) is first
// permuted as xcd_tile permutes a single grid's, so that consecutive tiles
// of a member run on one XCD; its member is the one whose span holds it.
// Returns false for the padding workgroups past the last member.  Uniform
// over the workgroup.
__device__ __forceinline__ bool ens_find(const EnsIndex &ix, int nm, int &m, unsigned &local) {
    const unsigned n = gridDim.x, b = blockIdx.x;
    const unsigned l = (b & 7u) * (n >> 3) + (b >> 3);
    if (l >= ix.first[nm]) return false;
    int lo = 0, hi = nm;  // first[lo] <= l < first[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (l >= ix.first[mid])
            lo = mid;
        else
            hi = mid;
    }
    m = __builtin_amdgcn_readfirstlane(lo);
    local = __builtin_amdgcn_readfirstlane(l - ix.first[lo]);
    return true;
}

// launchers of the batched stages (k_dyn_fused.hip, k_halo.hip); n_* are
// the launches' workgroup counts, rounded up to a multiple of 8
enum { ENS_HALO_C1 = 0, ENS_HALO_CH, ENS_HALO_D1, ENS_HALO_DH };
unsigned atm_ens_blocks_pair(const Geo &g);
unsigned atm_ens_blocks_wind(const Geo &g);
unsigned atm_ens_blocks_corner(const Geo &g);
unsigned atm_ens_blocks_scan(const Geo &g);
unsigned atm_ens_blocks_halo(const Geo &g, int kind);
void atm_ens_launch_pair(hipStream_t s, const EnsMember *tab, const EnsIndex &ix, int nm, unsigned long long par,
                         double dt, bool cgrid, bool mixed, unsigned n);
void atm_ens_launch_c2(hipStream_t s, const EnsMember *tab, const EnsIndex &ix, int nm, unsigned long long par,
                       double dt, bool mixed, unsigned n);
void atm_ens_launch_d2(hipStream_t s, const EnsMember *tab, const EnsIndex &ix, int nm, unsigned long long par,
                       double dt, bool mixed, unsigned n);
void atm_ens_launch_scan(hipStream_t s, const EnsMember *tab, const EnsIndex &ix, int nm, unsigned long long par,
                         bool dstep, bool mixed, unsigned n);
// The remap of every member (k_remap.hip): the member table's remap
// fields from a context, the workgroups a member needs in each of the four
// remap stages (grid, winds, energy inputs, energy), and the four launches.
bool atm_remap_member(const atm_ctx *c, EnsMember &m);
unsigned atm_ens_blocks_remap(const Geo &g, int stage);
void atm_ens_launch_remap(hipStream_t s, const EnsMember *tab, const EnsIndex *ix, int nm, unsigned long long par,
                          const unsigned *n);
// The radiation of every member (k_rad.hip): the member table's radiation
// fields, the workgroups of the three stages (profiles, the net flux's zero
// fill, columns) and the launches (lmax: the members' largest nlay).
bool atm_rad_member(const atm_ctx *c, EnsMember &m);
unsigned atm_ens_blocks_rad(const Geo &g, int stage);
void atm_ens_launch_rad(hipStream_t s, const EnsMember *tab, const EnsIndex *ix, int nm, unsigned long long par,
                        const unsigned *n, int lmax, bool mixed);
void atm_ens_launch_corners(hipStream_t s, const EnsMember *tab, const EnsIndex &ix, int nm, unsigned n);
void atm_ens_launch_halo(hipStream_t s, const EnsMember *tab, const EnsIndex &ix, int nm, unsigned long long par,
                         int kind, unsigned n);
bool atm_ens_scan_fits(const Geo &g);  // the batched scans need nlay <= 4 x their chunk
