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
int atm_fused_autotune(atm_ctx *c, int reps);
