// atm_fused.hpp -- declarations shared by the fused dynamics path and the API.
#pragma once

#include "atm_common.hpp"

// output tiles of one workgroup (one cell per thread): 32 x 8 for the dp/pt
// kernels, 32 x 16 for the wind and corner kernels
#define ATM_TILE_X 32
#define ATM_TILE_Y_PAIR 8
#define ATM_TILE_Y 16

struct FusedBufs {
    double *dp, *pt, *ud, *vd;
};

bool atm_fused_supported(const Geo &g);
int atm_fused_cgrid(atm_ctx *c, FusedBufs &b, double dt);
int atm_fused_cgrid_halo(atm_ctx *c);
int atm_fused_dgrid(atm_ctx *c, FusedBufs &cur, FusedBufs &nxt, double dt);
int atm_fused_autotune(atm_ctx *c, int reps);
