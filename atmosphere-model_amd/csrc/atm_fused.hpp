// atm_fused.hpp -- declarations shared by the fused dynamics path and the API.
#pragma once

#include "atm_common.hpp"

// output tile of one workgroup (one cell per thread)
#define ATM_TILE_X 32
#define ATM_TILE_Y 8

struct FusedBufs {
    double *dp, *pt, *ud, *vd;
};

bool atm_fused_supported(const Geo &g);
int atm_fused_cgrid(atm_ctx *c, FusedBufs &b, double dt);
int atm_fused_cgrid_halo(atm_ctx *c);
int atm_fused_dgrid(atm_ctx *c, FusedBufs &cur, FusedBufs &nxt, double dt);
