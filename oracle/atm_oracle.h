/*
 * atm_oracle.h -- CPU oracle for the atmosphere hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, single-threaded, FP64
 * restatement of the reference Fortran model's per-timestep hot path
 * (felix-goddard/atmosphere-model: src/mod_sw_dyn.f90, src/mod_remap.f90,
 * src/mod_radiation.f90, src/mod_gas_optics.f90, src/mod_solar.f90,
 * src/mod_physics.f90, src/mod_output.f90 accumulation, src/mod_sync.f90 on a
 * single image).  It is the checker the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py may load it; the product path never does.
 *
 * Layout follows the reference: every 3-D field is Fortran column-major
 * (i fastest, then j, then k) over the halo'd bounds is..ie x js..je with
 * is = 1-7, ie = nx+7 (one image, halo width 7), k = 1..nlay or 1..nlay+1.
 * 2-D fields (ts, radius) are (is:ie, js:je).
 */
#ifndef ATM_ORACLE_H
#define ATM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* field ids for orc_field() */
enum {
    ORC_DP = 0, ORC_PT, ORC_UD, ORC_VD, ORC_TS, ORC_PLAY, ORC_PLAYKAP,
    ORC_PLEV, ORC_PKAP, ORC_GZ, ORC_NET_FLUX, ORC_AK, ORC_BK, ORC_RADIUS,
    ORC_DPC, ORC_PTC, ORC_UC, ORC_VC, ORC_NFIELDS
};

orc_t *orc_create(int nx, int ny, int nlay, double dx, double dy,
                  double gravity, double coriolis, double divdamp_coeff,
                  double vortdamp_coeff);
void orc_destroy(orc_t *o);
double *orc_field(orc_t *o, int id);
long orc_field_size(orc_t *o, int id);

/* module procedures (reference names in comments) */
void orc_init_sw_dyn(orc_t *o);                 /* init_sw_dyn */
void orc_cgrid_step(orc_t *o, double dt);       /* cgrid_dynamics_step */
void orc_cgrid_halo(orc_t *o);                  /* cgrid_halo_exchange */
void orc_dgrid_step(orc_t *o, double dt);       /* dgrid_dynamics_step */
void orc_dgrid_halo(orc_t *o);                  /* dgrid_halo_exchange */
void orc_allocate_remapping(orc_t *o);          /* allocate_remapping_arrays */
void orc_remap(orc_t *o);                       /* perform_vertical_remapping */
void orc_allocate_radiation(orc_t *o);          /* allocate_radiation_arrays */
int  orc_radiation(orc_t *o, double time);      /* calculate_radiative_heating; !=0: table range */
void orc_radiation_halo(orc_t *o);              /* radiation_halo_exchange */
void orc_halo(orc_t *o, double *a, int nk);     /* halo_exchange (nk==0: 2-D field) */
void orc_allocate_physics(orc_t *o);            /* allocate_physics_arrays */
void orc_physics_tendencies(orc_t *o);          /* calculate_physics_tendencies */
void orc_apply_physics(orc_t *o, double dt);    /* apply_physics_tendencies */
void orc_physics_halo(orc_t *o);                /* physics_halo_exchange */
void orc_output_init(orc_t *o);                 /* init_output (accumulators only) */
void orc_accumulate_output(orc_t *o, double dt); /* accumulate_output */
void orc_output_mean(orc_t *o, double *out);    /* write_output's means, (nx, ny, nlay+1, 6), then reset */
void orc_physics_column_tend(orc_t *o, int i, int j, double *u_tend, double *v_tend,
                             double *pt_tend); /* tendencies of column (i, j), 1-based i, j */

/* unit-level entry points for the tests */
double orc_sw_abs(int g, double p, double T, int *err);
double orc_sw_sca(int g);
double orc_lw_abs(int g, double p, double T, int *err);
double orc_planck(int g, double T, int *err);
double orc_solar_fraction(int g);
double orc_solar_irradiance(double radius);
void orc_constants(double *out5); /* pi, R_dry, cp_dry, cv_dry, kappa */
int orc_set_threads(int n);       /* threads of the parallel loops (n <= 0: query); returns the count */
void orc_flux_pencil(double *f, const double *q, const double *c, int n,
                     int variant);
void orc_vertical_ppm(int nlay, const double *var, const double *coord,
                      double *top, double *btm, double *curv,
                      double *level_raw);
void orc_remap_column(int nlay, const double *src_var, double *dst_var,
                      const double *src_coord, const double *dst_coord);
void orc_calculate_fluxes(int nlay, const double *R, const double *T,
                          const double *Sup, const double *Sdn, double toa_dn,
                          double sfc_up, double sfc_alb, double *flux_up,
                          double *flux_dn);
void orc_delta_eddington(double *tau, double *ssa, double *g);
void orc_interpolate_to_corners(double *out, const double *q, int ni, int nj);
int orc_sw_column(int nlay, int g, const double *play, const double *dm,
                  const double *T, double irradiance, double *flux_up,
                  double *flux_dn);
int orc_lw_column(int nlay, int g, const double *play, const double *dm,
                  const double *T, const double *Tlev, double ts,
                  double *flux_up, double *flux_dn);

/* general two-stream columns (synthetic correlated-k workload); the _ld
   and _q variants run the same statements in long double and binary128 */
void orc_ts_set_jitter(int on);
#define ORC_TS_DECL(name)                                                                                          \
    void name(int L, int kind, int g, int nsw, int nlw, double pref, double tref, const double *optics,            \
              const double *play, const double *plev, const double *tlay, const double *tlev, const double *q,     \
              const double *ctau, const double *cssa, const double *casy, double mu0, double alb, double tsfc,    \
              double irr, double gravity, double guard, double *flux_up, double *flux_dn);
#define ORC_TS_DECL_N(name)                                                                                        \
    void name(long ncol, int L, int nsw, int nlw, double pref, double tref, const double *optics,                  \
              const double *play, const double *plev, const double *tlay, const double *tlev, const double *q,     \
              const double *ctau, const double *cssa, const double *casy, const double *mu0, const double *alb,    \
              const double *tsfc, const double *irr, double gravity, double guard, double *net_flux,             \
              double *fluxes);
ORC_TS_DECL(orc_ts_column_g)
ORC_TS_DECL(orc_ts_column_g_ld)
ORC_TS_DECL_N(orc_twostream)
ORC_TS_DECL_N(orc_twostream_ld)
ORC_TS_DECL(orc_ts_column_g_q)
ORC_TS_DECL_N(orc_twostream_q)

#ifdef __cplusplus
}
#endif
#endif
