/*
 * moloch_hip.h -- C ABI of the MI355X MOLOCH dynamical core.
 *
 * The reference exposes the timestep as the Fortran module mod_moloch with
 * three public procedures and no arguments beyond (istep, lprint); all data
 * flows through module state (nompi/mod_moloch.F90:93):
 *
 *   allocate_moloch()          nompi/mod_moloch.F90:183-217
 *   init_moloch()              nompi/mod_moloch.F90:219-268
 *   moloch(istep, lprint)      nompi/mod_moloch.F90:272-421
 *
 * This header is the C side of the drop-in replacement.  The Fortran module
 * dynamical_model_amd/fortran/mod_moloch.F90 keeps that API and calls these
 * functions through ISO_C_BINDING; Python (ctypes) calls them directly.
 *
 * Conventions
 *  - Plain C types only.  Every function returns 0 on success and a non-zero
 *    MLC_E* code on failure; mlc_last_error() then describes it.  No C++
 *    exception crosses this boundary.
 *  - Host arrays are Fortran-ordered (j fastest, then i, then k) and are
 *    described by their lower/upper bounds, exactly as the reference
 *    allocates them (mod_atmosphere.F90:104-162): e.g. u is
 *    (-1:jx+2, 0:iy+1, 1:kz).  A transfer copies the intersection of the
 *    host array's bounds with the device field's halo-2 extent
 *    (-1:jx+2, -1:iy+2, 1:nlev).  The library never keeps host pointers.
 *  - One HIP stream per handle.  mlc_step may return before the device has
 *    finished; mlc_get_field and mlc_sync wait.
 */
#ifndef MOLOCH_HIP_H
#define MOLOCH_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ error codes */
#define MLC_OK 0
#define MLC_EINVAL 1   /* bad argument (sizes, field id, bounds)            */
#define MLC_EHIP 2     /* a HIP runtime call failed                          */
#define MLC_ESTATE 3   /* call out of order (e.g. mlc_step before mlc_init)  */
#define MLC_ENOMEM 4   /* device allocation failed                           */

/* --------------------------------------------------------------- precision */
#define MLC_FP64 0       /* every field and operation in FP64                */
#define MLC_FP32 1       /* FP32 state and stencils; the Exner function, the  */
                         /* hydrostatic residuals and the tridiagonal w solve */
                         /* stay FP64                                          */

/* ------------------------------------------------------------- field ids */
enum mlc_field {
  /* prognostic / diagnostic state: mo_atm, sfs (mod_regcm_types.F90:80-104) */
  MLC_U = 0,    /* u      (jde1gb:jde2gb, ice1ga:ice2ga, 1:kz)   */
  MLC_V,        /* v      (jce1ga:jce2ga, ide1gb:ide2gb, 1:kz)   */
  MLC_W,        /* w      (jce1:jce2, ice1:ice2, 1:kz+1)         */
  MLC_PAI,      /* pai    (ga, 1:kz)  Exner function             */
  MLC_T,        /* t      (ga, 1:kz)                             */
  MLC_TETAV,    /* tetav  (ga, 1:kz)  virtual potential temp.    */
  MLC_QV,       /* qx(:,:,:,iqv)  (ga, 1:kz)                     */
  MLC_QC,       /* qx(:,:,:,iqc)                                 */
  MLC_QI,       /* qx(:,:,:,iqi)                                 */
  MLC_QR,       /* qx(:,:,:,iqr)                                 */
  MLC_QS,       /* qx(:,:,:,iqs)                                 */
  MLC_TVIRT,    /* tvirt  (jce, ice, 1:kz)                       */
  MLC_P,        /* p      (jce, ice, 1:kz)                       */
  MLC_RHO,      /* rho    (jce, ice, 1:kz)                       */
  MLC_QSAT,     /* qs     (jce, ice, 1:kz)  saturation mixing r. */
  MLC_UX,       /* ux     (jce1gb:jce2gb, ice1ga:ice2ga, 1:kz)   */
  MLC_VX,       /* vx     (jce1ga:jce2ga, ice1gb:ice2gb, 1:kz)   */
  MLC_PS,       /* sfs%ps (ga) 2-D                               */
  /* time-invariant geometry: mo_atm, mddom (mod_regcm_types.F90:44-79) */
  MLC_FMZ,      /* fmz    (ga, 1:kz)                             */
  MLC_FMZF,     /* fmzf   (jce, ice, 1:kz+1)                     */
  MLC_ZETA,     /* zeta   (gb, 1:kz)                             */
  MLC_HX,       /* mddom%hx   (jde1ga:jde2ga, ice1:ice2)  2-D     */
  MLC_HY,       /* mddom%hy   (jce1:jce2, ide1ga:ide2ga)  2-D     */
  MLC_MSFU,     /* mddom%msfu (jde1ga:jde2ga, ide1:ide2)  2-D     */
  MLC_MSFV,     /* mddom%msfv (jde1:jde2, ide1ga:ide2ga)  2-D     */
  MLC_MSFX,     /* mddom%msfx (jde1:jde2, ide1:ide2)      2-D     */
  MLC_ULAT,     /* mddom%ulat 2-D (degrees)                      */
  MLC_VLAT,     /* mddom%vlat 2-D (degrees)                      */
  MLC_NFIELDS
};

/* 1-D level arrays of mod_runparams.F90:60-62 */
enum mlc_levels {
  MLC_ZITA = 0, /* zita(1:kz+1)  */
  MLC_ZITAH,    /* zitah(1:kz)   */
  MLC_FFILT,    /* ffilt(1:kz)   */
  MLC_NLEVELS
};

typedef struct mlc_dims {
  int jx, iy, kz;   /* nompi/mod_dynparam.F90:30-35                        */
  int nqx;          /* number of water species, must be 5 (qv..qs)        */
} mlc_dims;

typedef struct mlc_params {
  /* run parameters (mod_runparams.F90:54-65, set in
   * nompi/regcm_dynamical_core.F90:99-113) */
  double dt;        /* model step (s), 30                                 */
  double dx;        /* grid spacing (m), 3000                             */
  double dzita;     /* mo_dzita = zita(kz)                                */
  double ztop;      /* mo_ztop                                            */
  double zh;        /* mo_h, decay height of the zita coordinate          */
  double a0;        /* mo_a0                                              */
  double anu2;      /* mo_anu2, divergence-filter coefficient             */
  int nadv;         /* mo_nadv, advection cycles per step (3)             */
  int nsound;       /* mo_nsound, acoustic substeps per cycle (5)         */
  int iqfrst;       /* first advected condensate species, 1-based (2)     */
  /* switches */
  int rotllr;       /* iproj == 'ROTLLR' (nompi/mod_dynparam.F90:39)      */
  int divfilter;    /* mo_divfilter (mod_runparams.F90:55)                */
  int halo_mpi;     /* 0: halo order of the nompi tree (default);         */
                    /* 1: order of the MPI tree (pai/deltaw after update) */
  int fulleq, divdamp, vadvtwice, filterpai, filtertheta;
                    /* module parameters do_* (nompi/mod_moloch.F90:100-104) */
  int precision;    /* MLC_FP64 or MLC_FP32                               */
  int device;       /* HIP device ordinal                                 */
  int unfused;      /* 1: one kernel per phase as in the reference's loop */
                    /* nests and the one-lane column solve (for           */
                    /* comparisons); 0: fused and segmented kernels       */
                    /* (default); both give bit-identical results         */
} mlc_params;

/* lprint diagnostics of nompi/mod_moloch.F90:395-419 */
typedef struct mlc_diag {
  double ps_max, ps_min, u_max, u_min, v_max, v_min;
  int cfl_violation;  /* ps holds a NaN, or its max/min is Inf: the       */
                      /* reference tests its maxval/minval (:402-403),    */
                      /* whose NaN handling is compiler-dependent; the    */
                      /* extrema here skip NaN values                     */
} mlc_diag;

typedef struct mlc_handle mlc_handle;

/* Fill *p with the reference defaults: dt=30, dx=3000, nadv=3, nsound=5,
 * anu2=0.6, a0=0, ztop=36000, zh=12000, iqfrst=2, rotllr=1, divfilter=0,
 * fulleq=divdamp=vadvtwice=1, filterpai=filtertheta=0, FP64, nompi halo
 * order (nompi/regcm_dynamical_core.F90:99-113, nompi/mod_moloch.F90:100-104).
 * dzita must still be set (= ztop/kz for the uniform zita levels). */
int mlc_default_params(mlc_params *p);

/* allocate_moloch (nompi/mod_moloch.F90:183-217) and the device mirrors of
 * the module state it works on (mod_atmosphere.F90:104-162). */
int mlc_create(const mlc_dims *dims, const mlc_params *params, mlc_handle **out);

/* Copy-in / copy-out of one field between a Fortran-ordered host array with
 * bounds lb[0..r-1]..ub[0..r-1] (r = 2 for 2-D fields, 3 otherwise) and the
 * device mirror.  mlc_get_field synchronises the handle's stream. */
int mlc_set_field(mlc_handle *h, int field, const double *host, const int *lb, const int *ub);
int mlc_get_field(mlc_handle *h, int field, double *host, const int *lb, const int *ub);

/* zita / zitah / ffilt (mod_runparams.F90:60-62, allocate_mod_runparams :71-81) */
int mlc_set_levels(mlc_handle *h, int which, const double *values, int n);

/* init_moloch (nompi/mod_moloch.F90:219-268): derives coru, corv, mx2, rmu,
 * rmv, gzitak, gzitakh, xknu and sets w(:,:,1) = 0, from the fields and
 * levels set so far.  Must be called after the geometry is in place and
 * before the first mlc_step. */
int mlc_init(mlc_handle *h);

/* moloch(istep, lprint) (nompi/mod_moloch.F90:272-421): one dt step on the
 * device.  When lprint != 0 the extrema and the CFL (NaN/Inf) check are
 * computed on the device and written to *diag (which may then not be NULL);
 * the call then synchronises.  A CFL violation is reported in
 * diag->cfl_violation and, as in the nompi tree, does not stop the model
 * (nompi/mod_dynparam.F90:121-138). */
int mlc_step(mlc_handle *h, int istep, int lprint, mlc_diag *diag);

/* Wait for all work queued on the handle's stream. */
int mlc_sync(mlc_handle *h);

/* Debug mode (no counterpart in the reference, which has no halo checks):
 * with enable != 0, every mlc_step first writes NaN into the two-point halo
 * ring of every state and scratch field (the geometry excepted), so that a
 * kernel reading a halo cell the step has not refilled puts a NaN into the
 * interior.  The step's results are otherwise unchanged. */
int mlc_debug_poison_halos(mlc_handle *h, int enable);

/* Per-kernel timing for benchmarks (no counterpart in the reference, whose
 * only timer is whole-run cpu_time, nompi/regcm_dynamical_core.F90:30,222-228).
 * mlc_profile(h, 1) brackets every subsequent kernel launch of the handle
 * with HIP events on its stream and clears the accumulators; mlc_profile(h, 0)
 * stops.  mlc_profile_get returns, for kernel kind 0 <= kind <
 * mlc_profile_kinds(), its name, the summed event time, the number of
 * launches and the summed algorithmic bytes (the fields the kernel must read
 * and write once each, kz levels x jx*iy columns). */
int mlc_profile(mlc_handle *h, int enable);
int mlc_profile_kinds(void);
int mlc_profile_get(mlc_handle *h, int kind, const char **name, double *total_ms, long long *launches,
                    double *bytes);

/* Release every device allocation of the handle. */
int mlc_destroy(mlc_handle *h);

/* Description of the most recent failure on the calling thread. */
const char *mlc_last_error(void);

/* ------------------------------------------------------------------------
 * Synthetic cases: geometry and initial state built on the host.
 *
 * The reference builds its geometry in `param`
 * (nompi/regcm_dynamical_core.F90:93-194) and its initial state in
 * fill_atmosphere (mod_atmosphere.F90:171-235).  MLC_CASE_REFERENCE
 * reproduces that state; MLC_CASE_STANDARD builds the benchmark
 * configurations (standard-atmosphere base state, Gaussian terrain, jet,
 * noise, warm bubble) with the same discrete hydrostatic Exner recursion
 * (mod_atmosphere.F90:210-234), so a resting flat atmosphere is an exact
 * steady state of the step.
 * ---------------------------------------------------------------------- */
#define MLC_CASE_STANDARD 0
#define MLC_CASE_REFERENCE 1

typedef struct mlc_case {
  int kind;                 /* MLC_CASE_STANDARD or MLC_CASE_REFERENCE        */
  int jx, iy, kz;
  double ztop, zh, a0;      /* zita coordinate (mod_zita.F90)                 */
  /* terrain: hills h*exp(-ln2*(r/hw)^2), r in cells.  nhills == 1: one hill
   * of hill_h metres and half-width hill_hw cells at the domain centre;
   * nhills > 1: seeded random hills (periodic distance) with heights in
   * [hill_hmin, hill_hmax] and half-widths in [hill_hwmin, hill_hwmax]. */
  int nhills;
  double hill_h, hill_hw;
  double hill_hmin, hill_hmax, hill_hwmin, hill_hwmax;
  /* MLC_CASE_STANDARD base state: T = max(tsfc - lapse*z, ttrop), sea-level
   * pressure psfc, relative humidity rh */
  double tsfc, lapse, ttrop, psfc, rh;
  /* wind: u = u0 + jet_amp*exp(-((z-jet_z)/jet_width)^2), v = v0, w = 0 */
  double u0, v0, jet_amp, jet_z, jet_width;
  /* perturbations of potential temperature (K): uniform noise in
   * [-noise_amp, noise_amp] on the lowest noise_levels levels, and a warm
   * bubble bubble_amp*cos^2(pi*r/2), r the ellipsoidal distance with
   * horizontal radius bubble_rcells cells and vertical radius bubble_rz (m)
   * around height bubble_z (m above ground) at the domain centre. */
  double noise_amp;
  int noise_levels;
  double bubble_amp, bubble_rcells, bubble_z, bubble_rz;
  int zero_coriolis;        /* 1: ulat = vlat = 0, i.e. f = 0                 */
  unsigned long long seed;  /* splitmix64 seed of terrain and noise           */
} mlc_case;

/* Defaults of the benchmark configurations 0..4: 0: 100 x 100 x 30, one
 * 500 m hill; 1: 512 x 512 x 40, one hill, noise and a warm bubble; 2, 3, 4:
 * 1536 x 1536 x 60, 512 x 512 x 40 and 2048 x 2048 x 60 with a 20 km top, 20
 * random hills, a jet and 60 % relative humidity. */
int mlc_case_config(int config, mlc_case *c);

/* Copy the case's ztop, zh, a0 and dzita = ztop/kz into *p. */
int mlc_case_params(const mlc_case *c, mlc_params *p);

/* Build the case on the host.  fields[f] (may be NULL) receives field f as a
 * Fortran array (-1:jx+2, -1:iy+2[, 1:nlev]) with nlev = kz+1 for MLC_W and
 * MLC_FMZF, kz for other 3-D fields, one plane for 2-D fields; halos are
 * filled periodically.  levels[l] (may be NULL) receives zita, zitah, ffilt. */
int mlc_case_build(const mlc_case *c, double *const *fields, double *const *levels);

/* Build the case, upload every field and level to the handle and call
 * mlc_init. */
int mlc_case_load(mlc_handle *h, const mlc_case *c);

/* The same case built on the device: the levels, the 2-D geometry and ps on
 * the host as in mlc_case_build, the 3-D fields by a kernel straight into
 * the handle's arrays, then mlc_init.  The 3-D fields agree with
 * mlc_case_build to the last few bits of the device exp/pow (FP64). */
int mlc_case_load_device(mlc_handle *h, const mlc_case *c);

#ifdef __cplusplus
}
#endif
#endif
