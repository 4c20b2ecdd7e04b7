/*
 * salsa_amd.h -- C ABI of the MI355X SALSA coagulation + condensation step.
 *
 * This is the drop-in boundary for UCLALES-SALSA's per-grid-cell microphysics
 * call (run_SALSA, src/src_salsa/mo_salsa_driver.f90:112-435, restricted to the
 * coagulation and condensation processes of salsa(), mo_salsa.f90:92-107).
 * The Fortran LES binds these symbols through ISO_C_BINDING
 * (uclales-salsa_amd/fortran/salsa_amd_cabi.f90); plain pointers and sizes
 * only, no torch or HIP types in any signature.
 *
 * Array layout (identical to the LES FieldArray storage):
 *   3-D fields      f(k,i,j)          -> f[k + nzp*(i + nxp*j)]   (Fortran order)
 *   number fields   n(k,i,j,b)        -> n[cell + ncell3d*b]
 *   mass fields     m(k,i,j,(s-1)*nb+b) -> m[cell + ncell3d*(s*nb + b)]
 *                   (util.f90:955-961 getMassIndex)
 * with ncell3d = nzp*nxp*nyp.  Interior cells are k = hz..nzp-1-hz,
 * i,j = hxy..n-1-hxy (0-based); the LES uses hz = 1, hxy = 2
 * (mo_salsa_driver.f90:180-182).  Categories are ordered
 * [aerosol, cloud, precipitation, ice] (classSection phase 1..4).
 * Units follow the LES: #/kg and kg/kg; conversion to #/m3 and m3/m3 with
 * the air density field dn happens inside the kernels.
 */
#ifndef SALSA_AMD_H
#define SALSA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SALSA_MAXBIN 64
#define SALSA_MAXSPEC 9

/* ---- status codes (the reference STOPs or prints instead) ---- */
enum {
  SALSA_OK = 0,
  SALSA_ERR_ARG = -1,          /* invalid argument / unsupported configuration     */
  SALSA_ERR_HIP = -2,          /* a HIP runtime call failed                         */
  SALSA_ERR_NOCONV = -3,       /* equilibration exceeded 100 iterations in some cell
                                  (mo_salsa_properties.f90:238-243, :352-357)       */
  SALSA_ERR_NONFINITE = -4,    /* NaN/Inf in the result (check_arrays analogue,
                                  mo_salsa.f90:157-192)                              */
  SALSA_ERR_NOMEM = -5,
  SALSA_ERR_DISTR = -6         /* distr_update failed in some cell (SALSA_CELL_DISTR)   */
};

/* per-cell fault bits returned by salsa_get_cell_flags() */
enum {
  SALSA_CELL_NEG_NUMC = 1,     /* numc < -1e-6 clamped after coagulation (mo_salsa_dynamics.f90:191) */
  SALSA_CELL_NOCONV = 2,       /* equilibration did not converge in 100 iterations                 */
  SALSA_CELL_SUBSTEPS = 4,     /* > 100 condensation substeps (mo_salsa_dynamics.f90:1051)         */
  SALSA_CELL_NONFINITE = 8,    /* NaN/Inf produced                                                  */
  SALSA_CELL_GRAD_UNSTABLE = 16, /* salsa_gradient_device / salsa_adjoint_device (whose bound is on J
                                   alone): the linearised trajectory of the cell is
                                   unstable -- a relative input change is amplified more than 1e8
                                   times in J, the vapour or an active bin, the explicit water
                                   update of one step amplifies a perturbation more than 1e4 times
                                   past its +-10 % limiter, or a derivative left the range of its
                                   type; the derivative then describes no perturbation resolvable
                                   in FP64 */
  SALSA_CELL_DISTR = 32        /* distr_update did not settle in 100 sweeps or moved a negative
                                  fraction (the reference STOPs, mo_salsa_update.f90:137, 216, 311) */
};

/* Process-rate diagnostic fields written by a step when `rates` is given:
 * rates[f*nzp*nxp*nyp + cell], interior cells overwritten, the other cells
 * untouched (the Diag fields of mo_salsa_driver.f90:392-419 with the names
 * given there).  Level 4 writes the first SALSA_NRATE_L4 fields, level 5 all
 * SALSA_NRATE_L5.  Ice nucleation is not part of the step; its fields are
 * written as 0, and so is the activation field unless lsactiv is on. */
enum {
  SALSA_RATE_AUTOC = 0,        /* s_m_autoc      Autoconversion%volc(iwa)*rhowa/dn            */
  SALSA_RATE_AUTOC50,          /* s_m_autoc50                                                  */
  SALSA_RATE_AUTOC80,          /* s_m_autoc80                                                  */
  SALSA_RATE_ACCR,             /* s_m_accr       Accretion%volc(iwa)*rhowa/dn                 */
  SALSA_RATE_ACCR50,           /* s_m_accr50                                                   */
  SALSA_RATE_ACCR80,           /* s_m_accr80                                                   */
  SALSA_RATE_ACCOLL,           /* s_m_ACcoll_dry SUM(ACcoll%volc(1:ndry)*rholiq)/dn           */
  SALSA_RATE_APCOLL,           /* s_m_APcoll_dry                                               */
  SALSA_RATE_ACTIV,            /* s_n_activ      Activation%numc/dn (0 without lsactiv)        */
  SALSA_RATE_CONDA,            /* s_m_conda      Cond_a%volc(iwa)/dn                           */
  SALSA_RATE_CONDC,            /* s_m_condc                                                    */
  SALSA_RATE_CONDP,            /* s_m_condp                                                    */
  SALSA_NRATE_L4,              /* 12 fields at level 4                                         */
  SALSA_RATE_AICOLL = SALSA_NRATE_L4, /* s_m_AIcoll_dry                                        */
  SALSA_RATE_ICEHOM,           /* s_n_icehom     (0)                                           */
  SALSA_RATE_ICEDEP,           /* s_n_icedep     (0)                                           */
  SALSA_RATE_ICEIMM,           /* s_n_iceimm     (0)                                           */
  SALSA_RATE_CONDI,            /* s_m_condi      (Cond_i%volc(iwa)*rhoic+%volc(irim)*rhori)/dn */
  SALSA_NRATE_L5               /* 17 fields at level 5                                         */
};

/* precision modes */
enum {
  SALSA_FP64 = 0,              /* plain FP64 everywhere                                             */
  SALSA_FP32_MIXED = 1         /* FP32 kernel coefficients and particle properties; FP64 state,
                                  gain/loss sums, equilibration and vapour/mass closure             */
};

/* step flags */
enum {
  SALSA_STEP_TENDENCY = 0,     /* run_SALSA semantics: p-fields read-only, t-fields accumulated     */
  SALSA_STEP_INPLACE = 1,      /* box-model / device-resident: write the new state into the p-fields */
  SALSA_STEP_NO_COMPACT = 2,   /* do not reorder cells by cost class (debug / comparison)          */
  SALSA_STEP_INIT = 4          /* initialization call, run_SALSA(initialize=.TRUE.) from init.f90:
                                  101-107: equilibration(init_rh = 0.3, init = .TRUE.) of every
                                  cell's aerosol bins before the step (mo_salsa_driver.f90:282-283),
                                  starting from the cell's own wet diameters                        */
};

/*
 * Bin layout and species table: what set_aerobins / set_cloudbins /
 * set_icebins / Species() derive from NAMELIST &salsa
 * (mo_salsa_init.f90:148-450, classSpecies.f90:205-299).  0-based indices.
 */
typedef struct {
  int32_t cur, par;            /* t_parallelbin, mo_submctl.f90:13-17 */
} salsa_parbin;

typedef struct {
  /* configuration */
  int32_t nbin[2];             /* nbin(1:2)                          */
  double reglim[4];            /* reglim(1:4) [m]                    */
  int32_t nprc; double prc_dlo, prc_vratio;   /* bloPrc            */
  int32_t nice; double ice_dlo, ice_vratio;   /* bloIce            */
  int32_t nspec_dry;
  int32_t dry_species[7];      /* 0 SO4, 1 OC, 2 BC, 3 DU, 4 SS, 5 NO, 6 NH (namelist order) */
  /* derived indices (0-based, category-local) */
  int32_t in1a, fn1a, in2a, fn2a, in2b, fn2b, nbins;
  salsa_parbin ica, fca, icb, fcb;
  int32_t ncld;
  int32_t ntotal;              /* nbins + ncld + nprc + nice         */
  /* species in mass-array order: dry species, H2O, rime */
  int32_t nspec, iwa, irim;
  double rholiq[SALSA_MAXSPEC], rhoice[SALSA_MAXSPEC], mm[SALSA_MAXSPEC], diss[SALSA_MAXSPEC];
  /* per bin, master order [aero | cloud | precip | ice] */
  int32_t phase[SALSA_MAXBIN];
  double vlolim[SALSA_MAXBIN], vhilim[SALSA_MAXBIN], dmid[SALSA_MAXBIN];
  double nlim[SALSA_MAXBIN], dlim[SALSA_MAXBIN];
  double precpbins[SALSA_MAXBIN];
} salsa_bin_desc;

/* The &salsa switches and constants that the coagulation + condensation path
 * reads (mo_submctl.f90:21-142, mo_ice_shape.f90:25-30), plus driver options. */
typedef struct {
  int32_t level;               /* thermodynamical level 4 (warm) or 5 (ice)      */
  int32_t lscoag, lscnd;       /* master switch states                           */
  int32_t lsauto, lsauto_mode; /* lsauto%state and %mode (1: coagulation-based precip formation) */
  int32_t lscgaa, lscgcc, lscgpp, lscgii, lscgca, lscgpa, lscgia, lscgpc, lscgic, lscgip;
  int32_t lscndh2oae, lscndh2ocl, lscndh2oic;
  int32_t lsfreeRH;            /* lsfreeRH%state; else rv limited to rhlim*rs    */
  double rhlim;
  double eddy_dis, Eiagg_max, Eiagg_min;
  double ice_alpha, ice_beta, ice_gamma, ice_sigma;
  int32_t lcharge;             /* charge enhancement with chargeTime = chargeTimeMax */
  double charge_coll_enh;      /* chargeCollEnh: kernel scaled by (1 + this)       */
  int32_t precision;           /* SALSA_FP64 or SALSA_FP32_MIXED                  */
  int32_t hz, hxy;             /* halo widths of the 3-D arrays (LES: 1, 2)       */
  int32_t lscoag_mode;         /* lscoag%mode: 1 kernels every step; 2 kernels every
                                  cgintvl seconds, stored per cell in between
                                  (mo_salsa_driver.f90:469-472, :601-697)          */
  double cgintvl;              /* kernel update interval [s] for lscoag_mode 2    */
  int32_t lsdistupdate;        /* size-distribution update after condensation
                                  (distr_update, mo_salsa.f90:148-149,
                                  mo_salsa_update.f90:21-433); the reference
                                  default is on, salsa_params_default leaves it
                                  off (coagulation + condensation only)         */
  int32_t lsactiv;             /* cloud activation after condensation: 0 off, 1
                                  interstitial (lsactiv%mode 1, mo_salsa.f90:115-119,
                                  actInterst mo_salsa_cloud.f90:859-1195), 2 at cloud
                                  base (mode 2, ActCloudBase :161-375; the state is
                                  unchanged, the activated droplets go to the fields
                                  of salsa_set_cloud_base_fields)                 */
} salsa_params;

typedef struct salsa_ctx salsa_ctx;

/* Fill the bin/species descriptor (replaces set_aerobins/set_cloudbins/
 * set_icebins/Species(), mo_salsa_init.f90:148-450, 792-858). */
int salsa_bin_desc_init(salsa_bin_desc *desc, const int32_t nbin[2], const double reglim[4],
                        int32_t nprc, double prc_dlo, double prc_vratio,
                        int32_t nice, double ice_dlo, double ice_vratio,
                        int32_t nspec_dry, const int32_t dry_species[7]);

/* Reference defaults for the hot-path switches (mo_submctl.f90) with the
 * level applied as define_salsa does (mo_salsa_init.f90:615-635). */
int salsa_params_default(salsa_params *p, int32_t level);

/* Create a context for one (nzp,nxp,nyp) grid on the current HIP device:
 * device tables, work buffers, a stream and the tuned launch shapes.
 * Replaces initialize_salsa_driver + the module state of run_SALSA
 * (mo_salsa_driver.f90:481-561). */
int salsa_ctx_create(const salsa_bin_desc *desc, const salsa_params *params,
                     int32_t nzp, int32_t nxp, int32_t nyp, salsa_ctx **out);

void salsa_ctx_destroy(salsa_ctx *ctx);

/*
 * One SALSA coagulation + condensation step over all interior cells
 * (run_SALSA with tstep = dt, mo_salsa_driver.f90:112-435).
 *   press, temp, rsl, rsi, dn : Diag fields, (nzp,nxp,nyp)
 *   rv_p : water vapour mixing ratio (Prog rp, tlev 1); rv_t : its tendency (+=)
 *   n_p[c], m_p[c] : category c number / mass mixing ratios (read-only unless
 *                    SALSA_STEP_INPLACE); n_t[c], m_t[c] : tendencies (+=).
 *   Category 3 (ice) pointers are ignored at level 4 and may be NULL.
 *   rates : NULL, or the process-rate fields (SALSA_RATE_*), same memory
 *           kind as the other pointers of the call
 *   time, istp : model time [s] and step number of run_SALSA.  With
 *           lscoag_mode 2 they select the kernel-update steps exactly as
 *           set_SALSA_runtime does (MOD(time,cgintvl) < tstep .OR. istp <= 1,
 *           mo_salsa_driver.f90:469-472); the process switch states (%state)
 *           are taken from salsa_params as the caller resolved them
 * salsa_step_host takes host pointers (copies in and out on the context
 * stream); salsa_step_device takes device pointers and only enqueues work on
 * the context stream (salsa_ctx_stream): work that produces its inputs must
 * be complete or ordered before that stream, and results are ready after
 * salsa_ctx_sync (or an event recorded on that stream).
 */
int salsa_step_host(salsa_ctx *ctx, const double *press, const double *temp,
                    const double *rsl, const double *rsi, const double *dn,
                    const double *rv_p, double *rv_t,
                    const double *const n_p[4], double *const n_t[4],
                    const double *const m_p[4], double *const m_t[4],
                    double *rates, double tstep, double time, int32_t istp, int32_t flags);

int salsa_step_device(salsa_ctx *ctx, const double *press, const double *temp,
                      const double *rsl, const double *rsi, const double *dn,
                      const double *rv_p, double *rv_t,
                      const double *const n_p[4], double *const n_t[4],
                      const double *const m_p[4], double *const m_t[4],
                      double *rates, double tstep, double time, int32_t istp, int32_t flags);

/*
 * Cloud-base activation (params.lsactiv = 2): the LES fields that run_SALSA
 * reads and writes for it (mo_salsa_driver.f90:360-366; cloud_activation's w
 * and pactd, mo_salsa.f90:115-119), device arrays of the caller used by every
 * following step of this context, salsa_step_host included:
 *   w     : vertical velocity [n3] (m/s); cells with w <= 0 activate nothing
 *   nactd : pa_nactd [ncld][n3], activated droplets per cloud bin (#/kg)
 *   vactd : pa_vactd [nwet*ncld][n3] (row s*ncld + b), their mass mixing
 *           ratios per wet species (kg/kg)
 * Every interior cell of a step overwrites its nactd / vactd entries.  All
 * three non-null, or all null (then a step with lsactiv = 2 is refused).
 */
int salsa_set_cloud_base_fields(salsa_ctx *ctx, const double *w, double *nactd, double *vactd);

/* Wait for the context stream; returns SALSA_ERR_NONFINITE / SALSA_ERR_NOCONV /
 * SALSA_ERR_DISTR if any cell of the last step raised the corresponding fault
 * (SALSA_OK on a context that has not stepped yet: its flags start at zero). */
int salsa_ctx_sync(salsa_ctx *ctx);

/* Copy the per-cell fault bits of the last step (one byte per interior cell,
 * interior order k fastest) to host memory. */
int salsa_get_cell_flags(salsa_ctx *ctx, uint8_t *flags_host, int64_t n);

/*
 * Gradient of a multi-step trajectory (no counterpart in the reference: the
 * sensitivity of precipitation formation to the initial size distribution
 * and to the collision-kernel enhancement).  Starting from the given fields,
 * the step of salsa_step_device is applied nsteps times in place (T, p, rs,
 * rsi and dn fixed; N, masses and rv evolving; lscoag_mode 1) to a copy of
 * the state, and
 *     J = sum over interior cells and precipitation bins of the bin volume
 *         concentration, all species [m^3 m^-3], after the last step
 * is differentiated in forward mode:
 *   djdn[c]   : dJ/dN0 for every bin of category c, in the layout of n_p[c]
 *               (device; interior cells written, halo untouched); N0 is the
 *               initial number mixing ratio [kg^-1], masses held fixed
 *   j_cell    : per interior cell J (device, ncells; may be NULL)
 *   djde_cell : per interior cell dJ/dE (device, ncells; may be NULL), E the
 *               collision-kernel enhancement factor K = E * coagc, i.e.
 *               1 + chargeCollEnh of mo_salsa_coagulation_kernels.f90:411-416
 *               with chargeTime = chargeTimeMax, at the context's value
 *   totals    : host double[2]: J summed over cells, and dJ/dE summed over
 *               the cells without SALSA_CELL_GRAD_UNSTABLE, in a fixed order
 * Per-cell fault bits (salsa_get_cell_flags) include SALSA_CELL_GRAD_UNSTABLE;
 * that bit is not an error of the call.
 * Inputs are device pointers and are not modified.  The call waits for the
 * context stream and returns its fault status like salsa_ctx_sync.
 */
int salsa_gradient_device(salsa_ctx *ctx, const double *press, const double *temp,
                          const double *rsl, const double *rsi, const double *dn,
                          const double *rv, const double *const n_p[4],
                          const double *const m_p[4], double tstep, int32_t nsteps,
                          double *const djdn[4], double *j_cell, double *djde_cell,
                          double totals[2]);

/*
 * salsa_gradient_device without any device memory allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory during the call (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Arguments and results as
 * salsa_gradient_device, which allocates its per-cell work buffer on first
 * use; this variant keeps the per-cell J and dJ/dE and the two totals in a
 * static device array of the library instead (16 MiB on every GPU the
 * library runs on, held until the process ends; its address is resolved at
 * salsa_ctx_create).  The context must have at most SALSA_NOALLOC_MAX_CELLS
 * interior cells (else SALSA_ERR_ARG) and its device must be current.  All
 * calls on one GPU share the array: calls that overlap on one GPU (other
 * threads or streams) race, and each call overwrites what the last left.
 */
/*
 * Reverse-mode (adjoint) gradient of the objective of salsa_gradient_device
 * (no counterpart in the reference).  Arguments and results as
 * salsa_gradient_device, plus
 *   checkpoint_every : the forward pass keeps the cell states every this many
 *                      steps (>= 1); the reverse pass recomputes the states
 *                      between two checkpoints from the earlier one, then runs
 *                      the steps' adjoints backwards.
 * One reverse sweep gives dJ/dN0 of every bin and dJ/dE at once.  Device
 * memory, kept by the context and grown on demand: ceil(nsteps /
 * checkpoint_every) + min(checkpoint_every, nsteps) + 2 cell states of
 * 8 (1 + nt (1 + nwet)) bytes per interior cell, and a tape of about 30 KiB
 * per cell slot (two per wavefront, at most 8192 slots).  Level 4 with lscoag_mode 1 and without autoconv2,
 * activation and the size-distribution update; SALSA_ERR_ARG otherwise, or
 * for a bin layout beyond the kernel's capacity (32 bins, 2 wet species).
 * SALSA_CELL_GRAD_UNSTABLE marks cells whose condensation substeps amplify a
 * water perturbation more than 1e4-fold in one step, whose J amplifies a
 * relative change of the state before some step, or of E, more than 1e8
 * times, or whose adjoint is not finite; dJ/dE totals exclude them.  Inputs are device pointers and are not
 * modified; the call waits for the context stream.
 */
int salsa_adjoint_device(salsa_ctx *ctx, const double *press, const double *temp,
                         const double *rsl, const double *rsi, const double *dn,
                         const double *rv, const double *const n_p[4],
                         const double *const m_p[4], double tstep, int32_t nsteps,
                         int32_t checkpoint_every, double *const djdn[4], double *j_cell,
                         double *djde_cell, double totals[2]);

#define SALSA_NOALLOC_MAX_CELLS (1 << 20)
int salsa_gradient_device_noalloc(salsa_ctx *ctx, const double *press, const double *temp,
                                  const double *rsl, const double *rsi, const double *dn,
                                  const double *rv, const double *const n_p[4],
                                  const double *const m_p[4], double tstep, int32_t nsteps,
                                  double *const djdn[4], double *j_cell, double *djde_cell,
                                  double totals[2]);

/* Launch-shape selection.  lanes_per_cell in {8,16,32,64} for each cost class
 * (0 = clear and dry, 1 = clear and humid, 2 = cloudy); 0 keeps the current
 * value.  salsa_autotune times the candidates on the given device-resident
 * inputs and keeps the fastest per class. */
/*
 * The LES side of a SALSA step for the SALSA tracers, on device: what cfll
 * runs after run_SALSA (step.f90:330-333) -- tend_constrain2
 * (constrain_SALSA.f90:18-36), update_sclrs (step.f90:450-489: a += a_t*dtlt
 * on the interior, sclrset 'mixd' util.f90:71-107, periodic lateral halos
 * of a process that owns the whole grid), tend0 (step.f90:379-396) and
 * SALSA_diagnostics(.FALSE., lcharge) (constrain_SALSA.f90:50-402).
 *   dtlt       : time step [s]
 *   dzt        : HOST array of nzp inverse level thicknesses (mo_aux_state
 *                dzt); the top boundary extrapolates with dzt(nzp-2)/dzt(nzp-1)
 *   temp, rh   : Diag temperature and relative humidity (a_temp, a_rh)
 *   rhi        : relative humidity over ice (a_rhi), level 5; may be NULL at 4
 *   rp         : water vapour mixing ratio (a_rp), for rtot
 *   n_p, m_p   : number / mass fields, updated in place; n_t, m_t : their
 *                tendencies, applied then zeroed (ice ignored at level 4)
 *   charge_time: NULL, or a_chargeTimep [nbins+ncld+nprc][n3]: zeroed where
 *                the reference zeroes it (lcharge)
 *   diag       : [6][n3] rc, srp, snrp, rtot, ri, riri (rows 4-5 at level 5)
 * Needs hz = 1, hxy = 2, nzp >= 4, nxp and nyp >= 6.  The gas-phase transfer
 * of tiny aerosols (lscndgas) and the a_indefp clamp (ice_theta_dist) run when
 * their fields are registered with salsa_set_les_extra_fields.  Device
 * pointers; work is enqueued on the context stream like salsa_step_device.
 */
int salsa_les_update_device(salsa_ctx *ctx, double dtlt, const double *dzt, const double *temp,
                            const double *rh, const double *rhi, const double *rp,
                            double *const n_p[4], double *const n_t[4], double *const m_p[4],
                            double *const m_t[4], double *charge_time, double *diag);

/* The two LES fields that constrain_SALSA touches besides the SALSA tracers,
 * for the following salsa_les_update_device calls (NULL unregisters):
 *   gaerop : a_gaerop [5][n3] (device), registered when lscndgas is on: the
 *            SO4, NO, NH and OC of aerosols removed below 0.1 nm go to gas
 *            rows 1, 2, 3 and 5 (1-based) as molecules, mass / molar mass *
 *            avog (SALSA_diagnostics, constrain_SALSA.f90:332-351)
 *   indefp : a_indefp [nindef][n3] (device), registered with ice_theta_dist:
 *            clamped to [0, 1] at level 5 (tend_constrain2, :33-34) */
int salsa_set_les_extra_fields(salsa_ctx *ctx, double *gaerop, double *indefp, int32_t nindef);

/*
 * Sedimentation on device (sedim_SALSA, mcrp.f90:745-921): DepositionSlow
 * (:926-1069) for aerosols (sed_aero) and cloud droplets (sed_cloud), with the
 * dry deposition velocity at the lowest level; DepositionFast (:1073-1275) for
 * precipitation (sed_precp) and, at level 5, for ice (sed_ice: flag 4, the
 * fall speed of non-spherical ice from getDiameter, getShapeCoefficients and
 * terminal_vel's ice branch, mo_ice_shape.f90, mo_particle_external_
 * properties.f90:91-158; water and rime in the rate, latent heat alvi).
 *   dzt        : HOST array of nzp inverse level thicknesses
 *   tk, th, adn: temperature, potential temperature, air density (device)
 *   ustar      : friction velocity, nxp*nyp (device)
 *   n_p, m_p   : fields (read; ice at index 3 at level 5); n_t, m_t, tlt :
 *                tendencies (added to)
 *   rrate      : rain-rate diagnostic [n3], W m-2 (overwritten)
 *   sfcrrate   : surface rain rate [nxp*nyp] (overwritten)
 *   vt_prc     : precipitation fall speed per bin [nprc][n3] (overwritten)
 *   irate, sfcirate, vt_ice : the ice rate [n3], surface ice rate [nxp*nyp]
 *                and ice fall speed [nice][n3] (level 5, overwritten; NULL
 *                at level 4)
 * Needs hz = 1, hxy = 2, 4 <= nzp <= 256, nice <= 16.  Work is enqueued on the
 * context stream.
 */
int salsa_sedim_device(salsa_ctx *ctx, double dtlt, const double *dzt, const double *tk, const double *th,
                       const double *adn, const double *ustar, const double *const n_p[4],
                       const double *const m_p[4], double *const n_t[4], double *const m_t[4], double *tlt,
                       double *rrate, double *sfcrrate, double *vt_prc, double *irate, double *sfcirate,
                       double *vt_ice, int32_t sed_aero, int32_t sed_cloud, int32_t sed_precp, int32_t sed_ice);

int salsa_set_launch_shape(salsa_ctx *ctx, int32_t cls, int32_t lanes_per_cell);
int salsa_get_launch_shape(salsa_ctx *ctx, int32_t cls);
int salsa_autotune(salsa_ctx *ctx, const double *press, const double *temp,
                   const double *rsl, const double *rsi, const double *dn,
                   const double *rv_p, const double *const n_p[4],
                   const double *const m_p[4], double tstep);

/* Kernel timing with HIP events on the context stream.  salsa_ctx_timing(ctx, 1)
 * resets and enables it: each step then records one event before its first and
 * one after its last step-kernel launch.  salsa_ctx_kernel_time waits for the
 * stream and returns the summed milliseconds and the number of steps timed.
 * (Instrumentation with no counterpart in the reference, which only prints
 * cpu_time per LES step, step.f90:134-141.) */
int salsa_ctx_timing(salsa_ctx *ctx, int32_t enable);
int salsa_ctx_kernel_time(salsa_ctx *ctx, double *ms, int64_t *nsteps);
/* The same events split by launch, summed over the timed steps: ms[0] gather
 * and cost classification, ms[1..3] the step kernel of cost class 0, 1, 2,
 * ms[4] scatter; cells[0..2] the cells each class kernel processed.  Each
 * figure spans only its own launches (the class-count read-back between the
 * classification and the class kernels is outside every span). */
int salsa_ctx_phase_time(salsa_ctx *ctx, double ms[5], int64_t cells[3], int64_t *nsteps);

/* Number of interior cells and of active bins in the context's layout. */
int64_t salsa_ctx_ncells(const salsa_ctx *ctx);
int32_t salsa_ctx_ntotal(const salsa_ctx *ctx);

/* The HIP stream the context enqueues on (as an opaque pointer). */
void *salsa_ctx_stream(salsa_ctx *ctx);

/* Last error message of the calling thread (static storage). */
const char *salsa_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
