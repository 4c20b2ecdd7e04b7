/*
 * salsa_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of UCLALES-SALSA's per-grid-cell coagulation +
 * water condensation step, used as the checker for the HIP implementation.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load or call it.  The product path never routes through this code.
 *
 * Every function in salsa_oracle.c cites the reference file:line it follows
 * (paths relative to the UCLALES-SALSA source tree).
 *
 * Indexing is 0-based throughout; the reference's 1-based ranges are
 * translated one-to-one (in1a = 0, ...).
 */
#ifndef SALSA_ORACLE_H
#define SALSA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define SO_MAXBIN 64   /* max bins over all categories */
#define SO_MAXSPEC 9   /* Section%volc(9), src/src_salsa/classSection.f90:22 */
#define SO_NRATE 20    /* classProcessRates.f90:41 nproc = 20 */

/* t_parallelbin, src/src_shared/mo_submctl.f90:13-17 */
typedef struct { int cur, par; } so_parbin;

/* Bin layout + species table (mo_submctl.f90:204-249, mo_salsa_init.f90:148-450,
 * classSpecies.f90:119-299).  Built by so_layout_init(). */
typedef struct {
  /* ---- configuration (NAMELIST &salsa) ---- */
  int nbin[2];          /* nbin(1:2)                        mo_submctl.f90:213 */
  double reglim[4];     /* reglim(1:4) [m]                  mo_submctl.f90:209 */
  int nprc;  double prc_dlo, prc_vratio;   /* bloPrc        mo_submctl.f90:232 */
  int nice;  double ice_dlo, ice_vratio;   /* bloIce        mo_submctl.f90:233 */
  int nspec_dry;        /* number of dry species            mo_submctl.f90:187 */
  int dry_species[7];   /* listspec as ids: 0 SO4,1 OC,2 BC,3 DU,4 SS,5 NO,6 NH */

  /* ---- derived indices (mo_salsa_init.f90:800-814, 255-284) ---- */
  int in1a, fn1a, in2a, fn2a, in2b, fn2b, nbins;
  so_parbin ica, fca, icb, fcb;
  int ncld;
  int iaero, icloud, iprecp, iice, ntotal;   /* offsets into the master array */

  /* ---- species (classSpecies.f90:205-299) ---- */
  int nspec;            /* total incl. H2O and rime          */
  int nwet, ndry;       /* nspec-1, nspec-2                  */
  int iwa, irim;        /* 0-based H2O and rime slots        */
  int spec_id[SO_MAXSPEC];      /* allNames index per slot (7 = H2O, 8 = rime) */
  double rholiq[SO_MAXSPEC], rhoice[SO_MAXSPEC], mm[SO_MAXSPEC], diss[SO_MAXSPEC];

  /* ---- per-bin constants (over the master array of ntotal bins) ---- */
  int phase[SO_MAXBIN];
  double vlolim[SO_MAXBIN], vhilim[SO_MAXBIN], dmid[SO_MAXBIN];
  double nlim[SO_MAXBIN], dlim[SO_MAXBIN];
  double precpbins[SO_MAXBIN];  /* precip low-limit diameters (mo_salsa_init.f90:347-350) */
} so_layout;

/* Switches and parameters read by the hot path (mo_submctl.f90:21-142,
 * mo_ice_shape.f90:25-30).  All logicals are 0/1. */
typedef struct {
  int lscoag, lscnd;                          /* master %state                 */
  int lscgaa, lscgcc, lscgpp, lscgii, lscgca, lscgpa, lscgia, lscgpc, lscgic, lscgip;
  int lscndh2oae, lscndh2ocl, lscndh2oic;
  int lsauto_state, lsauto_mode;
  double eddy_dis, Eiagg_max, Eiagg_min;
  double ice_alpha, ice_beta, ice_gamma, ice_sigma;
  int lcharge;                 /* charge enhancement on                         */
  double charge_coll_enh;      /* chargeCollEnh; with chargeTime = chargeTimeMax
                                  the kernel is scaled by (1 + charge_coll_enh) */
  /* Test hook with no counterpart in the reference: when kpert != 0 every
   * kernel value is multiplied by (1 + kpert*u), u in [-1,1) a hash of
   * (kseed, bin pair, T, p).  Used to measure how strongly the reference's
   * answer depends on rounding-sized kernel errors.  Default 0. */
  double kpert;
  unsigned long long kseed;
  int lsdistupdate;            /* size distribution update after condensation
                                  (mo_salsa.f90:148-149); default 0 here       */
  int level;                   /* thermodynamic level (4 or 5): distr_update
                                  skips the ice bins below 5 (:323)            */
  int lsactiv;                 /* cloud activation: 0 off, 1 interstitial
                                  (lsactiv%mode 1, mo_salsa.f90:115-119)       */
  /* Test hook with no counterpart in the reference: when set, the per-cell
   * status carries SO_DIAG_GROWTH3 / SO_DIAG_GROWTH5 when, within one step,
   * the explicit water update of the condensation substeps amplifies a
   * perturbation more than 1e3 / 1e5 times where its +-10 % limiter does
   * not bind (the product over the substeps of the largest per-bin
   * multiplier).  Default 0. */
  int diag_growth;
  int init;                    /* initialization call of run_SALSA (initialize =
                                  .TRUE.): equilibration(init_rh = 0.3, init =
                                  .TRUE.) before the step, mo_salsa_driver.f90:
                                  282-283; initial wet diameters from
                                  updateDiameter of the cell itself           */
} so_params;

/* One Rate (classProcessRates.f90:40-45): volc(9) and numc */
typedef struct { double volc[SO_MAXSPEC]; double numc; } so_rate;

/* Rate slots, same order as classProcessRates.f90:60-98 */
enum {
  SO_R_AUTOCONV = 0, SO_R_AUTOCONV50, SO_R_AUTOCONV80, SO_R_ACCR, SO_R_ACCR50,
  SO_R_ACCR80, SO_R_ACCOLL, SO_R_APCOLL, SO_R_AICOLL, SO_R_ACTIV, SO_R_ICEHOM,
  SO_R_ICEIMM, SO_R_ICEDEP, SO_R_CONDA, SO_R_CONDC, SO_R_CONDP, SO_R_CONDI,
  SO_R_DRFR, SO_R_RMSPL, SO_R_IIBR
};

/* Per-cell status bits (the reference prints or STOPs instead) */
enum {
  SO_OK = 0,
  SO_WARN_NEG_NUMC = 1,       /* numc < -1e-6 after coagulation (mo_salsa_dynamics.f90:191-193) */
  SO_ERR_EQUIL_NOCONV = 2,    /* equilibration > 100 iterations (mo_salsa_properties.f90:238-243) */
  SO_WARN_SUBSTEPS = 4,       /* > 100 condensation substeps (mo_salsa_dynamics.f90:1051) */
  SO_ERR_DISTR = 8,           /* distr_update: > 100 passes or a negative fraction (the
                                 reference STOPs, mo_salsa_update.f90:109, 137, 216, 311) */
  SO_DIAG_GROWTH3 = 16,       /* diag_growth: explicit water update grew > 1e3 in a step */
  SO_DIAG_GROWTH5 = 32        /* diag_growth: ... > 1e5 */
};

/* Fill the layout from nbin/reglim/bloPrc/bloIce/species; returns 0 on success */
int so_layout_init(so_layout *L, const int nbin[2], const double reglim[4],
                   int nprc, double prc_dlo, double prc_vratio,
                   int nice, double ice_dlo, double ice_vratio,
                   int nspec_dry, const int dry_species[]);

/* Reference defaults for the hot-path switches (mo_submctl.f90) with
 * `level` applied as define_salsa does (mo_salsa_init.f90:615-635). */
void so_params_default(so_params *P, int level);

/*
 * One SALSA coagulation + condensation step for one cell, SALSA units
 * (#/m3, m3/m3).  numc[ntotal], volc[ntotal*SO_MAXSPEC] (row = bin) are
 * updated in place; *rv (kg/kg) is updated in place.  rates (nullable)
 * receives SO_NRATE accumulators (reset here, like mo_salsa_driver.f90:293).
 * Returns a bitmask of SO_* status bits.
 */
int so_step_cell(const so_layout *L, const so_params *P,
                 double ptemp, double ppres, double prs, double prsi,
                 double *prv, double ptstep,
                 double *numc, double *volc, so_rate *rates);

/* The same with the kernel store of lscoag%mode = 2 (mo_salsa_driver.f90:
 * 601-697): kmode 1 computes the kernels and copies them to kstore
 * (ntotal*ntotal, row-major), kmode 2 uses kstore instead of computing them. */
int so_step_cell_ks(const so_layout *L, const so_params *P,
                    double ptemp, double ppres, double prs, double prsi,
                    double *prv, double ptstep,
                    double *numc, double *volc, so_rate *rates, double *kstore, int kmode);

/* n cells (numc [n*ntotal], volc [n*ntotal*SO_MAXSPEC]), nsteps each. */
int so_step_batch(const so_layout *L, const so_params *P, long n, const double *T, const double *p,
                  const double *rs, const double *rsi, double *rv, double ptstep, int nsteps,
                  double *numc, double *volc);

/* Individual pieces, exported so that tests can pin them separately. */
double so_coagc(const so_layout *L, const so_params *P,
                const double *numc, const double *volc,
                int b1, int b2, double temp, double pres);
double so_terminal_vel_liquid(double diam, double rhop, double rhoa,
                              double visc, double beta);
double so_acth2o(const so_layout *L, const double *volc_bin, const double *pcw);
double so_rslf(double p, double t);
double so_rsif(double p, double t);

/* Driver-level pack (mo_salsa_driver.f90:230-279) and tendency unpack
 * (:310-388) for one cell.  n_p[ntotal] (#/kg), m_p[ntotal*... ] in the LES
 * mass layout of each category: m(cat)[(nc)*nb_cat + b].  These operate on
 * one cell's vectors gathered by the caller. */
void so_pack_cell(const so_layout *L, int level, double dn,
                  const double *n_cat[4], const double *m_cat[4],
                  double *numc, double *volc);
void so_unpack_tend(const so_layout *L, int level, double dn, double tstep,
                    const double *numc, const double *volc,
                    const double *n_cat[4], const double *m_cat[4],
                    double *nt_cat[4], double *mt_cat[4]);

/* LES coupling after run_SALSA for one process owning the whole grid:
 * tend_constrain2, update_sclrs (update, sclrset 'mixd', periodic halos),
 * tend0 and SALSA_diagnostics (constrain_SALSA.f90:18-402, step.f90:330-333,
 * 379-489, util.f90:71-107, 1005-1024).  Updates n_p/m_p in place, zeroes
 * n_t/m_t, and writes diag[6][n3] = rc, srp, snrp, rtot, ri, riri (the last
 * two at level 5 only).  charge_time: NULL or [nbins+ncld+nprc][n3].
 * dzf2 = dzt(nzp-2)/dzt(nzp-1).  Returns 0, or -1 for a grid too small. */
int so_les_update(const so_layout *L, int level, int nzp, int nxp, int nyp, double dtlt, double dzf2,
                  const double *temp, const double *rh, const double *rhi, const double *rp,
                  double *const n_p[4], double *const n_t[4], double *const m_p[4], double *const m_t[4],
                  double *charge_time, double *diag);
int so_les_update_ext(const so_layout *L, int level, int nzp, int nxp, int nyp, double dtlt, double dzf2,
                      const double *temp, const double *rh, const double *rhi, const double *rp,
                      double *const n_p[4], double *const n_t[4], double *const m_p[4], double *const m_t[4],
                      double *charge_time, double *diag, double *gaerop, double *indefp, int nindef);
double so_les_corr(const double *x, const double *y, int n);
/* Sedimentation at level 4 (sedim_SALSA, mcrp.f90:745-921, DepositionSlow
 * :926-1069 for aerosols and cloud, DepositionFast :1073-1275 for
 * precipitation): adds to the tendencies n_t/m_t and tlt, writes rrate
 * [n3], sfcrrate [nxp*nyp] and vt_prc [nprc][n3] (zeroed first).  dzt, tk,
 * th, adn: LES fields; ustar [nxp*nyp].  Returns 0, or -1 for a grid too
 * small or nzp > 256. */
int so_sedim(const so_layout *L, int nzp, int nxp, int nyp, double dtlt, const double *dzt, const double *tk,
             const double *th, const double *adn, const double *ustar, double *const n_p[3], double *const m_p[3],
             double *const n_t[3], double *const m_t[3], double *tlt, double *rrate, double *sfcrrate, double *vt_prc,
             int sed_aero, int sed_cloud, int sed_precp);
double so_terminal_vel(double diam, double rhop, double rhoa, double visc, double beta);
int so_sedim5(const so_layout *L, const so_params *P, int nzp, int nxp, int nyp, double dtlt, const double *dzt,
              const double *tk, const double *th, const double *adn, const double *ustar, double *const n_p[4],
              double *const m_p[4], double *const n_t[4], double *const m_t[4], double *tlt, double *rrate,
              double *sfcrrate, double *irate, double *sfcirate, double *vt_prc, double *vt_ice, int sed_aero,
              int sed_cloud, int sed_precp, int sed_ice);
int so_step_cell_act(const so_layout *L, const so_params *P, double ptemp, double ppres, double prs, double prsi,
                     double *prv, double ptstep, double *numc, double *volc, so_rate *rates, double w, double *nact,
                     double *vact);
void so_act_cloud_base(const so_layout *L, double temp, double pres, double w, const double *numc, const double *volc,
                       double *nact, double *vact);
void so_sed_ice_vc(const so_layout *L, const so_params *P, double tk, double adn, double numc, const double *m,
                   double *out);

#ifdef __cplusplus
}
#endif
#endif
