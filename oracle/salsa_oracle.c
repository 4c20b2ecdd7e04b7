/*
 * salsa_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Plain-C FP64 restatement of the UCLALES-SALSA coagulation + condensation
 * step for one grid cell.  Paths cited as file:line are relative to the
 * reference source tree (src/...).  The reference runs all-FP64
 * (Makefile:52, -fdefault-real-8); literals below are the reference's decimal
 * literals read as doubles, exactly as -fdefault-real-8 reads them.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off).  Intentionally
 * unoptimised and sequential: each loop follows the reference's loop order.
 */
#include "salsa_oracle.h"

#include <math.h>
#include <string.h>

/* ---- physical constants, src/src_shared/mo_submctl.f90:258-338 ---------- */
static const double avog = 6.0221e+23, boltz = 1.3807e-23, grav = 9.81;
static const double pstand = 1.01325e+5, rg = 8.314, pi = 3.1415927;
static const double pi6 = 0.5235988, mair = 28.97e-3;
static const double rd = 287.04, alv = 2.5e6, als = 2.834e6;
static const double surfw0 = 0.073, surfi0 = 0.105, epsoc = 0.15;
static const double nlim_c = 1.0, prlim_c = 1.e-6;
static const double dlaero = 30.e-6, dlcloud = 100.e-6, dlprecp = 5.e-3, dlice = 10.e-3;
/* eps = epsilon(1.0) with default real kind 8 (mo_submctl.f90:277) */
static const double eps_r8 = 2.220446049250313080847e-16;
/* mvwa = mwa/avog/rhowa with mwa = 18.016e-3, rhowa = 1000. (mo_submctl.f90:310,323) */
#define MVWA ((18.016e-3 / avog) / 1000.)

/* ---- species master tables, src/src_shared/classSpecies.f90:97-110 ------ */
static const double allMM[9] = {98.08e-3, 150.e-3, 12.e-3, 100.e-3, 58.44e-3,
                                62.01e-3, 18.04e-3, 18.016e-3, 18.016e-3};
static const double allRho[9] = {1830., 1400., 2000., 2650., 2165.,
                                 1479., 1530., 1000., 400.};
static const double allDiss[9] = {3., 1., 0., 0., 2., 1., 1., 1., 1.};
/* named pointers, classSpecies.f90:266-294 */
static const double rhowa = 1000., rhoic = 917., rhori = 400.;
static const double mwa = 18.016e-3;
static const double cpa = 1005.;                    /* mo_submctl.f90:267 */
static const double rhosu = 1830., msu = 98.08e-3, rhooc = 1400., moc = 150.e-3;
static const double rhoss = 2165., mss = 58.44e-3;

static double cube(double x) { return x * x * x; }
static double cbrt_f(double x) { return pow(x, 1. / 3.); } /* Fortran x**(1./3.) */

/* ====================================================================== */
/* Layout: mo_salsa_init.f90:148-450 (bins), 792-858 (indices), classSpecies */
/* ====================================================================== */
int so_layout_init(so_layout *L, const int nbin[2], const double reglim[4],
                   int nprc, double prc_dlo, double prc_vratio,
                   int nice, double ice_dlo, double ice_vratio,
                   int nspec_dry, const int dry_species[]) {
  memset(L, 0, sizeof(*L));
  if (nbin[0] < 1 || nbin[1] < 4 || nprc < 1 || nice < 0 || nspec_dry < 1 || nspec_dry > 7)
    return -1;
  L->nbin[0] = nbin[0]; L->nbin[1] = nbin[1];
  for (int i = 0; i < 4; ++i) L->reglim[i] = reglim[i];
  L->nprc = nprc; L->prc_dlo = prc_dlo; L->prc_vratio = prc_vratio;
  L->nice = nice; L->ice_dlo = ice_dlo; L->ice_vratio = ice_vratio;
  L->nspec_dry = nspec_dry;
  for (int i = 0; i < nspec_dry; ++i) L->dry_species[i] = dry_species[i];

  /* salsa_initialize, mo_salsa_init.f90:800-810 */
  L->in1a = 0;
  L->in2a = L->in1a + nbin[0];
  L->in2b = L->in2a + nbin[1];
  L->fn1a = L->in2a - 1;
  L->fn2a = L->fn1a + nbin[1];
  L->fn2b = L->fn2a + nbin[1];
  L->nbins = L->fn2b + 1;

  /* set_cloudbins, mo_salsa_init.f90:270-284 */
  int nba = L->fn2a - L->in2a + 1, nbb = nba;
  L->ica.cur = 0;                 L->ica.par = L->in2a;
  L->fca.cur = L->ica.cur + nba - 1; L->fca.par = L->ica.par + nba - 1;
  L->icb.cur = L->fca.cur + 1;    L->icb.par = L->fn2b - nbb + 1;
  L->fcb.cur = L->icb.cur + nbb - 1; L->fcb.par = L->icb.par + nbb - 1;
  L->ncld = L->fcb.cur + 1;

  /* set_masterbins, mo_salsa_init.f90:48-100 */
  L->iaero = 0;
  L->icloud = L->nbins;
  L->iprecp = L->icloud + L->ncld;
  L->iice = L->iprecp + nprc;
  L->ntotal = L->iice + nice;
  if (L->ntotal > SO_MAXBIN) return -2;

  /* Species(nspec_dry, listspec), classSpecies.f90:205-299:
   * mass-array order = namelist order, then H2O, then rime. */
  L->nspec = nspec_dry + 2;
  L->nwet = L->nspec - 1;
  L->ndry = L->nspec - 2;
  L->iwa = nspec_dry;
  L->irim = nspec_dry + 1;
  for (int s = 0; s < nspec_dry; ++s) L->spec_id[s] = dry_species[s];
  L->spec_id[L->iwa] = 7;
  L->spec_id[L->irim] = 8;
  for (int s = 0; s < L->nspec; ++s) {
    int id = L->spec_id[s];
    L->mm[s] = allMM[id];
    L->diss[s] = allDiss[id];
    L->rholiq[s] = allRho[id];
    L->rhoice[s] = (id == 7) ? 917. : allRho[id];  /* classSpecies.f90:258 */
  }

  /* set_aerobins, mo_salsa_init.f90:148-241 (nbin2 = 4 hard-coded :158) */
  const int nbin2 = 4, nbin3 = nbin[1] - nbin2;
  double ratio = reglim[1] / reglim[0];
  for (int cc1 = L->in1a + 1; cc1 <= L->fn1a + 1; ++cc1) {   /* 1-based cc */
    int b = cc1 - 1;
    L->vlolim[b] = pi6 * cube(reglim[0] * pow(ratio, (double)(cc1 - 1) / nbin[0]));
    L->vhilim[b] = pi6 * cube(reglim[0] * pow(ratio, (double)(cc1) / nbin[0]));
    L->dmid[b] = cbrt_f((L->vhilim[b] + L->vlolim[b]) / (2. * pi6));
  }
  ratio = reglim[2] / reglim[1];
  for (int dd = L->in2a; dd <= L->in2a + nbin2 - 1; ++dd) {
    int cc = dd - L->in2a;
    L->vlolim[dd] = pi6 * cube(reglim[1] * pow(ratio, (double)cc / nbin2));
    L->vhilim[dd] = pi6 * cube(reglim[1] * pow(ratio, (double)(cc + 1) / nbin2));
    L->dmid[dd] = cbrt_f((L->vhilim[dd] + L->vlolim[dd]) / (2. * pi6));
  }
  ratio = reglim[3] / reglim[2];
  for (int dd = L->in2a + nbin2; dd <= L->fn2a; ++dd) {
    int cc = dd - (L->fn2a - (nbin3 - 1));
    L->vlolim[dd] = pi6 * cube(reglim[2] * pow(ratio, (double)cc / nbin3));
    L->vhilim[dd] = pi6 * cube(reglim[2] * pow(ratio, (double)(cc + 1) / nbin3));
    L->dmid[dd] = cbrt_f((L->vhilim[dd] + L->vlolim[dd]) / (2. * pi6));
  }
  for (int k = 0; k < nbin[1]; ++k) {   /* 2b copies 2a, :223-227 */
    L->vlolim[L->in2b + k] = L->vlolim[L->in2a + k];
    L->vhilim[L->in2b + k] = L->vhilim[L->in2a + k];
    L->dmid[L->in2b + k] = L->dmid[L->in2a + k];
  }
  for (int b = 0; b < L->nbins; ++b) {
    L->phase[b] = 1; L->nlim[b] = nlim_c; L->dlim[b] = dlaero;
  }
  /* cloud bins parallel to 2a/2b, :317-333 */
  for (int k = 0; k < nba; ++k) {
    int ca = L->icloud + L->ica.cur + k, pa = L->ica.par + k;
    int cb = L->icloud + L->icb.cur + k, pb = L->icb.par + k;
    L->vhilim[ca] = L->vhilim[pa]; L->vlolim[ca] = L->vlolim[pa]; L->dmid[ca] = L->dmid[pa];
    L->vhilim[cb] = L->vhilim[pb]; L->vlolim[cb] = L->vlolim[pb]; L->dmid[cb] = L->dmid[pb];
  }
  for (int c = 0; c < L->ncld; ++c) {
    int b = L->icloud + c;
    L->phase[b] = 2; L->nlim[b] = nlim_c; L->dlim[b] = dlcloud;
  }
  /* buildBinLimits, :429-450, and the precip/ice bins, :335-345, :405-415 */
  {
    double r3 = pow(prc_vratio, 1. / 3.);
    double dlo = prc_dlo, dhi = prc_dlo * r3;
    for (int k = 0; k < nprc; ++k) {
      if (k > 0) { dlo = dlo * r3; dhi = dlo * r3; }
      int b = L->iprecp + k;
      L->vhilim[b] = pi6 * cube(dhi);
      L->vlolim[b] = pi6 * cube(dlo);
      L->dmid[b] = cbrt_f((L->vlolim[b] + L->vhilim[b]) / (2. * pi6));
      L->phase[b] = 3; L->nlim[b] = prlim_c; L->dlim[b] = dlprecp;
      L->precpbins[k] = cbrt_f(L->vlolim[b] / pi6);
    }
  }
  if (nice > 0) {
    double r3 = pow(ice_vratio, 1. / 3.);
    double dlo = ice_dlo, dhi = ice_dlo * r3;
    for (int k = 0; k < nice; ++k) {
      if (k > 0) { dlo = dlo * r3; dhi = dlo * r3; }
      int b = L->iice + k;
      L->vhilim[b] = pi6 * cube(dhi);
      L->vlolim[b] = pi6 * cube(dlo);
      L->dmid[b] = cbrt_f((L->vlolim[b] + L->vhilim[b]) / (2. * pi6));
      L->phase[b] = 4; L->nlim[b] = prlim_c; L->dlim[b] = dlice;
    }
  }
  return 0;
}

/* mo_submctl.f90 defaults; define_salsa level-4 overrides (mo_salsa_init.f90:615-635) */
void so_params_default(so_params *P, int level) {
  memset(P, 0, sizeof(*P));
  P->lscoag = 1; P->lscnd = 1;
  P->lscgaa = P->lscgcc = P->lscgpp = P->lscgii = 1;
  P->lscgca = P->lscgpa = P->lscgia = 1;
  P->lscgpc = P->lscgic = P->lscgip = 1;
  P->lscndh2oae = P->lscndh2ocl = P->lscndh2oic = 1;
  P->lsauto_state = 1; P->lsauto_mode = 1;
  P->eddy_dis = 0.001; P->Eiagg_max = 0.3; P->Eiagg_min = 0.05;
  P->ice_alpha = 15.56999e-3; P->ice_beta = 2.02;   /* mo_ice_shape.f90:25-30 */
  P->ice_gamma = 0.55; P->ice_sigma = 1.97;
  P->lcharge = 0; P->charge_coll_enh = 0.;
  P->kpert = 0.; P->kseed = 0ull;
  P->lsdistupdate = 0; P->level = level; P->lsactiv = 0; P->diag_growth = 0;
  if (level == 4) {
    P->lscgia = P->lscgic = P->lscgii = P->lscgip = 0;
    P->lscndh2oic = 0;
  }
}

/* ====================================================================== */
/* Per-cell working state (the allSALSA Section array for one cell)       */
/* ====================================================================== */
typedef struct {
  double numc[SO_MAXBIN];
  double volc[SO_MAXBIN][SO_MAXSPEC];
  double dwet[SO_MAXBIN], ddry[SO_MAXBIN], dnsp[SO_MAXBIN], rhomean[SO_MAXBIN];
  so_rate *rates;
} so_cell;

static void rate_acc(so_cell *c, int r, const double *n, const double *v, int nv) {
  if (!c->rates) return;
  if (v) for (int s = 0; s < nv; ++s) c->rates[r].volc[s] += v[s];
  if (n) c->rates[r].numc += *n;
}

/* updateDiameter(limit=.TRUE., type="all"), classSection.f90:91-159.
 * dnsp is assigned MIN(dwet, dlim) by the limit branch (:148-150), which
 * discards the getDiameter() result computed at :118-123. */
static void update_diameter(const so_layout *L, so_cell *c, int b) {
  double numc = c->numc[b];
  double dwet, ddry;
  if (numc > L->nlim[b]) {
    double s = 0.;
    for (int k = 0; k < SO_MAXSPEC; ++k) s += c->volc[b][k];
    dwet = cbrt_f(s / numc / pi6);
    double sd = 0.;
    for (int k = 0; k < L->ndry; ++k) sd += c->volc[b][k];
    ddry = cbrt_f(sd / numc / pi6);
  } else {
    dwet = L->dmid[b];
    ddry = L->dmid[b];
  }
  c->dnsp[b] = fmin(dwet, L->dlim[b]);
  c->dwet[b] = fmin(dwet, L->dlim[b]);
  c->ddry[b] = fmin(ddry, L->dlim[b]);
}

/* updateRhomean, classSection.f90:169-195 */
static void update_rhomean(const so_layout *L, so_cell *c, int b) {
  double r = rhowa;
  if (L->phase[b] > 3) {
    if (c->numc[b] > L->nlim[b]) {
      double mp = rhoic * c->volc[b][L->iwa];
      double mr = rhori * c->volc[b][L->irim];
      double mt = mp + mr;
      r = (mp * rhoic + mr * rhori) / fmax(mt, 7.e-25);
      r = fmin(r, rhoic);
    } else {
      r = rhoic;
    }
  }
  c->rhomean[b] = r;
}

/* getRimeFraction, classSection.f90:202-224 */
static double rime_fraction(const so_layout *L, const so_cell *c, int b) {
  double f = 0.;
  if (L->phase[b] == 4 && c->volc[b][L->iwa] > 1.e-23 &&
      c->volc[b][L->irim] > 1.e-23 && c->numc[b] > L->nlim[b])
    f = c->volc[b][L->irim] * rhori /
        (c->volc[b][L->iwa] * rhoic + c->volc[b][L->irim] * rhori);
  return f;
}

/* ---- ice shape, src/src_shared/mo_ice_shape.f90 ---- */
typedef struct { double alpha, beta, gamma, sigma; } so_shape;

/* getShapeCoefficients, mo_ice_shape.f90:145-191 */
static so_shape shape_coeffs(const so_params *P, double mpri, double mrim, double numc) {
  so_shape s;
  s.alpha = pi6 * rhoic; s.beta = 3.; s.gamma = pi / 4.; s.sigma = 2.;
  if (numc < 1.e-6) return s;
  double Mtot = (mpri + mrim) / numc;
  double Fr = fmax(fmin(mrim / (Mtot * numc), 0.9999), 0.0001);
  double rho_b = Fr * rhori + (1. - Fr) * rhoic;           /* getBulkRho :113-116 */
  double Dth = pow(pi6 * rhoic / P->ice_alpha, 1. / (P->ice_beta - 3.));
  double Dgr = pow(P->ice_alpha / (pi6 * rho_b), 1. / (3. - P->ice_beta));
  double Dcr = pow((1. / (1. - Fr)) * P->ice_alpha / (pi6 * rho_b), 1. / (3. - P->ice_beta));
  double Mth = rhoic * pi6 * cube(Dth);
  double Mgr = rho_b * pi6 * cube(Dgr);
  double Mcr = rho_b * pi6 * cube(Dcr);
  if (Mtot >= Mth && Mtot < Mgr) {
    s.alpha = P->ice_alpha; s.beta = P->ice_beta;
    s.gamma = P->ice_gamma; s.sigma = P->ice_sigma;
  } else if (Mtot > Mcr) {
    double hlp = 1. / (1. - Fr);
    s.alpha = P->ice_alpha * hlp; s.beta = P->ice_beta;
    s.gamma = P->ice_gamma * (1. - Fr) + pi / 4. * Fr;
    s.sigma = P->ice_sigma * (1. - Fr) + 2. * Fr;
  }
  return s;
}

/* getDiameter, mo_ice_shape.f90:47-86, with getBulkRho, D_nonsphericalIce,
   D_nonsphericalRimed and D_spherical (:105-143): the maximum dimension of a
   non-spherical ice particle of pristine mass mpri and rime mass mrim */
static double ice_diameter(const so_params *P, double mpri, double mrim, double numc) {
  if (numc < 1.e-6) return 1.e-9;
  const double Mtot = (mpri + mrim) / numc;
  const double Fr = fmax(fmin(mrim / (Mtot * numc), 0.9999), 0.0001);
  const double rho_b = Fr * rhori + (1. - Fr) * rhoic;
  const double Dth = pow(pi6 * rhoic / P->ice_alpha, 1. / (P->ice_beta - 3.));
  const double Dgr = pow(P->ice_alpha / (pi6 * rho_b), 1. / (3. - P->ice_beta));
  const double Dcr = pow((1. / (1. - Fr)) * P->ice_alpha / (pi6 * rho_b), 1. / (3. - P->ice_beta));
  const double Mth = rhoic * pi6 * cube(Dth);
  const double Mgr = rho_b * pi6 * cube(Dgr);
  const double Mcr = rho_b * pi6 * cube(Dcr);
  if (Mtot < Mth) return pow(Mtot / (pi6 * rhoic), 1. / 3.);                  /* small spherical */
  if (Mtot >= Mth && Mtot < Mgr) return pow(Mtot / P->ice_alpha, 1. / P->ice_beta);   /* dense non-spherical */
  if (Mtot >= Mgr && Mtot < Mcr) return pow(Mtot / (pi6 * rho_b), 1. / 3.);   /* graupel */
  if (Mtot >= Mcr) {                                                         /* partially rimed */
    const double hlp = 1. / (1. - Fr);
    return pow(Mtot / (hlp * P->ice_alpha), 1. / P->ice_beta);
  }
  return 1.e-9;                                                              /* NaN masses */
}

/* terminal_vel, liquid branch (flag 1..3), mo_particle_external_properties.f90:61-90 */
double so_terminal_vel_liquid(double diam, double rhop, double rhoa, double visc, double beta) {
  double v;
  if (diam <= 19E-6) {
    v = (diam * diam) * (rhop - rhoa) * grav * beta / (18. * visc);
  } else if (diam > 19E-6 && diam < 1.07E-3) {
    double NDa = 4. / 3. * pow(diam, 3.) * grav * (rhop - rhoa) * rhoa / pow(visc, 2.);
    double X = log(NDa);
    double Y = -0.318657E1 + 0.992696 * X - 0.153193E-2 * X * X - 0.987059E-3 * X * X * X
               - 0.578878E-3 * (X * X) * (X * X) + 0.855176E-4 * pow(X, 5) - 0.327815E-5 * pow(X, 6);
    double Re2 = beta * exp(Y);
    v = visc / rhoa / diam * Re2;
  } else {
    double Np = pow(surfw0, 3.) * pow(rhoa, 2.) / (pow(visc, 4.) * (rhop - rhoa) * grav);
    double Bo = 4. / 3. * (rhop - rhoa) * grav / surfw0 * pow(diam, 2.);
    double X3 = log(Bo * pow(Np, 1. / 6.));
    double Y3 = -0.500015E1 + 0.523778E1 * X3 - 0.204914E1 * X3 * X3 + 0.475294 * X3 * X3 * X3
                - 0.542819E-1 * (X3 * X3) * (X3 * X3) + 0.238449E-2 * pow(X3, 5);
    double Re3 = pow(Np, 1. / 6.) * exp(Y3);
    v = fmin(visc / rhoa / diam * Re3, 9.2);
  }
  return v;
}

/* terminal_vel, ice branch (flag 4), :91-98 with kcVt/kcAv/kcBv/kcare/kcbre :103-158 */
static double terminal_vel_ice(double diam, double rhop, double rhoa, double visc,
                               so_shape sh, double dnsp) {
  const double c1 = 0.0902, delta0 = 9.06;
  double Vb = pi6 * cube(diam);
  double Ap = sh.gamma * pow(dnsp, sh.sigma);
  double X = (2. * Vb * (rhop - rhoa) * grav * dnsp * dnsp) / (Ap * rhoa * visc * visc);
  double k1213 = sqrt(1. + c1 * sqrt(X));
  double bre = 0.5 * c1 * sqrt(X);
  bre = bre / (k1213 - 1.);
  bre = bre / k1213;
  double are = 0.25 * delta0 * delta0;
  are = are * (k1213 - 1.) * (k1213 - 1.);
  are = are / pow(X, bre);
  double Av = are * pow(visc, 1. - 2. * bre);
  Av = Av * pow((2. * sh.alpha * grav) / (rhoa * sh.gamma), bre);
  double Bv = bre * (sh.beta - sh.sigma + 2.) - 1.;
  return Av * pow(dnsp, Bv);
}

/* uniform number in [-1,1) from (seed, pair, T, p): splitmix64 finaliser */
static double kpert_unit(unsigned long long seed, int b1, int b2, double temp, double pres) {
  unsigned long long x, bt, bp;
  memcpy(&bt, &temp, 8);
  memcpy(&bp, &pres, 8);
  x = seed ^ ((unsigned long long)b1 << 32) ^ (unsigned long long)(b2 * 2654435761u) ^ bt ^ (bp << 1);
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (double)(x >> 11) * (2. / 9007199254740992.) - 1.;
}

/* coagc(pp1,pp2,temp,pres,lcharge,kernel=2),
 * src/src_salsa/mo_salsa_coagulation_kernels.f90:185-420 */
static double coagc_cell(const so_layout *L, const so_params *P, const so_cell *c,
                         int b1, int b2, double temp, double pres) {
  const int ns = L->nspec;
  int ph1 = L->phase[b1], ph2 = L->phase[b2];
  double diam1 = (ph1 == 4) ? c->dnsp[b1] : c->dwet[b1];
  double diam2 = (ph2 == 4) ? c->dnsp[b2] : c->dwet[b2];
  double dsph1 = c->dwet[b1], dsph2 = c->dwet[b2];
  double mass1 = c->rhomean[b1] * pi6 * cube(dsph1);
  double mass2 = c->rhomean[b2] * pi6 * cube(dsph2);
  /* self-coagulation offset, :273-276 */
  if (fabs(L->dmid[b1] - L->dmid[b2]) / L->dmid[b1] < 1.e-2) {
    diam1 = 1.1 * diam2;
    mass1 = 1.33 * mass2;
  }
  double diam[2] = {diam1, diam2}, mpart[2] = {mass1, mass2};
  double visc = (7.44523e-3 * sqrt(cube(temp))) / (5093. * (temp + 110.4));
  double mfp = (1.656e-10 * temp + 1.828e-8) * pstand / pres;
  double knud[2], beta[2], dfpart[2], mtvel[2], omega[2], tva[2], termv[2];
  for (int k = 0; k < 2; ++k) {
    knud[k] = 2. * mfp / diam[k];
    beta[k] = 1. + knud[k] * (1.142 + 0.558 * exp(-0.999 / knud[k]));
    dfpart[k] = beta[k] * boltz * temp / (3. * pi * visc * diam[k]);
    mtvel[k] = sqrt((8. * boltz * temp) / (pi * mpart[k]));
    omega[k] = 8. * dfpart[k] / (pi * mtvel[k]);
  }
  double mdiam = 0.5 * (diam[0] + diam[1]);
  double flux1 = 4. * pi * mdiam * (dfpart[0] + dfpart[1]);
  double flux2 = pi * sqrt((mtvel[0] * mtvel[0]) + (mtvel[1] * mtvel[1])) * (mdiam * mdiam);
  for (int k = 0; k < 2; ++k) {
    double a = mdiam * mdiam + omega[k] * omega[k];
    tva[k] = (cube(mdiam + omega[k]) - a * sqrt(a)) / (3. * mdiam * omega[k]) - mdiam;
  }
  double fmdist = sqrt(tva[0] * tva[0] + tva[1] * tva[1]);

  int sml = 0, lrg = 1;
  if (diam[0] >= diam[1]) { lrg = 0; sml = 1; }
  so_shape sh1 = {0, 0, 0, 0}, sh2 = {0, 0, 0, 0};
  if (ph1 == 4) {
    double s = 0.; for (int k = 0; k < ns - 1; ++k) s += c->volc[b1][k] * rhoic;
    sh1 = shape_coeffs(P, s, c->volc[b1][ns - 1] * rhori, c->numc[b1]);
  }
  if (ph2 == 4) {
    double s = 0.; for (int k = 0; k < ns - 1; ++k) s += c->volc[b2][k] * rhoic;
    sh2 = shape_coeffs(P, s, c->volc[b2][ns - 1] * rhori, c->numc[b2]);
  }
  double zrhoa = pres / (rd * temp);
  double vkin = visc / zrhoa;
  if (ph1 < 4) termv[0] = so_terminal_vel_liquid(diam1, c->rhomean[b1], zrhoa, visc, beta[0]);
  else termv[0] = terminal_vel_ice(dsph1, c->rhomean[b1], zrhoa, visc, sh1, diam1);
  if (ph2 < 4) termv[1] = so_terminal_vel_liquid(diam2, c->rhomean[b2], zrhoa, visc, beta[1]);
  else termv[1] = terminal_vel_ice(dsph2, c->rhomean[b2], zrhoa, visc, sh2, diam2);

  double reyn[2], schm[2];
  for (int k = 0; k < 2; ++k) { reyn[k] = diam[k] * termv[k] / vkin; schm[k] = vkin / dfpart[k]; }
  double dv = fabs(termv[0] - termv[1]);
  double stok = 2. * termv[sml] * dv / (diam[lrg] * grav);
  double zbrown = flux1 / (mdiam / (mdiam + fmdist) + flux1 / flux2);
  double zbrconv = 0.;
  if (reyn[lrg] <= 1.) zbrconv = 0.45 * zbrown * pow(reyn[lrg] * schm[sml], 1. / 3.);
  else if (reyn[lrg] > 1.) zbrconv = 0.45 * zbrown * sqrt(reyn[lrg]) * pow(schm[sml], 1. / 3.);
  double dsum = 0.5 * (diam[0] + diam[1]);
  double ztshear = sqrt(8. * pi * P->eddy_dis / (15. * vkin)) * cube(dsum);
  double zturbinert = pi * pow(P->eddy_dis, 0.75) / (grav * sqrt(sqrt(vkin))) * (dsum * dsum) * dv;
  double zea = stok * stok / ((stok + 0.5) * (stok + 0.5));
  double zev = 0.;
  if (stok > 1.214) {
    zev = 0.75 * log(2. * stok) / (stok - 1.214);
    zev = 1. / ((1. + zev) * (1. + zev));
  }
  double zecoll = (60. * zev + zea * reyn[lrg]) / (60. + reyn[lrg]);
  double zgrav = zecoll * pi * (mdiam * mdiam);
  zgrav = zgrav * dv;
  double k = zbrown + zbrconv + sqrt(zgrav * zgrav + ztshear * ztshear + zturbinert * zturbinert);
  /* charge enhancement :411-416 with chargeTime = chargeTimeMax for both bins */
  if (P->lcharge) k = k + k * P->charge_coll_enh * 0.5 * (1000. + 1000.) / 1000.;
  if (P->kpert != 0.) k *= 1. + P->kpert * kpert_unit(P->kseed, b1, b2, temp, pres);
  return k;
}

double so_coagc(const so_layout *L, const so_params *P, const double *numc,
                const double *volc, int b1, int b2, double temp, double pres) {
  so_cell c;
  memset(&c, 0, sizeof(c));
  for (int b = 0; b < L->ntotal; ++b) {
    c.numc[b] = numc[b];
    for (int s = 0; s < SO_MAXSPEC; ++s) c.volc[b][s] = volc[b * SO_MAXSPEC + s];
    update_diameter(L, &c, b);
    update_rhomean(L, &c, b);
  }
  return coagc_cell(L, P, &c, b1, b2, temp, pres);
}

/* ====================================================================== */
/* Kernel matrices: update_coagulation_kernels / buildKernel[Self]        */
/* mo_salsa_coagulation_kernels.f90:24-180.  Stored as one symmetric      */
/* ntotal x ntotal matrix: every zcc block is read with both index orders */
/* and holds a single value per unordered pair.                           */
/* ====================================================================== */
typedef struct { double k[SO_MAXBIN][SO_MAXBIN]; } so_kmat;

static void build_self(const so_layout *L, const so_params *P, so_cell *c, so_kmat *K,
                       int off, int nb, double T, double p) {
  for (int m = 0; m < nb; ++m) { update_diameter(L, c, off + m); update_rhomean(L, c, off + m); }
  for (int m = 0; m < nb; ++m) {
    if (c->numc[off + m] < L->nlim[off + m]) continue;
    for (int n = m; n < nb; ++n) {
      if (c->numc[off + n] < L->nlim[off + n]) continue;
      double v = coagc_cell(L, P, c, off + m, off + n, T, p);
      K->k[off + m][off + n] = v; K->k[off + n][off + m] = v;
    }
  }
}

static void build_cross(const so_layout *L, const so_params *P, so_cell *c, so_kmat *K,
                        int off1, int nb1, int off2, int nb2, double T, double p) {
  for (int m = 0; m < nb1; ++m) { update_diameter(L, c, off1 + m); update_rhomean(L, c, off1 + m); }
  for (int n = 0; n < nb2; ++n) { update_diameter(L, c, off2 + n); update_rhomean(L, c, off2 + n); }
  for (int m = 0; m < nb1; ++m) {
    if (c->numc[off1 + m] < L->nlim[off1 + m]) continue;
    for (int n = 0; n < nb2; ++n) {
      if (c->numc[off2 + n] < L->nlim[off2 + n]) continue;
      double v = coagc_cell(L, P, c, off1 + m, off2 + n, T, p);
      K->k[off1 + m][off2 + n] = v; K->k[off2 + n][off1 + m] = v;
    }
  }
}

static void update_kernels(const so_layout *L, const so_params *P, so_cell *c, so_kmat *K,
                           double T, double p) {
  memset(K, 0, sizeof(*K));
  int A = L->iaero, C = L->icloud, R = L->iprecp, I = L->iice;
  if (P->lscgaa) build_self(L, P, c, K, A, L->nbins, T, p);
  if (P->lscgcc) build_self(L, P, c, K, C, L->ncld, T, p);
  if (P->lscgpp) build_self(L, P, c, K, R, L->nprc, T, p);
  if (P->lscgii && L->nice) build_self(L, P, c, K, I, L->nice, T, p);
  if (P->lscgca) build_cross(L, P, c, K, A, L->nbins, C, L->ncld, T, p);
  if (P->lscgpa) build_cross(L, P, c, K, A, L->nbins, R, L->nprc, T, p);
  if (P->lscgia && L->nice) build_cross(L, P, c, K, A, L->nbins, I, L->nice, T, p);
  if (P->lscgpc) build_cross(L, P, c, K, C, L->ncld, R, L->nprc, T, p);
  if (P->lscgic && L->nice) build_cross(L, P, c, K, C, L->ncld, I, L->nice, T, p);
  if (P->lscgip && L->nice) build_cross(L, P, c, K, R, L->nprc, I, L->nice, T, p);
}

/* ====================================================================== */
/* Accumulators, mo_salsa_coagulation_processes.f90:707-1611.             */
/* Category-local bin indices; `co` / `to` are master-array offsets.      */
/* ====================================================================== */
static int count_precpbins_lt(const so_layout *L, double lim) {
  int n = 0;
  for (int k = 0; k < L->nprc; ++k) n += (L->precpbins[k] < lim);
  return n;
}

/* accumulateSink (direct), :707-1000 (secondary-ice branch omitted: lssecice off) */
static void acc_sink(const so_layout *L, const so_params *P, so_cell *c, const so_kmat *K,
                     int to, int trgtphase, int itrgt, int co, int collphase,
                     int istr, int iend, double *sink, double xx) {
  double nrate = 0., nrate80 = 0., nrate50 = 0., nrate_au80 = 0., nrate_au50 = 0.;
  double Eagg[SO_MAXBIN];
  for (int ll = 0; ll < SO_MAXBIN; ++ll) Eagg[ll] = 1.;
  if (collphase == 4 && trgtphase == 4) {
    for (int ll = istr; ll <= iend; ++ll) {
      double r1 = rime_fraction(L, c, co + ll), r2 = rime_fraction(L, c, co + itrgt);
      double rimfr = fmax(r1, r2);
      Eagg[ll] = P->Eiagg_max - rimfr * (P->Eiagg_max - P->Eiagg_min);
    }
  }
  for (int ll = istr; ll <= iend; ++ll) {
    double dnum = Eagg[ll] * xx * K->k[to + itrgt][co + ll] * c->numc[co + ll];
    *sink = *sink + dnum;
    nrate = nrate + dnum;
  }
  if (!c->rates) return;
  /* 80 um and 50 um diagnostics, :784-863 */
  if (collphase == 3) {
    const double lims[2] = {80.e-6, 50.e-6};
    for (int q = 0; q < 2; ++q) {
      double lim = lims[q];
      int cnt = count_precpbins_lt(L, lim);
      int tmin = itrgt < L->nprc - 1 ? itrgt : L->nprc - 1;
      if (trgtphase == 2 || (trgtphase == 3 && L->precpbins[tmin] < lim)) {
        double acc = 0., au = 0.;
        for (int ll = cnt; ll <= iend; ++ll)
          acc += Eagg[ll] * xx * K->k[to + itrgt][co + ll] * c->numc[co + ll];
        int ex = istr > cnt - 1 ? istr : cnt - 1;
        double D = (trgtphase == 2) ? c->dwet[L->icloud + itrgt] : c->dwet[L->iprecp + itrgt];
        for (int ll = istr; ll <= ex; ++ll) {
          if (cube(D) + cube(c->dwet[co + ll]) >= cube(lim))
            au += Eagg[ll] * xx * K->k[to + itrgt][co + ll] * c->numc[co + ll];
        }
        if (q == 0) { nrate80 = acc; nrate_au80 = au; } else { nrate50 = acc; nrate_au50 = au; }
      }
    }
  }
  if (trgtphase == 2 || trgtphase == 3) {
    double nt = c->numc[to + itrgt];
    double a80 = nrate80 * nt, a50 = nrate50 * nt, u80 = nrate_au80 * nt, u50 = nrate_au50 * nt;
    rate_acc(c, SO_R_ACCR80, &a80, 0, 0);
    rate_acc(c, SO_R_ACCR50, &a50, 0, 0);
    rate_acc(c, SO_R_AUTOCONV80, &u80, 0, 0);
    rate_acc(c, SO_R_AUTOCONV50, &u50, 0, 0);
  }
  double nt = c->numc[to + itrgt];
  double nr = nrate * nt;
  if (collphase == 3 && trgtphase == 2) rate_acc(c, SO_R_ACCR, &nr, 0, 0);
  else if (collphase == 2 && trgtphase == 1) rate_acc(c, SO_R_ACCOLL, &nr, 0, 0);
  else if (collphase == 3 && trgtphase == 1) rate_acc(c, SO_R_APCOLL, &nr, 0, 0);
  else if (collphase == 4 && trgtphase == 1) rate_acc(c, SO_R_AICOLL, &nr, 0, 0);
}

/* accumulateSinkReverse, :1004-1027 */
static void acc_sink_rev(so_cell *c, const so_kmat *K, int to, int itrgt, int co,
                         int istr, int iend, double *sink) {
  for (int ll = istr; ll <= iend; ++ll)
    *sink = *sink + K->k[co + ll][to + itrgt] * c->numc[co + ll];
}

/* accumulateSource (direct), :1031-1210 */
static void acc_source(const so_layout *L, so_cell *c, const so_kmat *K, int nspec,
                       int to, int trgtphase, int itrgt, int co, int collphase,
                       int istr, int iend, double *source) {
  double vrate[SO_MAXSPEC] = {0}, v80[SO_MAXSPEC] = {0}, v50[SO_MAXSPEC] = {0};
  double au80[SO_MAXSPEC] = {0}, au50[SO_MAXSPEC] = {0};
  for (int ll = istr; ll <= iend; ++ll) {
    double kk = K->k[co + ll][to + itrgt];
    for (int s = 0; s < nspec; ++s) {
      double dvol = kk * c->volc[co + ll][s];
      vrate[s] = vrate[s] + dvol;
      source[s] = source[s] + dvol;
    }
  }
  if (!c->rates) return;
  const double lims[2] = {80.e-6, 50.e-6};
  for (int q = 0; q < 2; ++q) {
    double lim = lims[q];
    double *au = q == 0 ? au80 : au50;
    int cnt = count_precpbins_lt(L, lim);
    double D = 0.;
    int ex = iend;
    if (trgtphase == 3) D = c->dwet[L->iprecp + itrgt];
    if (collphase == 3) ex = iend < cnt - 1 ? iend : cnt - 1;
    if (D > 0. && D < lim && (collphase == 2 || collphase == 3)) {
      for (int ll = istr; ll <= ex; ++ll)
        if (cube(D) + cube(c->dwet[co + ll]) > cube(lim))
          for (int s = 0; s < nspec; ++s) au[s] += K->k[co + ll][to + itrgt] * c->volc[co + ll][s];
    }
  }
  for (int q = 0; q < 2; ++q) {
    double lim = lims[q];
    double *vv = q == 0 ? v80 : v50;
    if (trgtphase == 3 && L->precpbins[itrgt] > lim) {
      int cnt = count_precpbins_lt(L, lim);
      int ex = iend;
      if (collphase == 3) ex = iend < cnt - 1 ? iend : cnt - 1;
      if (collphase == 2 || collphase == 3)
        for (int ll = istr; ll <= ex; ++ll)
          for (int s = 0; s < nspec; ++s) vv[s] += K->k[co + ll][to + itrgt] * c->volc[co + ll][s];
    }
  }
  if (trgtphase == 3) {
    double nt = c->numc[to + itrgt];
    for (int s = 0; s < nspec; ++s) { v80[s] *= nt; v50[s] *= nt; au80[s] *= nt; au50[s] *= nt; }
    rate_acc(c, SO_R_ACCR80, 0, v80, nspec);
    rate_acc(c, SO_R_ACCR50, 0, v50, nspec);
    rate_acc(c, SO_R_AUTOCONV80, 0, au80, nspec);
    rate_acc(c, SO_R_AUTOCONV50, 0, au50, nspec);
  }
  double nt = c->numc[to + itrgt];
  for (int s = 0; s < nspec; ++s) vrate[s] *= nt;
  if (trgtphase == 3 && collphase == 2) rate_acc(c, SO_R_ACCR, 0, vrate, nspec);
  else if (trgtphase == 2 && collphase == 1) rate_acc(c, SO_R_ACCOLL, 0, vrate, nspec);
  else if (trgtphase == 3 && collphase == 1) rate_acc(c, SO_R_APCOLL, 0, vrate, nspec);
  else if (trgtphase == 4 && collphase == 1) rate_acc(c, SO_R_AICOLL, 0, vrate, nspec);
}

/* accumulateSourceReverse, :1212-1234 */
static void acc_source_rev(so_cell *c, const so_kmat *K, int nspec, int to, int itrgt,
                           int co, int istr, int iend, double *source) {
  for (int ll = istr; ll <= iend; ++ll)
    for (int s = 0; s < nspec; ++s)
      source[s] = source[s] + K->k[to + itrgt][co + ll] * c->volc[co + ll][s];
}

/* accumulateSourcePhaseChange, :1236-1266 (SIP diagnostics omitted) */
static void acc_source_phase(const so_layout *L, so_cell *c, const so_kmat *K, int iice,
                             int to, int itrgt, int co, int istr, int iend,
                             double rhotrgt, double rhocoll, double *source) {
  for (int ll = istr; ll <= iend; ++ll) {
    double kk = K->k[co + ll][to + itrgt];
    for (int s = 0; s < L->ndry; ++s) source[s] = source[s] + kk * c->volc[co + ll][s];
    source[iice] = source[iice] + kk * c->volc[co + ll][L->iwa] * rhocoll / rhotrgt;
  }
}

/* accumulateSourceIce, :1373-1404 */
static void acc_source_ice(const so_layout *L, const so_params *P, so_cell *c, const so_kmat *K,
                           int nspec, int to, int itrgt, int istr, int iend, double *source) {
  for (int ll = istr; ll <= iend; ++ll) {
    double rimfr = fmax(rime_fraction(L, c, to + ll), rime_fraction(L, c, to + itrgt));
    double E = P->Eiagg_max - rimfr * (P->Eiagg_max - P->Eiagg_min);
    for (int s = 0; s < nspec; ++s)
      source[s] = source[s] + E * K->k[to + ll][to + itrgt] * c->volc[to + ll][s];
  }
}

/* accumulatePrecipFormation, :1430-1584 */
typedef struct {
  double volsink_slf[SO_MAXSPEC];
  double vol_prc[SO_MAXBIN][SO_MAXSPEC];
  double num_prc[SO_MAXBIN];
} so_pf;

static void acc_precip_formation(const so_layout *L, so_cell *c, const so_kmat *K, int nspec,
                                 int itrgt, int istr, int iend, double fix_coag,
                                 double *source, double *sink, so_pf *pf) {
  const int C = L->icloud;
  int selfcoll = (itrgt == istr && itrgt == iend);
  for (int ll = istr; ll <= iend; ++ll) {
    double D_new = cbrt_f(cube(c->dwet[C + itrgt]) + cube(c->dwet[C + ll]));
    int trgt_prc = 0;
    for (int q = 0; q < L->nprc; ++q) trgt_prc += (D_new > L->precpbins[q]);
    double dvol[SO_MAXSPEC] = {0}, dnum = 0.;
    double kk = K->k[C + ll][C + itrgt];
    if (D_new > L->precpbins[0]) {
      int tp = trgt_prc - 1;    /* 1-based COUNT -> 0-based bin */
      if (selfcoll) {
        for (int s = 0; s < nspec; ++s) {
          dvol[s] = c->volc[C + ll][s] * c->numc[C + itrgt] * kk * fix_coag;
          pf->vol_prc[tp][s] = pf->vol_prc[tp][s] + dvol[s];
        }
        dnum = (0.5 * kk * c->numc[C + ll] * c->numc[C + itrgt]) * fix_coag;
        pf->num_prc[tp] = pf->num_prc[tp] + dnum;
      } else {
        for (int s = 0; s < nspec; ++s) {
          dvol[s] = (c->volc[C + ll][s] * c->numc[C + itrgt] +
                     c->volc[C + itrgt][s] * c->numc[C + ll]) * kk * fix_coag;
          pf->vol_prc[tp][s] = pf->vol_prc[tp][s] + dvol[s];
        }
        dnum = (kk * c->numc[C + ll] * c->numc[C + itrgt]) * fix_coag;
        pf->num_prc[tp] = pf->num_prc[tp] + dnum;
      }
      rate_acc(c, SO_R_AUTOCONV, &dnum, dvol, nspec);
      if (D_new > 50.e-6) rate_acc(c, SO_R_AUTOCONV50, &dnum, dvol, nspec);
      if (D_new > 80.e-6) rate_acc(c, SO_R_AUTOCONV80, &dnum, dvol, nspec);
      if (selfcoll) {
        for (int s = 0; s < nspec; ++s)
          pf->volsink_slf[s] = pf->volsink_slf[s] +
                               c->volc[C + ll][s] * c->numc[C + itrgt] * kk * fix_coag;
        *sink = *sink + 0.5 * kk * c->numc[C + itrgt];
      } else {
        *sink = *sink + K->k[C + itrgt][C + ll] * c->numc[C + ll];
      }
    } else {
      if (selfcoll) {
        *sink = *sink + 0.5 * K->k[C + itrgt][C + ll] * c->numc[C + ll];
      } else {
        for (int s = 0; s < nspec; ++s) source[s] = source[s] + kk * c->volc[C + ll][s];
      }
    }
  }
}

/* applyCoag, :539-559 */
static void apply_coag(so_cell *c, int b, int nspec, double dt, const double *source,
                       double sink, double sink_self) {
  for (int s = 0; s < nspec; ++s)
    c->volc[b][s] = (c->volc[b][s] + dt * source[s] * c->numc[b]) / (1. + dt * sink);
  c->numc[b] = c->numc[b] / (1. + dt * (sink + sink_self));
}

/* ---- category sweeps ---- */

/* coag_ice, :476-535 ; applyCoagIce :606-645 */
static void coag_ice(const so_layout *L, const so_params *P, so_cell *c, const so_kmat *K, double dt) {
  const int nspec = L->nspec, I = L->iice;
  for (int kk = 0; kk < L->nice; ++kk) {
    if (c->numc[I + kk] < L->nlim[I + kk]) continue;
    double plus[SO_MAXSPEC] = {0}, minus = 0., minus_self = 0.;
    if (P->lscgii) {
      if (kk < L->nice - 1) acc_sink(L, P, c, K, I, 4, kk, I, 4, kk + 1, L->nice - 1, &minus, 1.);
      acc_sink(L, P, c, K, I, 4, kk, I, 4, kk, kk, &minus_self, 0.5);
    }
    if (P->lscgia)
      acc_source_phase(L, c, K, L->iwa, I, kk, L->iaero, L->in1a, L->fn2b, rhoic, rhowa, plus);
    if (P->lscgic)
      acc_source_phase(L, c, K, L->irim, I, kk, L->icloud, L->ica.cur, L->fcb.cur, rhori, rhowa, plus);
    if (P->lscgip)
      acc_source_phase(L, c, K, L->irim, I, kk, L->iprecp, 0, L->nprc - 1, rhori, rhowa, plus);
    if (P->lscgii && kk > 0) acc_source_ice(L, P, c, K, nspec, I, kk, 0, kk - 1, plus);
    int b = I + kk;
    for (int s = 0; s < L->ndry; ++s)
      c->volc[b][s] = (c->volc[b][s] + dt * plus[s] * c->numc[b]) / (1. + dt * minus);
    c->volc[b][L->irim] = (c->volc[b][L->irim] + dt * plus[L->irim] * c->numc[b]) / (1. + dt * minus);
    c->volc[b][L->iwa] = (c->volc[b][L->iwa] + dt * plus[L->iwa] * c->numc[b]) / (1. + dt * minus);
    c->numc[b] = c->numc[b] / (1. + dt * (minus + minus_self));
    update_rhomean(L, c, b);
  }
}

/* coag_precp, :418-470 ; precipSelfCoag :1588-1611 ; applyCoagPrecp :563-602 */
static void coag_precp(const so_layout *L, const so_params *P, so_cell *c, const so_kmat *K, double dt) {
  const int nspec = L->nwet, R = L->iprecp;
  for (int kk = 0; kk < L->nprc; ++kk) {
    if (c->numc[R + kk] < L->nlim[R + kk]) continue;
    double plus[SO_MAXSPEC] = {0}, minus = 0., minus_self = 0., num_slf = 0.;
    double vol_slf[SO_MAXSPEC] = {0};
    if (P->lscgpp) {
      if (kk < L->nprc - 1) {
        acc_sink(L, P, c, K, R, 3, kk, R, 3, kk + 1, L->nprc - 1, &minus, 1.);
        double kself = K->k[R + kk][R + kk];
        for (int s = 0; s < nspec; ++s)
          vol_slf[s] = vol_slf[s] + c->volc[R + kk][s] * c->numc[R + kk] * kself;
        num_slf = num_slf + (0.5 * kself * c->numc[R + kk] * c->numc[R + kk]);
      } else {
        acc_sink(L, P, c, K, R, 3, kk, R, 3, kk, kk, &minus_self, 0.5);
      }
    }
    if (P->lscgip && L->nice) acc_sink(L, P, c, K, R, 3, kk, L->iice, 4, 0, L->nice - 1, &minus, 1.);
    if (P->lscgpa) acc_source(L, c, K, nspec, R, 3, kk, L->iaero, 1, L->in1a, L->fn2b, plus);
    if (P->lscgpc) acc_source(L, c, K, nspec, R, 3, kk, L->icloud, 2, 0, L->ncld - 1, plus);
    if (P->lscgpp && kk > 0) acc_source(L, c, K, nspec, R, 3, kk, R, 3, 0, kk - 1, plus);
    apply_coag(c, R + kk, nspec, dt, plus, minus, minus_self);
    if (kk < L->nprc - 1) {
      int b = R + kk, b1 = R + kk + 1;
      for (int s = 0; s < nspec; ++s) c->volc[b1][s] = c->volc[b1][s] + dt * vol_slf[s];
      c->numc[b1] = c->numc[b1] + dt * num_slf;
      for (int s = 0; s < nspec; ++s) c->volc[b][s] = c->volc[b][s] - dt * vol_slf[s];
      c->numc[b] = c->numc[b] - dt * num_slf;
    }
  }
}

/* coag_aero, :24-215 */
static void coag_aero(const so_layout *L, const so_params *P, so_cell *c, const so_kmat *K, double dt) {
  const int nspec = L->nwet, A = L->iaero, C = L->icloud;
  const so_parbin ica = L->ica, fca = L->fca, icb = L->icb, fcb = L->fcb;
  /* regime 1a, :37-90 */
  for (int kk = L->in1a; kk <= L->fn1a; ++kk) {
    if (c->numc[A + kk] < L->nlim[A + kk]) continue;
    int icla = ica.cur + (kk - ica.par > 0 ? kk - ica.par : 0);
    int iclb = icb.cur + (icla - ica.cur);
    double plus[SO_MAXSPEC] = {0}, minus = 0., minus_self = 0.;
    if (P->lscgaa) {
      acc_sink(L, P, c, K, A, 1, kk, A, 1, kk + 1, L->fn2b, &minus, 1.);
      acc_sink(L, P, c, K, A, 1, kk, A, 1, kk, kk, &minus_self, 0.5);
    }
    if (P->lscgca) {
      acc_sink(L, P, c, K, A, 1, kk, C, 2, icla, fca.cur, &minus, 1.);
      acc_sink(L, P, c, K, A, 1, kk, C, 2, iclb, fcb.cur, &minus, 1.);
    }
    if (P->lscgpa) acc_sink(L, P, c, K, A, 1, kk, L->iprecp, 3, 0, L->nprc - 1, &minus, 1.);
    if (P->lscgia && L->nice) acc_sink(L, P, c, K, A, 1, kk, L->iice, 4, 0, L->nice - 1, &minus, 1.);
    if (P->lscgaa && kk > L->in1a) acc_source(L, c, K, nspec, A, 1, kk, A, 1, L->in1a, kk - 1, plus);
    if (P->lscgca) {
      icla = ica.cur + (kk - ica.par) - 1;
      iclb = icb.cur + (icla - ica.cur);
      if (icla >= ica.cur) acc_source_rev(c, K, nspec, A, kk, C, ica.cur, icla, plus);
      if (iclb >= icb.cur) acc_source_rev(c, K, nspec, A, kk, C, icb.cur, iclb, plus);
    }
    apply_coag(c, A + kk, nspec, dt, plus, minus, minus_self);
  }
  /* regime 2a, :95-152 */
  for (int kk = L->in2a; kk <= L->fn2a; ++kk) {
    if (c->numc[A + kk] < L->nlim[A + kk]) continue;
    int index_b = L->in2b + (kk - L->in2a);
    int d = kk - ica.par > 0 ? kk - ica.par : 0;
    int icla = ica.cur + d, iclb = icb.cur + d;
    double plus[SO_MAXSPEC] = {0}, minus = 0., minus_self = 0.;
    if (P->lscgaa) {
      if (kk < L->fn2a) {
        acc_sink(L, P, c, K, A, 1, kk, A, 1, kk + 1, L->fn2a, &minus, 1.);
        acc_sink(L, P, c, K, A, 1, kk, A, 1, index_b, L->fn2b, &minus, 1.);
      }
      acc_sink(L, P, c, K, A, 1, kk, A, 1, kk, kk, &minus_self, 0.5);
    }
    if (P->lscgca) {
      acc_sink(L, P, c, K, A, 1, kk, C, 2, icla, fca.cur, &minus, 1.);
      acc_sink(L, P, c, K, A, 1, kk, C, 2, iclb, fcb.cur, &minus, 1.);
    }
    if (P->lscgpa) acc_sink(L, P, c, K, A, 1, kk, L->iprecp, 3, 0, L->nprc - 1, &minus, 1.);
    if (P->lscgia && L->nice) acc_sink(L, P, c, K, A, 1, kk, L->iice, 4, 0, L->nice - 1, &minus, 1.);
    if (P->lscgaa) {
      acc_source(L, c, K, nspec, A, 1, kk, A, 1, L->in1a, kk - 1, plus);
      acc_source(L, c, K, nspec, A, 1, kk, A, 1, L->in2b, index_b, plus);
    }
    if (P->lscgca) {
      icla = ica.cur + (kk - ica.par) - 1;
      iclb = icb.cur + (icla - ica.cur);
      if (icla >= ica.cur) acc_source_rev(c, K, nspec, A, kk, C, ica.cur, icla, plus);
      if (iclb >= icb.cur) acc_source_rev(c, K, nspec, A, kk, C, icb.cur, iclb, plus);
    }
    apply_coag(c, A + kk, nspec, dt, plus, minus, minus_self);
  }
  /* regime 2b, :157-213 */
  for (int kk = L->in2b; kk <= L->fn2b; ++kk) {
    if (c->numc[A + kk] < L->nlim[A + kk]) continue;
    int index_a = L->in2a + (kk - L->in2b);
    int d = kk - icb.par > 0 ? kk - icb.par : 0;
    int icla = ica.cur + d, iclb = icb.cur + d;
    double plus[SO_MAXSPEC] = {0}, minus = 0., minus_self = 0.;
    if (P->lscgaa) {
      if (kk < L->fn2b) {
        acc_sink(L, P, c, K, A, 1, kk, A, 1, kk + 1, L->fn2b, &minus, 1.);
        acc_sink(L, P, c, K, A, 1, kk, A, 1, index_a, L->fn2a, &minus, 1.);
      }
      acc_sink(L, P, c, K, A, 1, kk, A, 1, kk, kk, &minus_self, 0.5);
    }
    if (P->lscgca) {
      acc_sink(L, P, c, K, A, 1, kk, C, 2, icla, fca.cur, &minus, 1.);
      acc_sink(L, P, c, K, A, 1, kk, C, 2, iclb, fcb.cur, &minus, 1.);
    }
    if (P->lscgpa) acc_sink(L, P, c, K, A, 1, kk, L->iprecp, 3, 0, L->nprc - 1, &minus, 1.);
    if (P->lscgia && L->nice) acc_sink(L, P, c, K, A, 1, kk, L->iice, 4, 0, L->nice - 1, &minus, 1.);
    if (P->lscgaa) {
      acc_source(L, c, K, nspec, A, 1, kk, A, 1, L->in2b, kk - 1, plus);
      acc_source(L, c, K, nspec, A, 1, kk, A, 1, L->in1a, index_a, plus);
    }
    if (P->lscgca) {
      iclb = icb.cur + (kk - icb.par) - 1;
      icla = ica.cur + (iclb - icb.cur);
      if (icla >= ica.cur) acc_source_rev(c, K, nspec, A, kk, C, ica.cur, icla, plus);
      if (iclb >= icb.cur) acc_source_rev(c, K, nspec, A, kk, C, icb.cur, iclb, plus);
    }
    apply_coag(c, A + kk, nspec, dt, plus, minus, minus_self);
  }
}

/* applyCoagPrecipFormation, :649-703 (IN-fraction bookkeeping omitted: INdef is
 * a level-5 ice-nucleation diagnostic that feeds nothing in this step) */
static void apply_precip_formation(const so_layout *L, so_cell *c, int nspec, int itrgt,
                                   double dt, const so_pf *pf) {
  int b = L->icloud + itrgt;
  for (int s = 0; s < nspec; ++s) c->volc[b][s] = c->volc[b][s] - dt * pf->volsink_slf[s];
  for (int q = 0; q < L->nprc; ++q) {
    int r = L->iprecp + q;
    for (int s = 0; s < nspec; ++s) c->volc[r][s] = c->volc[r][s] + dt * pf->vol_prc[q][s];
    c->numc[r] = c->numc[r] + dt * pf->num_prc[q];
  }
}

/* coag_cloud, :221-412; the a- and b-series loops differ only in index maps */
static void coag_cloud(const so_layout *L, const so_params *P, so_cell *c, const so_kmat *K, double dt) {
  const int nspec = L->nwet, A = L->iaero, C = L->icloud;
  const so_parbin ica = L->ica, fca = L->fca, icb = L->icb, fcb = L->fcb;
  const int pfmode = P->lsauto_state && P->lsauto_mode == 1;
  for (int series = 0; series < 2; ++series) {
    int cur0 = series == 0 ? ica.cur : icb.cur;     /* first bin of this series   */
    int cur1 = series == 0 ? fca.cur : fcb.cur;     /* last bin of this series    */
    int oth0 = series == 0 ? icb.cur : ica.cur;     /* first bin of other series  */
    int oth1 = series == 0 ? fcb.cur : fca.cur;
    for (int kk = cur0; kk <= cur1; ++kk) {
      if (c->numc[C + kk] < L->nlim[C + kk]) continue;
      double plus[SO_MAXSPEC] = {0}, minus = 0., minus_self = 0.;
      so_pf pf;
      memset(&pf, 0, sizeof(pf));
      int index_o = oth0 + (kk - cur0);             /* index_b (a-loop) / index_a (b-loop) */
      if (P->lscgca && kk < cur1) {
        int iaa = ica.par + (kk - cur0) + 1, iab = icb.par + (kk - cur0) + 1;
        acc_sink_rev(c, K, C, kk, A, iaa, L->fn2a, &minus);
        acc_sink_rev(c, K, C, kk, A, iab, L->fn2b, &minus);
      }
      if (P->lscgcc) {
        if (kk < cur1) {
          acc_sink(L, P, c, K, C, 2, kk, C, 2, kk + 1, cur1, &minus, 1.);
          acc_sink(L, P, c, K, C, 2, kk, C, 2, index_o, oth1, &minus, 1.);
        }
        if (pfmode) {
          double nk = c->numc[C + kk];
          double fix = fmax(1. - dt * (0.5 * K->k[C + kk][C + kk] * (nk * nk) / (nk)), 0.1);
          acc_precip_formation(L, c, K, nspec, kk, kk, kk, fix, plus, &minus_self, &pf);
        } else {
          acc_sink(L, P, c, K, C, 2, kk, C, 2, kk, kk, &minus_self, 0.5);
        }
      }
      if (P->lscgpc) acc_sink(L, P, c, K, C, 2, kk, L->iprecp, 3, 0, L->nprc - 1, &minus, 1.);
      if (P->lscgic && L->nice) acc_sink(L, P, c, K, C, 2, kk, L->iice, 4, 0, L->nice - 1, &minus, 1.);
      if (P->lscgca) {
        int iaa = ica.par + (kk - cur0), iab = icb.par + (kk - cur0);
        acc_source(L, c, K, nspec, C, 2, kk, A, 1, L->in1a, iaa, plus);
        acc_source(L, c, K, nspec, C, 2, kk, A, 1, L->in2b, iab, plus);
      }
      if (P->lscgcc && kk > cur0) {
        if (pfmode) {
          double s1 = 0.;
          for (int q = cur0; q <= cur1; ++q) s1 += K->k[C + kk][C + q] * c->numc[C + q];
          double fix = fmax(1. - dt * s1, 0.1);
          acc_precip_formation(L, c, K, nspec, kk, cur0, kk - 1, fix, plus, &minus, &pf);
          double s2 = 0.;
          for (int q = oth0; q <= oth1; ++q) s2 += K->k[C + kk][C + q] * c->numc[C + q];
          fix = fmax(1. - dt * s2, 0.1);
          acc_precip_formation(L, c, K, nspec, kk, oth0, index_o, fix, plus, &minus, &pf);
        } else {
          acc_source(L, c, K, nspec, C, 2, kk, C, 2, cur0, kk - 1, plus);
          acc_source(L, c, K, nspec, C, 2, kk, C, 2, oth0, index_o, plus);
        }
      }
      apply_coag(c, C + kk, nspec, dt, plus, minus, minus_self);
      if (pfmode) apply_precip_formation(L, c, nspec, kk, dt, &pf);
    }
  }
}

/* coagulation, src/src_salsa/mo_salsa_dynamics.f90:92-202.  kmode 0: kernels
 * computed (lscoag%mode 1); 1: computed and copied to kstore (mode 2 update
 * step, store_coag mo_salsa_driver.f90:649-673); 2: taken from kstore (mode 2
 * other steps, fetch_coag :677-697, no update_coagulation_kernels :146-147).
 * kstore is ntotal x ntotal, row-major k[i][j]. */
static int coagulation(const so_layout *L, const so_params *P, so_cell *c,
                       double dt, double T, double p, double *kstore, int kmode) {
  int status = 0;
  so_kmat K;
  const int nt = L->ntotal;
  for (int b = 0; b < L->ntotal; ++b) update_diameter(L, c, b);
  if (kmode == 2) {
    memset(&K, 0, sizeof(K));
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j) K.k[i][j] = kstore[i * nt + j];
  } else {
    update_kernels(L, P, c, &K, T, p);
    if (kmode == 1)
      for (int i = 0; i < nt; ++i)
        for (int j = 0; j < nt; ++j) kstore[i * nt + j] = K.k[i][j];
  }
  int any_aero = P->lscgaa || P->lscgca || P->lscgpa || P->lscgia;
  int any_cloud = P->lscgcc || P->lscgca || P->lscgpc || P->lscgic;
  int any_precp = P->lscgpp || P->lscgpa || P->lscgpc || P->lscgip;
  int any_ice = P->lscgii || P->lscgia || P->lscgic || P->lscgip;
  if (any_ice && T < 273.15 && L->nice) coag_ice(L, P, c, &K, dt);
  if (any_precp) coag_precp(L, P, c, &K, dt);
  if (any_aero) coag_aero(L, P, c, &K, dt);
  if (any_cloud) coag_cloud(L, P, c, &K, dt);
  for (int b = 0; b < L->ntotal; ++b) {
    if (c->numc[b] < -1.e-6) status |= SO_WARN_NEG_NUMC;
    c->numc[b] = fmax(0., c->numc[b]);
    for (int s = 0; s < SO_MAXSPEC; ++s) c->volc[b][s] = fmax(0., c->volc[b][s]);
  }
  return status;
}

/* ====================================================================== */
/* Condensation                                                           */
/* ====================================================================== */

/* acth2o, mo_salsa_dynamics.f90:1127-1157 */
static double acth2o_bin(const so_layout *L, const double *volc, const double *pcw) {
  double zns = 0.;
  for (int nn = 0; nn < L->ndry; ++nn)
    zns = zns + L->diss[nn] * volc[nn] * L->rholiq[nn] / L->mm[nn];
  double znw = pcw ? *pcw : volc[L->iwa] * L->rholiq[L->iwa] / L->mm[L->iwa];
  return fmax(0.1, znw / fmax(eps_r8, (znw + zns)));
}

/* Test hook (so_params.diag_growth), no counterpart in the reference: the
 * factor by which one explicit substep multiplies a perturbation of bin b's
 * water where the +-10 % limiter does not bind: the bin's own update,
 * 1 - a csu w' with a = adt zmt, w' = d zwsat / d zcw (Raoult term of
 * acth2o_bin times the Kelvin factor), times the share 1 - a / (1 + adt S)
 * that the implicit vapour predictor leaves. */
static double update_multiplier(const so_layout *L, const double *volc, double zcw, double d, double a,
                                double csu, double zkelb, double zhlp2) {
  if (!(fabs(d) < 1.e-1 * zcw)) return 1.;
  double zns = 0.;
  for (int nn = 0; nn < L->ndry; ++nn) zns += L->diss[nn] * volc[nn] * L->rholiq[nn] / L->mm[nn];
  const double den = zcw + zns;
  const double wp = den > eps_r8 && zcw / den > 0.1 ? zkelb * zns / (den * den) : 0.;
  return fabs((1. - a * csu * wp) * (1. - a / zhlp2));
}

double so_acth2o(const so_layout *L, const double *volc_bin, const double *pcw) {
  return acth2o_bin(L, volc_bin, pcw);
}

/* equilibration, src/src_salsa/mo_salsa_properties.f90:73-380 */
static int equilibration(const so_layout *L, so_cell *c, double prh, double ptemp, int init) {
  const int A = L->iaero;
  int status = 0;
  /* massindx, :128-137; only slots 1..5 are copied into zvolc (:140-153) */
  int massindx[7] = {-1, -1, -1, -1, -1, -1, -1};
  for (int s = 0; s < L->ndry; ++s) massindx[L->spec_id[s]] = s;
  double zvolc[SO_MAXBIN][7];
  for (int kk = L->in1a; kk <= L->fn2b; ++kk)
    for (int cc = 0; cc < 7; ++cc)
      zvolc[kk][cc] = (cc < 5 && massindx[cc] >= 0) ? c->volc[A + kk][massindx[cc]] : 0.;

  double zke = 1.001;   /* carried across 1a bins (:165) */
  for (int kk = L->in1a; kk <= L->fn1a; ++kk) {
    double zdold = 1.;
    int b = A + kk;
    if (c->numc[b] > nlim_c) {
      double zvpart1 = zvolc[kk][0] / c->numc[b];
      double zvpart2 = zvolc[kk][1] / c->numc[b];
      double zvp67 = zvolc[kk][5] / c->numc[b] + zvolc[kk][6] / c->numc[b];
      double zcore = zvpart1 + zvpart2;
      double zdwet = c->dwet[b];
      double zrh = prh;
      zrh = fmax(zrh, 0.05);
      zrh = fmin(zrh, 0.98);
      double zlwc = 0.;
      int count = 0;
      while (fabs(zdwet / zdold - 1.) > 1.e-2) {
        zdold = fmax(zdwet, 1.e-20);
        double zaw = zrh / zke;
        double bm1 = +1.1065495e+2 - 3.6759197e+2 * zaw + 5.0462934e+2 * (zaw * zaw)
                     - 3.1543839e+2 * (zaw * zaw * zaw) + 6.770824e+1 * ((zaw * zaw) * (zaw * zaw));
        double bm2 = 1. / (zaw * mwa) - 1. / mwa;
        zlwc = (zvolc[kk][0] * (rhosu / msu)) / bm1 +
               epsoc * zvolc[kk][1] * (rhooc / moc) / bm2;
        zdwet = cbrt_f(zlwc / c->numc[b] / rhowa / pi6 + zvp67 / pi6 + zcore / pi6);
        zke = exp(2. * surfw0 * MVWA / (boltz * ptemp * zdwet));
        count = count + 1;
        if (count > 100) { status |= SO_ERR_EQUIL_NOCONV; break; }
      }
      c->volc[b][L->iwa] = zlwc / rhowa;
      if (init) c->dwet[b] = zdwet;
    } else {
      if (init) c->dwet[b] = L->dmid[b];
    }
  }
  if (init) {
    for (int kk = L->in2a; kk <= L->fn2b; ++kk) {
      int b = A + kk;
      double zke2 = 1.02, zdold = 1.;
      if (c->numc[b] > nlim_c) {
        double zvpart[7];
        for (int q = 0; q < 7; ++q) zvpart[q] = zvolc[kk][q] / c->numc[b];
        double zcore = zvpart[0] + zvpart[1] + zvpart[2] + zvpart[3] + zvpart[4];
        double zvp67 = zvpart[5] + zvpart[6];
        double zdwet = c->dwet[b];
        double zrh = fmin(fmax(prh, 0.37), 0.98);
        double zlwc = 0.;
        int count = 0;
        while (fabs(zdwet / zdold - 1.) > 1.e-12) {
          zdold = fmax(zdwet, 1.e-20);
          double zaw = zrh / zke2;
          double bm1 = +1.1065495e+2 - 3.6759197e+2 * zaw + 5.0462934e+2 * (zaw * zaw)
                       - 3.1543839e+2 * (zaw * zaw * zaw) + 6.770824e+1 * ((zaw * zaw) * (zaw * zaw));
          double bm2 = 1. / (zaw * mwa) - 1. / mwa;
          double bm5 = +5.875248e+1 - 1.8781997e+2 * zaw + 2.7211377e+2 * (zaw * zaw)
                       - 1.8458287e+2 * (zaw * zaw * zaw) + 4.153689e+1 * ((zaw * zaw) * (zaw * zaw));
          zlwc = (zvolc[kk][0] * (rhosu / msu)) / bm1 +
                 epsoc * (zvolc[kk][1] * (rhooc / moc)) / bm2 +
                 (zvolc[kk][4] * (rhoss / mss)) / bm5;
          zdwet = cbrt_f(zlwc / c->numc[b] / rhowa / pi6 + zvp67 / pi6 + zcore / pi6);
          zke2 = exp(2. * surfw0 * MVWA / (boltz * ptemp * zdwet));
          count = count + 1;
          if (count > 100) { status |= SO_ERR_EQUIL_NOCONV; break; }
        }
        c->volc[b][L->iwa] = zlwc / rhowa;
        c->dwet[b] = zdwet;
      } else {
        c->dwet[b] = L->dmid[b];
      }
    }
  }
  return status;
}

/* gpparth2o, mo_salsa_dynamics.f90:691-1125 */
static int gpparth2o(const so_layout *L, const so_params *P, so_cell *c,
                     double ptemp, double ppres, double prs, double prsi,
                     double *prv, double ptstep) {
  const int A = L->iaero, C = L->icloud, R = L->iprecp, I = L->iice;
  const int iwa = L->iwa, irim = L->irim, nspec = L->nspec;
  const int nstr = L->in2a;
  int status = 0;
  double zrh = *prv / prs;

  double zaelwc1 = 0.;
  for (int kk = L->in1a; kk <= L->fn2b; ++kk) zaelwc1 += c->volc[A + kk][iwa];
  zaelwc1 = zaelwc1 * rhowa;
  status |= equilibration(L, c, zrh, ptemp, 0);
  if (zrh < 0.98 || !P->lscndh2oae) status |= equilibration(L, c, zrh, ptemp, 1);
  double zaelwc2 = 0.;
  for (int kk = L->in1a; kk <= L->fn2b; ++kk) zaelwc2 += c->volc[A + kk][iwa];
  zaelwc2 = zaelwc2 * rhowa;
  *prv = *prv - (zaelwc2 - zaelwc1) / (ppres * mair / (rg * ptemp));

  /* cell-skip predicate, :784-788 (.AND. binds before .OR.) */
  int anyc = 0, anyp = 0, anyi = 0, anya = 0;
  for (int q = 0; q < L->ncld; ++q) anyc |= c->numc[C + q] > nlim_c;
  for (int q = 0; q < L->nprc; ++q) anyp |= c->numc[R + q] > prlim_c;
  for (int q = 0; q < L->nice; ++q) anyi |= c->numc[I + q] > prlim_c;
  for (int q = 0; q < L->nbins; ++q) anya |= c->numc[A + q] > nlim_c;
  if (!((anyc || (anyp && P->lscndh2ocl)) || (anyi && P->lscndh2oic) ||
        (anya && zrh > 0.98 && P->lscndh2oae)))
    return status;

  double rhoair = mair * ppres / (rg * ptemp);
  double zdfh2o = (5. / (16. * avog * rhoair * 1.e-3 * ((3.11e-8) * (3.11e-8)))) *
                  sqrt(rg * 1.e7 * ptemp * mair * 1.e3 * (mwa + mair) * 1.e3 / (2. * pi * mwa * 1.e3));
  zdfh2o = zdfh2o * 1.e-4;
  double zmfph2o = 3. * zdfh2o * sqrt(pi * mwa / (8. * rg * ptemp));
  double zthcond = 0.023807 + 7.1128e-5 * (ptemp - 273.16);

  double zkel[SO_MAXBIN], zcwsurf[SO_MAXBIN], zmt[SO_MAXBIN], zwsat[SO_MAXBIN];
  double zcwc_b[SO_MAXBIN], zcwint_b[SO_MAXBIN], zorgic[SO_MAXBIN], zorgri[SO_MAXBIN];
  for (int b = 0; b < L->ntotal; ++b) {
    zkel[b] = 1.; zcwsurf[b] = 0.; zmt[b] = 0.; zwsat[b] = 0.;
    zcwc_b[b] = 0.; zcwint_b[b] = 0.; zorgic[b] = 0.; zorgri[b] = 0.;
  }
  for (int b = 0; b < L->ntotal; ++b) { update_diameter(L, c, b); update_rhomean(L, c, b); }

  /* liquid hydrometeors, :826-890 */
  for (int cat = 0; cat < 2; ++cat) {
    int off = cat == 0 ? C : R, nb = cat == 0 ? L->ncld : L->nprc;
    for (int cc = 0; cc < nb; ++cc) {
      int b = off + cc;
      if (c->numc[b] > L->nlim[b] && P->lscndh2ocl) {
        double dwet = c->dwet[b];
        double zact = acth2o_bin(L, c->volc[b], 0);
        zkel[b] = exp(4. * surfw0 * mwa / (rg * ptemp * rhowa * dwet));
        zcwsurf[b] = prs * rhoair / mwa;
        zwsat[b] = zact * zkel[b];
        double zknud = 2. * zmfph2o / dwet;
        double zbeta = (zknud + 1.) / (0.377 * zknud + 1. + 4. / (3.) * (zknud + zknud * zknud));
        double h1 = c->numc[b] * 2. * pi * dwet * zdfh2o * zbeta;
        double h2 = mwa * zdfh2o * alv * zwsat[b] * zcwsurf[b] / (zthcond * ptemp);
        double h3 = ((alv * mwa) / (rg * ptemp)) - 1.;
        zmt[b] = h1 / (h2 * h3 + 1.);
      }
    }
  }
  /* ice, :892-930 */
  for (int cc = 0; cc < L->nice; ++cc) {
    int b = I + cc;
    if (c->numc[b] > L->nlim[b] && P->lscndh2oic && ptemp < 273.15) {
      double dwet = c->dnsp[b];
      double cap = 0.5 * dwet;
      double zact = 1.0;
      zkel[b] = exp(4. * surfi0 * mwa / (rg * ptemp * rhowa * dwet));
      zcwsurf[b] = prsi * rhoair / mwa;
      zwsat[b] = zact * zkel[b];
      double zknud = 2. * zmfph2o / dwet;
      double zbeta = (zknud + 1.) / (0.377 * zknud + 1. + 4. / (3.) * (zknud + zknud * zknud));
      double h1 = c->numc[b] * 4. * pi * cap * zdfh2o * zbeta;
      double h2 = mwa * zdfh2o * als * zwsat[b] * zcwsurf[b] / (zthcond * ptemp);
      double h3 = ((als * mwa) / (rg * ptemp)) - 1.;
      zmt[b] = h1 / (h2 * h3 + 1.);
    }
  }
  /* aerosol, :932-955 (massacc = 1) */
  for (int cc = 0; cc < L->nbins; ++cc) {
    int b = A + cc;
    if (c->numc[b] > L->nlim[b] && zrh > 0.98 && P->lscndh2oae) {
      double dwet = c->dwet[b];
      double zact = acth2o_bin(L, c->volc[b], 0);
      zkel[b] = exp(4. * surfw0 * mwa / (rg * ptemp * rhowa * dwet));
      zcwsurf[b] = prs * rhoair / mwa;
      zwsat[b] = zact * zkel[b];
      double zknud = 2. * zmfph2o / dwet;
      double zbeta = (zknud + 1.) / (0.377 * zknud + 1. + 4. / (3. * 1.) * (zknud + zknud * zknud));
      double h1 = c->numc[b] * 2. * pi * dwet * zdfh2o * zbeta;
      double h2 = mwa * zdfh2o * alv * zwsat[b] * zcwsurf[b] / (zthcond * ptemp);
      double h3 = ((alv * mwa) / (rg * ptemp)) - 1.;
      zmt[b] = h1 / (h2 * h3 + 1.);
    }
  }

  /* current mole concentrations, :958-978 */
  double zcwc = *prv * rhoair / mwa;
  for (int b = A; b < A + L->nbins; ++b) zcwc_b[b] = c->volc[b][iwa] * rhowa / mwa;
  for (int b = C; b < C + L->ncld; ++b) zcwc_b[b] = c->volc[b][iwa] * rhowa / mwa;
  for (int b = R; b < R + L->nprc; ++b) zcwc_b[b] = c->volc[b][iwa] * rhowa / mwa;
  for (int b = I; b < I + L->nice; ++b) {
    zcwc_b[b] = c->volc[b][iwa] * rhoic / mwa;
    zcwc_b[b] = zcwc_b[b] + c->volc[b][irim] * rhori / mwa;
    zorgic[b] = c->volc[b][iwa] * rhoic / mwa;
    zorgri[b] = c->volc[b][irim] * rhori / mwa;
  }
  /* zcwtot = zcwc + SUM(ae) + SUM(cd) + SUM(pd) + SUM(it), :980-984 */
  double sae = 0., scd = 0., spd = 0., sit = 0.;
  for (int b = A; b < A + L->nbins; ++b) sae += zcwc_b[b];
  for (int b = C; b < C + L->ncld; ++b) scd += zcwc_b[b];
  for (int b = R; b < R + L->nprc; ++b) spd += zcwc_b[b];
  for (int b = I; b < I + L->nice; ++b) sit += zcwc_b[b];
  double zcwtot = zcwc + sae + scd + spd + sit;
  for (int b = 0; b < L->ntotal; ++b) zcwint_b[b] = zcwc_b[b];

  int act_a = 0, act_c = 0, act_p = 0, act_i = 0;
  for (int q = 0; q < L->nbins; ++q) act_a |= c->numc[A + q] > L->nlim[A + q];
  for (int q = 0; q < L->ncld; ++q) act_c |= c->numc[C + q] > L->nlim[C + q];
  for (int q = 0; q < L->nprc; ++q) act_p |= c->numc[R + q] > L->nlim[R + q];
  for (int q = 0; q < L->nice; ++q) act_i |= c->numc[I + q] > L->nlim[I + q];

  double ttot = 0., zcwint = 0.;
  int counter = 0;
  double grow = 1.;                         /* diag_growth */
  while (ttot < ptstep) {
    double adt = 2.e-2;
    /* vapour predictor, :992-1000 */
    double s1a = 0., s1c = 0., s1p = 0., s1i = 0., s2a = 0., s2c = 0., s2p = 0., s2i = 0.;
    for (int b = A + nstr; b < A + L->nbins; ++b) s1a += zmt[b] * zwsat[b] * zcwsurf[b];
    for (int b = C; b < C + L->ncld; ++b) s1c += zmt[b] * zwsat[b] * zcwsurf[b];
    for (int b = R; b < R + L->nprc; ++b) s1p += zmt[b] * zwsat[b] * zcwsurf[b];
    for (int b = I; b < I + L->nice; ++b) s1i += zmt[b] * zwsat[b] * zcwsurf[b];
    for (int b = A + nstr; b < A + L->nbins; ++b) s2a += zmt[b];
    for (int b = C; b < C + L->ncld; ++b) s2c += zmt[b];
    for (int b = R; b < R + L->nprc; ++b) s2p += zmt[b];
    for (int b = I; b < I + L->nice; ++b) s2i += zmt[b];
    double zhlp1 = zcwc + adt * (s1a + s1c + s1p + s1i);
    double zhlp2 = 1. + adt * (s2a + s2c + s2p + s2i);
    zcwint = zhlp1 / zhlp2;
    zcwint = fmin(zcwint, zcwtot);
    double mmax = 1.;
    /* per-bin correctors, :1002-1031 */
    if (act_a && zrh > 0.98) {
      for (int b = A + nstr; b < A + L->nbins; ++b) {
        double d = adt * zmt[b] * (zcwint - zwsat[b] * zcwsurf[b]);
        if (P->diag_growth)
          mmax = fmax(mmax, update_multiplier(L, c->volc[b], zcwc_b[b], d, adt * zmt[b], zcwsurf[b], zkel[b], zhlp2));
        zcwint_b[b] = zcwc_b[b] + fmin(fmax(d, -1.e-1 * zcwc_b[b]), 1.e-1 * zcwc_b[b]);
        zwsat[b] = acth2o_bin(L, c->volc[b], &zcwint_b[b]) * zkel[b];
      }
    }
    if (act_c) {
      for (int b = C; b < C + L->ncld; ++b) {
        double d = adt * zmt[b] * (zcwint - zwsat[b] * zcwsurf[b]);
        if (P->diag_growth)
          mmax = fmax(mmax, update_multiplier(L, c->volc[b], zcwc_b[b], d, adt * zmt[b], zcwsurf[b], zkel[b], zhlp2));
        zcwint_b[b] = zcwc_b[b] + fmin(fmax(d, -1.e-1 * zcwc_b[b]), 1.e-1 * zcwc_b[b]);
        zwsat[b] = acth2o_bin(L, c->volc[b], &zcwint_b[b]) * zkel[b];
      }
    }
    if (act_p) {
      for (int b = R; b < R + L->nprc; ++b) {
        double d = adt * zmt[b] * (zcwint - zwsat[b] * zcwsurf[b]);
        if (P->diag_growth)
          mmax = fmax(mmax, update_multiplier(L, c->volc[b], zcwc_b[b], d, adt * zmt[b], zcwsurf[b], zkel[b], zhlp2));
        zcwint_b[b] = zcwc_b[b] + fmin(fmax(d, -1.e-1 * zcwc_b[b]), 1.e-1 * zcwc_b[b]);
        zwsat[b] = acth2o_bin(L, c->volc[b], &zcwint_b[b]) * zkel[b];
      }
    }
    if (act_i) {
      for (int b = I; b < I + L->nice; ++b) {
        double d = adt * zmt[b] * (zcwint - zwsat[b] * zcwsurf[b]);
        zcwint_b[b] = zcwc_b[b] + fmin(fmax(d, -1.e-1 * zcwc_b[b]), 1.e-1 * zcwc_b[b]);
        zwsat[b] = zkel[b];
      }
    }
    for (int b = A + nstr; b < A + L->nbins; ++b) zcwint_b[b] = fmax(zcwint_b[b], 0.);
    for (int b = C; b < L->ntotal; ++b) zcwint_b[b] = fmax(zcwint_b[b], 0.);
    /* closure, :1034-1038 */
    double ta = 0., tc = 0., tp = 0., ti = 0.;
    for (int b = A; b < A + L->nbins; ++b) ta += zcwint_b[b];
    for (int b = C; b < C + L->ncld; ++b) tc += zcwint_b[b];
    for (int b = R; b < R + L->nprc; ++b) tp += zcwint_b[b];
    for (int b = I; b < I + L->nice; ++b) ti += zcwint_b[b];
    zcwint = zcwtot - ta - tc - tp - ti;
    for (int b = 0; b < L->ntotal; ++b) zcwc_b[b] = zcwint_b[b];
    zcwc = zcwint;
    ttot = ttot + adt;
    counter = counter + 1;
    if (counter > 100) status |= SO_WARN_SUBSTEPS;
    grow = fmin(grow * mmax, 1.e300);
  }
  if (P->diag_growth && grow > 1.e3) status |= SO_DIAG_GROWTH3;
  if (P->diag_growth && grow > 1.e5) status |= SO_DIAG_GROWTH5;
  *prv = zcwint * mwa / rhoair;

  /* write-back and diagnostics, :1058-1117 */
  double vr[SO_MAXSPEC];
  for (int cat = 0; cat < 3; ++cat) {
    int off = cat == 0 ? A : (cat == 1 ? C : R);
    int nb = cat == 0 ? L->nbins : (cat == 1 ? L->ncld : L->nprc);
    int rslot = cat == 0 ? SO_R_CONDA : (cat == 1 ? SO_R_CONDC : SO_R_CONDP);
    for (int b = off; b < off + nb; ++b) {
      for (int s = 0; s < nspec; ++s) vr[s] = 0.;
      double nv = fmax(0., zcwint_b[b] * mwa / rhowa);
      vr[iwa] = nv - c->volc[b][iwa];
      rate_acc(c, rslot, 0, vr, nspec);
      c->volc[b][iwa] = fmax(0., zcwint_b[b] * mwa / rhowa);
    }
  }
  for (int s = 0; s < nspec; ++s) vr[s] = 0.;
  for (int b = I; b < I + L->nice; ++b) {
    double dvitot = zcwint_b[b] - (zorgic[b] + zorgri[b]);
    if (dvitot > 0.) {
      double dvice = dvitot;
      c->volc[b][iwa] = c->volc[b][iwa] + dvice * mwa / rhoic;
      vr[iwa] = dvice * mwa / rhoic;
      rate_acc(c, SO_R_CONDI, 0, vr, nspec);
    } else if (dvitot < 0.) {
      double dvice = (zcwint_b[b] / (zorgic[b] + zorgri[b])) - 1.;
      dvice = dvice * zorgic[b];
      double dvrime = (zcwint_b[b] / (zorgic[b] + zorgri[b])) - 1.;
      dvrime = dvrime * zorgri[b];
      vr[iwa] = dvice * mwa / rhoic;
      vr[irim] = dvrime * mwa / rhori;
      rate_acc(c, SO_R_CONDI, 0, vr, nspec);
      c->volc[b][iwa] = c->volc[b][iwa] + dvice * mwa / rhoic;
      c->volc[b][irim] = c->volc[b][irim] + dvrime * mwa / rhori;
    }
  }
  return status;
}

/* ====================================================================== */
/* distr_update: moving-centre re-binning of one cell,                    */
/* src/src_salsa/mo_salsa_update.f90:21-433.  Each category is swept from */
/* its largest bin down, repeatedly, until the LAST bin examined in a     */
/* sweep lies within its limits (within_bins is overwritten per bin, not  */
/* accumulated, :119-122), at most 100 sweeps.  The bins whose test is    */
/* skipped by CYCLE do not touch within_bins.                             */
/* ====================================================================== */
static int distr_update(const so_layout *L, so_cell *c, int level) {
  const int ndry = L->ndry, nwet = L->nwet, nnsp = L->nspec;
  int status = 0;
  /* ---- aerosols, :50-140 ---- */
  {
    const int A = L->iaero;
    int within = 0, count = 0;
    while (!within) {
      within = 1;
      for (int kk = L->fn2b - 1; kk >= L->in1a; --kk) {
        const int b = A + kk;
        if (c->numc[b] > nlim_c) {
          if (kk == L->fn2a) continue;
          double sd = 0.;
          for (int s = 0; s < ndry; ++s) sd += c->volc[b][s];
          double zvpart = sd / c->numc[b];
          if (L->vlolim[b] > zvpart && kk == L->in1a) continue;   /* smallest bin cannot decrease */
          if (L->vlolim[b] > zvpart) {                           /* decreasing: whole bin down */
            int mm = kk - 1;
            if (kk == L->in2b) mm = L->fn1a;                     /* 2b goes to 1a */
            const int m = A + mm;
            c->numc[m] = c->numc[m] + c->numc[b];
            c->numc[b] = 0.;
            for (int s = 0; s < nwet; ++s) {
              c->volc[m][s] = c->volc[m][s] + c->volc[b][s];
              c->volc[b][s] = 0.;
            }
            continue;
          }
          if (zvpart <= pi6 * cube(L->dmid[b])) continue;       /* not grown */
          const double zVrat = L->vhilim[b] / L->vlolim[b];
          const double zVilo = 2. * zvpart / (1. + zVrat);
          const double zVihi = zVrat * zVilo;
          const double zVexc = 0.5 * (zVihi + L->vhilim[b]);
          const double znfrac = fmin(1., (zVihi - L->vhilim[b]) / (zVihi - zVilo));
          const double zvfrac = fmin(0.99, znfrac * zVexc / zvpart);
          if (zvfrac < 0.) status |= SO_ERR_DISTR;
          const int m = b + 1;
          double add[SO_MAXSPEC];
          for (int s = 0; s < nwet; ++s) add[s] = znfrac * c->numc[b] * zVexc * c->volc[b][s] / sd;
          for (int s = 0; s < nwet; ++s) c->volc[m][s] = c->volc[m][s] + add[s];
          for (int s = 0; s < nwet; ++s) c->volc[b][s] = c->volc[b][s] - add[s];
          c->numc[m] = c->numc[m] + znfrac * c->numc[b];
          c->numc[b] = c->numc[b] * (1. - znfrac);
        }
        if (c->numc[b] > nlim_c) {
          double sd = 0.;
          for (int s = 0; s < ndry; ++s) sd += c->volc[b][s];
          const double zvpart = sd / c->numc[b];
          within = L->vlolim[b] < zvpart && zvpart < L->vhilim[b];
        }
      }
      if (++count > 100) { status |= SO_ERR_DISTR; break; }
    }
  }
  /* ---- cloud droplets (dry volume), :145-225; the category ends stay ---- */
  {
    const int C = L->icloud;
    int within = 0, count = 0;
    while (!within) {
      within = 1;
      for (int kk = L->ncld - 1; kk >= L->ica.cur; --kk) {
        const int b = C + kk;
        double sd = 0.;
        for (int s = 0; s < ndry; ++s) sd += c->volc[b][s];
        if (c->numc[b] > nlim_c && sd > 1.e-30) {
          const double zvpart = sd / c->numc[b];
          const double zVrat = L->vhilim[b] / L->vlolim[b];
          const double zVilo = 2. * zvpart / (1. + zVrat);
          const double zVihi = zVrat * zVilo;
          double zVexc, znfrac;
          int mm;
          if (zvpart < L->vlolim[b] && kk != L->ica.cur && kk != L->icb.cur) {
            zVexc = 0.5 * (zVilo + L->vlolim[b]);
            znfrac = fmin(1., (L->vlolim[b] - zVilo) / (zVihi - zVilo));
            mm = kk - 1;
          } else if (zvpart > L->vhilim[b] && kk != L->fca.cur && kk != L->fcb.cur) {
            zVexc = 0.5 * (zVihi + L->vhilim[b]);
            znfrac = fmin(1., (zVihi - L->vhilim[b]) / (zVihi - zVilo));
            mm = kk + 1;
          } else {
            continue;
          }
          const double zvfrac = fmin(0.99, znfrac * zVexc / zvpart);
          if (zvfrac < 0.) status |= SO_ERR_DISTR;
          const int m = C + mm;
          for (int s = 0; s < nwet; ++s) c->volc[m][s] = c->volc[m][s] + zvfrac * c->volc[b][s];
          for (int s = 0; s < nwet; ++s) c->volc[b][s] = c->volc[b][s] - zvfrac * c->volc[b][s];
          c->numc[m] = c->numc[m] + znfrac * c->numc[b];
          c->numc[b] = c->numc[b] - znfrac * c->numc[b];
        }
        sd = 0.;
        for (int s = 0; s < ndry; ++s) sd += c->volc[b][s];
        if (c->numc[b] > nlim_c && sd > 1.e-30) {
          const double zvpart = sd / c->numc[b];
          within = L->vlolim[b] < zvpart && zvpart < L->vhilim[b];
        }
      }
      if (++count > 100) { status |= SO_ERR_DISTR; break; }
    }
  }
  /* ---- precipitation (wet volume), :227-320 ---- */
  {
    const int R = L->iprecp;
    int within = 0, count = 0;
    while (!within) {
      within = 1;
      for (int kk = L->nprc - 1; kk >= 0; --kk) {
        const int b = R + kk;
        if (c->numc[b] > prlim_c) {
          double sw = 0.;
          for (int s = 0; s < nwet; ++s) sw += c->volc[b][s];
          const double zvpart = sw / c->numc[b];
          const double zVrat = L->vhilim[b] / L->vlolim[b];
          const double zVilo = 2. * zvpart / (1. + zVrat);
          const double zVihi = zVrat * zVilo;
          const double vmid = pi6 * cube(L->dmid[b]);
          const double zvdec = vmid - 0.2 * (vmid - L->vlolim[b]);
          double zVexc, znfrac;
          int mm;
          if (zvpart < zvdec && kk != 0) {
            zVexc = 0.5 * (zVilo + L->vlolim[b]);
            znfrac = fmin(1., (L->vlolim[b] - zVilo) / (zVihi - zVilo));
            mm = kk - 1;
          } else if (zvpart > vmid && kk != L->nprc - 1) {
            zVexc = 0.5 * (zVihi + L->vhilim[b]);
            znfrac = fmin(.99, (zVihi - L->vhilim[b]) / (zVihi - zVilo));
            mm = kk + 1;
          } else {
            continue;
          }
          const double zvfrac = fmin(0.99, znfrac * zVexc / zvpart);
          if (zvfrac < 0.) status |= SO_ERR_DISTR;
          const int m = R + mm;
          for (int s = 0; s < nwet; ++s) c->volc[m][s] = c->volc[m][s] + zvfrac * c->volc[b][s];
          for (int s = 0; s < nwet; ++s) c->volc[b][s] = c->volc[b][s] - zvfrac * c->volc[b][s];
          c->numc[m] = c->numc[m] + znfrac * c->numc[b];
          c->numc[b] = c->numc[b] - znfrac * c->numc[b];
        }
        if (c->numc[b] > prlim_c) {
          double sw = 0.;
          for (int s = 0; s < nwet; ++s) sw += c->volc[b][s];
          const double zvpart = sw / c->numc[b];
          within = L->vlolim[b] < zvpart && zvpart < L->vhilim[b];
        }
      }
      if (++count > 100) { status |= SO_ERR_DISTR; break; }
    }
  }
  if (level < 5) return status;
  /* ---- ice (all species incl. rime), :330-430 ---- */
  {
    const int I = L->iice;
    int within = 0, count = 0;
    while (!within) {
      within = 1;
      for (int kk = L->nice - 1; kk >= 0; --kk) {
        const int b = I + kk;
        if (c->numc[b] > L->nlim[b]) {
          double st = 0.;
          for (int s = 0; s < nnsp; ++s) st += c->volc[b][s];
          const double zvpart = st / c->numc[b];
          const double zVrat = L->vhilim[b] / L->vlolim[b];
          const double zVilo = 2. * zvpart / (1. + zVrat);
          const double zVihi = zVrat * zVilo;
          const double vmid = pi6 * cube(L->dmid[b]);
          const double zvdec = vmid - 0.2 * (vmid - L->vlolim[b]);
          double zVexc, znfrac;
          int mm;
          if (zvpart < zvdec && kk != 0) {
            zVexc = 0.5 * (zVilo + L->vlolim[b]);
            znfrac = fmin(1.0, (L->vlolim[b] - zVilo) / (zVihi - zVilo));
            mm = kk - 1;
          } else if (zvpart > vmid && kk != L->nice - 1) {
            zVexc = 0.5 * (zVihi + L->vhilim[b]);
            znfrac = fmin(0.99, (zVihi - L->vhilim[b]) / (zVihi - zVilo));
            mm = kk + 1;
          } else {
            continue;
          }
          const double zvfrac = fmin(0.99, znfrac * zVexc / zvpart);
          if (zvfrac < 0.) status |= SO_ERR_DISTR;
          const int m = I + mm;
          for (int s = 0; s < nnsp; ++s) c->volc[m][s] = c->volc[m][s] + zvfrac * c->volc[b][s];
          for (int s = 0; s < nnsp; ++s) c->volc[b][s] = c->volc[b][s] - zvfrac * c->volc[b][s];
          c->numc[m] = c->numc[m] + znfrac * c->numc[b];
          c->numc[b] = c->numc[b] - znfrac * c->numc[b];
        }
        double st = 0.;
        for (int s = 0; s < nnsp; ++s) st += c->volc[b][s];
        if (c->numc[b] > L->nlim[b] && st > 1.e-30) {
          const double zvpart = st / c->numc[b];
          within = L->vlolim[b] < zvpart && zvpart < L->vhilim[b];
        }
      }
      if (++count > 100) { status |= SO_ERR_DISTR; break; }
    }
  }
  return status;
}

/* ====================================================================== */
/* autoconv2: cloud drops beyond 50 um into the first precipitation bin,  */
/* src/src_salsa/mo_salsa_cloud.f90:1401-1477.  Each cloud bin, largest   */
/* first, is read as a lognormal of width sigmag = 1.2 around its mean    */
/* volume; the number and water volume above zd0 = 50 um move to precip   */
/* bin 1 (the dry species with the number fraction).  cumlognorm:         */
/* src/src_shared/math_functions.f90:64-76.  The loop over bins stops     */
/* once the converted number exceeds 1e10 m^-3 s^-1 * dt (:1459-1462).    */
/* ====================================================================== */
static double cumlognorm(double dg, double sigmag, double dpart) {
  const double hlp1 = log(dpart) - log(dg);
  const double hlp2 = sqrt(2.) * log(sigmag);
  return 0.5 + 0.5 * erf(hlp1 / hlp2);
}

static void autoconv2(const so_layout *L, so_cell *c, double ptstep) {
  const double zd0 = 50.e-6, sigmag = 1.2, max_rate_autoc = 1.0e10;
  const int nwet = L->nwet, ndry = L->ndry, iwa = L->iwa;
  const int P1 = L->iprecp;
  const double ls = log(sigmag);
  double tot = 0.;
  for (int cc = L->ncld - 1; cc >= 0; --cc) {
    const int b = L->icloud + cc;
    const double Ntot = c->numc[b];
    double Vtot = 0.;
    for (int s = 0; s < nwet; ++s) Vtot += c->volc[b][s];
    if (Ntot > L->nlim[b] && Vtot > 0.) {
      const double dvg = pow(Vtot / Ntot / pi6, 1. / 3.) * exp((3. * (ls * ls)) / 2.);
      const double dg = dvg * exp(-(3. * (ls * ls)));
      const double Vrem = Vtot * (1. - cumlognorm(dvg, sigmag, zd0));
      const double Nrem = Ntot * (1. - cumlognorm(dg, sigmag, zd0));
      if (Vrem > 0. && Nrem > L->nlim[P1]) {
        for (int s = 0; s < ndry; ++s) {
          c->volc[P1][s] = c->volc[P1][s] + c->volc[b][s] * (Nrem / Ntot);
          c->volc[b][s] = c->volc[b][s] * (1. - (Nrem / Ntot));
        }
        c->volc[P1][iwa] = c->volc[P1][iwa] + c->volc[b][iwa] * (Vrem / Vtot);
        c->volc[b][iwa] = c->volc[b][iwa] * (1. - (Vrem / Vtot));
        c->numc[P1] = c->numc[P1] + Nrem;
        c->numc[b] = c->numc[b] - Nrem;
        tot = tot + Nrem;
        if (tot > max_rate_autoc * ptstep) break;
      }
    }
  }
}

/* ====================================================================== */
/* Interstitial cloud activation (cloud_activation with lsactiv%mode 1,   */
/* src/src_salsa/mo_salsa_cloud.f90:46-106, actInterst :859-1195).  When  */
/* the cell is supersaturated, each 2a/2b aerosol bin is given piecewise  */
/* linear profiles of wet volume and of number density over its dry       */
/* volume, from its own and its neighbours' mean values; the part of the  */
/* bin whose wet volume exceeds the critical volume of the cell's         */
/* supersaturation moves to the parallel cloud bin.  intgN / intgV:       */
/* :1381-1397.  The chargeTime and INdef bookkeeping (:1161-1187) is not  */
/* carried: the kernels assume chargeTime = chargeTimeMax, and ice        */
/* nucleation is not on the path.                                         */
/* ====================================================================== */
static double intgN(double ikk, double icc, double ilow, double ihigh) {
  return 0.5 * ikk * fmax(ihigh * ihigh - ilow * ilow, 0.) + icc * fmax(ihigh - ilow, 0.);
}
static double intgV(double ikk, double icc, double ilow, double ihigh) {
  return (1. / 3.) * ikk * fmax(ihigh * ihigh * ihigh - ilow * ilow * ilow, 0.) +
         0.5 * icc * fmax(ihigh * ihigh - ilow * ilow, 0.);
}

static void act_interst(const so_layout *L, so_cell *c, double prv, double prs, double temp) {
  const double THvol = 1.e-28;
  const int ndry = L->ndry, nwet = L->nwet, iwa = L->iwa;
  if (prv / prs <= 1.000) return;
  const double zkelvin = 4. * mwa * surfw0 / (rg * rhowa * temp);
  const double zdcstar = 2. * zkelvin / (3. * ((prv / prs) - 1.));
  const double zvcstar = pi6 * (zdcstar * zdcstar * zdcstar);
  const int A = L->iaero, C = L->icloud;
  double act_n[SO_MAXBIN], act_v[SO_MAXBIN][SO_MAXSPEC];
  for (int cb = L->ica.cur; cb <= L->fcb.cur; ++cb) {
    const int ab = cb <= L->fca.cur ? L->ica.par + (cb - L->ica.cur) : L->icb.par + (cb - L->icb.cur);
    act_n[cb] = 0.;
    for (int s = 0; s < SO_MAXSPEC; ++s) act_v[cb][s] = 0.;
    const int a = A + ab;
    double vol_sol = 0., vol_insol = 0.;
    for (int s = 0; s < ndry; ++s) {
      const int id = L->spec_id[s];
      if (id == 2 || id == 3) vol_insol += c->volc[a][s];       /* BC, DU (classSpecies.f90:99-100) */
      else vol_sol += c->volc[a][s];
    }
    (void)vol_insol;
    if (c->numc[a] < L->nlim[a]) continue;
    if (vol_sol < c->numc[a] * THvol) continue;
    if (c->volc[a][iwa] < c->numc[a] * THvol) continue;
    const double vrlo = L->vlolim[a] / (pi6 * cube(L->dmid[a]));   /* mo_salsa_init.f90:186-227 */
    const double vrhi = L->vhilim[a] / (pi6 * cube(L->dmid[a]));
    const double Nmid = c->numc[a];
    double sw = 0., sd = 0.;
    for (int s = 0; s < nwet; ++s) sw += c->volc[a][s];
    for (int s = 0; s < ndry; ++s) sd += c->volc[a][s];
    const double Vwmid = sw / Nmid, Vmid = sd / Nmid;
    const double Vlo = Vmid * vrlo, Vhi = Vmid * vrhi;
    double Nim1, Vim1, Vlom1, Vhim1, Vwim1, Nip1, Vip1, Vlop1, Vhip1, Vwip1;
    if (ab == L->in1a || ab == L->in2b) {                 /* first bin of a regime, :982-988 */
      Nim1 = L->nlim[a]; Vim1 = Vlo / 2.; Vlom1 = 0.; Vhim1 = Vlo; Vwim1 = Vim1;
    } else {
      const int m = a - 1;
      Nim1 = c->numc[m];
      if (Nim1 > L->nlim[m]) {
        double d = 0., w = 0.;
        for (int s = 0; s < ndry; ++s) d += c->volc[m][s];
        for (int s = 0; s < nwet; ++s) w += c->volc[m][s];
        Vim1 = d / Nim1; Vwim1 = w / Nim1;
      } else {
        Vim1 = pi6 * cube(L->dmid[m]); Vwim1 = pi6 * cube(L->dmid[m]);
      }
      Vlom1 = Vim1 * (L->vlolim[m] / (pi6 * cube(L->dmid[m])));
      Vhim1 = Vim1 * (L->vhilim[m] / (pi6 * cube(L->dmid[m])));
    }
    if (ab == L->fn2a || ab == L->fn2b) {                 /* last bin of a regime, :1001-1007 */
      Nip1 = L->nlim[a]; Vip1 = Vhi + 0.5 * (Vhi - Vlo); Vlop1 = Vhi; Vhip1 = Vhi + (Vhi - Vlo); Vwip1 = Vip1;
    } else {
      const int m = a + 1;
      Nip1 = c->numc[m];
      if (Nip1 > L->nlim[m]) {
        double d = 0., w = 0.;
        for (int s = 0; s < ndry; ++s) d += c->volc[m][s];
        for (int s = 0; s < nwet; ++s) w += c->volc[m][s];
        Vip1 = d / Nip1; Vwip1 = w / Nip1;
      } else {
        Vip1 = pi6 * cube(L->dmid[m]); Vwip1 = pi6 * cube(L->dmid[m]);
      }
      Vlop1 = Vip1 * (L->vlolim[m] / (pi6 * cube(L->dmid[m])));
      Vhip1 = Vip1 * (L->vhilim[m] / (pi6 * cube(L->dmid[m])));
    }
    Vip1 = fmax(Vhi, Vip1);
    Vim1 = fmin(Vlo, Vim1);
    /* wet volume profile, :1027-1037 */
    double zs1 = (Vwmid - fmax(Vwim1, 0.)) / (Vmid - Vim1);
    double zs2 = (fmax(Vwip1, 0.) - Vwmid) / (Vip1 - Vmid);
    const double V01 = Vwmid - zs1 * Vmid, V02 = Vwmid - zs2 * Vmid;
    const double Vwlo = fmax(V01 + zs1 * Vlo, Vlo), Vwhi = fmax(V02 + zs2 * Vhi, Vhi);
    /* integration limits, :1050-1097 */
    double vint1, vint2;
    int ir[4];
    if (zvcstar < Vwlo && zvcstar < Vwmid && zvcstar < Vwhi) {
      vint1 = Vlo; vint2 = Vhi; ir[0] = ir[1] = ir[2] = ir[3] = 1;
    } else if (zvcstar > Vwlo && zvcstar > Vwmid && zvcstar > Vwhi) {
      vint1 = 999.; vint2 = 999.; ir[0] = ir[1] = ir[2] = ir[3] = 0;
    } else {
      double vcut = (zvcstar - V01) / zs1;
      vint1 = (vcut < Vlo || vcut > Vmid) ? Vlo : vcut;
      vcut = (zvcstar - V02) / zs2;
      vint2 = (vcut < Vmid || vcut > Vhi) ? Vmid : vcut;
      ir[0] = Vwlo > zvcstar; ir[1] = Vwmid > zvcstar; ir[2] = Vwmid > zvcstar; ir[3] = Vwhi > zvcstar;
    }
    /* number density profile and the integrals, :1099-1139 */
    const double dNim1 = Nim1 / (Vhim1 - Vlom1), dNip1 = Nip1 / (Vhip1 - Vlop1), dNmid = Nmid / (Vhi - Vlo);
    zs1 = (dNmid - dNim1) / (Vmid - Vim1);
    zs2 = (dNip1 - dNmid) / (Vip1 - Vmid);
    const double N01 = dNmid - zs1 * Vmid, N02 = dNmid - zs2 * Vmid;
    const double Nnorm = intgN(zs1, N01, Vlo, Vmid) + intgN(zs2, N02, Vmid, Vhi);
    const double Vnorm = intgV(zs1, N01, Vlo, Vmid) + intgV(zs2, N02, Vmid, Vhi);
    double zactvol = 0., Nact = 0.;
    if (ir[0]) {
      Nact = Nact + (Nmid / Nnorm) * intgN(zs1, N01, Vlo, vint1);
      zactvol = zactvol + (Nmid * Vmid / Vnorm) * intgV(zs1, N01, Vlo, vint1);
    }
    if (ir[1]) {
      Nact = Nact + (Nmid / Nnorm) * intgN(zs1, N01, vint1, Vmid);
      zactvol = zactvol + (Nmid * Vmid / Vnorm) * intgV(zs1, N01, vint1, Vmid);
    }
    if (ir[2]) {
      Nact = Nact + (Nmid / Nnorm) * intgN(zs2, N02, Vmid, vint2);
      zactvol = zactvol + (Nmid * Vmid / Vnorm) * intgV(zs2, N02, Vmid, vint2);
    }
    if (ir[3]) {
      Nact = Nact + (Nmid / Nnorm) * intgN(zs2, N02, vint2, Vhi);
      zactvol = zactvol + (Nmid * Vmid / Vnorm) * intgV(zs2, N02, vint2, Vhi);
    }
    act_n[cb] = fmin(Nact, Nmid);
    for (int s = 0; s < nwet; ++s) act_v[cb][s] = fmin(zactvol * (c->volc[a][s] / (Vmid * Nmid)), c->volc[a][s]);
    if (c->rates) {                                       /* rateDiag%Activation, :1143-1144 */
      c->rates[SO_R_ACTIV].numc += act_n[cb];
      for (int s = 0; s < nwet; ++s) c->rates[SO_R_ACTIV].volc[s] += act_v[cb][s];
    }
  }
  /* apply, :1150-1189 (cloud bin k and aerosol bin ica%par + k) */
  for (int k = 0; k <= L->fcb.cur - L->ica.cur; ++k) {
    const int a = A + L->ica.par + k, cl = C + L->ica.cur + k, cb = L->ica.cur + k;
    c->numc[a] = fmax(0., c->numc[a] - act_n[cb]);
    c->numc[cl] = c->numc[cl] + act_n[cb];
    for (int s = 0; s < nwet; ++s) {
      c->volc[a][s] = fmax(0., c->volc[a][s] - act_v[cb][s]);
      c->volc[cl][s] = c->volc[cl][s] + act_v[cb][s];
    }
  }
}

/* ====================================================================== */
/* Cloud-base activation (cloud_activation with lsactiv%mode 2,           */
/* mo_salsa_cloud.f90:46-106; ActCloudBase :161-375 after Abdul-Razzak and */
/* Ghan 2002, getSolute :113-157, activate3 :1199-1378): for an updraft w */
/* > 0 the maximum supersaturation of the rising parcel, the fraction of  */
/* each aerosol bin above its critical supersaturation, the critical dry  */
/* volume in the first partly activated 1a/2a bin, and per cloud bin the  */
/* number and volume of its parallel aerosol bin above that volume from   */
/* piecewise linear number-density profiles.  The state is not changed:   */
/* the result is the activated droplets (pactd) that run_SALSA hands to   */
/* the LES as pa_nactd / pa_vactd (mo_salsa_driver.f90:360-366).          */
/* Per-bin quantities are formed where they are used (getSolute's moles,  */
/* the critical diameters) instead of being kept in arrays.               */
/* ====================================================================== */
/* moles of solute per particle of aerosol bin b (getSolute :136-151) */
static double cb_solute(const so_layout *L, const so_cell *c, int b) {
  const int a = L->iaero + b;
  if (!(c->numc[a] > L->nlim[a])) return 0.;
  double ns = 0.;
  for (int nn = 0; nn < L->ndry; ++nn) ns = ns + L->diss[nn] * c->volc[a][nn] * L->rholiq[nn] / L->mm[nn];
  return ns / c->numc[a];
}

/* ActCloudBase's per-bin test (:327-351): whether aerosol bin b is counted
   (number at the limit and s_max above its upper-edge critical
   supersaturation), with its activated fraction and critical diameters */
static int cb_bin(const so_layout *L, const so_cell *c, int b, double x, double aa, double bb, double s_max,
                  double *frac, double *dcrit, double *dcrlo, double *dcrhi) {
  const int a = L->iaero + b;
  *frac = 0.; *dcrit = 0.; *dcrlo = 0.; *dcrhi = 0.;
  if (c->numc[a] < L->nlim[a]) return 0;
  const double ns = cb_solute(L, c, b);
  const double dm3 = L->dmid[a] * L->dmid[a] * L->dmid[a];
  const double nshi = ns * (L->vhilim[a] / (pi6 * dm3));     /* vratiohi, mo_salsa_init.f90:186 */
  const double sil = exp(sqrt(x / nshi)) - 1.;
  if (s_max < sil) return 0;
  const double nslo = ns * (L->vlolim[a] / (pi6 * dm3));     /* vratiolo, :187 */
  const double siu = exp(sqrt(x / nslo)) - 1.;
  *frac = fmin(1., log(s_max / sil) / log(siu / sil));
  *dcrlo = 2. * sqrt(3. * nslo * bb / aa);
  *dcrhi = 2. * sqrt(3. * nshi * bb / aa);
  *dcrit = 2. * sqrt(3. * ns * bb / aa);
  return 1;
}

/* nact[cb] (#/m3) and vact[cb][s] (m3/m3) of cloud bin cb; zero unless w > 0 */
static void act_cloud_base(const so_layout *L, const so_cell *c, double temp, double pres, double w, double *nact,
                           double (*vact)[SO_MAXSPEC]) {
  const int A = L->iaero, ndry = L->ndry, nwet = L->nwet, iwa = L->iwa;
  for (int cb = 0; cb < L->ncld; ++cb) {
    nact[cb] = 0.;
    for (int s = 0; s < SO_MAXSPEC; ++s) vact[cb][s] = 0.;
  }
  const double bb = 6. * mwa / (pi * rhowa);
  if (w <= 0.) return;
  const double aa = 4. * mwa * surfw0 / (rg * rhowa * temp);
  const double x = 4. * (aa * aa * aa) / (27. * bb);
  double ntot = 0., sum1 = 0.;
  for (int b = L->in1a; b <= L->fn2b; ++b) ntot = ntot + c->numc[A + b];
  for (int b = L->in1a; b <= L->fn2b; ++b) {
    const double scrit = exp(sqrt(x / fmax(eps_r8, cb_solute(L, c, b)))) - 1.;
    sum1 = sum1 + c->numc[A + b] / pow(scrit, 2. / 3.);
  }
  if (ntot < L->nlim[A]) return;
  const double V = w;
  const double Lh = 2.501e6 - 2370. * (temp - 273.15);
  const double a1 = 1. - (373.15 / temp);
  const double a12 = a1 * a1;
  const double ps = 101325. * exp(13.3185 * a1 - 1.976 * a12 - 0.6445 * (a12 * a1) - 0.1299 * (a12 * a12));
  const double alpha = grav * mwa * Lh / (cpa * rg * (temp * temp)) - grav * mair / (rg * temp);
  const double gamma = rg * temp / (ps * mwa) + mwa * (Lh * Lh) / (cpa * pres * mair * temp);
  const double x1 = pres / 101325.;
  const double dv1 = 1.e-4 * (0.211 / x1) * pow(temp / 273., 1.94);
  const double ka1 = 1.e-3 * (4.39 + 0.071 * temp);
  const double Gc = 1. / (rhowa * rg * temp / (ps * dv1 * mwa) + Lh * rhowa / (ka1 * temp) * (Lh * mwa / (temp * rg) - 1.));
  const double r = ntot / sum1;
  const double s_eff = sqrt(r * r * r);
  const double avg = alpha * V / Gc;
  const double theta = (sqrt(avg * avg * avg)) / (2. * pi * rhowa * gamma * ntot);
  const double khi = (2. / 3.) * aa * sqrt(avg);
  const double kt = khi / theta;
  const double s_max = s_eff / sqrt(0.5 * sqrt(kt * kt * kt) + pow((s_eff * s_eff) / (theta + 3. * khi), 0.75));
  const double zdcstar = 2. * aa / (3. * s_max);
  /* the first partly activated 1a/2a bin (:355-360), and whether any bin is counted */
  int bcrita = L->fn2a, any = 0;
  double f, d0, d1, d2;
  for (int b = L->in1a; b <= L->fn2b; ++b)
    if (cb_bin(L, c, b, x, aa, bb, s_max, &f, &d0, &d1, &d2) && d0 >= eps_r8) any = 1;
  for (int b = L->in1a; b <= L->fn2a; ++b) {
    cb_bin(L, c, b, x, aa, bb, s_max, &f, &d0, &d1, &d2);
    if (f < 1. && f > 0.) { bcrita = b; break; }
  }
  /* activate3 (:1239-1377) */
  if (!any) return;                              /* ALL(pdcrit < epsilon) */
  double zvcstar = 0.;
  {
    const int a = A + bcrita;
    double dcrit, dcrlo, dcrhi;
    cb_bin(L, c, bcrita, x, aa, bb, s_max, &f, &dcrit, &dcrlo, &dcrhi);
    double Vmid;
    if (c->numc[a] < L->nlim[a]) {
      Vmid = pi6 * (L->dmid[a] * L->dmid[a] * L->dmid[a]);
    } else {
      double sd = 0.;
      for (int s = 0; s < ndry; ++s) sd = sd + c->volc[a][s];
      Vmid = sd / fmax(L->nlim[a], c->numc[a]);
    }
    const double Vhi = L->vhilim[a], Vlo = L->vlolim[a];
    const double cs3 = zdcstar * zdcstar * zdcstar, cr3 = dcrit * dcrit * dcrit;
    if (zdcstar >= dcrit) {
      const double zhlp = (pi6 * cs3 - pi6 * cr3) / fmax(eps_r8, pi6 * (dcrhi * dcrhi * dcrhi) - pi6 * cr3);
      zvcstar = Vmid + zhlp * (Vhi - Vmid);
    } else if (zdcstar < dcrit) {
      const double zhlp = (pi6 * cr3 - pi6 * cs3) / fmax(eps_r8, pi6 * cr3 - pi6 * (dcrlo * dcrlo * dcrlo));
      zvcstar = Vmid - zhlp * (Vmid - Vlo);
    }
    zvcstar = fmax(zvcstar, L->vlolim[a]);
    zvcstar = fmin(zvcstar, L->vhilim[a]);
  }
  for (int cb = L->ica.cur; cb <= L->fcb.cur; ++cb) {
    const int ab = cb <= L->fca.cur ? L->ica.par + (cb - L->ica.cur) : L->icb.par + (cb - L->icb.cur);
    const int a = A + ab;
    if (c->numc[a] < L->nlim[a]) continue;
    const double Nmid = fmax(c->numc[a], L->nlim[a]);
    double sd = 0.;
    for (int s = 0; s < ndry; ++s) sd = sd + c->volc[a][s];
    const double Vmid = sd / Nmid;
    const double Vlo = L->vlolim[a], Vhi = L->vhilim[a];
    double Nim1, Vim1, Vlom1, Vhim1, Nip1, Vip1, Vlop1, Vhip1;
    if (ab == L->in1a || ab == L->in2b) {
      Nim1 = L->nlim[a]; Vim1 = Vlo / 2.; Vlom1 = 0.; Vhim1 = Vlo;
    } else {
      Nim1 = fmax(c->numc[a - 1], L->nlim[a - 1]);
      if (Nim1 > L->nlim[a - 1]) {
        double s1 = 0.;
        for (int s = 0; s < ndry; ++s) s1 = s1 + c->volc[a - 1][s];
        Vim1 = s1 / Nim1;
      } else {
        Vim1 = pi6 * (L->dmid[a - 1] * L->dmid[a - 1] * L->dmid[a - 1]);
      }
      Vlom1 = L->vlolim[a - 1]; Vhim1 = L->vhilim[a - 1];
    }
    if (ab == L->fn2a || ab == L->fn2b) {
      Nip1 = L->nlim[a]; Vip1 = Vhi + 0.5 * (Vhi - Vlo); Vlop1 = Vhi; Vhip1 = Vhi + (Vhi - Vlo);
    } else {
      Nip1 = fmax(c->numc[a + 1], L->nlim[a + 1]);
      if (Nip1 > L->nlim[a + 1]) {
        double s1 = 0.;
        for (int s = 0; s < ndry; ++s) s1 = s1 + c->volc[a + 1][s];
        Vip1 = s1 / Nip1;
      } else {
        Vip1 = pi6 * (L->dmid[a + 1] * L->dmid[a + 1] * L->dmid[a + 1]);
      }
      Vlop1 = L->vlolim[a + 1]; Vhip1 = L->vhilim[a + 1];
    }
    Vip1 = fmax(Vlop1, fmin(Vip1, Vhip1));
    Vim1 = fmax(Vlom1, fmin(Vim1, Vhim1));
    const double dNim1 = Nim1 / (Vhim1 - Vlom1), dNip1 = Nip1 / (Vhip1 - Vlop1), dNmid = Nmid / (Vhi - Vlo);
    const double zs1 = (dNmid - dNim1) / (Vmid - Vim1), zs2 = (dNip1 - dNmid) / (Vip1 - Vmid);
    const double N01 = dNmid - zs1 * Vmid, N02 = dNip1 - zs2 * Vip1;
    const double Nnorm = intgN(zs1, N01, Vlo, Vmid) + intgN(zs2, N02, Vmid, Vhi);
    const double Vnorm = intgV(zs1, N01, Vlo, Vmid) + intgV(zs2, N02, Vmid, Vhi);
    double n = 0., zactvol = 0.;
    if (zvcstar < Vmid) {
      const double zvcint = fmax(zvcstar, Vlo);
      n = (Nmid / Nnorm) * (intgN(zs1, N01, zvcint, Vmid) + intgN(zs2, N02, Vmid, Vhi));
      zactvol = (Nmid * Vmid / Vnorm) * (intgV(zs1, N01, zvcint, Vmid) + intgV(zs2, N02, Vmid, Vhi));
    } else if (zvcstar >= Vmid) {
      const double zvcint = fmin(zvcstar, Vhi);
      n = (Nmid / Nnorm) * (intgN(zs2, N02, zvcint, Vhi));
      zactvol = (Nmid * Vmid / Vnorm) * (intgV(zs2, N02, zvcint, Vhi));
    }
    if (zvcstar < Vmid || zvcstar >= Vmid)
      for (int s = 0; s < nwet; ++s) vact[cb][s] = zactvol * (c->volc[a][s] / (Vmid * Nmid));
    nact[cb] = fmax(0., n);
    for (int s = 0; s < nwet; ++s) vact[cb][s] = fmax(0., vact[cb][s]);
    double dcrit;
    cb_bin(L, c, ab, x, aa, bb, s_max, &f, &dcrit, &d1, &d2);
    const double q = 3.e-6 / fmax(eps_r8, dcrit);
    vact[cb][iwa] = nact[cb] * pi6 * (dcrit * dcrit * dcrit) * fmin(2., q * q);
  }
}

/* the activated droplets of one cell in SALSA units (a test entry) */
void so_act_cloud_base(const so_layout *L, double temp, double pres, double w, const double *numc, const double *volc,
                       double *nact, double *vact) {
  so_cell c;
  memset(&c, 0, sizeof(c));
  for (int b = 0; b < L->ntotal; ++b) {
    c.numc[b] = numc[b];
    for (int s = 0; s < SO_MAXSPEC; ++s) c.volc[b][s] = volc[b * SO_MAXSPEC + s];
  }
  act_cloud_base(L, &c, temp, pres, w, nact, (double (*)[SO_MAXSPEC])vact);
}

/* ====================================================================== */
/* One cell: salsa() with coagulation and condensation enabled,           */
/* src/src_salsa/mo_salsa.f90:92-107                                      */
/* ====================================================================== */
static int step_cell(const so_layout *L, const so_params *P, double ptemp, double ppres, double prs, double prsi,
                     double *prv, double ptstep, double *numc, double *volc, so_rate *rates, double *kstore,
                     int kmode, double w, double *nact, double *vact);

int so_step_cell_ks(const so_layout *L, const so_params *P,
                    double ptemp, double ppres, double prs, double prsi,
                    double *prv, double ptstep,
                    double *numc, double *volc, so_rate *rates, double *kstore, int kmode) {
  return step_cell(L, P, ptemp, ppres, prs, prsi, prv, ptstep, numc, volc, rates, kstore, kmode, 0., NULL, NULL);
}

/* A step with cloud-base activation (lsactiv%mode 2): w is the cell's
   updraft; nact[ncld] and vact[ncld][SO_MAXSPEC] receive the activated
   droplets in SALSA units (zero with another lsactiv). */
int so_step_cell_act(const so_layout *L, const so_params *P, double ptemp, double ppres, double prs, double prsi,
                     double *prv, double ptstep, double *numc, double *volc, so_rate *rates, double w, double *nact,
                     double *vact) {
  return step_cell(L, P, ptemp, ppres, prs, prsi, prv, ptstep, numc, volc, rates, 0, 0, w, nact, vact);
}

static int step_cell(const so_layout *L, const so_params *P, double ptemp, double ppres, double prs, double prsi,
                     double *prv, double ptstep, double *numc, double *volc, so_rate *rates, double *kstore,
                     int kmode, double w, double *nact, double *vact) {
  so_cell c;
  memset(&c, 0, sizeof(c));
  c.rates = rates;
  if (rates) memset(rates, 0, sizeof(so_rate) * SO_NRATE);
  for (int b = 0; b < L->ntotal; ++b) {
    c.numc[b] = numc[b];
    for (int s = 0; s < SO_MAXSPEC; ++s) c.volc[b][s] = volc[b * SO_MAXSPEC + s];
    c.dwet[b] = L->dmid[b];
    update_rhomean(L, &c, b);   /* mo_salsa_driver.f90:272 */
  }
  int status = 0;
  if (P->init) {
    /* run_SALSA initialization call, mo_salsa_driver.f90:282-283 with
     * init_rh = 0.3 (:39).  The reference's initial guess is the Section's
     * dwet as left by the cell processed before (updateDiameter is commented
     * out at :270); the cell's own updateDiameter value stands in for it. */
    for (int b = 0; b < L->ntotal; ++b) update_diameter(L, &c, b);
    status |= equilibration(L, &c, 0.3, ptemp, 1);
  }
  if (P->lscoag) status |= coagulation(L, P, &c, ptstep, ptemp, ppres, kstore, kstore ? kmode : 0);
  else for (int b = 0; b < L->ntotal; ++b) update_diameter(L, &c, b);
  if (P->lscnd && (P->lscndh2ocl || P->lscndh2oae || P->lscndh2oic))
    status |= gpparth2o(L, P, &c, ptemp, ppres, prs, prsi, prv, ptstep);
  if (P->lsauto_state && P->lsauto_mode == 2) autoconv2(L, &c, ptstep);   /* mo_salsa.f90:110-112 */
  if (P->lsactiv == 1) act_interst(L, &c, *prv, prs, ptemp);                /* :115-119, mode 1 */
  if (nact && vact) {                       /* cloud_activation resets pactd; mode 2 fills it */
    if (P->lsactiv == 2) {
      act_cloud_base(L, &c, ptemp, ppres, w, nact, (double (*)[SO_MAXSPEC])vact);
    } else {
      for (int i = 0; i < L->ncld * SO_MAXSPEC; ++i) vact[i] = 0.;
      for (int i = 0; i < L->ncld; ++i) nact[i] = 0.;
    }
  }
  if (P->lsdistupdate) status |= distr_update(L, &c, P->level);
  for (int b = 0; b < L->ntotal; ++b) {
    numc[b] = c.numc[b];
    for (int s = 0; s < SO_MAXSPEC; ++s) volc[b * SO_MAXSPEC + s] = c.volc[b][s];
  }
  return status;
}

int so_step_cell(const so_layout *L, const so_params *P,
                 double ptemp, double ppres, double prs, double prsi,
                 double *prv, double ptstep,
                 double *numc, double *volc, so_rate *rates) {
  return so_step_cell_ks(L, P, ptemp, ppres, prs, prsi, prv, ptstep, numc, volc, rates, 0, 0);
}

/* ---- thermodynamics used to build inputs, src/src_LES/thrm.f90:345-387 ---- */
double so_rslf(double p, double t) {
  const double c0 = 0.6105851e+03, c1 = 0.4440316e+02, c2 = 0.1430341e+01,
               c3 = 0.2641412e-01, c4 = 0.2995057e-03, c5 = 0.2031998e-05,
               c6 = 0.6936113e-08, c7 = 0.2564861e-11, c8 = -.3704404e-13;
  double x = fmin(fmax(-80., t - 273.16), 50.);
  double esl = c0 + x * (c1 + x * (c2 + x * (c3 + x * (c4 + x * (c5 + x * (c6 + x * (c7 + x * c8)))))));
  return .622 * esl / (p - esl);
}

double so_rsif(double p, double t) {
  const double c0 = 0.6114327e+03, c1 = 0.5027041e+02, c2 = 0.1875982e+01,
               c3 = 0.4158303e-01, c4 = 0.5992408e-03, c5 = 0.5743775e-05,
               c6 = 0.3566847e-07, c7 = 0.1306802e-09, c8 = 0.2152144e-12;
  double x = fmin(fmax(-80., t - 273.16), 0.);
  double esi = c0 + x * (c1 + x * (c2 + x * (c3 + x * (c4 + x * (c5 + x * (c6 + x * (c7 + x * c8)))))));
  return .622 * esi / (p - esi);
}

/* ---- driver pack / unpack, mo_salsa_driver.f90:230-279 and :310-357 ---- */
static int cat_of(const so_layout *L, int b, int *loc, int *nb) {
  if (b < L->icloud) { *loc = b - L->iaero; *nb = L->nbins; return 0; }
  if (b < L->iprecp) { *loc = b - L->icloud; *nb = L->ncld; return 1; }
  if (b < L->iice) { *loc = b - L->iprecp; *nb = L->nprc; return 2; }
  *loc = b - L->iice; *nb = L->nice; return 3;
}

void so_pack_cell(const so_layout *L, int level, double dn,
                  const double *n_cat[4], const double *m_cat[4],
                  double *numc, double *volc) {
  memset(numc, 0, sizeof(double) * L->ntotal);
  memset(volc, 0, sizeof(double) * L->ntotal * SO_MAXSPEC);
  for (int b = 0; b < L->ntotal; ++b) {
    int loc, nb, cat = cat_of(L, b, &loc, &nb);
    if (cat == 3 && level < 5) break;
    if (cat < 3) {
      for (int nc = 0; nc < L->nwet; ++nc)
        volc[b * SO_MAXSPEC + nc] = m_cat[cat][nc * nb + loc] * dn / L->rholiq[nc];
    } else {
      for (int nc = 0; nc < L->nwet; ++nc)
        volc[b * SO_MAXSPEC + nc] = m_cat[cat][nc * nb + loc] * dn / L->rhoice[nc];
      volc[b * SO_MAXSPEC + L->irim] = m_cat[cat][L->irim * nb + loc] * dn / rhori;
    }
    numc[b] = n_cat[cat][loc] * dn;
  }
}

void so_unpack_tend(const so_layout *L, int level, double dn, double tstep,
                    const double *numc, const double *volc,
                    const double *n_cat[4], const double *m_cat[4],
                    double *nt_cat[4], double *mt_cat[4]) {
  for (int b = 0; b < L->ntotal; ++b) {
    int loc, nb, cat = cat_of(L, b, &loc, &nb);
    if (cat == 3 && level < 5) break;
    if (cat < 3) {
      for (int nc = 0; nc < L->nwet; ++nc) {
        int str = nc * nb + loc;
        mt_cat[cat][str] = mt_cat[cat][str] +
            ((volc[b * SO_MAXSPEC + nc] * L->rholiq[nc] / dn) - m_cat[cat][str]) / tstep;
      }
    } else {
      for (int nc = 0; nc < L->nwet; ++nc) {
        int str = nc * nb + loc;
        mt_cat[cat][str] = mt_cat[cat][str] +
            ((volc[b * SO_MAXSPEC + nc] * L->rhoice[nc] / dn) - m_cat[cat][str]) / tstep;
      }
      int str = L->irim * nb + loc;
      mt_cat[cat][str] = mt_cat[cat][str] +
          ((volc[b * SO_MAXSPEC + L->irim] * rhori / dn) - m_cat[cat][str]) / tstep;
    }
    nt_cat[cat][loc] = nt_cat[cat][loc] + ((numc[b] / dn) - n_cat[cat][loc]) / tstep;
  }
}

/* Batch driver for timing: n cells, each stepped nsteps times (SALSA units). */
int so_step_batch(const so_layout *L, const so_params *P, long n, const double *T, const double *p,
                  const double *rs, const double *rsi, double *rv, double ptstep, int nsteps,
                  double *numc, double *volc) {
  int st = 0;
  for (long i = 0; i < n; ++i)
    for (int k = 0; k < nsteps; ++k)
      st |= so_step_cell(L, P, T[i], p[i], rs[i], rsi[i], &rv[i], ptstep,
                         numc + i * L->ntotal, volc + i * L->ntotal * SO_MAXSPEC, 0);
  return st;
}

/* ====================================================================== */
/* LES coupling after run_SALSA for the SALSA tracers of one process that  */
/* owns the whole (nzp,nxp,nyp) grid (halo 1 in z, 2 laterally):          */
/*   tend_constrain2   constrain_SALSA.f90:18-36                           */
/*   update_sclrs      step.f90:450-489 (update; sclrset 'mixd',           */
/*                     util.f90:71-107, with periodic lateral halos)       */
/*   tend0             step.f90:379-396                                    */
/*   SALSA_diagnostics constrain_SALSA.f90:50-402, findDry4Wet :408-470,   */
/*                     calc_correlation util.f90:1005-1024                 */
/* in the order of cfll (step.f90:330-333).  Arrays are the LES ones:      */
/* n(c)[b][n3], m(c)[s*nb+b][n3], n3 index k + nzp*(i + nxp*j).            */
/* ====================================================================== */
static double les_corr(const double *x, const double *y, int n) {
  /* calc_correlation: sxy keeps the last product only (util.f90:1017) */
  double sx = 0., sy = 0., sx2 = 0., sy2 = 0., sxy = 0.;
  for (int i = 0; i < n; ++i) {
    sx = sx + x[i];
    sy = sy + y[i];
    sx2 = sx2 + x[i] * x[i];
    sy2 = sy2 + y[i] * y[i];
    sxy = x[i] * y[i];
  }
  if (sx2 * n - sx * sx < eps_r8 || sy2 * n - sy * sy < eps_r8) return 0.;
  return (sxy * n - sx * sy) / (sqrt(sx2 * n - sx * sx) * sqrt(sy2 * n - sy * sy));
}

double so_les_corr(const double *x, const double *y, int n) { return les_corr(x, y, n); }

/* findDry4Wet for an evaporating bin with number ne and dry masses md[ndry];
 * na / ma: the cell's aerosol numbers and dry masses ma[s][nbins]. */
static int les_dry4wet(const so_layout *L, double ne, const double *md, const double *na, double ma[][SO_MAXBIN]) {
  const int ndry = L->ndry, A = L->iaero;
  double cd = 0.;
  for (int s = 0; s < ndry; ++s) cd = cd + md[s] / L->rholiq[s];
  cd = cd / ne;
  cd = pow(cd / pi6, 1. / 3.);
  int cnt = 0;
  for (int b = L->in2a; b <= L->fn2a; ++b) cnt += cd > pow(L->vlolim[A + b] / pi6, 1. / 3.);
  int ba = L->in2a + cnt - 1;
  ba = ba < L->in2a ? L->in2a : ba;
  ba = ba > L->fn2a ? L->fn2a : ba;
  const int bb = L->fn2a + (ba - L->in2a);     /* as written, :443 */
  if (na[bb] <= nlim_c) return ba;
  if (na[ba] <= nlim_c) return bb;
  double xa[SO_MAXSPEC], xb[SO_MAXSPEC];
  for (int s = 0; s < ndry; ++s) { xa[s] = ma[s][ba]; xb[s] = ma[s][bb]; }
  const double ra = les_corr(xa, md, ndry), rb = les_corr(xb, md, ndry);
  return ra < rb ? bb : ba;
}

int so_les_update(const so_layout *L, int level, int nzp, int nxp, int nyp, double dtlt, double dzf2,
                  const double *temp, const double *rh, const double *rhi, const double *rp,
                  double *const n_p[4], double *const n_t[4], double *const m_p[4], double *const m_t[4],
                  double *charge_time, double *diag) {
  return so_les_update_ext(L, level, nzp, nxp, nyp, dtlt, dzf2, temp, rh, rhi, rp, n_p, n_t, m_p, m_t, charge_time,
                           diag, 0, 0, 0);
}

/* gas row (0-based) that a dry species' mass goes to when a tiny aerosol is
   removed with lscndgas on (constrain_SALSA.f90:332-351): SO4 -> 1, NO -> 2,
   NH -> 3, OC -> 5 (1-based rows of a_gaerop); -1 for the others */
static int les_gas_row(int species_id) {
  switch (species_id) {
    case 0: return 0;
    case 1: return 4;
    case 5: return 1;
    case 6: return 2;
    default: return -1;
  }
}

/* so_les_update with the two optional fields of constrain_SALSA that the
   SALSA tracers do not cover: gaerop[5][n3] (a_gaerop; when given, the dry
   mass of aerosols removed below 0.1 nm goes to the gas phase as molecules,
   the lscndgas branch of :332-351) and indefp[nindef][n3] (a_indefp; when
   given and level 5, clamped to [0, 1] in tend_constrain2, :33-34, the
   ice_theta_dist branch). */
int so_les_update_ext(const so_layout *L, int level, int nzp, int nxp, int nyp, double dtlt, double dzf2,
                      const double *temp, const double *rh, const double *rhi, const double *rp,
                      double *const n_p[4], double *const n_t[4], double *const m_p[4], double *const m_t[4],
                      double *charge_time, double *diag, double *gaerop, double *indefp, int nindef) {
  const long n3 = (long)nzp * nxp * nyp;
  const int ncat = level == 5 ? 4 : 3;
  const int nb[4] = {L->nbins, L->ncld, L->nprc, L->nice};
  const int nwet = L->nwet, ndry = L->ndry, iwa = L->iwa, irim = L->irim;
  const int nsw[4] = {nwet, nwet, nwet, L->nspec};
  const int nx = nxp - 4, ny = nyp - 4;
  const double massTH = 1.e-25;
  if (nzp < 4 || nx < 2 || ny < 2) return -1;     /* the periodic images need two interior columns */
#define IDX(k, i, j) ((long)(k) + (long)nzp * ((long)(i) + (long)nxp * (long)(j)))
  /* tend_constrain2's a_indefp clamp (:33-34) */
  if (indefp && level > 4)
    for (long e = 0; e < (long)nindef * n3; ++e) indefp[e] = fmin(fmax(indefp[e], 0.), 1.);
  /* tend_constrain2, update (interior), tend0 */
  for (int c = 0; c < ncat; ++c) {
    const int nrow = nb[c] * (1 + nsw[c]);
    for (int r = 0; r < nrow; ++r) {
      double *p = r < nb[c] ? n_p[c] + (long)r * n3 : m_p[c] + (long)(r - nb[c]) * n3;
      double *t = r < nb[c] ? n_t[c] + (long)r * n3 : m_t[c] + (long)(r - nb[c]) * n3;
      for (int j = 0; j < nyp; ++j)
        for (int i = 0; i < nxp; ++i)
          for (int k = 0; k < nzp; ++k) {
            const long x = IDX(k, i, j);
            const double tc = fmax((-1. + 1.e-8) * p[x] / dtlt, t[x]);
            if (k >= 1 && k <= nzp - 2 && i >= 2 && i <= nxp - 3 && j >= 2 && j <= nyp - 3) p[x] = p[x] + tc * dtlt;
            t[x] = 0.;
          }
      /* sclrset 'mixd' on every column, then the periodic halos */
      for (int j = 0; j < nyp; ++j)
        for (int i = 0; i < nxp; ++i) {
          p[IDX(0, i, j)] = p[IDX(1, i, j)] - 0. * (p[IDX(2, i, j)] - p[IDX(1, i, j)]);
          p[IDX(nzp - 1, i, j)] = p[IDX(nzp - 2, i, j)] + dzf2 * (p[IDX(nzp - 2, i, j)] - p[IDX(nzp - 3, i, j)]);
        }
      for (int j = 0; j < nyp; ++j)
        for (int i = 0; i < nxp; ++i) {
          if (i >= 2 && i <= nxp - 3 && j >= 2 && j <= nyp - 3) continue;
          const int ii = i < 2 ? i + nx : (i > nxp - 3 ? i - nx : i);
          const int jj = j < 2 ? j + ny : (j > nyp - 3 ? j - ny : j);
          for (int k = 0; k < nzp; ++k) p[IDX(k, i, j)] = p[IDX(k, ii, jj)];
        }
    }
  }
  /* SALSA_diagnostics: remove negative values everywhere */
  for (int c = 0; c < ncat; ++c) {
    for (long x = 0; x < (long)nb[c] * n3; ++x) n_p[c][x] = fmax(0., n_p[c][x]);
    for (long x = 0; x < (long)nb[c] * nsw[c] * n3; ++x) m_p[c][x] = fmax(0., m_p[c][x]);
  }
  const int A = 0;
  (void)A;
  for (int j = 2; j <= nyp - 3; ++j)
    for (int i = 2; i <= nxp - 3; ++i)
      for (int k = 0; k < nzp; ++k) {
        const long x = IDX(k, i, j);
#define NP(c, b) n_p[c][(long)(b) * n3 + x]
#define MP(c, s, b) m_p[c][((long)(s) * nb[c] + (b)) * n3 + x]
        /* particles with number but no mass (:103-154) */
        for (int b = 0; b < nb[0]; ++b) {
          double sm = 0.;
          for (int s = 0; s < nwet; ++s) sm = sm + MP(0, s, b);
          if (NP(0, b) > 0. && sm <= 0.) {
            NP(0, b) = 0.;
            for (int s = 0; s < nwet; ++s) MP(0, s, b) = 0.;
            if (charge_time) charge_time[(long)b * n3 + x] = 0.;
          }
        }
        for (int b = 0; b < nb[1]; ++b) {
          double sm = 0.;
          for (int s = 0; s < nwet; ++s) sm = sm + MP(1, s, b);
          if (NP(1, b) > 0. && sm <= 0.) {
            NP(1, b) = 0.;
            for (int s = 0; s < nwet; ++s) MP(1, s, b) = 0.;
            if (charge_time) charge_time[(long)(nb[0] + b) * n3 + x] = 0.;
          }
        }
        for (int b = 0; b < nb[2]; ++b) {
          if (NP(2, b) > 0. && MP(2, iwa, b) <= 0.) {
            NP(2, b) = 0.;
            for (int s = 0; s < nwet; ++s) MP(2, s, b) = 0.;
            if (charge_time) charge_time[(long)(nb[0] + nb[1] + b) * n3 + x] = 0.;
          }
        }
        if (level == 5)
          for (int b = 0; b < nb[3]; ++b) {
            const double mt = MP(3, iwa, b) + MP(3, irim, b);
            if (NP(3, b) > 0. && mt < massTH) {
              NP(3, b) = 0.;
              for (int s = 0; s <= irim; ++s) MP(3, s, b) = 0.;
            }
          }
        /* aerosol numbers and dry masses of the cell for findDry4Wet */
        double na[SO_MAXBIN], ma[SO_MAXSPEC][SO_MAXBIN];
        /* evaporating cloud droplets (:156-207) */
        for (int bc = L->ica.cur; bc <= L->fcb.cur; ++bc) {
          if (!(NP(1, bc) > nlim_c && rh[x] < 0.999 && MP(1, iwa, bc) < 1.e-5)) continue;
          double ns = 0.;
          for (int s = 0; s < ndry; ++s) ns = ns + L->diss[s] * MP(1, s, bc) / L->mm[s];
          ns = ns / NP(1, bc);
          const double zbb = 6. * mwa * ns / (pi * rhowa);
          const double zaa = 4. * mwa * surfw0 / (rg * rhowa * temp[x]);
          const double cdcld = sqrt(3. * zbb / zaa);
          double zvol = 0.;
          for (int s = 0; s < nwet; ++s) zvol = zvol + MP(1, s, bc) / L->rholiq[s];
          zvol = zvol / NP(1, bc);
          const double zdh2o = pow(zvol / pi6, 1. / 3.);
          if (zdh2o < fmin(0.7 * cdcld, 1.3e-5) || MP(1, iwa, bc) < massTH * NP(1, bc)) {
            const int ba = bc <= L->fca.cur ? L->ica.par + (bc - L->ica.cur) : L->icb.par + (bc - L->icb.cur);
            NP(0, ba) = NP(0, ba) + NP(1, bc);
            NP(1, bc) = 0.;
            for (int s = 0; s < nwet; ++s) {
              MP(0, s, ba) = MP(0, s, ba) + MP(1, s, bc);
              MP(1, s, bc) = 0.;
            }
            if (charge_time) charge_time[(long)(nb[0] + bc) * n3 + x] = 0.;
          }
        }
        /* evaporating precipitation (:209-258) */
        for (int bc = 0; bc < nb[2]; ++bc) {
          if (!(NP(2, bc) > prlim_c && rh[x] < 0.999 && MP(2, iwa, bc) < 1.e-6)) continue;
          double ns = 0.;
          for (int s = 0; s < ndry; ++s) ns = ns + L->diss[s] * MP(2, s, bc) / L->mm[s];
          ns = ns / NP(2, bc);
          const double zbb = 6. * mwa * ns / (pi * rhowa);
          const double zaa = 4. * mwa * surfw0 / (rg * rhowa * temp[x]);
          const double cdprc = sqrt(3. * zbb / zaa);
          double zvol = 0.;
          for (int s = 0; s < nwet; ++s) zvol = zvol + MP(2, s, bc) / L->rholiq[s];
          zvol = zvol / NP(2, bc);
          const double zdh2o = pow(zvol / pi6, 1. / 3.);
          if (zdh2o < fmin(0.2 * cdprc, 10.e-6) || MP(2, iwa, bc) < massTH * NP(2, bc)) {
            double md[SO_MAXSPEC];
            for (int s = 0; s < ndry; ++s) md[s] = MP(2, s, bc);
            for (int b = 0; b < nb[0]; ++b) {
              na[b] = NP(0, b);
              for (int s = 0; s < ndry; ++s) ma[s][b] = MP(0, s, b);
            }
            const int ba = les_dry4wet(L, NP(2, bc), md, na, ma);
            NP(0, ba) = NP(0, ba) + NP(2, bc);
            NP(2, bc) = 0.;
            for (int s = 0; s < nwet; ++s) {
              MP(0, s, ba) = MP(0, s, ba) + MP(2, s, bc);
              MP(2, s, bc) = 0.;
            }
            if (charge_time) charge_time[(long)(nb[0] + nb[1] + bc) * n3 + x] = 0.;
          }
        }
        /* sublimating ice (:260-315) */
        if (level == 5)
          for (int bc = 0; bc < nb[3]; ++bc) {
            const double mt = MP(3, iwa, bc) + MP(3, irim, bc);
            double vt = 0.;
            for (int s = 0; s <= irim; ++s) vt = vt + MP(3, s, bc) / L->rhoice[s];
            if (!(NP(3, bc) > prlim_c && rhi[x] < 0.999 && mt < 1.e-15)) continue;
            double cdice = vt / NP(3, bc);
            cdice = pow(cdice / pi6, 1. / 3.);
            double zdrms = 0., zwams = 0.;
            for (int s = 0; s < ndry; ++s) zdrms = zdrms + MP(3, s, bc);
            for (int s = 0; s <= irim; ++s) zwams = zwams + MP(3, s, bc);
            const double zvol = zdrms / zwams;
            if (zvol > 0.5 || cdice < 2.e-6) {
              double md[SO_MAXSPEC];
              for (int s = 0; s < ndry; ++s) md[s] = MP(3, s, bc);
              for (int b = 0; b < nb[0]; ++b) {
                na[b] = NP(0, b);
                for (int s = 0; s < ndry; ++s) ma[s][b] = MP(0, s, b);
              }
              const int ba = les_dry4wet(L, NP(3, bc), md, na, ma);
              NP(0, ba) = NP(0, ba) + NP(3, bc);
              NP(3, bc) = 0.;
              for (int s = 0; s < nwet; ++s) {
                MP(0, s, ba) = MP(0, s, ba) + MP(3, s, bc);
                MP(3, s, bc) = 0.;
              }
              MP(0, iwa, ba) = MP(0, iwa, ba) + MP(3, irim, bc);
              MP(3, irim, bc) = 0.;
            }
          }
        /* aerosols below 0.1 nm dry diameter (:317-366); with gaerop their
           SO4, OC, NO and NH masses go to the gas phase as molecules */
        for (int b = 0; b < nb[0]; ++b) {
          if (!(NP(0, b) > nlim_c)) continue;
          double zvol = 0.;
          for (int s = 0; s < ndry; ++s) zvol = zvol + MP(0, s, b) / L->rholiq[s];
          zvol = zvol / NP(0, b);
          if (zvol < pi6 * 1.e-30) {
            if (gaerop)
              for (int s = 0; s < ndry; ++s) {
                const int g = les_gas_row(L->spec_id[s]);
                if (g >= 0) gaerop[g * n3 + x] = gaerop[g * n3 + x] + MP(0, s, b) / L->mm[s] * avog;
              }
            for (int s = 0; s < nwet; ++s) MP(0, s, b) = 0.;
            NP(0, b) = 0.;
          }
        }
#undef NP
#undef MP
      }
  /* diagnostic tracers over the whole grid (:370-400):
     diag[0] rc, [1] srp, [2] snrp, [3] rtot, [4] ri, [5] riri */
  for (long x = 0; x < n3; ++x) {
    double sa = 0., sc = 0., sp = 0., sn = 0.;
    for (int b = L->in1a; b <= L->fn2b; ++b) sa = sa + m_p[0][((long)iwa * nb[0] + b) * n3 + x];
    for (int b = L->ica.cur; b <= L->fcb.cur; ++b) sc = sc + m_p[1][((long)iwa * nb[1] + b) * n3 + x];
    for (int b = 0; b < nb[2]; ++b) sp = sp + m_p[2][((long)iwa * nb[2] + b) * n3 + x];
    for (int b = 0; b < nb[2]; ++b) sn = sn + n_p[2][(long)b * n3 + x];
    const double rc = sa + sc;
    double rtot = rc + sp + rp[x];
    diag[x] = rc;
    diag[n3 + x] = sp;
    diag[2 * n3 + x] = sn;
    if (level == 5) {
      double si = 0., sr = 0.;
      for (int b = 0; b < nb[3]; ++b) si = si + m_p[3][((long)iwa * nb[3] + b) * n3 + x];
      for (int b = 0; b < nb[3]; ++b) sr = sr + m_p[3][((long)irim * nb[3] + b) * n3 + x];
      rtot = rtot + si + sr;
      diag[4 * n3 + x] = si;
      diag[5 * n3 + x] = sr;
    }
    diag[3 * n3 + x] = rtot;
  }
#undef IDX
  return 0;
}

/* ====================================================================== */
/* Sedimentation of the liquid categories at level 4 (sedim_SALSA,        */
/* mcrp.f90:745-921): DepositionSlow for aerosols and cloud droplets      */
/* (:926-1069, with the dry deposition velocity at the lowest level) and  */
/* DepositionFast for precipitation (:1073-1275, fall distance per bin    */
/* and time step, with the reference's numerical-diffusion switch);       */
/* calcDiamLES and terminal_vel (mo_particle_external_properties.f90:     */
/* 43-100, 169-201) for liquid particles.  Constants of defs.f90: pi with */
/* full precision, kb, g = 9.81, cp = 1005, alvl = 2.5e6; of mo_submctl:  */
/* pi6, grav, surfw0.                                                     */
/* ====================================================================== */
static const double sd_pi = 3.14159265358979323846, sd_kb = 1.3807e-23, sd_g = 9.81, sd_cp = 1005.,
                    sd_alvl = 2.5e6;

static double sd_terminal_vel(double diam, double rhop, double rhoa, double visc, double beta) {
  if (diam <= 19e-6) return (diam * diam) * (rhop - rhoa) * grav * beta / (18. * visc);
  if (diam > 19e-6 && diam < 1.07e-3) {
    const double NDa = 4. / 3. * pow(diam, 3.) * grav * (rhop - rhoa) * rhoa / pow(visc, 2.);
    const double X = log(NDa);
    const double X2 = X * X, X3 = X2 * X, X4 = X2 * X2, X5 = X4 * X, X6 = X3 * X3;
    const double Y = -0.318657e1 + 0.992696 * X - 0.153193e-2 * X2 - 0.987059e-3 * X3 - 0.578878e-3 * X4 +
                     0.855176e-4 * X5 - 0.327815e-5 * X6;
    const double Re2 = beta * exp(Y);
    return visc / rhoa / diam * Re2;
  }
  const double Np = pow(surfw0, 3.) * pow(rhoa, 2.) / (pow(visc, 4.) * (rhop - rhoa) * grav);
  const double Bo = 4. / 3. * (rhop - rhoa) * grav / surfw0 * pow(diam, 2.);
  const double X3 = log(Bo * pow(Np, 1. / 6.));
  const double Z2 = X3 * X3, Z3 = Z2 * X3, Z4 = Z2 * Z2, Z5 = Z4 * X3;
  const double Y3 = -0.500015e1 + 0.523778e1 * X3 - 0.204914e1 * Z2 + 0.475294 * Z3 - 0.542819e-1 * Z4 +
                    0.238449e-2 * Z5;
  const double Re3 = pow(Np, 1. / 6.) * exp(Y3);
  return fmin(visc / rhoa / diam * Re3, 9.2);
}

double so_terminal_vel(double diam, double rhop, double rhoa, double visc, double beta) {
  return sd_terminal_vel(diam, rhop, rhoa, visc, beta);
}

/* air properties of a level: viscosity (atm. modelling Eq. 4.54) and the
   mean free path */
static void sd_air(double tk, double adn, double *avis, double *lambda) {
  *avis = 1.8325e-5 * (416.16 / (tk + 120.0)) * pow(tk / 296.16, 1.5);
  const double va = sqrt(8 * sd_kb * tk / (sd_pi * 4.8096e-26));
  *lambda = 2 * (*avis) / (adn * va);
}

/* wet diameter, bulk density and Cunningham-corrected terminal velocity
   of one bin (the bin's ns masses at m[s*mstride]) */
static double sd_vc(const so_layout *L, int ns, double numc, const double *m, long mstride, double adn, double avis,
                    double lambda) {
  double v = 0., ms = 0.;
  for (int s = 0; s < ns; ++s) v = v + m[s * mstride] / L->rholiq[s];
  for (int s = 0; s < ns; ++s) ms = ms + m[s * mstride];
  const double dwet = numc < 1.e-15 ? 0. : pow(v / numc / pi6, 1. / 3.);
  const double zdn = ms / numc / (pi6 * (dwet * dwet * dwet));
  const double Kn = 2. * lambda / dwet;
  const double GG = 1. + Kn * (1.249 + 0.42 * exp(-0.87 / Kn));
  return sd_terminal_vel(dwet, zdn, adn, avis, GG);
}

/* DepositionFast's fall speed of one ice bin (flag 4, mcrp.f90:1146-1176):
   calcDiamLES with sph = .TRUE. (volume over the ice densities) and
   .FALSE. (getDiameter of the pristine and rime masses), getShapeCoefficients,
   the bulk density over the spherical diameter, and terminal_vel's ice
   branch (Khvorostyanov and Curry 2002; the Cunningham factor is formed but
   not used there).  The bin's ns masses (rime last) at m[s*mstride].  out,
   if given: dwet, dnsp, alpha, beta, gamma, sigma, zdn, GG. */
static double sd_vc_ice(const so_layout *L, const so_params *P, int ns, double numc, const double *m, long mstride,
                        double adn, double avis, double lambda, double *out) {
  double v = 0., mpri = 0., ms = 0.;
  for (int s = 0; s < ns; ++s) v = v + m[s * mstride] / L->rhoice[s];
  for (int s = 0; s < ns - 1; ++s) mpri = mpri + m[s * mstride];
  const double mrim = m[(long)(ns - 1) * mstride];
  const double dwet = numc < 1.e-15 ? 0. : pow(v / numc / pi6, 1. / 3.);
  const double dnsp = numc < 1.e-15 ? 0. : ice_diameter(P, mpri, mrim, numc);
  const so_shape sh = shape_coeffs(P, mpri, mrim, numc);
  for (int s = 0; s < ns; ++s) ms = ms + m[s * mstride];
  const double zdn = ms / numc / (pi6 * (dwet * dwet * dwet));
  const double Kn = 2. * lambda / dwet;
  const double GG = 1. + Kn * (1.249 + 0.42 * exp(-0.87 / Kn));
  if (out) {
    out[0] = dwet; out[1] = dnsp; out[2] = sh.alpha; out[3] = sh.beta; out[4] = sh.gamma; out[5] = sh.sigma;
    out[6] = zdn; out[7] = GG;
  }
  return terminal_vel_ice(dwet, zdn, adn, avis, sh, dnsp);
}

/* the ice fall-speed chain of one bin for a level of temperature tk and
   air density adn (masses: ns values, rime last); out[9]: the eight values
   of sd_vc_ice, then the fall speed */
void so_sed_ice_vc(const so_layout *L, const so_params *P, double tk, double adn, double numc, const double *m,
                   double *out) {
  double avis, lambda;
  sd_air(tk, adn, &avis, &lambda);
  out[8] = sd_vc_ice(L, P, L->nspec, numc, m, 1, adn, avis, lambda, out);
}

/* DepositionFast (mcrp.f90:1073-1275) for one category on every interior
   column and its part of sedim_SALSA: tendencies += change / dtlt (:862-863,
   :884-885), tlt (:871-878, :897-904).  ice = 0: precipitation (flag 3,
   water the last of ns species, alvl); ice = 1: ice (flag 4, ns species
   with water and rime last, their sum in the rate, alvi). */
static int sd_fast(const so_layout *L, const so_params *P, int ice, int nzp, int nxp, int nyp, double dtlt,
                   const double *dzt, const double *tk, const double *th, const double *adn, int nb, int ns,
                   const double *N, const double *M, double *n_t, double *m_t, double *tlt, double *rate,
                   double *srate, double *vt) {
  const long n3 = (long)nzp * nxp * nyp;
  const double alv = ice ? 2.834e6 : sd_alvl;
  double prn[SO_MAXBIN][256], prv[SO_MAXBIN * SO_MAXSPEC][256];
  double rem[SO_MAXBIN * SO_MAXSPEC];
  if (nzp > 256 || nb * ns > SO_MAXBIN * SO_MAXSPEC) return -1;
#define IX(k, i, j) ((long)(k) + (long)nzp * ((long)(i) + (long)nxp * (long)(j)))
  for (int j = 2; j <= nyp - 3; ++j)
    for (int i = 2; i <= nxp - 3; ++i) {
      for (int b = 0; b < nb; ++b)
        for (int k = 0; k < nzp; ++k) prn[b][k] = 0.;
      for (int q = 0; q < nb * ns; ++q) {
        for (int k = 0; k < nzp; ++k) prv[q][k] = 0.;
        rem[q] = 0.;
      }
      for (int k = nzp - 2; k >= 1; --k) {
        const long x = IX(k, i, j);
        double avis, lambda;
        sd_air(tk[x], adn[x], &avis, &lambda);
        for (int b = 0; b < nb; ++b) {
          const double zpn = N[(long)b * n3 + x];
          if (zpn < prlim_c) continue;
          const double *mb = M + (long)b * n3 + x;
          const double vc = ice ? sd_vc_ice(L, P, ns, zpn, mb, (long)nb * n3, adn[x], avis, lambda, NULL)
                                : sd_vc(L, ns, zpn, mb, (long)nb * n3, adn[x], avis, lambda);
          vt[(long)b * n3 + x] = vc;
          const double fdmax = dtlt * vc;
          double fd = 0.;
          int fi = 0, dep = 0;
          while (fd < fdmax) {
            fd = fd + (1. / dzt[k - fi]);
            fi = fi + 1;
            if (k - fi <= 0) { dep = 1; break; }     /* 1-based k-fi <= 1 */
          }
          fi = fi - 1;
          const int kf = k - fi;
          fd = fd - (1. / dzt[kf]);
          const double fdos = fmin(fmax(fdmax - fd, 0.), 1. / dzt[kf]);
          prn[b][k] = prn[b][k] - zpn;
          for (int s = 0; s < ns; ++s) prv[s * nb + b][k] = prv[s * nb + b][k] - mb[(long)s * nb * n3];
          if (dep)
            for (int s = 0; s < ns; ++s) rem[s * nb + b] = rem[s * nb + b] + mb[(long)s * nb * n3] * adn[x] * vc;
          if (ice)                                     /* pristine and rimed ice (:1208-1210) */
            rate[x] = rate[x] + (mb[(long)(ns - 2) * nb * n3] + mb[(long)(ns - 1) * nb * n3]) * adn[x] * vc * alv;
          else
            rate[x] = rate[x] + mb[(long)(ns - 1) * nb * n3] * adn[x] * vc * alv;
          if (fdos * dzt[kf] > 0.5) {
            prn[b][kf - 1] = prn[b][kf - 1] + zpn * dzt[kf - 1] / dzt[k];
            for (int s = 0; s < ns; ++s)
              prv[s * nb + b][kf - 1] = prv[s * nb + b][kf - 1] + mb[(long)s * nb * n3] * dzt[kf - 1] / dzt[k];
          } else {
            prn[b][kf - 1] = prn[b][kf - 1] + (fdos * dzt[kf]) * zpn * dzt[kf - 1] / dzt[kf];
            prn[b][kf] = prn[b][kf] + (1. - fdos * dzt[kf]) * zpn;
            for (int s = 0; s < ns; ++s) {
              prv[s * nb + b][kf - 1] =
                  prv[s * nb + b][kf - 1] + (fdos * dzt[kf]) * mb[(long)s * nb * n3] * dzt[kf - 1] / dzt[kf];
              prv[s * nb + b][kf] = prv[s * nb + b][kf] + (1. - fdos * dzt[kf]) * mb[(long)s * nb * n3];
            }
          }
        }
      }
      /* surface rate: SUM(remprc(i,j,(ns-1)*nb:ns*nb)) for precipitation,
         (ns-2)*nb:ns*nb for ice (1-based, one entry before the water rows, as
         written; :1256-1260) */
      double s2 = 0.;
      for (int q = (ice ? ns - 2 : ns - 1) * nb - 1; q < ns * nb; ++q) s2 = s2 + rem[q];
      srate[(long)i + (long)nxp * j] = srate[(long)i + (long)nxp * j] + s2 * alv;
      for (int k = 0; k < nzp; ++k) {
        const long x = IX(k, i, j);
        for (int b = 0; b < nb; ++b) n_t[(long)b * n3 + x] = n_t[(long)b * n3 + x] + prn[b][k] / dtlt;
        for (int q = 0; q < nb * ns; ++q) m_t[(long)q * n3 + x] = m_t[(long)q * n3 + x] + prv[q][k] / dtlt;
        if (k <= nzp - 2) {                            /* 1-based k = 1 .. nzp-1 */
          /* water rows in bin order; for ice the rime rows follow */
          double sw = 0.;
          for (int q = (ice ? ns - 2 : ns - 1) * nb; q < ns * nb; ++q) sw = sw + prv[q][k];
          tlt[x] = tlt[x] - (sw / dtlt) * (alv / sd_cp) * th[x] / tk[x];
        }
      }
    }
#undef IX
  return 0;
}

int so_sedim(const so_layout *L, int nzp, int nxp, int nyp, double dtlt, const double *dzt, const double *tk,
             const double *th, const double *adn, const double *ustar, double *const n_p[3], double *const m_p[3],
             double *const n_t[3], double *const m_t[3], double *tlt, double *rrate, double *sfcrrate, double *vt_prc,
             int sed_aero, int sed_cloud, int sed_precp) {
  const long n3 = (long)nzp * nxp * nyp;
  const int ns = L->nwet, iw = L->iwa;
  const int nbs[3] = {L->nbins, L->ncld, L->nprc};
  if (nzp < 4 || nxp < 5 || nyp < 5) return -1;
#define IX(k, i, j) ((long)(k) + (long)nzp * ((long)(i) + (long)nxp * (long)(j)))
  for (long x = 0; x < n3; ++x) rrate[x] = 0.;
  for (long c = 0; c < (long)nxp * nyp; ++c) sfcrrate[c] = 0.;
  for (long x = 0; x < (long)L->nprc * n3; ++x) vt_prc[x] = 0.;
  /* slow deposition: aerosols (flag 1), cloud droplets (flag 2) */
  for (int cat = 0; cat < 2; ++cat) {
    if (!(cat == 0 ? sed_aero : sed_cloud)) continue;
    const int nb = nbs[cat];
    double *N = n_p[cat], *M = m_p[cat];
    double rfln[SO_MAXBIN], rflm[SO_MAXBIN * SO_MAXSPEC], rfln1[SO_MAXBIN], rflm1[SO_MAXBIN * SO_MAXSPEC];
    for (int j = 2; j <= nyp - 3; ++j)
      for (int i = 2; i <= nxp - 3; ++i) {
        for (int q = 0; q < nb * ns; ++q) rflm1[q] = 0.;
        for (int b = 0; b < nb; ++b) rfln1[b] = 0.;
        for (int k = nzp - 2; k >= 1; --k) {          /* 1-based k = nzp-1 .. 2 */
          const long x = IX(k, i, j), x1 = IX(k + 1, i, j);
          double avis, lambda;
          sd_air(tk[x], adn[x], &avis, &lambda);
          const double kvis = avis / adn[x];
          for (int q = 0; q < nb * ns; ++q) rflm[q] = 0.;
          for (int b = 0; b < nb; ++b) rfln[b] = 0.;
          for (int b = 0; b < nb; ++b) {
            const double zpn = N[(long)b * n3 + x];
            if (zpn < nlim_c) continue;
            const double *mb = M + (long)b * n3 + x;
            double vc;
            {
              double v = 0., ms = 0.;
              for (int s = 0; s < ns; ++s) v = v + mb[(long)s * nb * n3] / L->rholiq[s];
              for (int s = 0; s < ns; ++s) ms = ms + mb[(long)s * nb * n3];
              const double dwet = zpn < 1.e-15 ? 0. : pow(v / zpn / pi6, 1. / 3.);
              const double zdn = ms / zpn / (pi6 * (dwet * dwet * dwet));
              const double Kn = 2. * lambda / dwet;
              const double GG = 1. + Kn * (1.249 + 0.42 * exp(-0.87 / Kn));
              vc = sd_terminal_vel(dwet, zdn, adn[x], avis, GG);
              vc = fmin(vc, fmin(0.5 * (1. / dzt[k]) / dtlt, 2.0));
              if (k == 1) {                           /* the level just above the surface */
                const double mdiff = (sd_kb * tk[x] * GG) / (3.0 * sd_pi * avis * dwet);
                const double Sc = kvis / mdiff;
                double St = vc * pow(ustar[(long)i + (long)nxp * j], 2.0) / sd_g * kvis;
                if (St < 0.01) St = 0.01;
                const double rt = 1.0 / fmax(eps_r8, (ustar[(long)i + (long)nxp * j] *
                                                      (pow(Sc, -2.0 / 3.0) + pow(10., -3.0 / St))));
                vc = (1. / rt) + vc;
              }
            }
            for (int s = 0; s < ns; ++s) rflm[s * nb + b] = -mb[(long)s * nb * n3] * vc;
            rfln[b] = -zpn * vc;
          }
          double sw = 0.;
          for (int q = ns * nb - nb - 1; q < ns * nb; ++q) sw = sw + rflm[q];   /* (ns-1)*nb : ns*nb, as written */
          rrate[x] = rrate[x] - sw * adn[x] * sd_alvl;
          for (int b = 0; b < nb; ++b) {
            const double dn = (rfln1[b] * adn[x1] - rfln[b] * adn[x]) * dzt[k] / adn[x];
            n_t[cat][(long)b * n3 + x] = n_t[cat][(long)b * n3 + x] - dn;
            for (int s = 0; s < ns; ++s) {
              const double dm = (rflm1[s * nb + b] * adn[x1] - rflm[s * nb + b] * adn[x]) * dzt[k] / adn[x];
              m_t[cat][((long)s * nb + b) * n3 + x] = m_t[cat][((long)s * nb + b) * n3 + x] - dm;
            }
          }
          /* SUM over the water rows in bin order (mcrp.f90:810-816) */
          {
            double sdw = 0.;
            for (int b = 0; b < nb; ++b)
              sdw = sdw + (rflm1[iw * nb + b] * adn[x1] - rflm[iw * nb + b] * adn[x]) * dzt[k] / adn[x];
            tlt[x] = tlt[x] + sdw * (sd_alvl / sd_cp) * th[x] / tk[x];
          }
          if (k == 1) {
            double s2 = 0.;
            for (int q = ns * nb - nb - 1; q < ns * nb; ++q) s2 = s2 + rflm[q];
            sfcrrate[(long)i + (long)nxp * j] = sfcrrate[(long)i + (long)nxp * j] - s2 * adn[x] * sd_alvl;
          }
          for (int q = 0; q < nb * ns; ++q) rflm1[q] = rflm[q];
          for (int b = 0; b < nb; ++b) rfln1[b] = rfln[b];
        }
      }
  }
  /* fast deposition: precipitation (flag 3) */
  if (sed_precp && sd_fast(L, NULL, 0, nzp, nxp, nyp, dtlt, dzt, tk, th, adn, L->nprc, ns, n_p[2], m_p[2], n_t[2],
                           m_t[2], tlt, rrate, sfcrrate, vt_prc))
    return -1;
#undef IX
  return 0;
}

/* sedim_SALSA at level 5 (mcrp.f90:745-921): the liquid categories as at
   level 4 (their ns = the wet species), then ice with DepositionFast flag 4
   (ns = the wet species and rime) into irate, sfcirate and VtIce, which are
   zeroed first (:790-795, :1127). */
int so_sedim5(const so_layout *L, const so_params *P, int nzp, int nxp, int nyp, double dtlt, const double *dzt,
              const double *tk, const double *th, const double *adn, const double *ustar, double *const n_p[4],
              double *const m_p[4], double *const n_t[4], double *const m_t[4], double *tlt, double *rrate,
              double *sfcrrate, double *irate, double *sfcirate, double *vt_prc, double *vt_ice, int sed_aero,
              int sed_cloud, int sed_precp, int sed_ice) {
  const long n3 = (long)nzp * nxp * nyp;
  int rc = so_sedim(L, nzp, nxp, nyp, dtlt, dzt, tk, th, adn, ustar, n_p, m_p, n_t, m_t, tlt, rrate, sfcrrate,
                    vt_prc, sed_aero, sed_cloud, sed_precp);
  if (rc) return rc;
  for (long x = 0; x < n3; ++x) irate[x] = 0.;
  for (long c = 0; c < (long)nxp * nyp; ++c) sfcirate[c] = 0.;
  for (long x = 0; x < (long)L->nice * n3; ++x) vt_ice[x] = 0.;
  if (sed_ice)
    rc = sd_fast(L, P, 1, nzp, nxp, nyp, dtlt, dzt, tk, th, adn, L->nice, L->nspec, n_p[3], m_p[3], n_t[3], m_t[3],
                 tlt, irate, sfcirate, vt_ice);
  return rc;
}
