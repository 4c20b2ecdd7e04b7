// salsa_layout.cpp -- bin descriptor, default switches and the device tables
// (pair list + coagulation sweep program) derived from them.
//
// The bin geometry follows set_aerobins / set_cloudbins / set_icebins /
// buildBinLimits (src/src_salsa/mo_salsa_init.f90:148-450); the sweep program
// encodes the collector and source index ranges of coag_ice, coag_precp,
// coag_aero and coag_cloud (src/src_salsa/mo_salsa_coagulation_processes.f90:
// 24-535) as bit masks over the master bin array.
#include <cmath>
#include <cstring>

#include "salsa_internal.h"

namespace salsa {
// Thread-local last error, also used by salsa_capi.cpp.
thread_local char g_err[512] = "";
void set_error(const char *msg) {
  std::strncpy(g_err, msg, sizeof(g_err) - 1);
  g_err[sizeof(g_err) - 1] = 0;
}
}  // namespace salsa

using namespace salsa;

// Species master tables (classSpecies.f90:97-110): SO4 OC BC DU SS NO NH H2O rime.
static const double kAllMM[9] = {98.08e-3, 150.e-3, 12.e-3, 100.e-3, 58.44e-3,
                                 62.01e-3, 18.04e-3, 18.016e-3, 18.016e-3};
static const double kAllRho[9] = {1830., 1400., 2000., 2650., 2165., 1479., 1530., 1000., 400.};
static const double kAllDiss[9] = {3., 1., 0., 0., 2., 1., 1., 1., 1.};

static inline double cube(double x) { return x * x * x; }
static inline double cbrt3(double x) { return std::pow(x, 1. / 3.); }

extern "C" int salsa_bin_desc_init(salsa_bin_desc *d, const int32_t nbin[2], const double reglim[4],
                                   int32_t nprc, double prc_dlo, double prc_vratio,
                                   int32_t nice, double ice_dlo, double ice_vratio,
                                   int32_t nspec_dry, const int32_t dry_species[7]) {
  if (!d || !nbin || !reglim || !dry_species) { set_error("null argument"); return SALSA_ERR_ARG; }
  std::memset(d, 0, sizeof(*d));
  // the 2a regime hard-codes 4 bins below reglim(3) (mo_salsa_init.f90:158)
  if (nbin[0] < 1 || nbin[1] < 4 || nprc < 1 || nice < 0 || nspec_dry < 1 || nspec_dry > 7) {
    set_error("unsupported bin configuration (need nbin(1)>=1, nbin(2)>=4, nprc>=1)");
    return SALSA_ERR_ARG;
  }
  for (int s = 0; s < nspec_dry; ++s)
    if (dry_species[s] < 0 || dry_species[s] > 6) { set_error("bad species id"); return SALSA_ERR_ARG; }
  d->nbin[0] = nbin[0]; d->nbin[1] = nbin[1];
  for (int i = 0; i < 4; ++i) d->reglim[i] = reglim[i];
  d->nprc = nprc; d->prc_dlo = prc_dlo; d->prc_vratio = prc_vratio;
  d->nice = nice; d->ice_dlo = ice_dlo; d->ice_vratio = ice_vratio;
  d->nspec_dry = nspec_dry;
  for (int s = 0; s < 7; ++s) d->dry_species[s] = s < nspec_dry ? dry_species[s] : 0;

  // indices (mo_salsa_init.f90:800-810), 0-based
  d->in1a = 0;
  d->in2a = nbin[0];
  d->in2b = d->in2a + nbin[1];
  d->fn1a = d->in2a - 1;
  d->fn2a = d->fn1a + nbin[1];
  d->fn2b = d->fn2a + nbin[1];
  d->nbins = d->fn2b + 1;
  // cloud bins parallel to 2a and 2b (mo_salsa_init.f90:270-284)
  const int nba = d->fn2a - d->in2a + 1;
  d->ica.cur = 0;                    d->ica.par = d->in2a;
  d->fca.cur = nba - 1;              d->fca.par = d->in2a + nba - 1;
  d->icb.cur = nba;                  d->icb.par = d->fn2b - nba + 1;
  d->fcb.cur = 2 * nba - 1;          d->fcb.par = d->fn2b;
  d->ncld = 2 * nba;
  d->ntotal = d->nbins + d->ncld + nprc + nice;
  if (d->ntotal > SALSA_MAXBIN) { set_error("too many bins"); return SALSA_ERR_ARG; }

  // species in mass-array order: namelist dry species, H2O, rime (classSpecies.f90:205-299)
  d->nspec = nspec_dry + 2;
  d->iwa = nspec_dry;
  d->irim = nspec_dry + 1;
  for (int s = 0; s < d->nspec; ++s) {
    int id = s < nspec_dry ? dry_species[s] : (s == d->iwa ? 7 : 8);
    d->mm[s] = kAllMM[id];
    d->diss[s] = kAllDiss[id];
    d->rholiq[s] = kAllRho[id];
    d->rhoice[s] = id == 7 ? 917. : kAllRho[id];
  }

  // aerosol bins: log-spaced dry volume limits per regime (mo_salsa_init.f90:173-227)
  const int nbin2 = 4, nbin3 = nbin[1] - nbin2;
  double ratio = reglim[1] / reglim[0];
  for (int b = d->in1a; b <= d->fn1a; ++b) {
    d->vlolim[b] = kPi6 * cube(reglim[0] * std::pow(ratio, (double)b / nbin[0]));
    d->vhilim[b] = kPi6 * cube(reglim[0] * std::pow(ratio, (double)(b + 1) / nbin[0]));
    d->dmid[b] = cbrt3((d->vhilim[b] + d->vlolim[b]) / (2. * kPi6));
  }
  ratio = reglim[2] / reglim[1];
  for (int b = d->in2a; b < d->in2a + nbin2; ++b) {
    int c = b - d->in2a;
    d->vlolim[b] = kPi6 * cube(reglim[1] * std::pow(ratio, (double)c / nbin2));
    d->vhilim[b] = kPi6 * cube(reglim[1] * std::pow(ratio, (double)(c + 1) / nbin2));
    d->dmid[b] = cbrt3((d->vhilim[b] + d->vlolim[b]) / (2. * kPi6));
  }
  ratio = reglim[3] / reglim[2];
  for (int b = d->in2a + nbin2; b <= d->fn2a; ++b) {
    int c = b - (d->fn2a - (nbin3 - 1));
    d->vlolim[b] = kPi6 * cube(reglim[2] * std::pow(ratio, (double)c / nbin3));
    d->vhilim[b] = kPi6 * cube(reglim[2] * std::pow(ratio, (double)(c + 1) / nbin3));
    d->dmid[b] = cbrt3((d->vhilim[b] + d->vlolim[b]) / (2. * kPi6));
  }
  for (int k = 0; k < nbin[1]; ++k) {
    d->vlolim[d->in2b + k] = d->vlolim[d->in2a + k];
    d->vhilim[d->in2b + k] = d->vhilim[d->in2a + k];
    d->dmid[d->in2b + k] = d->dmid[d->in2a + k];
  }
  for (int b = 0; b < d->nbins; ++b) { d->phase[b] = 1; d->nlim[b] = kNlim; d->dlim[b] = 30.e-6; }
  const int C = d->nbins, R = C + d->ncld, I = R + nprc;
  for (int k = 0; k < nba; ++k) {
    int pa = d->ica.par + k, pb = d->icb.par + k;
    int ca = C + d->ica.cur + k, cb = C + d->icb.cur + k;
    d->vlolim[ca] = d->vlolim[pa]; d->vhilim[ca] = d->vhilim[pa]; d->dmid[ca] = d->dmid[pa];
    d->vlolim[cb] = d->vlolim[pb]; d->vhilim[cb] = d->vhilim[pb]; d->dmid[cb] = d->dmid[pb];
  }
  for (int c = 0; c < d->ncld; ++c) { d->phase[C + c] = 2; d->nlim[C + c] = kNlim; d->dlim[C + c] = 100.e-6; }
  // precipitation and ice: geometric in volume (buildBinLimits, mo_salsa_init.f90:429-450)
  {
    const double r3 = std::pow(prc_vratio, 1. / 3.);
    double lo = prc_dlo, hi = prc_dlo * r3;
    for (int k = 0; k < nprc; ++k) {
      if (k) { lo = lo * r3; hi = lo * r3; }
      int b = R + k;
      d->vhilim[b] = kPi6 * cube(hi); d->vlolim[b] = kPi6 * cube(lo);
      d->dmid[b] = cbrt3((d->vlolim[b] + d->vhilim[b]) / (2. * kPi6));
      d->phase[b] = 3; d->nlim[b] = kPrlim; d->dlim[b] = 5.e-3;
      d->precpbins[k] = cbrt3(d->vlolim[b] / kPi6);
    }
  }
  {
    const double r3 = nice ? std::pow(ice_vratio, 1. / 3.) : 0.;
    double lo = ice_dlo, hi = ice_dlo * r3;
    for (int k = 0; k < nice; ++k) {
      if (k) { lo = lo * r3; hi = lo * r3; }
      int b = I + k;
      d->vhilim[b] = kPi6 * cube(hi); d->vlolim[b] = kPi6 * cube(lo);
      d->dmid[b] = cbrt3((d->vlolim[b] + d->vhilim[b]) / (2. * kPi6));
      d->phase[b] = 4; d->nlim[b] = kPrlim; d->dlim[b] = 10.e-3;
    }
  }
  return SALSA_OK;
}

extern "C" int salsa_params_default(salsa_params *p, int32_t level) {
  if (!p || (level != 4 && level != 5)) { set_error("level must be 4 or 5"); return SALSA_ERR_ARG; }
  std::memset(p, 0, sizeof(*p));
  p->level = level;
  p->lscoag = 1; p->lscnd = 1; p->lsauto = 1; p->lsauto_mode = 1;
  p->lscgaa = p->lscgcc = p->lscgpp = p->lscgii = p->lscgca = 1;
  p->lscgpa = p->lscgia = p->lscgpc = p->lscgic = p->lscgip = 1;
  p->lscndh2oae = p->lscndh2ocl = p->lscndh2oic = 1;
  p->lsfreeRH = 1; p->rhlim = 1.20;
  p->eddy_dis = 0.001; p->Eiagg_max = 0.3; p->Eiagg_min = 0.05;
  p->ice_alpha = 15.56999e-3; p->ice_beta = 2.02; p->ice_gamma = 0.55; p->ice_sigma = 1.97;
  p->lcharge = 0; p->charge_coll_enh = 0.;
  p->precision = SALSA_FP32_MIXED;
  p->hz = 1; p->hxy = 2;
  p->lscoag_mode = 1; p->cgintvl = 10.;     // mo_submctl.f90:23-24, :90
  p->lsdistupdate = 0;                       // reference default on (mo_submctl.f90:118): opt in
  p->lsactiv = 0;
  if (level == 4) {   // define_salsa, mo_salsa_init.f90:615-635
    p->lscgia = p->lscgic = p->lscgii = p->lscgip = 0;
    p->lscndh2oic = 0;
  }
  return SALSA_OK;
}

namespace salsa {

static inline uint64_t bits(int off, int lo, int hi) {   // master bins off+lo..off+hi
  uint64_t m = 0;
  for (int k = lo; k <= hi; ++k) m |= 1ull << (off + k);
  return m;
}

// Build the device tables; returns 0 or an error (message set).
int build_dev_layout(const salsa_bin_desc &d, const salsa_params &p, DevLayout &L) {
  std::memset(&L, 0, sizeof(L));
  const int level = p.level;
  if (level != 4 && level != 5) { set_error("level must be 4 or 5"); return SALSA_ERR_ARG; }
  L.level = level;
  L.ncat = level == 5 ? 4 : 3;
  // Level 4 has no ice prognostics and every ice process is switched off
  // (mo_salsa_init.f90:615-635, mo_salsa_driver.f90:235, 312): the ice bins
  // stay empty and inert, so they are left out.
  L.nb[0] = d.nbins; L.nb[1] = d.ncld; L.nb[2] = d.nprc; L.nb[3] = level == 5 ? d.nice : 0;
  L.off[0] = 0; L.off[1] = d.nbins; L.off[2] = d.nbins + d.ncld; L.off[3] = L.off[2] + d.nprc;
  L.nt = L.off[3] + L.nb[3];
  L.in1a = d.in1a; L.fn1a = d.fn1a; L.in2a = d.in2a; L.fn2a = d.fn2a; L.in2b = d.in2b; L.fn2b = d.fn2b;
  L.nspec = d.nspec; L.nwet = d.nspec - 1; L.ndry = d.nspec - 2; L.iwa = d.iwa; L.irim = d.irim;
  for (int s = 0; s < kMaxSpec; ++s) {
    L.rholiq[s] = d.rholiq[s]; L.rhoice[s] = d.rhoice[s]; L.mm[s] = d.mm[s]; L.diss[s] = d.diss[s];
  }
  for (int s = 0; s < d.nspec_dry; ++s) L.dry_id[s] = d.dry_species[s];
  for (int b = 0; b < L.nt; ++b) {
    L.phase[b] = d.phase[b]; L.dmid[b] = d.dmid[b]; L.nlim[b] = d.nlim[b]; L.dlim[b] = d.dlim[b];
    L.vlo[b] = d.vlolim[b]; L.vhi[b] = d.vhilim[b];
    int c = b < L.off[1] ? 0 : b < L.off[2] ? 1 : b < L.off[3] ? 2 : 3;
    L.cat_of[b] = c; L.loc_of[b] = b - L.off[c];
  }
  for (int k = 0; k < d.nprc; ++k) L.precpbins[k] = d.precpbins[k];
  // cell-record slots
  {
    int sl = kThermo;
    for (int r = 0; r < kThermo; ++r) { L.slot_cat[r] = -1; L.slot_row[r] = 0; L.slot_mass[r] = 0; }
    for (int b = 0; b < L.nt; ++b) {
      const int c = L.cat_of[b], loc = L.loc_of[b], nbc = L.nb[c];
      const int nsw = c < 3 ? L.nwet : L.nspec;
      L.sbase[b] = sl;
      L.slot_cat[sl] = (int16_t)c; L.slot_row[sl] = loc; L.slot_mass[sl] = 0; ++sl;
      for (int q = 0; q < nsw; ++q) {
        L.slot_cat[sl] = (int16_t)c; L.slot_row[sl] = q * nbc + loc; L.slot_mass[sl] = 1; ++sl;
      }
    }
    L.nslot = sl;
    L.nout = sl - kThermo + 1;
  }

  L.cld_a0 = d.ica.cur; L.cld_a1 = d.fca.cur; L.cld_b0 = d.icb.cur; L.cld_b1 = d.fcb.cur;
  L.lscoag = p.lscoag; L.lscnd = p.lscnd;
  L.lsdistupdate = p.lsdistupdate;
  L.autoconv2 = p.lsauto && p.lsauto_mode == 2;
  L.lsactiv = p.lsactiv;
  L.lscndh2oae = p.lscndh2oae; L.lscndh2ocl = p.lscndh2ocl;
  L.lscndh2oic = level == 5 ? p.lscndh2oic : 0;
  L.eddy_dis = p.eddy_dis; L.Eiagg_max = p.Eiagg_max; L.Eiagg_min = p.Eiagg_min;
  L.ice_alpha = p.ice_alpha; L.ice_beta = p.ice_beta; L.ice_gamma = p.ice_gamma; L.ice_sigma = p.ice_sigma;
  L.charge_factor = p.lcharge ? 1. + p.charge_coll_enh : 1.;
  L.eddy_075 = std::pow(p.eddy_dis, 0.75);

  const int ice = level == 5 && d.nice > 0;
  const int cgaa = p.lscgaa, cgcc = p.lscgcc, cgpp = p.lscgpp, cgca = p.lscgca, cgpa = p.lscgpa,
            cgpc = p.lscgpc;
  const int cgii = ice && p.lscgii, cgia = ice && p.lscgia, cgic = ice && p.lscgic, cgip = ice && p.lscgip;
  L.any_ice_coag = cgii || cgia || cgic || cgip;

  // ---- pairs: blocks and canonical orientation of update_coagulation_kernels
  //      (mo_salsa_coagulation_kernels.f90:24-90): self blocks (m <= n), then cross.
  for (int i = 0; i < kMaxBin; ++i)
    for (int j = 0; j < kMaxBin; ++j) L.pidx[i][j] = -1;
  int np = 0;
  auto add = [&](int a, int b) {
    L.pb1[np] = (int16_t)a; L.pb2[np] = (int16_t)b;
    L.poff[np] = std::fabs(d.dmid[a] - d.dmid[b]) / d.dmid[a] < 1.e-2;   // kernels.f90:273
    L.pidx[a][b] = L.pidx[b][a] = (int16_t)np;
    L.pw[np] = (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)L.poff[np] << 16);
    ++np;
  };
  auto self_block = [&](int off, int n) { for (int m = 0; m < n; ++m) for (int k = m; k < n; ++k) add(off + m, off + k); };
  auto cross_block = [&](int o1, int n1, int o2, int n2) {
    for (int m = 0; m < n1; ++m) for (int k = 0; k < n2; ++k) add(o1 + m, o2 + k);
  };
  const int A = L.off[0], C = L.off[1], R = L.off[2], I = L.off[3];
  const int nb = d.nbins, nc = d.ncld, nr = d.nprc, ni = L.nb[3];
  if (cgaa) self_block(A, nb);
  L.npair_aa = np;
  L.pf0 = np;
  if (cgcc) self_block(C, nc);
  L.ncc = np - L.pf0;
  if (cgpp) self_block(R, nr);
  if (cgii) self_block(I, ni);
  if (cgca) cross_block(A, nb, C, nc);
  if (cgpa) cross_block(A, nb, R, nr);
  if (cgia) cross_block(A, nb, I, ni);
  if (cgpc) cross_block(C, nc, R, nr);
  if (cgic) cross_block(C, nc, I, ni);
  if (cgip) cross_block(R, nr, I, ni);
  L.npair = np;
  // The per-bin "offset particle" tables assume the self-offset fires only
  // between liquid bins or on an ice diagonal.
  for (int q = 0; q < np; ++q) {
    if (!L.poff[q]) continue;
    int a = L.pb1[q], b = L.pb2[q];
    if ((d.phase[a] == 4 || d.phase[b] == 4) && a != b) {
      set_error("bin layout places an ice bin within 1% of another bin's dmid (unsupported)");
      return SALSA_ERR_ARG;
    }
  }

  // ---- sweep program (coag_ice, coag_precp, coag_aero, coag_cloud) ----
  int nt = 0;
  const salsa_parbin ica = d.ica, fca = d.fca, icb = d.icb, fcb = d.fcb;
  const int pfmode = p.lsauto && p.lsauto_mode == 1;
  const int nwet = L.nwet;
  const int any_aero = cgaa || cgca || cgpa || cgia;
  const int any_cloud = cgcc || cgca || cgpc || cgic;
  const int any_precp = cgpp || cgpa || cgpc || cgip;
  if (p.lscoag) {
    if (L.any_ice_coag) {          // coag_ice, processes.f90:476-535
      for (int kk = 0; kk < ni; ++kk) {
        DevTarget &T = L.tgt[nt++];
        std::memset(&T, 0, sizeof(T));
        T.t = I + kk; T.cat = 3; T.flags = TF_ICE; T.nupd = L.nspec;
        if (cgii) {
          if (kk < ni - 1) T.sink_eagg |= bits(I, kk + 1, ni - 1);
          T.flags |= TF_SELF_HALF_EAGG;
        }
        if (cgia) T.src_pc_wa |= bits(A, d.in1a, d.fn2b);
        if (cgic) T.src_pc_ri |= bits(C, ica.cur, fcb.cur);
        if (cgip) T.src_pc_ri |= bits(R, 0, nr - 1);
        if (cgii && kk > 0) T.src_eagg |= bits(I, 0, kk - 1);
      }
    }
    if (any_precp) {               // coag_precp, processes.f90:418-470
      for (int kk = 0; kk < nr; ++kk) {
        DevTarget &T = L.tgt[nt++];
        std::memset(&T, 0, sizeof(T));
        T.t = R + kk; T.cat = 2; T.nupd = nwet;
        if (cgpp) {
          if (kk < nr - 1) { T.sink |= bits(R, kk + 1, nr - 1); T.flags |= TF_PRECIP_SELFCOLL; }
          else T.flags |= TF_SELF_HALF;
        }
        if (cgip) T.sink |= bits(I, 0, ni - 1);
        if (cgpa) T.src |= bits(A, d.in1a, d.fn2b);
        if (cgpc) T.src |= bits(C, 0, nc - 1);
        if (cgpp && kk > 0) T.src |= bits(R, 0, kk - 1);
      }
    }
    L.it_aero0 = nt;
    if (any_aero) {                // coag_aero, processes.f90:24-215
      for (int kk = d.in1a; kk <= d.fn1a; ++kk) {
        DevTarget &T = L.tgt[nt++];
        std::memset(&T, 0, sizeof(T));
        T.t = A + kk; T.cat = 0; T.nupd = nwet;
        int icla = ica.cur + std::max(kk - ica.par, 0);
        int iclb = icb.cur + (icla - ica.cur);
        if (cgaa) { T.sink |= bits(A, kk + 1, d.fn2b); T.flags |= TF_SELF_HALF; }
        if (cgca) { T.sink |= bits(C, icla, fca.cur) | bits(C, iclb, fcb.cur); }
        if (cgpa) T.sink |= bits(R, 0, nr - 1);
        if (cgia) T.sink |= bits(I, 0, ni - 1);
        if (cgaa && kk > d.in1a) T.src |= bits(A, d.in1a, kk - 1);
        if (cgca) {
          icla = ica.cur + (kk - ica.par) - 1;
          iclb = icb.cur + (icla - ica.cur);
          if (icla >= ica.cur) T.src |= bits(C, ica.cur, icla);
          if (iclb >= icb.cur) T.src |= bits(C, icb.cur, iclb);
        }
      }
      for (int kk = d.in2a; kk <= d.fn2a; ++kk) {
        DevTarget &T = L.tgt[nt++];
        std::memset(&T, 0, sizeof(T));
        T.t = A + kk; T.cat = 0; T.nupd = nwet;
        int index_b = d.in2b + (kk - d.in2a);
        int dd = std::max(kk - ica.par, 0);
        int icla = ica.cur + dd, iclb = icb.cur + dd;
        if (cgaa) {
          if (kk < d.fn2a) T.sink |= bits(A, kk + 1, d.fn2a) | bits(A, index_b, d.fn2b);
          T.flags |= TF_SELF_HALF;
        }
        if (cgca) T.sink |= bits(C, icla, fca.cur) | bits(C, iclb, fcb.cur);
        if (cgpa) T.sink |= bits(R, 0, nr - 1);
        if (cgia) T.sink |= bits(I, 0, ni - 1);
        if (cgaa) T.src |= bits(A, d.in1a, kk - 1) | bits(A, d.in2b, index_b);
        if (cgca) {
          icla = ica.cur + (kk - ica.par) - 1;
          iclb = icb.cur + (icla - ica.cur);
          if (icla >= ica.cur) T.src |= bits(C, ica.cur, icla);
          if (iclb >= icb.cur) T.src |= bits(C, icb.cur, iclb);
        }
      }
      for (int kk = d.in2b; kk <= d.fn2b; ++kk) {
        DevTarget &T = L.tgt[nt++];
        std::memset(&T, 0, sizeof(T));
        T.t = A + kk; T.cat = 0; T.nupd = nwet;
        int index_a = d.in2a + (kk - d.in2b);
        int dd = std::max(kk - icb.par, 0);
        int icla = ica.cur + dd, iclb = icb.cur + dd;
        if (cgaa) {
          if (kk < d.fn2b) T.sink |= bits(A, kk + 1, d.fn2b) | bits(A, index_a, d.fn2a);
          T.flags |= TF_SELF_HALF;
        }
        if (cgca) T.sink |= bits(C, icla, fca.cur) | bits(C, iclb, fcb.cur);
        if (cgpa) T.sink |= bits(R, 0, nr - 1);
        if (cgia) T.sink |= bits(I, 0, ni - 1);
        if (cgaa) T.src |= bits(A, d.in2b, kk - 1) | bits(A, d.in1a, index_a);
        if (cgca) {
          iclb = icb.cur + (kk - icb.par) - 1;
          icla = ica.cur + (iclb - icb.cur);
          if (icla >= ica.cur) T.src |= bits(C, ica.cur, icla);
          if (iclb >= icb.cur) T.src |= bits(C, icb.cur, iclb);
        }
      }
    }
    L.it_aero1 = nt;
    if (any_cloud) {               // coag_cloud, processes.f90:221-412
      for (int series = 0; series < 2; ++series) {
        const int cur0 = series ? icb.cur : ica.cur, cur1 = series ? fcb.cur : fca.cur;
        const int oth0 = series ? ica.cur : icb.cur, oth1 = series ? fca.cur : fcb.cur;
        for (int kk = cur0; kk <= cur1; ++kk) {
          DevTarget &T = L.tgt[nt++];
          std::memset(&T, 0, sizeof(T));
          T.t = C + kk; T.cat = 1; T.nupd = nwet;
          const int index_o = oth0 + (kk - cur0);
          if (cgca && kk < cur1) {
            T.sink |= bits(A, ica.par + (kk - cur0) + 1, d.fn2a);
            T.sink |= bits(A, icb.par + (kk - cur0) + 1, d.fn2b);
          }
          if (cgcc) {
            if (kk < cur1) T.sink |= bits(C, kk + 1, cur1) | bits(C, index_o, oth1);
            T.flags |= pfmode ? TF_PF_SELF : TF_SELF_HALF;
          }
          if (cgpc) T.sink |= bits(R, 0, nr - 1);
          if (cgic) T.sink |= bits(I, 0, ni - 1);
          if (cgca) {
            T.src |= bits(A, d.in1a, ica.par + (kk - cur0));
            T.src |= bits(A, d.in2b, icb.par + (kk - cur0));
          }
          if (cgcc && kk > cur0) {
            if (pfmode) {
              T.flags |= TF_PF;
              T.pf_own = bits(C, cur0, kk - 1);  T.fix_own = bits(C, cur0, cur1);
              T.pf_oth = bits(C, oth0, index_o); T.fix_oth = bits(C, oth0, oth1);
            } else {
              T.src |= bits(C, cur0, kk - 1) | bits(C, oth0, index_o);
            }
          }
        }
      }
    }
  }
  L.ntarget = nt;
  // entry lists: every bin named by a mask of the target, with its roles
  {
    int ne = 0;
    for (int it = 0; it < nt; ++it) {
      DevTarget &T = L.tgt[it];
      T.ent0 = ne;
      const uint64_t m[10] = {T.sink, T.sink_eagg, T.src, T.src_eagg, T.src_pc_wa, T.src_pc_ri,
                              T.pf_own, T.pf_oth, T.fix_own, T.fix_oth};
      for (int pass = 0; pass < 2; ++pass) {        // aerosol collectors first
        for (int j = 0; j < L.nt; ++j) {
          if ((L.cat_of[j] == 0) != (pass == 0)) continue;
          uint32_t role = 0;
          for (int r = 0; r < 10; ++r)
            if ((m[r] >> j) & 1ull) role |= 1u << r;
          if (!role) continue;
          const int pid = L.pidx[T.t][j];
          if (pid < 0) continue;                    // kernel block switched off: K = 0
          L.ent[ne++] = (uint32_t)j | (role << 6) | ((uint32_t)pid << 16);
        }
        if (pass == 0) T.nent_aero = ne - T.ent0;
      }
      T.nent = ne - T.ent0;
    }
    L.nent_total = ne;
  }
  return SALSA_OK;
}

}  // namespace salsa
