// salsa_activation.h -- cloud activation of one cell: interstitial
// (cloud_activation with lsactiv%mode 1, mo_salsa.f90:115-119; actInterst,
// mo_salsa_cloud.f90:859-1195), serial over its 2a/2b aerosol bins, and at
// cloud base (mode 2; ActCloudBase and activate3, at the end of the file).  R is the state
// type (double in the step kernel, a dual number in the derivative kernel);
// N[b] and V[b * vs + s] are the cell's numbers and volumes in SALSA units,
// rv and rs the vapour and saturation mixing ratios after condensation.
//
// When the cell is supersaturated, each aerosol bin gets piecewise linear
// profiles of wet volume and of number density over its dry volume, from its
// own and its neighbours' means (a regime's first and last bins get
// extrapolated neighbours); the part of the bin whose wet volume exceeds the
// critical volume of the cell's supersaturation moves to the parallel cloud
// bin (cloud bin k <-> aerosol bin in2a + k), every bin's share computed
// from the state before activation.  The chargeTime and INdef
// bookkeeping of the reference (:1161-1187) is not carried: the collision
// kernels assume chargeTime = chargeTimeMax, and ice nucleation is not on the
// path.  tmp holds 2 * kMaxSpec values of R (work space: LDS in the step
// kernel, which keeps the kernel free of private arrays).  act_interst_cell
// returns the activated number (rateDiag%Activation%numc).
#pragma once

#include "salsa_dual.h"
#include "salsa_internal.h"

namespace salsa {

template <class R> struct ActOps {
  static __device__ __forceinline__ R max_(R a, R b) { return fmax(a, b); }
  static __device__ __forceinline__ R min_(R a, R b) { return fmin(a, b); }
  static __device__ __forceinline__ R c(double x) { return x; }
  static __device__ __forceinline__ double val(R x) { return x; }
};
template <> struct ActOps<Dual<double>> {
  using D = Dual<double>;
  static __device__ __forceinline__ D max_(D a, D b) { return dmax(a, b); }
  static __device__ __forceinline__ D min_(D a, D b) { return dmin(a, b); }
  static __device__ __forceinline__ D c(double x) { return dconst(x); }
  static __device__ __forceinline__ double val(D x) { return x.v; }
};

// intgN / intgV, mo_salsa_cloud.f90:1381-1397
template <class R> __device__ __forceinline__ R act_intg_n(R k, R c, R lo, R hi) {
  using O = ActOps<R>;
  return 0.5 * k * O::max_(hi * hi - lo * lo, O::c(0.)) + c * O::max_(hi - lo, O::c(0.));
}
template <class R> __device__ __forceinline__ R act_intg_v(R k, R c, R lo, R hi) {
  using O = ActOps<R>;
  return (1. / 3.) * k * O::max_(hi * hi * hi - lo * lo * lo, O::c(0.)) + 0.5 * c * O::max_(hi * hi - lo * lo, O::c(0.));
}

template <class R>
__device__ double act_interst_cell(const DevLayout &D, R *N, R *V, int vs, R rv, double rs, double T, R *tmp) {
  using O = ActOps<R>;
  const double THvol = 1.e-28;
  const int ndry = D.ndry, nwet = D.nwet, iwa = D.iwa;
  if (!(rv / rs > 1.000)) return 0.;
  const double zkelvin = 4. * kMwa * kSurfw0 / (kRg * kRhowa * T);
  const R zdcstar = 2. * zkelvin / (3. * ((rv / rs) - 1.));
  const R zvcstar = kPi6 * (zdcstar * zdcstar * zdcstar);
  const int A = D.off[0], C = D.off[1], ncld = D.nb[1];
  double ntot = 0.;
  // The reference computes every bin's activation from the state before any
  // of it is applied (:1142-1189).  Bin k reads aerosol bins k-1, k and k+1,
  // so bin k's transfer is applied after bin k+1 has been computed.
  int pk = -1;
  R pn = O::c(0.);
  R *pv = tmp, *av = tmp + kMaxSpec;
  auto apply = [&]() {
    if (pk < 0) return;
    const int a = A + D.in2a + pk, cl = C + pk;
    N[a] = O::max_(O::c(0.), N[a] - pn);
    N[cl] = N[cl] + pn;
    for (int s = 0; s < nwet; ++s) {
      V[a * vs + s] = O::max_(O::c(0.), V[a * vs + s] - pv[s]);
      V[cl * vs + s] = V[cl * vs + s] + pv[s];
    }
    pk = -1;
  };
  auto mean = [&](int b, int ns) {
    R t = V[b * vs];
    for (int s = 1; s < ns; ++s) t = t + V[b * vs + s];
    return t;
  };
  for (int k = 0; k < ncld; ++k) {
    const int ab = D.in2a + k, a = A + ab;
    double vol_sol = 0.;
    for (int s = 0; s < ndry; ++s)
      if (D.dry_id[s] != 2 && D.dry_id[s] != 3) vol_sol += O::val(V[a * vs + s]);   // BC, DU insoluble
    if (N[a] < D.nlim[a] || vol_sol < O::val(N[a]) * THvol || V[a * vs + iwa] < O::val(N[a]) * THvol) {
      apply();
      continue;
    }
    const double vm = kPi6 * (D.dmid[a] * D.dmid[a] * D.dmid[a]);
    const double vrlo = D.vlo[a] / vm, vrhi = D.vhi[a] / vm;   // mo_salsa_init.f90:186-227
    const R Nmid = N[a];
    const R Vwmid = mean(a, nwet) / Nmid, Vmid = mean(a, ndry) / Nmid;
    const R Vlo = Vmid * vrlo, Vhi = Vmid * vrhi;
    R Nim1, Vim1, Vlom1, Vhim1, Vwim1, Nip1, Vip1, Vlop1, Vhip1, Vwip1;
    if (ab == D.in1a || ab == D.in2b) {                       // first bin of a regime, :982-988
      Nim1 = O::c(D.nlim[a]); Vim1 = Vlo / 2.; Vlom1 = O::c(0.); Vhim1 = Vlo; Vwim1 = Vim1;
    } else {
      const int m = a - 1;
      const double vmm = kPi6 * (D.dmid[m] * D.dmid[m] * D.dmid[m]);
      Nim1 = N[m];
      if (Nim1 > D.nlim[m]) {
        Vim1 = mean(m, ndry) / Nim1; Vwim1 = mean(m, nwet) / Nim1;
      } else {
        Vim1 = O::c(vmm); Vwim1 = O::c(vmm);
      }
      Vlom1 = Vim1 * (D.vlo[m] / vmm);
      Vhim1 = Vim1 * (D.vhi[m] / vmm);
    }
    if (ab == D.fn2a || ab == D.fn2b) {                       // last bin of a regime, :1001-1007
      Nip1 = O::c(D.nlim[a]); Vip1 = Vhi + 0.5 * (Vhi - Vlo); Vlop1 = Vhi; Vhip1 = Vhi + (Vhi - Vlo); Vwip1 = Vip1;
    } else {
      const int m = a + 1;
      const double vmm = kPi6 * (D.dmid[m] * D.dmid[m] * D.dmid[m]);
      Nip1 = N[m];
      if (Nip1 > D.nlim[m]) {
        Vip1 = mean(m, ndry) / Nip1; Vwip1 = mean(m, nwet) / Nip1;
      } else {
        Vip1 = O::c(vmm); Vwip1 = O::c(vmm);
      }
      Vlop1 = Vip1 * (D.vlo[m] / vmm);
      Vhip1 = Vip1 * (D.vhi[m] / vmm);
    }
    Vip1 = O::max_(Vhi, Vip1);
    Vim1 = O::min_(Vlo, Vim1);
    // wet volume profile, :1027-1037
    R zs1 = (Vwmid - O::max_(Vwim1, O::c(0.))) / (Vmid - Vim1);
    R zs2 = (O::max_(Vwip1, O::c(0.)) - Vwmid) / (Vip1 - Vmid);
    const R V01 = Vwmid - zs1 * Vmid, V02 = Vwmid - zs2 * Vmid;
    const R Vwlo = O::max_(V01 + zs1 * Vlo, Vlo), Vwhi = O::max_(V02 + zs2 * Vhi, Vhi);
    // integration limits, :1050-1097
    R vint1, vint2;
    bool ir[4];
    if (zvcstar < Vwlo && zvcstar < Vwmid && zvcstar < Vwhi) {
      vint1 = Vlo; vint2 = Vhi;
      ir[0] = ir[1] = ir[2] = ir[3] = true;
    } else if (zvcstar > Vwlo && zvcstar > Vwmid && zvcstar > Vwhi) {
      vint1 = O::c(999.); vint2 = O::c(999.);
      ir[0] = ir[1] = ir[2] = ir[3] = false;
    } else {
      R vcut = (zvcstar - V01) / zs1;
      vint1 = (vcut < Vlo || vcut > Vmid) ? Vlo : vcut;
      vcut = (zvcstar - V02) / zs2;
      vint2 = (vcut < Vmid || vcut > Vhi) ? Vmid : vcut;
      ir[0] = Vwlo > zvcstar; ir[1] = Vwmid > zvcstar; ir[2] = Vwmid > zvcstar; ir[3] = Vwhi > zvcstar;
    }
    // number density profile and the integrals, :1099-1139
    const R dNim1 = Nim1 / (Vhim1 - Vlom1), dNip1 = Nip1 / (Vhip1 - Vlop1), dNmid = Nmid / (Vhi - Vlo);
    zs1 = (dNmid - dNim1) / (Vmid - Vim1);
    zs2 = (dNip1 - dNmid) / (Vip1 - Vmid);
    const R N01 = dNmid - zs1 * Vmid, N02 = dNmid - zs2 * Vmid;
    const R Nnorm = act_intg_n(zs1, N01, Vlo, Vmid) + act_intg_n(zs2, N02, Vmid, Vhi);
    const R Vnorm = act_intg_v(zs1, N01, Vlo, Vmid) + act_intg_v(zs2, N02, Vmid, Vhi);
    R zactvol = O::c(0.), Nact = O::c(0.);
    if (ir[0]) {
      Nact = Nact + (Nmid / Nnorm) * act_intg_n(zs1, N01, Vlo, vint1);
      zactvol = zactvol + (Nmid * Vmid / Vnorm) * act_intg_v(zs1, N01, Vlo, vint1);
    }
    if (ir[1]) {
      Nact = Nact + (Nmid / Nnorm) * act_intg_n(zs1, N01, vint1, Vmid);
      zactvol = zactvol + (Nmid * Vmid / Vnorm) * act_intg_v(zs1, N01, vint1, Vmid);
    }
    if (ir[2]) {
      Nact = Nact + (Nmid / Nnorm) * act_intg_n(zs2, N02, Vmid, vint2);
      zactvol = zactvol + (Nmid * Vmid / Vnorm) * act_intg_v(zs2, N02, Vmid, vint2);
    }
    if (ir[3]) {
      Nact = Nact + (Nmid / Nnorm) * act_intg_n(zs2, N02, vint2, Vhi);
      zactvol = zactvol + (Nmid * Vmid / Vnorm) * act_intg_v(zs2, N02, vint2, Vhi);
    }
    // this bin's transfer, :1142-1144; the previous bin's is applied now
    const R an = O::min_(Nact, Nmid);
    for (int s = 0; s < nwet; ++s) av[s] = O::min_(zactvol * (V[a * vs + s] / (Vmid * Nmid)), V[a * vs + s]);
    ntot += O::val(an);
    apply();
    pk = k;
    pn = an;
    for (int s = 0; s < nwet; ++s) pv[s] = av[s];
  }
  apply();
  return ntot;
}

// ---------------------------------------------------------------------------
// Cloud-base activation (cloud_activation with lsactiv%mode 2; ActCloudBase,
// mo_salsa_cloud.f90:161-375 after Abdul-Razzak and Ghan 2002, getSolute
// :113-157, activate3 :1199-1378) of one cell, serial on one lane: for an
// updraft w > 0 the parcel's maximum supersaturation, the critical dry volume
// in the first partly activated 1a/2a bin, and per cloud bin the number and
// volume of its parallel aerosol bin above that volume from piecewise linear
// number-density profiles.  The state is not changed; the activated droplets
// go to the LES fields pa_nactd / pa_vactd (mo_salsa_driver.f90:360-366):
// nout[cb * n3] = numc / dn, vout[(s * ncld + cb) * n3] = volc_s rholiq_s / dn
// for the wet species s (zero without updraft).  Per-bin quantities (moles of
// solute, critical diameters) are formed where they are used, so the lane
// keeps no arrays; contraction is off so that every product rounds as the
// reference's.
// ---------------------------------------------------------------------------
struct CbConst { double x, aa, bb, s_max; };

// moles of solute per particle of aerosol master bin a (getSolute :136-151)
__device__ inline double cb_solute(const DevLayout &D, const double *N, const double *V, int vs, int a) {
#pragma clang fp contract(off)
  if (!(N[a] > D.nlim[a])) return 0.;
  double ns = 0.;
  for (int nn = 0; nn < D.ndry; ++nn) ns = ns + D.diss[nn] * V[a * vs + nn] * D.rholiq[nn] / D.mm[nn];
  return ns / N[a];
}

// ActCloudBase's per-bin test (:327-351) on aerosol master bin a: counted or
// not, the activated fraction and the critical diameters
__device__ inline bool cb_bin(const DevLayout &D, const double *N, const double *V, int vs, int a, const CbConst &k,
                              double &frac, double &dcrit, double &dcrlo, double &dcrhi) {
#pragma clang fp contract(off)
  frac = dcrit = dcrlo = dcrhi = 0.;
  if (N[a] < D.nlim[a]) return false;
  const double ns = cb_solute(D, N, V, vs, a);
  const double dm3 = D.dmid[a] * D.dmid[a] * D.dmid[a];
  const double nshi = ns * (D.vhi[a] / (kPi6 * dm3));        // vratiohi, mo_salsa_init.f90:186
  const double sil = ::exp(::sqrt(k.x / nshi)) - 1.;
  if (k.s_max < sil) return false;
  const double nslo = ns * (D.vlo[a] / (kPi6 * dm3));        // vratiolo, :187
  const double siu = ::exp(::sqrt(k.x / nslo)) - 1.;
  frac = fmin(1., ::log(k.s_max / sil) / ::log(siu / sil));
  dcrlo = 2. * ::sqrt(3. * nslo * k.bb / k.aa);
  dcrhi = 2. * ::sqrt(3. * nshi * k.bb / k.aa);
  dcrit = 2. * ::sqrt(3. * ns * k.bb / k.aa);
  return true;
}

__device__ inline void act_cloud_base_cell(const DevLayout &D, const double *N, const double *V, int vs, double T,
                                           double P, double w, double dn, double *nout, double *vout, int64_t n3) {
#pragma clang fp contract(off)
  const int A = D.off[0], ncld = D.nb[1], ndry = D.ndry, nwet = D.nwet, iwa = D.iwa;
  const double cpa = 1005.;                                  // mo_submctl.f90:267
  auto put_zero = [&](int cb) {                              // pa_nactd / pa_vactd of a bin without activation
    nout[cb * n3] = 0. / dn;
    for (int s = 0; s < nwet; ++s) vout[((int64_t)s * ncld + cb) * n3] = 0. * D.rholiq[s] / dn;
  };
  CbConst k;
  k.bb = 6. * kMwa / (kPi * kRhowa);
  k.aa = 4. * kMwa * kSurfw0 / (kRg * kRhowa * T);
  k.x = 4. * (k.aa * k.aa * k.aa) / (27. * k.bb);
  double ntot = 0., sum1 = 0.;
  bool go = w > 0.;
  if (go) {
    for (int b = D.in1a; b <= D.fn2b; ++b) ntot = ntot + N[A + b];
    for (int b = D.in1a; b <= D.fn2b; ++b) {
      const double scrit = ::exp(::sqrt(k.x / fmax(kEpsR8, cb_solute(D, N, V, vs, A + b)))) - 1.;
      sum1 = sum1 + N[A + b] / ::pow(scrit, 2. / 3.);
    }
    go = !(ntot < D.nlim[A]);
  }
  double zdcstar = 0.;
  int bcrita = D.fn2a;
  if (go) {
    const double Lh = 2.501e6 - 2370. * (T - 273.15);
    const double a1 = 1. - (373.15 / T);
    const double a12 = a1 * a1;
    const double ps = 101325. * ::exp(13.3185 * a1 - 1.976 * a12 - 0.6445 * (a12 * a1) - 0.1299 * (a12 * a12));
    const double alpha = kGrav * kMwa * Lh / (cpa * kRg * (T * T)) - kGrav * kMair / (kRg * T);
    const double gamma = kRg * T / (ps * kMwa) + kMwa * (Lh * Lh) / (cpa * P * kMair * T);
    const double x1 = P / 101325.;
    const double dv1 = 1.e-4 * (0.211 / x1) * ::pow(T / 273., 1.94);
    const double ka1 = 1.e-3 * (4.39 + 0.071 * T);
    const double Gc =
        1. / (kRhowa * kRg * T / (ps * dv1 * kMwa) + Lh * kRhowa / (ka1 * T) * (Lh * kMwa / (T * kRg) - 1.));
    const double r = ntot / sum1;
    const double s_eff = ::sqrt(r * r * r);
    const double avg = alpha * w / Gc;
    const double theta = (::sqrt(avg * avg * avg)) / (2. * kPi * kRhowa * gamma * ntot);
    const double khi = (2. / 3.) * k.aa * ::sqrt(avg);
    const double kt = khi / theta;
    k.s_max = s_eff / ::sqrt(0.5 * ::sqrt(kt * kt * kt) + ::pow((s_eff * s_eff) / (theta + 3. * khi), 0.75));
    zdcstar = 2. * k.aa / (3. * k.s_max);
    // activate3 only acts if some bin has a critical diameter (:1239-1240)
    bool any = false;
    double f, d0, d1, d2;
    for (int b = D.in1a; b <= D.fn2b; ++b)
      if (cb_bin(D, N, V, vs, A + b, k, f, d0, d1, d2) && d0 >= kEpsR8) any = true;
    for (int b = D.in1a; b <= D.fn2a; ++b) {
      cb_bin(D, N, V, vs, A + b, k, f, d0, d1, d2);
      if (f < 1. && f > 0.) { bcrita = b; break; }
    }
    go = any;
  }
  if (!go) {
    for (int cb = 0; cb < ncld; ++cb) put_zero(cb);
    return;
  }
  double zvcstar = 0.;
  {
    const int a = A + bcrita;
    double f, dcrit, dcrlo, dcrhi;
    cb_bin(D, N, V, vs, a, k, f, dcrit, dcrlo, dcrhi);
    double Vmid;
    if (N[a] < D.nlim[a]) {
      Vmid = kPi6 * (D.dmid[a] * D.dmid[a] * D.dmid[a]);
    } else {
      double sd = 0.;
      for (int s = 0; s < ndry; ++s) sd = sd + V[a * vs + s];
      Vmid = sd / fmax(D.nlim[a], N[a]);
    }
    const double Vhi = D.vhi[a], Vlo = D.vlo[a];
    const double cs3 = zdcstar * zdcstar * zdcstar, cr3 = dcrit * dcrit * dcrit;
    if (zdcstar >= dcrit) {
      const double zhlp = (kPi6 * cs3 - kPi6 * cr3) / fmax(kEpsR8, kPi6 * (dcrhi * dcrhi * dcrhi) - kPi6 * cr3);
      zvcstar = Vmid + zhlp * (Vhi - Vmid);
    } else if (zdcstar < dcrit) {
      const double zhlp = (kPi6 * cr3 - kPi6 * cs3) / fmax(kEpsR8, kPi6 * cr3 - kPi6 * (dcrlo * dcrlo * dcrlo));
      zvcstar = Vmid - zhlp * (Vmid - Vlo);
    }
    zvcstar = fmax(zvcstar, D.vlo[a]);
    zvcstar = fmin(zvcstar, D.vhi[a]);
  }
  for (int cb = 0; cb < ncld; ++cb) {
    const int ab = D.in2a + cb, a = A + ab;                 // ica%par + cb (a-series), icb%par + (cb - icb%cur)
    if (N[a] < D.nlim[a]) {
      put_zero(cb);
      continue;
    }
    const double Nmid = fmax(N[a], D.nlim[a]);
    double sd = 0.;
    for (int s = 0; s < ndry; ++s) sd = sd + V[a * vs + s];
    const double Vmid = sd / Nmid;
    const double Vlo = D.vlo[a], Vhi = D.vhi[a];
    double Nim1, Vim1, Vlom1, Vhim1, Nip1, Vip1, Vlop1, Vhip1;
    if (ab == D.in1a || ab == D.in2b) {
      Nim1 = D.nlim[a]; Vim1 = Vlo / 2.; Vlom1 = 0.; Vhim1 = Vlo;
    } else {
      Nim1 = fmax(N[a - 1], D.nlim[a - 1]);
      if (Nim1 > D.nlim[a - 1]) {
        double s1 = 0.;
        for (int s = 0; s < ndry; ++s) s1 = s1 + V[(a - 1) * vs + s];
        Vim1 = s1 / Nim1;
      } else {
        Vim1 = kPi6 * (D.dmid[a - 1] * D.dmid[a - 1] * D.dmid[a - 1]);
      }
      Vlom1 = D.vlo[a - 1]; Vhim1 = D.vhi[a - 1];
    }
    if (ab == D.fn2a || ab == D.fn2b) {
      Nip1 = D.nlim[a]; Vip1 = Vhi + 0.5 * (Vhi - Vlo); Vlop1 = Vhi; Vhip1 = Vhi + (Vhi - Vlo);
    } else {
      Nip1 = fmax(N[a + 1], D.nlim[a + 1]);
      if (Nip1 > D.nlim[a + 1]) {
        double s1 = 0.;
        for (int s = 0; s < ndry; ++s) s1 = s1 + V[(a + 1) * vs + s];
        Vip1 = s1 / Nip1;
      } else {
        Vip1 = kPi6 * (D.dmid[a + 1] * D.dmid[a + 1] * D.dmid[a + 1]);
      }
      Vlop1 = D.vlo[a + 1]; Vhip1 = D.vhi[a + 1];
    }
    Vip1 = fmax(Vlop1, fmin(Vip1, Vhip1));
    Vim1 = fmax(Vlom1, fmin(Vim1, Vhim1));
    const double dNim1 = Nim1 / (Vhim1 - Vlom1), dNip1 = Nip1 / (Vhip1 - Vlop1), dNmid = Nmid / (Vhi - Vlo);
    const double zs1 = (dNmid - dNim1) / (Vmid - Vim1), zs2 = (dNip1 - dNmid) / (Vip1 - Vmid);
    const double N01 = dNmid - zs1 * Vmid, N02 = dNip1 - zs2 * Vip1;
    const double Nnorm = act_intg_n<double>(zs1, N01, Vlo, Vmid) + act_intg_n<double>(zs2, N02, Vmid, Vhi);
    const double Vnorm = act_intg_v<double>(zs1, N01, Vlo, Vmid) + act_intg_v<double>(zs2, N02, Vmid, Vhi);
    double n = 0., zactvol = 0.;
    const bool lo_side = zvcstar < Vmid, hi_side = zvcstar >= Vmid;
    if (lo_side) {
      const double zvcint = fmax(zvcstar, Vlo);
      n = (Nmid / Nnorm) * (act_intg_n<double>(zs1, N01, zvcint, Vmid) + act_intg_n<double>(zs2, N02, Vmid, Vhi));
      zactvol =
          (Nmid * Vmid / Vnorm) * (act_intg_v<double>(zs1, N01, zvcint, Vmid) + act_intg_v<double>(zs2, N02, Vmid, Vhi));
    } else if (hi_side) {
      const double zvcint = fmin(zvcstar, Vhi);
      n = (Nmid / Nnorm) * (act_intg_n<double>(zs2, N02, zvcint, Vhi));
      zactvol = (Nmid * Vmid / Vnorm) * (act_intg_v<double>(zs2, N02, zvcint, Vhi));
    }
    n = fmax(0., n);
    double f, dcrit, d1, d2;
    cb_bin(D, N, V, vs, a, k, f, dcrit, d1, d2);
    const double q = 3.e-6 / fmax(kEpsR8, dcrit);
    // the water of the new droplets is set from the critical diameter (:1370-1371)
    const double vw = n * kPi6 * (dcrit * dcrit * dcrit) * fmin(2., q * q);
    nout[cb * n3] = n / dn;
    for (int s = 0; s < nwet; ++s) {
      double v = 0.;
      if (lo_side || hi_side) v = zactvol * (V[a * vs + s] / (Vmid * Nmid));
      v = fmax(0., v);
      if (s == iwa) v = vw;
      vout[((int64_t)s * ncld + cb) * n3] = v * D.rholiq[s] / dn;
    }
  }
}

}  // namespace salsa
