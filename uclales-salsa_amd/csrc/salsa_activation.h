// salsa_activation.h -- interstitial cloud activation (cloud_activation with
// lsactiv%mode 1, mo_salsa.f90:115-119; actInterst, mo_salsa_cloud.f90:
// 859-1195) for one cell, serial over its 2a/2b aerosol bins.  R is the state
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
// kernel, which keeps the kernel free of private arrays).  Returns the
// activated number (rateDiag%Activation%numc).
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

}  // namespace salsa
