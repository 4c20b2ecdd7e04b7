// salsa_autoconv.h -- autoconv2, the lognormal-fit autoconversion that
// follows condensation in salsa() when lsauto%mode = 2 (mo_salsa.f90:110-112,
// mo_salsa_cloud.f90:1401-1477), for one cell, serial over its cloud bins.
// R is the state type (double in the step kernel, a dual number in the
// derivative kernel); N[b] and V[b * vs + s] are the cell's numbers and
// volumes in SALSA units.
//
// Each cloud bin, largest first, is read as a lognormal of geometric width
// 1.2 around its mean volume; the number and water volume above 50 um move
// to the first precipitation bin, the dry species with the number fraction.
// The bin loop stops once the number converted in this cell exceeds
// 1e10 m^-3 s^-1 * dt (:1459-1462, where the reference also prints it).
#pragma once

#include "salsa_dual.h"
#include "salsa_internal.h"

namespace salsa {

template <class R> struct AcMath;
template <> struct AcMath<double> {
  static __device__ __forceinline__ double log_(double x) { return ::log(x); }
  static __device__ __forceinline__ double erf_(double x) { return ::erf(x); }
  // x**(1./3.) of the reference as a cube root: the two differ by a few ulp
  // (1./3. is not exact), and the device pow routine would give the whole
  // step kernel a stack frame
  static __device__ __forceinline__ double cbrt3_(double x) { return ::cbrt(x); }
  static __device__ __forceinline__ double val(double x) { return x; }
};
template <> struct AcMath<Dual<double>> {
  using D = Dual<double>;
  static __device__ __forceinline__ D log_(D x) { return dlog(x); }
  static __device__ __forceinline__ D erf_(D x) { return derf(x); }
  static __device__ __forceinline__ D cbrt3_(D x) { return dcbrt(x); }
  static __device__ __forceinline__ double val(D x) { return x.v; }
};

template <class R>
__device__ void autoconv2_cell(const DevLayout &D, R *N, R *V, int vs, double dt) {
  using M = AcMath<R>;
  const double zd0 = 50.e-6, sigmag = 1.2, max_rate_autoc = 1.0e10;
  const int nwet = D.nwet, ndry = D.ndry, iwa = D.iwa;
  const int P1 = D.off[2];
  const double ls = ::log(sigmag);
  const double hlp2 = ::sqrt(2.) * ls;                      // math_functions.f90:64-76
  const double lzd0 = ::log(zd0);
  const double up = ::exp((3. * (ls * ls)) / 2.), down = ::exp(-(3. * (ls * ls)));
  double tot = 0.;
  for (int cc = D.nb[1] - 1; cc >= 0; --cc) {
    const int b = D.off[1] + cc;
    const R Ntot = N[b];
    R Vtot = V[b * vs];
    for (int s = 1; s < nwet; ++s) Vtot = Vtot + V[b * vs + s];
    if (!(Ntot > D.nlim[b] && Vtot > 0.)) continue;
    const R dvg = M::cbrt3_(Vtot / Ntot / kPi6) * up;
    const R dg = dvg * down;
    // 1 - cumlognorm(d, sigmag, zd0)
    const R Vrem = Vtot * (1. - (0.5 + 0.5 * M::erf_((lzd0 - M::log_(dvg)) / hlp2)));
    const R Nrem = Ntot * (1. - (0.5 + 0.5 * M::erf_((lzd0 - M::log_(dg)) / hlp2)));
    if (!(Vrem > 0. && Nrem > D.nlim[P1])) continue;
    for (int s = 0; s < ndry; ++s) {
      V[P1 * vs + s] = V[P1 * vs + s] + V[b * vs + s] * (Nrem / Ntot);
      V[b * vs + s] = V[b * vs + s] * (1. - (Nrem / Ntot));
    }
    V[P1 * vs + iwa] = V[P1 * vs + iwa] + V[b * vs + iwa] * (Vrem / Vtot);
    V[b * vs + iwa] = V[b * vs + iwa] * (1. - (Vrem / Vtot));
    N[P1] = N[P1] + Nrem;
    N[b] = N[b] - Nrem;
    tot += M::val(Nrem);
    if (tot > max_rate_autoc * dt) break;
  }
}

}  // namespace salsa
