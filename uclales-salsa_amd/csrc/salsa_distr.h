// salsa_distr.h -- distr_update, the moving-centre re-binning that follows
// condensation in salsa() (mo_salsa.f90:148-149, mo_salsa_update.f90:21-433),
// for one cell, serial over its bins.  R is the state type (double in the
// step kernel, a dual number in the derivative kernel); N[b] and
// V[b * vs + s] are the cell's numbers and volumes in SALSA units.
//
// Each category is swept from its largest bin down, repeatedly, until the
// last bin examined in a sweep lies within its limits -- within_bins is
// overwritten per bin, not accumulated (:119-122) -- at most 100 sweeps.
// Bins left by CYCLE do not touch within_bins.  Returns SALSA_CELL_DISTR
// where the reference STOPs (no convergence, or a negative moved fraction).
// tmp holds kMaxSpec values of R (work space: LDS in the step kernel).
#pragma once

#include "salsa_dual.h"
#include "salsa_internal.h"

namespace salsa {

template <class R> struct DistrConst { static __device__ __forceinline__ R make(double x) { return x; } };
template <> struct DistrConst<Dual<double>> {
  static __device__ __forceinline__ Dual<double> make(double x) { return dconst(x); }
};

template <class R>
__device__ uint32_t distr_update_cell(const DevLayout &D, R *N, R *V, int vs, R *tmp) {
  const int ndry = D.ndry, nwet = D.nwet, nnsp = D.nspec;
  uint32_t flags = 0;
  auto vsum = [&](int b, int ns) {
    R t = V[b * vs];
    for (int s = 1; s < ns; ++s) t = t + V[b * vs + s];
    return t;
  };
  auto rmin = [](R a, double b) -> R { return b < a ? DistrConst<R>::make(b) : a; };
  const R zero = DistrConst<R>::make(0.);
  // ---- aerosols (dry volume), :50-140 ----
  {
    const int A = D.off[0];
    bool within = false;
    int count = 0;
    while (!within) {
      within = true;
      for (int kk = D.fn2b - 1; kk >= D.in1a; --kk) {
        const int b = A + kk;
        if (N[b] > kNlim) {
          if (kk == D.fn2a) continue;
          const R sd = vsum(b, ndry);
          const R zvpart = sd / N[b];
          if (D.vlo[b] > zvpart && kk == D.in1a) continue;      // the smallest bin cannot decrease
          if (D.vlo[b] > zvpart) {                              // the whole bin goes down
            const int m = A + (kk == D.in2b ? D.fn1a : kk - 1);  // 2b goes to 1a
            N[m] = N[m] + N[b];
            N[b] = zero;
            for (int s = 0; s < nwet; ++s) {
              V[m * vs + s] = V[m * vs + s] + V[b * vs + s];
              V[b * vs + s] = zero;
            }
            continue;
          }
          const double vmid = kPi6 * (D.dmid[b] * D.dmid[b] * D.dmid[b]);
          if (zvpart <= vmid) continue;                         // not grown
          const double zVrat = D.vhi[b] / D.vlo[b];
          const R zVilo = 2. * zvpart / (1. + zVrat);
          const R zVihi = zVrat * zVilo;
          const R zVexc = 0.5 * (zVihi + D.vhi[b]);
          const R znfrac = rmin((zVihi - D.vhi[b]) / (zVihi - zVilo), 1.);
          const R zvfrac = rmin(znfrac * zVexc / zvpart, 0.99);
          if (zvfrac < 0.) flags |= SALSA_CELL_DISTR;
          const int m = b + 1;
          R *add = tmp;
          for (int s = 0; s < nwet; ++s) add[s] = znfrac * N[b] * zVexc * V[b * vs + s] / sd;
          for (int s = 0; s < nwet; ++s) V[m * vs + s] = V[m * vs + s] + add[s];
          for (int s = 0; s < nwet; ++s) V[b * vs + s] = V[b * vs + s] - add[s];
          N[m] = N[m] + znfrac * N[b];
          N[b] = N[b] * (1. - znfrac);
        }
        if (N[b] > kNlim) {
          const R zvpart = vsum(b, ndry) / N[b];
          within = D.vlo[b] < zvpart && zvpart < D.vhi[b];
        }
      }
      if (++count > 100) { flags |= SALSA_CELL_DISTR; break; }
    }
  }
  // ---- cloud (dry volume), :145-225, and precipitation (wet volume),
  //      :227-320: the same moves, with category-specific thresholds ----
  for (int cat = 1; cat <= 2; ++cat) {
    const int O = D.off[cat], nb = D.nb[cat];
    const bool cloud = cat == 1;
    const double lim = cloud ? kNlim : kPrlim;
    const int nsum = cloud ? ndry : nwet;
    bool within = false;
    int count = 0;
    while (!within) {
      within = true;
      for (int kk = nb - 1; kk >= 0; --kk) {
        const int b = O + kk;
        const R s0 = vsum(b, nsum);
        if (N[b] > lim && (!cloud || s0 > 1.e-30)) {
          const R zvpart = s0 / N[b];
          const double zVrat = D.vhi[b] / D.vlo[b];
          const R zVilo = 2. * zvpart / (1. + zVrat);
          const R zVihi = zVrat * zVilo;
          const double vmid = kPi6 * (D.dmid[b] * D.dmid[b] * D.dmid[b]);
          bool down, up;
          if (cloud) {
            // the first bin of each series does not shrink, the last does not grow
            down = zvpart < D.vlo[b] && kk != D.cld_a0 && kk != D.cld_b0;
            up = !down && zvpart > D.vhi[b] && kk != D.cld_a1 && kk != D.cld_b1;
          } else {
            const double zvdec = vmid - 0.2 * (vmid - D.vlo[b]);
            down = zvpart < zvdec && kk != 0;
            up = !down && zvpart > vmid && kk != nb - 1;
          }
          if (!down && !up) continue;                           // size unchanged
          R zVexc, znfrac;
          int mm;
          if (down) {
            zVexc = 0.5 * (zVilo + D.vlo[b]);
            znfrac = rmin((D.vlo[b] - zVilo) / (zVihi - zVilo), 1.);
            mm = kk - 1;
          } else {
            zVexc = 0.5 * (zVihi + D.vhi[b]);
            znfrac = rmin((zVihi - D.vhi[b]) / (zVihi - zVilo), cloud ? 1. : .99);
            mm = kk + 1;
          }
          const R zvfrac = rmin(znfrac * zVexc / zvpart, 0.99);
          if (zvfrac < 0.) flags |= SALSA_CELL_DISTR;
          const int m = O + mm;
          for (int s = 0; s < nwet; ++s) V[m * vs + s] = V[m * vs + s] + zvfrac * V[b * vs + s];
          for (int s = 0; s < nwet; ++s) V[b * vs + s] = V[b * vs + s] - zvfrac * V[b * vs + s];
          N[m] = N[m] + znfrac * N[b];
          N[b] = N[b] - znfrac * N[b];
        }
        const R s1 = vsum(b, nsum);
        if (N[b] > lim && (!cloud || s1 > 1.e-30)) {
          const R zvpart = s1 / N[b];
          within = D.vlo[b] < zvpart && zvpart < D.vhi[b];
        }
      }
      if (++count > 100) { flags |= SALSA_CELL_DISTR; break; }
    }
  }
  if (D.level < 5) return flags;                                // :323
  // ---- ice (all species incl. rime), :330-430 ----
  {
    const int I = D.off[3], nb = D.nb[3];
    bool within = false;
    int count = 0;
    while (!within) {
      within = true;
      for (int kk = nb - 1; kk >= 0; --kk) {
        const int b = I + kk;
        if (N[b] > D.nlim[b]) {
          const R zvpart = vsum(b, nnsp) / N[b];
          const double zVrat = D.vhi[b] / D.vlo[b];
          const R zVilo = 2. * zvpart / (1. + zVrat);
          const R zVihi = zVrat * zVilo;
          const double vmid = kPi6 * (D.dmid[b] * D.dmid[b] * D.dmid[b]);
          const double zvdec = vmid - 0.2 * (vmid - D.vlo[b]);
          R zVexc, znfrac;
          int mm;
          if (zvpart < zvdec && kk != 0) {
            zVexc = 0.5 * (zVilo + D.vlo[b]);
            znfrac = rmin((D.vlo[b] - zVilo) / (zVihi - zVilo), 1.0);
            mm = kk - 1;
          } else if (zvpart > vmid && kk != nb - 1) {
            zVexc = 0.5 * (zVihi + D.vhi[b]);
            znfrac = rmin((zVihi - D.vhi[b]) / (zVihi - zVilo), 0.99);
            mm = kk + 1;
          } else {
            continue;
          }
          const R zvfrac = rmin(znfrac * zVexc / zvpart, 0.99);
          if (zvfrac < 0.) flags |= SALSA_CELL_DISTR;
          const int m = I + mm;
          for (int s = 0; s < nnsp; ++s) V[m * vs + s] = V[m * vs + s] + zvfrac * V[b * vs + s];
          for (int s = 0; s < nnsp; ++s) V[b * vs + s] = V[b * vs + s] - zvfrac * V[b * vs + s];
          N[m] = N[m] + znfrac * N[b];
          N[b] = N[b] - znfrac * N[b];
        }
        const R st = vsum(b, nnsp);
        if (N[b] > D.nlim[b] && st > 1.e-30) {
          const R zvpart = st / N[b];
          within = D.vlo[b] < zvpart && zvpart < D.vhi[b];
        }
      }
      if (++count > 100) { flags |= SALSA_CELL_DISTR; break; }
    }
  }
  return flags;
}

}  // namespace salsa
