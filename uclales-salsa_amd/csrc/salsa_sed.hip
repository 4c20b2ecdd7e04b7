// salsa_sed.hip -- sedimentation of the liquid SALSA categories at level 4
// (sedim_SALSA, mcrp.f90:745-921) on device.
//
//   k_sed_zero   rrate, sfcrrate, VtPrc and the work arrays to zero
//                (sedim_SALSA :792, DepositionFast :1129-1132)
//   k_sed_vslow  DepositionSlow's (:926-1069) fall speeds of one category per
//                (cell, bin), with the dry deposition velocity at the lowest
//                level, into a work array
//   k_sed_slow   one thread per interior cell: the bin fluxes at the cell's
//                level and the level above, their divergence into the
//                tendencies, the
//                water divergence into tlt (:810-816, :833-839), the rain
//                rate and, at the lowest level, the dry deposition velocity
//                and the surface rate.  One launch per category, so the
//                accumulation into rrate / tlt keeps the reference's order.
//   k_sed_vt     DepositionFast's fall speed per (cell, precipitation bin)
//                into VtPrc.
//   k_sed_fall   the fall-distance redistribution (:1188-1253) per (column,
//                bin, moved quantity), levels from the top down as in the
//                reference, into the work arrays; deposited mass into the
//                removal array.
//   k_sed_apply  per cell: rain rate (bins in order), tendencies += change /
//                dtlt, tlt (:862-868); the surface rate from the removal
//                array at the lowest level.
// Terminal velocities: terminal_vel (mo_particle_external_properties.f90:
// 43-100), liquid branches.  Contraction is off so that the products and
// sums round as the oracle's.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "salsa_internal.h"

#pragma clang fp contract(off)

namespace salsa {

namespace {

constexpr double kPiFull = 3.14159265358979323846, kKb = 1.3807e-23, kG = 9.81, kCp = 1005., kAlvl = 2.5e6;

__device__ double sd_terminal_vel(double diam, double rhop, double rhoa, double visc, double beta) {
  if (diam <= 19e-6) return (diam * diam) * (rhop - rhoa) * kGrav * beta / (18. * visc);
  if (diam > 19e-6 && diam < 1.07e-3) {
    const double NDa = 4. / 3. * ::pow(diam, 3.) * kGrav * (rhop - rhoa) * rhoa / ::pow(visc, 2.);
    const double X = ::log(NDa);
    const double X2 = X * X, X3 = X2 * X, X4 = X2 * X2, X5 = X4 * X, X6 = X3 * X3;
    const double Y = -0.318657e1 + 0.992696 * X - 0.153193e-2 * X2 - 0.987059e-3 * X3 - 0.578878e-3 * X4 +
                     0.855176e-4 * X5 - 0.327815e-5 * X6;
    const double Re2 = beta * ::exp(Y);
    return visc / rhoa / diam * Re2;
  }
  const double Np = ::pow(kSurfw0, 3.) * ::pow(rhoa, 2.) / (::pow(visc, 4.) * (rhop - rhoa) * kGrav);
  const double Bo = 4. / 3. * (rhop - rhoa) * kGrav / kSurfw0 * ::pow(diam, 2.);
  const double X3 = ::log(Bo * ::pow(Np, 1. / 6.));
  const double Z2 = X3 * X3, Z3 = Z2 * X3, Z4 = Z2 * Z2, Z5 = Z4 * X3;
  const double Y3 = -0.500015e1 + 0.523778e1 * X3 - 0.204914e1 * Z2 + 0.475294 * Z3 - 0.542819e-1 * Z4 +
                    0.238449e-2 * Z5;
  const double Re3 = ::pow(Np, 1. / 6.) * ::exp(Y3);
  return fmin(visc / rhoa / diam * Re3, 9.2);
}

struct Air { double avis, lambda; };
__device__ __forceinline__ Air sd_air(double tk, double adn) {
  Air a;
  a.avis = 1.8325e-5 * (416.16 / (tk + 120.0)) * ::pow(tk / 296.16, 1.5);
  const double va = ::sqrt(8 * kKb * tk / (kPiFull * 4.8096e-26));
  a.lambda = 2 * a.avis / (adn * va);
  return a;
}

// the bin's wet diameter, density, slip correction and fall speed
__device__ double sd_vc(const SedArgs &a, double zpn, const double *mb, int64_t mstride, double adn, const Air &air,
                        double *dwet_out, double *gg_out) {
  double v = 0., ms = 0.;
  for (int s = 0; s < a.ns; ++s) v = v + mb[s * mstride] / a.rholiq[s];
  for (int s = 0; s < a.ns; ++s) ms = ms + mb[s * mstride];
  const double dwet = zpn < 1.e-15 ? 0. : ::pow(v / zpn / kPi6, 1. / 3.);
  const double zdn = ms / zpn / (kPi6 * (dwet * dwet * dwet));
  const double Kn = 2. * air.lambda / dwet;
  const double GG = 1. + Kn * (1.249 + 0.42 * ::exp(-0.87 / Kn));
  if (dwet_out) *dwet_out = dwet;
  if (gg_out) *gg_out = GG;
  return sd_terminal_vel(dwet, zdn, adn, air.avis, GG);
}

__device__ __forceinline__ bool sd_interior(const SedArgs &a, int i, int j) {
  return i >= 2 && i <= a.nxp - 3 && j >= 2 && j <= a.nyp - 3;
}

__global__ void __launch_bounds__(256) k_sed_zero(SedArgs a) {
  const int64_t n3 = a.n3, ncol = (int64_t)a.nxp * a.nyp, nb = a.nb[2];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n3; e += (int64_t)gridDim.x * blockDim.x) {
    a.rrate[e] = 0.;
    for (int b = 0; b < nb; ++b) {
      a.vt_prc[b * n3 + e] = 0.;
      a.prn[b * n3 + e] = 0.;
    }
    for (int q = 0; q < nb * a.ns; ++q) a.prv[q * n3 + e] = 0.;
    if (e < ncol) {
      a.sfcrrate[e] = 0.;
      for (int q = 0; q < nb * a.ns; ++q) a.rem[e * nb * a.ns + q] = 0.;
    }
  }
}

// DepositionSlow fall speeds of category cat per (interior cell, bin), with
// the dry deposition velocity at the lowest level; 0 below the number limit
__global__ void __launch_bounds__(128) k_sed_vslow(SedArgs a, int cat) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (x >= a.n3) return;
  const int k = (int)(x % a.nzp);
  const int64_t col = x / a.nzp;
  const int i = (int)(col % a.nxp), j = (int)(col / a.nxp);
  if (k < 1 || k > a.nzp - 2 || !sd_interior(a, i, j)) return;
  const int64_t n3 = a.n3;
  const double zpn = a.n_p[cat][(int64_t)b * n3 + x];
  double vc = 0.;
  if (!(zpn < kNlim)) {
    const double ad = a.adn[x];
    const Air ar = sd_air(a.tk[x], ad);
    double dwet, GG;
    vc = sd_vc(a, zpn, a.m_p[cat] + (int64_t)b * n3 + x, (int64_t)a.nb[cat] * n3, ad, ar, &dwet, &GG);
    vc = fmin(vc, fmin(0.5 * (1. / a.dzt[k]) / a.dtlt, 2.0));
    if (k == 1) {                                 // the level just above the surface
      const double ust = a.ustar[col];
      const double mdiff = (kKb * a.tk[x] * GG) / (3.0 * kPiFull * ar.avis * dwet);
      const double Sc = (ar.avis / ad) / mdiff;
      double St = vc * ::pow(ust, 2.0) / kG * (ar.avis / ad);
      if (St < 0.01) St = 0.01;
      const double rt = 1.0 / fmax(kEpsR8, (ust * (::pow(Sc, -2.0 / 3.0) + ::pow(10., -3.0 / St))));
      vc = (1. / rt) + vc;
    }
  }
  a.vslow[(int64_t)b * n3 + x] = vc;
}

// DepositionSlow for category cat at one interior cell: flux divergence from
// the fall speeds of the cell's level and the level above
__global__ void __launch_bounds__(128) k_sed_slow(SedArgs a, int cat) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.n3) return;
  const int k = (int)(x % a.nzp);
  const int64_t col = x / a.nzp;
  const int i = (int)(col % a.nxp), j = (int)(col / a.nxp);
  if (k < 1 || k > a.nzp - 2 || !sd_interior(a, i, j)) return;
  const int nb = a.nb[cat], ns = a.ns, iw = a.iwa;
  const int64_t n3 = a.n3;
  const double *__restrict__ N = a.n_p[cat];
  const double *__restrict__ M = a.m_p[cat];
  const double *__restrict__ V = a.vslow;
  const double dz = a.dzt[k];
  const int64_t x1 = x + 1;
  const bool top = k == a.nzp - 2;              // rflm(nzp) = 0
  const double adn = a.adn[x], adn1 = a.adn[x1];
  double sdw = 0.;
  for (int b = 0; b < nb; ++b) {
    const double zpn = N[(int64_t)b * n3 + x], zpn1 = top ? 0. : N[(int64_t)b * n3 + x1];
    const bool on = !(zpn < kNlim), on1 = !top && !(zpn1 < kNlim);
    const double vc = V[(int64_t)b * n3 + x], vc1 = on1 ? V[(int64_t)b * n3 + x1] : 0.;
    const double rfn = on ? -zpn * vc : 0.;
    const double rfn1 = on1 ? -zpn1 * vc1 : 0.;
    const double dn = (rfn1 * adn1 - rfn * adn) * dz / adn;
    a.n_t[cat][(int64_t)b * n3 + x] = a.n_t[cat][(int64_t)b * n3 + x] - dn;
    for (int s = 0; s < ns; ++s) {
      const double m = M[((int64_t)s * nb + b) * n3 + x], m1 = on1 ? M[((int64_t)s * nb + b) * n3 + x1] : 0.;
      const double rfm = on ? -m * vc : 0.;
      const double rfm1 = on1 ? -m1 * vc1 : 0.;
      const double dm = (rfm1 * adn1 - rfm * adn) * dz / adn;
      a.m_t[cat][((int64_t)s * nb + b) * n3 + x] = a.m_t[cat][((int64_t)s * nb + b) * n3 + x] - dm;
      if (s == iw) sdw = sdw + dm;
    }
  }
  // rain rate over the flux entries (ns-1)*nb .. ns*nb (1-based), as written
  double sw = 0.;
  for (int q = ns * nb - nb - 1; q < ns * nb; ++q) {
    const int s = q / nb, b = q % nb;
    if (N[(int64_t)b * n3 + x] < kNlim) continue;
    sw = sw + (-M[((int64_t)s * nb + b) * n3 + x] * V[(int64_t)b * n3 + x]);
  }
  a.rrate[x] = a.rrate[x] - sw * adn * kAlvl;
  a.tlt[x] = a.tlt[x] + sdw * (kAlvl / kCp) * a.th[x] / a.tk[x];
  if (k == 1) a.sfcrrate[col] = a.sfcrrate[col] - sw * adn * kAlvl;
}

// precipitation fall speed per (cell, bin)
__global__ void __launch_bounds__(128) k_sed_vt(SedArgs a) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (x >= a.n3) return;
  const int k = (int)(x % a.nzp);
  const int64_t col = x / a.nzp;
  const int i = (int)(col % a.nxp), j = (int)(col / a.nxp);
  if (k < 1 || k > a.nzp - 2 || !sd_interior(a, i, j)) return;
  const int64_t n3 = a.n3;
  const double zpn = a.n_p[2][(int64_t)b * n3 + x];
  if (zpn < kPrlim) return;
  const Air air = sd_air(a.tk[x], a.adn[x]);
  a.vt_prc[(int64_t)b * n3 + x] =
      sd_vc(a, zpn, a.m_p[2] + (int64_t)b * n3 + x, (int64_t)a.nb[2] * n3, a.adn[x], air, nullptr, nullptr);
}

// fall-distance redistribution per (interior column, bin), top level down;
// blockIdx.y selects the quantity moved: 0 the number, 1 + s mass species s
// (each thread repeats the bin's fall-distance decisions)
__global__ void __launch_bounds__(64) k_sed_fall(SedArgs a) {
  const int nx = a.nxp - 4, ny = a.nyp - 4, nb = a.nb[2], ns = a.ns;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int qy = blockIdx.y;                     // 0: number, 1 + s: species s
  if (t >= nx * ny * nb) return;
  const int b = t % nb, c = t / nb;
  const int i = 2 + c % nx, j = 2 + c / nx;
  const int64_t col = (int64_t)i + (int64_t)a.nxp * j, base = col * a.nzp, n3 = a.n3;
  const double *__restrict__ Nb = a.n_p[2] + (int64_t)b * n3 + base;
  const double *__restrict__ Q = qy == 0 ? Nb : a.m_p[2] + ((int64_t)(qy - 1) * nb + b) * n3 + base;
  const double *__restrict__ vt = a.vt_prc + (int64_t)b * n3 + base;
  const double *__restrict__ adn = a.adn + base;
  double *__restrict__ out = qy == 0 ? a.prn + (int64_t)b * n3 + base : a.prv + ((int64_t)(qy - 1) * nb + b) * n3 + base;
  double *rem = qy > 0 ? a.rem + col * nb * ns + (qy - 1) * nb + b : nullptr;   // deposited mass (species only)
  const double dtlt = a.dtlt;
  // The changes of the four levels at and below the current source level are
  // kept in registers (w[0] = level k, ..., w[3] = level k-3): a level
  // receives only from sources at or above it, so it is final once its own
  // source is done and is written once; landings further down go to memory,
  // before their level enters the window.  The additions happen in the
  // reference's order (sources from the top down).
  const int ktop = a.nzp - 2;
  double w[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) w[u] = ktop - u >= 0 ? out[ktop - u] : 0.;
  // the next level's inputs are loaded one iteration ahead
  double zpn_n = Nb[ktop], q_n = Q[ktop], vc_n = vt[ktop], adn_n = adn[ktop];
  for (int k = ktop; k >= 1; --k) {
    const double zpn = zpn_n, q_k = q_n, vc_k = vc_n, adn_k = adn_n;
    if (k > 1) {
      zpn_n = Nb[k - 1];
      q_n = Q[k - 1];
      vc_n = vt[k - 1];
      adn_n = adn[k - 1];
    }
    if (!(zpn < kPrlim)) {
      const double q = q_k;
      const double vc = vc_k;
      const double fdmax = dtlt * vc;
      double fd = 0.;
      int fi = 0;
      bool dep = false;
      while (fd < fdmax) {                    // bounded: ends at the lowest level
        fd = fd + (1. / a.dzt[k - fi]);
        fi = fi + 1;
        if (k - fi <= 0) { dep = true; break; }
      }
      fi = fi - 1;
      const int kf = k - fi;
      fd = fd - (1. / a.dzt[kf]);
      const double fdos = fmin(fmax(fdmax - fd, 0.), 1. / a.dzt[kf]);
      w[0] = w[0] - q;
      if (dep && qy > 0) *rem = *rem + q * adn_k * vc;
      const int d = k - kf;                   // kf = k - d, landing at kf - 1 and kf
      double c1, c0;                          // to kf - 1, to kf
      bool two;
      if (fdos * a.dzt[kf] > 0.5) {
        c1 = q * a.dzt[kf - 1] / a.dzt[k];
        c0 = 0.;
        two = false;
      } else {
        c1 = (fdos * a.dzt[kf]) * q * a.dzt[kf - 1] / a.dzt[kf];
        c0 = (1. - fdos * a.dzt[kf]) * q;
        two = true;
      }
      // level kf - 1 first, then kf, as the reference adds them
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (u == d + 1) w[u] = w[u] + c1;
      if (d + 1 > 3) out[kf - 1] = out[kf - 1] + c1;
      if (two) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (u == d) w[u] = w[u] + c0;
        if (d > 3) out[kf] = out[kf] + c0;
      }
    }
    out[k] = w[0];                            // level k is final
    w[0] = w[1];
    w[1] = w[2];
    w[2] = w[3];
    w[3] = k - 4 >= 0 ? out[k - 4] : 0.;
  }
  out[0] = w[0];                              // the level below the surface
}

// per cell: rain rate, tendencies and tlt of the precipitation; surface rate
__global__ void __launch_bounds__(256) k_sed_apply(SedArgs a) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.n3) return;
  const int k = (int)(x % a.nzp);
  const int64_t col = x / a.nzp;
  const int i = (int)(col % a.nxp), j = (int)(col / a.nxp);
  if (!sd_interior(a, i, j)) return;
  const int nb = a.nb[2], ns = a.ns, iw = a.iwa;
  const int64_t n3 = a.n3;
  if (k >= 1 && k <= a.nzp - 2) {
    double r = a.rrate[x];
    for (int b = 0; b < nb; ++b) {
      if (a.n_p[2][(int64_t)b * n3 + x] < kPrlim) continue;
      r = r + a.m_p[2][((int64_t)iw * nb + b) * n3 + x] * a.adn[x] * a.vt_prc[(int64_t)b * n3 + x] * kAlvl;
    }
    a.rrate[x] = r;
  }
  for (int b = 0; b < nb; ++b)
    a.n_t[2][(int64_t)b * n3 + x] = a.n_t[2][(int64_t)b * n3 + x] + a.prn[(int64_t)b * n3 + x] / a.dtlt;
  for (int q = 0; q < nb * ns; ++q)
    a.m_t[2][(int64_t)q * n3 + x] = a.m_t[2][(int64_t)q * n3 + x] + a.prv[(int64_t)q * n3 + x] / a.dtlt;
  if (k <= a.nzp - 2) {
    double sw = 0.;
    for (int b = 0; b < nb; ++b) sw = sw + a.prv[((int64_t)iw * nb + b) * n3 + x];
    a.tlt[x] = a.tlt[x] - (sw / a.dtlt) * (kAlvl / kCp) * a.th[x] / a.tk[x];
  }
  if (k == 0) {
    double s2 = 0.;
    const double *rem = a.rem + col * nb * ns;
    for (int q = ns * nb - nb - 1; q < ns * nb; ++q) s2 = s2 + rem[q];
    a.sfcrrate[col] = a.sfcrrate[col] + s2 * kAlvl;
  }
}

}  // namespace

int launch_sedim(const SedArgs &a, int sed_aero, int sed_cloud, int sed_precp, hipStream_t s) {
  const unsigned g256 = (unsigned)((a.n3 + 255) / 256), g128 = (unsigned)((a.n3 + 127) / 128);
  hipLaunchKernelGGL(k_sed_zero, dim3(g256 < 65536 ? g256 : 65536), dim3(256), 0, s, a);
  for (int cat = 0; cat < 2; ++cat) {
    if (!(cat == 0 ? sed_aero : sed_cloud)) continue;
    hipLaunchKernelGGL(k_sed_vslow, dim3(g128, a.nb[cat]), dim3(128), 0, s, a, cat);
    hipLaunchKernelGGL(k_sed_slow, dim3(g128), dim3(128), 0, s, a, cat);
  }
  if (sed_precp) {
    hipLaunchKernelGGL(k_sed_vt, dim3(g128, a.nb[2]), dim3(128), 0, s, a);
    const int nthr = (a.nxp - 4) * (a.nyp - 4) * a.nb[2];
    hipLaunchKernelGGL(k_sed_fall, dim3((nthr + 63) / 64, 1 + a.ns), dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_sed_apply, dim3(g256), dim3(256), 0, s, a);
  }
  return (int)hipGetLastError();
}

}  // namespace salsa
