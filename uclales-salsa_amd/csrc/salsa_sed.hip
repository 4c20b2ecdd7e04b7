// salsa_sed.hip -- sedimentation of the SALSA categories (sedim_SALSA,
// mcrp.f90:745-921) on device: the liquid categories at levels 4 and 5, ice
// at level 5.
//
//   k_sed_zero   rrate, sfcrrate, VtPrc (level 5: irate, sfcirate, VtIce) to
//                zero (sedim_SALSA :792-795, DepositionFast :1127)
//   k_sed_slow   DepositionSlow (:926-1069) of one category, one workgroup
//                per interior column and one lane per level, one pass over
//                the bins: each bin's fall speeds (with the dry deposition
//                velocity at the lowest level) and fluxes, the level above's
//                fluxes through LDS, their divergence into the tendencies,
//                the water divergence into tlt (:810-816, :833-839), the rain
//                rate and at the lowest level the surface rate.  One launch
//                per category, so the accumulation into rrate / tlt keeps the
//                reference's order.
//   k_sed_prc    DepositionFast (:1073-1275) and its part of sedim_SALSA
//                (:862-878, :884-904) for precipitation (flag 3) or ice
//                (flag 4), one workgroup per interior column and one lane
//                per level: fall speeds into VtPrc / VtIce, the fall-distance
//                redistribution of every bin and moved quantity with each
//                level's changes summed in the reference's top-down order,
//                tendencies += change / dtlt, rate (bins in order), tlt and
//                the surface rate from the deposited mass.
// Terminal velocities: terminal_vel (mo_particle_external_properties.f90:
// 43-100), liquid and ice branches.  Contraction is off so that the products and
// sums round as the oracle's; the reference's x**(1./3.), x**1.5 and x**3. are
// formed as a cube root (FP32 seed and one Halley step), x*sqrt(x) and x*x*x,
// which differ from pow in the last bit only.
#include <hip/hip_runtime.h>

#include <stdint.h>

// (before the shared headers, so that their inline physics rounds the same way here)
#pragma clang fp contract(off)

#include "salsa_internal.h"
#include "salsa_physics.h"

namespace salsa {

namespace {

constexpr int kSedMaxZ = 256;                 // levels per column (SedArgs::dzt)
constexpr double kPiFull = 3.14159265358979323846, kKb = 1.3807e-23, kG = 9.81, kCp = 1005., kAlvl = 2.5e6;
constexpr double kLn10 = 2.302585092994045684;

__device__ double sd_terminal_vel(double diam, double rhop, double rhoa, double visc, double beta) {
  if (diam <= 19e-6) return (diam * diam) * (rhop - rhoa) * kGrav * beta / (18. * visc);
  if (diam > 19e-6 && diam < 1.07e-3) {
    const double NDa = 4. / 3. * (diam * diam * diam) * kGrav * (rhop - rhoa) * rhoa / (visc * visc);
    const double X = ::log(NDa);
    const double X2 = X * X, X3 = X2 * X, X4 = X2 * X2, X5 = X4 * X, X6 = X3 * X3;
    const double Y = -0.318657e1 + 0.992696 * X - 0.153193e-2 * X2 - 0.987059e-3 * X3 - 0.578878e-3 * X4 +
                     0.855176e-4 * X5 - 0.327815e-5 * X6;
    const double Re2 = beta * ::exp(Y);
    return visc / rhoa / diam * Re2;
  }
  // Np**(1/6) through log(Np) / 6: X3 = log(Bo) + log(Np) / 6 and
  // Re3 = exp(log(Np) / 6 + Y3), within a few 1e-16 of the powers
  const double v2 = visc * visc;
  const double Np = (kSurfw0 * kSurfw0 * kSurfw0) * (rhoa * rhoa) / ((v2 * v2) * (rhop - rhoa) * kGrav);
  const double L6 = ::log(Np) / 6.;
  const double Bo = 4. / 3. * (rhop - rhoa) * kGrav / kSurfw0 * (diam * diam);
  const double X3 = ::log(Bo) + L6;
  const double Z2 = X3 * X3, Z3 = Z2 * X3, Z4 = Z2 * Z2, Z5 = Z4 * X3;
  const double Y3 = -0.500015e1 + 0.523778e1 * X3 - 0.204914e1 * Z2 + 0.475294 * Z3 - 0.542819e-1 * Z4 +
                    0.238449e-2 * Z5;
  const double Re3 = ::exp(L6 + Y3);
  return fmin(visc / rhoa / diam * Re3, 9.2);
}

struct Air { double avis, lambda; };
__device__ __forceinline__ Air sd_air(double tk, double adn) {
  Air a;
  const double tr = tk / 296.16;
  a.avis = 1.8325e-5 * (416.16 / (tk + 120.0)) * (tr * ::sqrt(tr));
  const double va = ::sqrt(8 * kKb * tk / (kPiFull * 4.8096e-26));
  a.lambda = 2 * a.avis / (adn * va);
  return a;
}

// the bin's wet diameter, density, slip correction and fall speed
// from the bin's volume v = sum m_s / rho_s and mass ms = sum m_s
__device__ __forceinline__ double sd_vc_sums(double zpn, double v, double ms, double adn, const Air &air,
                                             double *dwet_out, double *gg_out) {
  const double dwet = zpn < 1.e-15 ? 0. : cbrt_fast(v / zpn / kPi6);
  const double zdn = ms / zpn / (kPi6 * (dwet * dwet * dwet));
  const double Kn = 2. * air.lambda / dwet;
  const double GG = 1. + Kn * (1.249 + 0.42 * ::exp(-0.87 / Kn));
  if (dwet_out) *dwet_out = dwet;
  if (gg_out) *gg_out = GG;
  return sd_terminal_vel(dwet, zdn, adn, air.avis, GG);
}

__device__ double sd_vc(const SedArgs &a, double zpn, const double *mb, int64_t mstride, double adn, const Air &air,
                        double *dwet_out, double *gg_out) {
  double v = 0., ms = 0.;
  for (int s = 0; s < a.ns; ++s) v = v + mb[s * mstride] / a.rholiq[s];
  for (int s = 0; s < a.ns; ++s) ms = ms + mb[s * mstride];
  return sd_vc_sums(zpn, v, ms, adn, air, dwet_out, gg_out);
}

__device__ __forceinline__ bool sd_interior(const SedArgs &a, int i, int j) {
  return i >= 2 && i <= a.nxp - 3 && j >= 2 && j <= a.nyp - 3;
}

__global__ void __launch_bounds__(256) k_sed_zero(SedArgs a) {
  const int64_t n3 = a.n3, ncol = (int64_t)a.nxp * a.nyp, nb = a.nb[2];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n3; e += (int64_t)gridDim.x * blockDim.x) {
    a.rrate[e] = 0.;
    for (int b = 0; b < nb; ++b) a.vt_prc[b * n3 + e] = 0.;
    if (e < ncol) a.sfcrrate[e] = 0.;
    if (a.level == 5) {                         // irate, sfcirate, VtIce (:793-795, :1127)
      a.irate[e] = 0.;
      for (int b = 0; b < a.nb[3]; ++b) a.vt_ice[b * n3 + e] = 0.;
      if (e < ncol) a.sfcirate[e] = 0.;
    }
  }
}

// DepositionSlow (:926-1069) for category cat (0 aerosol, 1 cloud) on one
// interior column: one workgroup per column, one lane per level, one pass over
// the bins in order.  Per bin, every interior level reads its number and mass
// rows once, evaluates the fall speed (0 below the number limit; capped at
// half a level per step and 2 m/s; at the lowest level plus the dry
// deposition velocity) and publishes its fluxes times the air density in LDS
// (two buffers by bin parity, one barrier per bin); each level then takes the
// level above's from there (:1036-1049; 0 above the top interior level) and
// forms their divergence into the tendencies and, for the water row, into
// tlt (:810-816, :833-839).  The rain rate sums the flux entries
// (ns-1)*nb .. ns*nb (1-based, one entry early, as written) at the lane's
// level: the water rows' terms wait in LDS until the bin loop has produced
// the early entry (the last dry row of the last bin), which comes first; at
// the lowest level also the surface rate.  One launch per category keeps the
// reference's category order of the rrate / tlt accumulation.
template <int MS>
__global__ void __launch_bounds__(kSedMaxZ, MS <= 3 ? 3 : 1) k_sed_slow(SedArgs a, int cat) {
  extern __shared__ double sm[];
  const int Z = (int)blockDim.x;
  double *flx = sm;                             // [2][1 + MS][Z]: flux * adn of each level, by bin parity
  double *rt = sm + 2 * (1 + MS) * Z;           // [nb][Z]: the water rows' rain-rate terms of each level
  const int nx = a.nxp - 4;
  const int i = 2 + (int)blockIdx.x % nx, j = 2 + (int)blockIdx.x / nx;
  const int64_t col = (int64_t)i + (int64_t)a.nxp * j, n3 = a.n3;
  const int k = threadIdx.x, nzp = a.nzp;
  const int nb = a.nb[cat], ns = a.ns, iw = a.iwa;
  const bool src = k >= 1 && k <= nzp - 2;
  const int64_t x = col * nzp + k;
  const double *__restrict__ N = a.n_p[cat];
  const double *__restrict__ M = a.m_p[cat];
  double *__restrict__ Nt = a.n_t[cat];
  double *__restrict__ Mt = a.m_t[cat];
  double ad = 1., dz = 0., cap = 0., tkx = 0., ust = 0.;
  Air ar{0., 0.};
  if (src) {
    ad = a.adn[x];
    tkx = a.tk[x];
    ar = sd_air(tkx, ad);
    dz = a.dzt[k];
    cap = fmin(0.5 * (1. / dz) / a.dtlt, 2.0);
    if (k == 1) ust = a.ustar[col];
  }
  // per-lane factor of the flux divergence (x * dz / adn rounded as x * (dz / adn))
  const double dza = dz / ad;
  double sdw = 0., t0 = 0.;
  uint64_t onm = 0;                             // bins at or above the number limit
  // the rows of bin b (number, masses, their tendencies), loaded one bin
  // ahead so that a bin's loads are in flight while the previous bin computes
  double qn = 0., qnt = 0., qm[MS], qmt[MS];
  auto load = [&](int b, double &n_, double &nt_, double *m_, double *mt_) {
    n_ = N[(int64_t)b * n3 + x];
    nt_ = Nt[(int64_t)b * n3 + x];
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      const int64_t r = ((int64_t)s * nb + b) * n3 + x;
      m_[s] = s < ns ? M[r] : 0.;
      mt_[s] = s < ns ? Mt[r] : 0.;
    }
  };
#pragma unroll
  for (int s = 0; s < MS; ++s) qm[s] = qmt[s] = 0.;
  if (src) load(0, qn, qnt, qm, qmt);
  for (int b = 0; b < nb; ++b) {
    double *F = flx + (b & 1) * (1 + MS) * Z;
    const double zpn = qn, znt = qnt;
    double m[MS], mt[MS];
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      m[s] = qm[s];
      mt[s] = qmt[s];
    }
    if (src && b + 1 < nb) load(b + 1, qn, qnt, qm, qmt);
    double vc = 0.;
    const bool on = src && !(zpn < kNlim);
    if (on) {
      double v = 0., ms = 0.;
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (s < ns) v = v + m[s] / a.rholiq[s];
#pragma unroll
      for (int s = 0; s < MS; ++s)
        if (s < ns) ms = ms + m[s];
      double dwet, GG;
      vc = sd_vc_sums(zpn, v, ms, ad, ar, &dwet, &GG);
      vc = fmin(vc, cap);
      if (k == 1) {                               // the level just above the surface
        // Sc**(-2/3) and 10**(-3/St) as exponentials of logarithms (a few
        // 1e-15 from the powers; Sc is the Schmidt number, 10**(-3/St) >= 1e-300)
        const double mdiff = (kKb * tkx * GG) / (3.0 * kPiFull * ar.avis * dwet);
        const double Sc = (ar.avis / ad) / mdiff;
        double St = vc * (ust * ust) / kG * (ar.avis / ad);
        if (St < 0.01) St = 0.01;
        const double rtr = 1.0 / fmax(kEpsR8, (ust * (::exp(-2.0 / 3.0 * ::log(Sc)) + ::exp(-3.0 / St * kLn10))));
        vc = (1. / rtr) + vc;
      }
      onm |= (uint64_t)1 << b;
#pragma unroll
      for (int s = 0; s < MS; ++s) {
        if (s == ns - 1) rt[b * Z + k] = -m[s] * vc;
        if (s == ns - 2 && b == nb - 1) t0 = -m[s] * vc;
      }
    }
    // the fluxes rfn, rfm (0 below the limit and outside the interior) times adn
    const double Pn = (on ? -zpn * vc : 0.) * ad;
    double Pm[MS];
#pragma unroll
    for (int s = 0; s < MS; ++s) Pm[s] = (on ? -m[s] * vc : 0.) * ad;
    F[k] = Pn;
#pragma unroll
    for (int s = 0; s < MS; ++s)
      if (s < ns) F[(1 + s) * Z + k] = Pm[s];
    __syncthreads();
    if (src) {
      const double dn = (F[k + 1] - Pn) * dza;
      Nt[(int64_t)b * n3 + x] = znt - dn;
#pragma unroll
      for (int s = 0; s < MS; ++s) {
        if (s < ns) {
          const double dm = (F[(1 + s) * Z + k + 1] - Pm[s]) * dza;
          Mt[((int64_t)s * nb + b) * n3 + x] = mt[s] - dm;
          if (s == iw) sdw = sdw + dm;
        }
      }
    }
  }
  if (!src) return;                             // no barrier below
  // rain rate over the flux entries (ns-1)*nb .. ns*nb (1-based), as written
  double sw = 0.;
  if ((onm >> (nb - 1)) & 1) sw = sw + t0;
  for (int b = 0; b < nb; ++b)
    if ((onm >> b) & 1) sw = sw + rt[b * Z + k];
  a.rrate[x] = a.rrate[x] - sw * ad * kAlvl;
  a.tlt[x] = a.tlt[x] + sdw * (kAlvl / kCp) * a.th[x] / a.tk[x];
  if (k == 1) a.sfcrrate[col] = a.sfcrrate[col] - sw * ad * kAlvl;
}

// LDS of k_sed_slow<MS> for blockDim Z and nb bins: the flux buffers and the
// rain-rate terms
size_t sed_slow_lds_bytes(int Z, int nb, int MS) {
  return ((size_t)2 * (1 + MS) * Z + (size_t)nb * Z) * sizeof(double);
}

// DepositionFast for one category on one interior column: one workgroup per
// column, one lane per level (blockDim = nzp rounded up to 64).  Per bin, in
// order:
//   1. every source level k (1 <= k <= nzp-2 with N >= prlim) evaluates its
//      fall speed (VtPrc / VtIce) and its fall-distance decision
//      (:1188-1225): the landing level kf, one or two receiving levels and
//      their shares;
//   2. per moved quantity (the number, then the mass species), every level L
//      sums what lands on it from the sources above it, in decreasing source
//      level as the reference's top-down walk adds them, then its own
//      departure (-q) and, when the particles stay within the level, its own
//      share; the change goes straight into the tendency (/ dtlt).  Only
//      sources within the column's largest fall distance of L are scanned.
// Deposited mass (sources whose walk reaches the surface) is summed per
// species and bin by lane 0 over those sources, top down.  Rate (bins in
// order), tlt and the surface rate as in sedim_SALSA (:862-878, :884-904).
// ICE: flag 4 -- the fall speed of non-spherical ice (calcDiamLES both ways,
// getShapeCoefficients, terminal_vel's ice branch), water and rime in the
// rate, the surface rate and tlt, latent heat alvi; MS: the most mass rows
// per bin (registers per level).
constexpr int kSedMaxIce = 16;                // ice bins whose rime changes a workgroup keeps in LDS

// The ice fall speed of one bin (DepositionFast flag 4: calcDiamLES,
// getShapeCoefficients and getDiameter, mo_ice_shape.f90:47-86, 145-191;
// terminal_vel's ice branch, mo_particle_external_properties.f90:91-98)
// with its powers taken in the log domain: the mass thresholds Mth, Mgr,
// Mcr and the non-spherical diameter as logarithms (Mgr = rho_b pi6 Dgr**3,
// Dgr = (alpha / (pi6 rho_b))**(1/(3-beta)), Dcr = Dgr h**(1/(3-beta))),
// the Best number X = 2 vb (rhop-rhoa) g / (gamma rhoa visc**2) D**(2-sigma)
// and the speed Are / X**bre * visc**(1-2 bre) * (2 alpha g / (rhoa gamma))**bre
// * D**bv as one exponential: five logarithms and two exponentials per bin
// instead of eight powers, within a few 1e-14 of the power form.  lvisc:
// log(visc) of the level; c: the logs of the constant thresholds.
struct IceLogC { double lmth, lia, lpi6ic; };
__device__ inline double ice_vc_log(const SedArgs &a, const IceLogC &c, double dsph, double rhop, double rhoa,
                                    double visc, double lvisc, double mpri, double mrim, double numc) {
  const double ia = a.ice_alpha, ib = a.ice_beta;
  double al = kPi6 * kRhoic, be = 3., ga = kPi / 4., si = 2.;
  double Ld = -20.72326583694641;               // log(1e-9): getDiameter's value below the number limit
  if (!(numc < 1.e-6)) {
    const double mtot = (mpri + mrim) / numc;
    const double fr = fmax(fmin(mrim / (mtot * numc), 0.9999), 0.0001);
    const double rho_b = fr * kRhori + (1. - fr) * kRhoic;
    const double Lm = ::log(mtot), Lrb = ::log(kPi6 * rho_b), inv = 1. / (3. - ib);
    const double lmgr = Lrb + 3. * ((c.lia - Lrb) * inv);
    const double h = 1. / (1. - fr), Lh = ::log(h);
    const double lmcr = lmgr + 3. * (Lh * inv);
    if (Lm >= c.lmth && Lm < lmgr) {
      al = ia; be = ib; ga = a.ice_gamma; si = a.ice_sigma;
    } else if (Lm > lmcr) {
      al = ia * h; be = ib; ga = a.ice_gamma * (1. - fr) + kPi / 4. * fr; si = a.ice_sigma * (1. - fr) + 2. * fr;
    }
    if (Lm < c.lmth) Ld = (Lm - c.lpi6ic) / 3.;
    else if (Lm < lmgr) Ld = (Lm - c.lia) / ib;
    else if (Lm < lmcr) Ld = (Lm - Lrb) / 3.;
    else if (Lm >= lmcr) Ld = (Lm - Lh - c.lia) / ib;
  }
  const double c1 = 0.0902, delta0 = 9.06;
  const double vb = kPi6 * cubed(dsph);
  const double LX = ::log(2. * vb * (rhop - rhoa) * kGrav / (ga * rhoa * visc * visc)) + (2. - si) * Ld;
  const double sX = ::sqrt(::exp(LX));
  const double k1213 = ::sqrt(1. + c1 * sX);
  const double bre = 0.5 * c1 * sX / (k1213 - 1.) / k1213;
  const double Lc = ::log((2. * al * kGrav) / (rhoa * ga));
  const double bv = bre * (be - si + 2.) - 1.;
  return 0.25 * delta0 * delta0 * (k1213 - 1.) * (k1213 - 1.) *
         ::exp(-bre * LX + (1. - 2. * bre) * lvisc + bre * Lc + bv * Ld);
}

// LDS of one column (Z = blockDim.x levels), carved from the dynamic buffer
// in this order: dz, rdz, adn [Z]; vc [2][Z] (source fall speed, by bin
// parity); c1, c0, q [2][Z] (by quantity parity); rem [ns*nb] (deposited
// mass per species and bin); rout [nb][Z] (ice: rime change per bin); then
// the ints dec [2][Z] (bit 0 source, 1 two levels, 2 deposits; (d + 1) << 8)
// and wmax [2][4] (largest fall distance per wavefront, by bin parity).
__host__ __device__ inline size_t sed_prc_lds_bytes(int Z, int nb, int ns, bool ice) {
  return (size_t)(11 * Z + ns * nb + (ice ? nb * Z : 0)) * sizeof(double) + (size_t)(2 * Z + 8) * sizeof(int);
}

template <int MS, bool ICE>
__global__ void __launch_bounds__(kSedMaxZ, (MS <= 3 && !ICE) ? 4 : 1) k_sed_prc(SedArgs a) {
  extern __shared__ double smem[];
  constexpr int cat = ICE ? 3 : 2;
  constexpr double alv = ICE ? kAls : kAlvl;
  const int nx = a.nxp - 4;
  const int i = 2 + (int)blockIdx.x % nx, j = 2 + (int)blockIdx.x / nx;
  const int64_t col = (int64_t)i + (int64_t)a.nxp * j, n3 = a.n3;
  const int k = threadIdx.x, nzp = a.nzp, ktop = nzp - 2, nw = blockDim.x / 64;
  const int nb = a.nb[cat], ns = ICE ? a.ns_ice : a.ns, iw = a.iwa, ir = a.irim;
  const int Z = blockDim.x;
  double *const Sdz = smem, *const Srdz = smem + Z, *const Sadn = smem + 2 * Z, *const Svc = smem + 3 * Z;
  double *const Sc1 = smem + 5 * Z, *const Sc0 = smem + 7 * Z, *const Sq = smem + 9 * Z, *const Srem = smem + 11 * Z;
  double *const Srout = Srem + ns * nb;
  int *const Sdec = reinterpret_cast<int *>(Srout + (ICE ? nb * Z : 0));
  int *const Swmax = Sdec + 2 * Z;
  const double *__restrict__ Np = a.n_p[cat];
  const double *__restrict__ Mp = a.m_p[cat];
  double *__restrict__ Nt = a.n_t[cat];
  double *__restrict__ Mt = a.m_t[cat];
  double *__restrict__ vt = ICE ? a.vt_ice : a.vt_prc;
  double *__restrict__ rate = ICE ? a.irate : a.rrate;
  double *__restrict__ srate = ICE ? a.sfcirate : a.sfcrrate;
  const bool lev = k < nzp, src = k >= 1 && k <= ktop;
  const int64_t x = col * nzp + k;
  const double dtlt = a.dtlt;
  // the bin's rows at this level, quantity u: 0 the number, 1 + s species s
  // (the next bin's are loaded while the current one is worked on)
  double qv[MS + 1], tv[MS + 1], qn[MS + 1], tn[MS + 1];
  auto load = [&](int b, double *q, double *t) {
#pragma unroll
    for (int u = 0; u <= MS; ++u) {
      const bool use = lev && b < nb && u <= ns;
      const int64_t r = u == 0 ? (int64_t)b * n3 : ((int64_t)(u - 1) * nb + b) * n3;
      q[u] = use ? (u == 0 ? Np : Mp)[r + x] : 0.;
      t[u] = use ? (u == 0 ? Nt : Mt)[r + x] : 0.;
    }
  };
  load(0, qv, tv);
  Air air = {0., 0.};
  double rr = 0., sw = 0.;
  IceLogC ilc{0., 0., 0.};                      // ice: log Mth (pristine threshold), log alpha, log(pi6 rhoic)
  if (ICE) {
    const double dth = ::pow(kPi6 * kRhoic / a.ice_alpha, 1. / (a.ice_beta - 3.));
    ilc.lmth = ::log(kRhoic * kPi6 * cubed(dth));
    ilc.lia = ::log(a.ice_alpha);
    ilc.lpi6ic = ::log(kPi6 * kRhoic);
  }
  double lvisc = 0.;
  if (lev) {
    Sdz[k] = a.dzt[k];
    Srdz[k] = 1. / a.dzt[k];
    Sadn[k] = a.adn[x];
  }
  if (src) {
    air = sd_air(a.tk[x], a.adn[x]);
    if (ICE) lvisc = ::log(air.avis);
    rr = rate[x];
  }
  __syncthreads();
  for (int b = 0; b < nb; ++b) {
    const int pb = b & 1;
    load(b + 1, qn, tn);
    const double zpn = qv[0];
    const bool on = src && !(zpn < kPrlim);
    double vc = 0., pre = 0., dA = 0., dB = 0.;
    int d = -1;
    bool two = false, dep = false;
    if (on) {
      if (ICE) {
        // calcDiamLES (mo_particle_external_properties.f90:169-201) both
        // ways, getShapeCoefficients, the bulk density, terminal_vel flag 4
        double v = 0., mpri = 0., ms = 0., mrim = 0.;
#pragma unroll
        for (int s = 0; s < MS; ++s)
          if (s < ns) v = v + qv[1 + s] / a.rhoice[s];
#pragma unroll
        for (int s = 0; s < MS; ++s)
          if (s < ns - 1) mpri = mpri + qv[1 + s];
#pragma unroll
        for (int s = 0; s < MS; ++s)
          if (s == ns - 1) mrim = qv[1 + s];
        const double dwet = zpn < 1.e-15 ? 0. : cbrt_fast(v / zpn / kPi6);
#pragma unroll
        for (int s = 0; s < MS; ++s)
          if (s < ns) ms = ms + qv[1 + s];
        const double zdn = ms / zpn / (kPi6 * (dwet * dwet * dwet));
        vc = ice_vc_log(a, ilc, dwet, zdn, Sadn[k], air.avis, lvisc, mpri, mrim, zpn);
      } else {
        double v = 0., ms = 0.;
#pragma unroll
        for (int s = 0; s < MS; ++s)
          if (s < ns) v = v + qv[1 + s] / a.rholiq[s];
#pragma unroll
        for (int s = 0; s < MS; ++s)
          if (s < ns) ms = ms + qv[1 + s];
        vc = sd_vc_sums(zpn, v, ms, Sadn[k], air, nullptr, nullptr);
      }
      vt[(int64_t)b * n3 + x] = vc;
      // the rate: water (and, for ice, rime) falling out of the level (:1206-1212)
      double wv = 0., rv = 0.;
#pragma unroll
      for (int s = 0; s < MS; ++s) {
        if (s == iw) wv = qv[1 + s];
        if (s == ir) rv = qv[1 + s];
      }
      rr = ICE ? rr + (wv + rv) * Sadn[k] * vc * alv : rr + wv * Sadn[k] * vc * alv;
      const double fdmax = dtlt * vc;
      double fd = 0.;
      int fi = 0;
      while (fd < fdmax) {                      // bounded: ends at the lowest level
        fd = fd + Srdz[k - fi];
        fi = fi + 1;
        if (k - fi <= 0) { dep = true; break; }
      }
      fi = fi - 1;
      const int kf = k - fi;                    // kf - 1 >= 0; kf <= k + 1 <= nzp - 1
      fd = fd - Srdz[kf];
      const double fdos = fmin(fmax(fdmax - fd, 0.), Srdz[kf]);
      pre = fdos * Sdz[kf];
      two = !(pre > 0.5);
      dA = Sdz[kf - 1];
      dB = two ? Sdz[kf] : Sdz[k];
      d = k - kf;                               // >= -1
      Svc[pb * Z + k] = vc;
    }
    Sdec[pb * Z + k] = on ? (1 | (two ? 2 : 0) | (dep ? 4 : 0) | ((d + 1) << 8)) : 0;
    int wm = d;
    for (int o = 32; o > 0; o >>= 1) wm = max(wm, __shfl_xor(wm, o));
    if ((k & 63) == 0) Swmax[pb * 4 + (k >> 6)] = wm;
    int dmax = -1;
#pragma unroll
    for (int qy = 0; qy <= MS; ++qy) {
      if (qy > ns) break;
      const int p = (b * (ns + 1) + qy) & 1;    // parity of the c1 / c0 / q buffers
      const double q = qv[qy];
      double c1 = 0., c0 = 0.;
      if (on) {
        // the reference's expressions, left to right (:1229-1246)
        if (two) {
          c1 = pre * q * dA / dB;
          c0 = (1. - pre) * q;
        } else {
          c1 = q * dA / dB;
        }
        Sc1[p * Z + k] = c1;
        Sc0[p * Z + k] = c0;
        if (dep) Sq[p * Z + k] = q;
      }
      __syncthreads();
      if (qy == 0)
        for (int w = 0; w < nw; ++w) dmax = max(dmax, Swmax[pb * 4 + (w)]);
      if (lev) {
        double out = 0.;
        if (k <= ktop) {
          for (int s = min(ktop, k + dmax + 1); s > k; --s) {
            const int ds = Sdec[pb * Z + s];
            if (!(ds & 1)) continue;
            const int kf = s - ((ds >> 8) - 1);
            if (kf - 1 == k) out = out + Sc1[p * Z + s];
            if ((ds & 2) && kf == k) out = out + Sc0[p * Z + s];
          }
          if (on) {
            out = out - q;
            if (d == -1) out = out + c1;        // particles that do not leave the level
            if (two && d == 0) out = out + c0;
          }
        }
        const int64_t r = qy == 0 ? (int64_t)b * n3 : ((int64_t)(qy - 1) * nb + b) * n3;
        (qy == 0 ? Nt : Mt)[r + x] = tv[qy] + out / dtlt;
        if (qy - 1 == iw) sw = sw + out;
        if (ICE && qy - 1 == ir) Srout[b * Z + k] = out;
      }
      if (qy > 0 && k == 0) {                   // lane 0 (level 0 is never a source)
        double r = 0.;
        for (int s = min(ktop, dmax + 1); s >= 1; --s)
          if ((Sdec[pb * Z + s] & 5) == 5) r = r + Sq[p * Z + s] * Sadn[s] * Svc[pb * Z + s];
        Srem[(qy - 1) * nb + b] = r;
      }
    }
#pragma unroll
    for (int u = 0; u <= MS; ++u) {
      qv[u] = qn[u];
      tv[u] = tn[u];
    }
  }
  if (src) rate[x] = rr;
  if (lev && k <= ktop) {
    // SUM over the water rows in bin order, for ice then the rime rows
    if (ICE)
      for (int b = 0; b < nb; ++b) sw = sw + Srout[b * Z + k];
    a.tlt[x] = a.tlt[x] - (sw / dtlt) * (alv / kCp) * a.th[x] / a.tk[x];
  }
  if (k == 0) {
    // SUM(remprc((ns-1)*nb:ns*nb)) (ice: (ns-2)*nb:ns*nb), 1-based, one entry
    // before the water rows as written (:1256-1260)
    double s2 = 0.;
    for (int q = (ICE ? ns - 2 : ns - 1) * nb - 1; q < ns * nb; ++q) s2 = s2 + Srem[q];
    srate[col] = srate[col] + s2 * alv;
  }
}

}  // namespace

int launch_sedim(const SedArgs &a, int sed_aero, int sed_cloud, int sed_precp, int sed_ice, hipStream_t s) {
  const unsigned g256 = (unsigned)((a.n3 + 255) / 256);
  hipLaunchKernelGGL(k_sed_zero, dim3(g256 < 65536 ? g256 : 65536), dim3(256), 0, s, a);
  for (int cat = 0; cat < 2; ++cat) {
    if (!(cat == 0 ? sed_aero : sed_cloud)) continue;
    // one workgroup per interior column, one lane per level; the fluxes of
    // one bin and the column's rain-rate terms in LDS
    const unsigned ncol = (unsigned)((a.nxp - 4) * (a.nyp - 4)), nthr = (unsigned)((a.nzp + 63) / 64 * 64);
    const int ms = a.ns <= 3 ? 3 : a.ns <= 5 ? 5 : kMaxSpec;
    const size_t lds = sed_slow_lds_bytes((int)nthr, a.nb[cat], ms);
    if (a.nzp > kSedMaxZ || a.nb[cat] > 64 || a.ns < 2 || lds > 65536) return (int)hipErrorInvalidValue;
    if (ms == 3)
      hipLaunchKernelGGL(k_sed_slow<3>, dim3(ncol), dim3(nthr), lds, s, a, cat);
    else if (ms == 5)
      hipLaunchKernelGGL(k_sed_slow<5>, dim3(ncol), dim3(nthr), lds, s, a, cat);
    else
      hipLaunchKernelGGL(k_sed_slow<kMaxSpec>, dim3(ncol), dim3(nthr), lds, s, a, cat);
  }
  if (sed_precp) {
    // one workgroup per interior column, one lane per level
    const unsigned ncol = (unsigned)((a.nxp - 4) * (a.nyp - 4)), nthr = (unsigned)((a.nzp + 63) / 64 * 64);
    const size_t lds = sed_prc_lds_bytes((int)nthr, a.nb[2], a.ns, false);
    if (a.nzp > kSedMaxZ || lds > 65536) return (int)hipErrorInvalidValue;
    if (a.ns <= 3)
      hipLaunchKernelGGL((k_sed_prc<3, false>), dim3(ncol), dim3(nthr), lds, s, a);
    else if (a.ns <= 5)
      hipLaunchKernelGGL((k_sed_prc<5, false>), dim3(ncol), dim3(nthr), lds, s, a);
    else
      hipLaunchKernelGGL((k_sed_prc<kMaxSpec, false>), dim3(ncol), dim3(nthr), lds, s, a);
  }
  if (sed_ice && a.level == 5) {
    const unsigned ncol = (unsigned)((a.nxp - 4) * (a.nyp - 4)), nthr = (unsigned)((a.nzp + 63) / 64 * 64);
    const size_t lds = sed_prc_lds_bytes((int)nthr, a.nb[3], a.ns_ice, true);
    if (a.nzp > kSedMaxZ || a.nb[3] > kSedMaxIce || lds > 65536) return (int)hipErrorInvalidValue;
    if (a.ns_ice <= 3)
      hipLaunchKernelGGL((k_sed_prc<3, true>), dim3(ncol), dim3(nthr), lds, s, a);
    else if (a.ns_ice <= 5)
      hipLaunchKernelGGL((k_sed_prc<5, true>), dim3(ncol), dim3(nthr), lds, s, a);
    else
      hipLaunchKernelGGL((k_sed_prc<kMaxSpec, true>), dim3(ncol), dim3(nthr), lds, s, a);
  }
  return (int)hipGetLastError();
}

}  // namespace salsa
