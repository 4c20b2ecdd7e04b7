// salsa_physics.h -- device physics shared by the step kernel
// (salsa_kernels.hip) and the adjoint kernel (salsa_adjoint.hip): precision
// helpers and cross-lane sums, particle terms and terminal velocities, the
// collision kernel of one pair (coagc), water activity and the aerosol
// equilibration.  Each function cites the reference lines it follows.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include "salsa_internal.h"

namespace salsa {

// ---------------------------------------------------------------------------
// precision helpers
// ---------------------------------------------------------------------------
template <class T> struct Mth;
template <> struct Mth<double> {
  static __device__ __forceinline__ double sqrt_(double x) { return ::sqrt(x); }
  static __device__ __forceinline__ double log_(double x) { return ::log(x); }
  static __device__ __forceinline__ double cbrt_(double x) { return ::cbrt(x); }
  static __device__ __forceinline__ double div_(double a, double b) { return a / b; }
};
template <> struct Mth<float> {
  // v_sqrt_f32 / v_log_f32 / v_exp_f32 / v_rcp_f32: ~1 ulp, far inside the
  // 1e-5 output tolerance of the mixed path.
  static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float log_(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945f; }
  static __device__ __forceinline__ float cbrt_(float x) {
    return x > 0.f ? __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.f / 3.f)) : 0.f;
  }
  static __device__ __forceinline__ float div_(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ float exp_(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
  static __device__ __forceinline__ float pow_(float x, float y) {
    return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * y);
  }
};

// Sum / OR over the L lanes of one cell; every lane receives the same bits.
// Within a 16-lane row: DPP quad permutes and row mirrors; across rows the
// gfx950 v_permlane16_swap / v_permlane32_swap exchanges.  A cell never spans
// more lanes than its own group, so rows of other cells are never read.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) { return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, false); }
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  return __hiloint2double(dpp_i<CTRL>(__double2hiint(x)), dpp_i<CTRL>(__double2loint(x)));
}
__device__ __forceinline__ double swap16_sum(double x) {
  auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(x), __double2loint(x), false, false);
  auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(x), __double2hiint(x), false, false);
  return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}
__device__ __forceinline__ double swap32_sum(double x) {
  auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(x), __double2loint(x), false, false);
  auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(x), __double2hiint(x), false, false);
  return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}
template <int L>
__device__ __forceinline__ double cell_sum(double x) {
  if constexpr (L >= 2) x += dpp_d<0xB1>(x);     // quad_perm [1,0,3,2]
  if constexpr (L >= 4) x += dpp_d<0x4E>(x);     // quad_perm [2,3,0,1]
  if constexpr (L >= 8) x += dpp_d<0x141>(x);    // row_half_mirror
  if constexpr (L >= 16) x += dpp_d<0x140>(x);   // row_mirror
  if constexpr (L >= 32) x = swap16_sum(x);
  if constexpr (L >= 64) x = swap32_sum(x);
  return x;
}
template <int L>
__device__ __forceinline__ uint32_t cell_or(uint32_t x) {
  int v = (int)x;
  if constexpr (L >= 2) v |= dpp_i<0xB1>(v);
  if constexpr (L >= 4) v |= dpp_i<0x4E>(v);
  if constexpr (L >= 8) v |= dpp_i<0x141>(v);
  if constexpr (L >= 16) v |= dpp_i<0x140>(v);
  if constexpr (L >= 32) { auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false); v = p[0] | p[1]; }
  if constexpr (L >= 64) { auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false); v = p[0] | p[1]; }
  return (uint32_t)v;
}

__device__ __forceinline__ double cubed(double x) { return x * x * x; }

// FP64 quotient from the hardware reciprocal refined by two Newton steps
// (within about one unit in the last place of the correctly rounded quotient).
__device__ __forceinline__ double rcp_nr(double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = r * (2. - b * r);
  return r * (2. - b * r);
}
__device__ __forceinline__ double div_nr(double a, double b) { return a * rcp_nr(b); }

// x**(1/3) for positive x: an FP32 seed and one Halley step in FP64, which
// converges cubically from ~1e-7 to the FP64 rounding level.
__device__ __forceinline__ double cbrt_fast(double x) {
  if (!(x > 1.e-36 && x < 1.e36)) return ::cbrt(x);
  const double y0 = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)x) * (1.f / 3.f));
  const double y3 = y0 * y0 * y0;
  return y0 * div_nr(y3 + 2. * x, 2. * y3 + x);
}

// ---------------------------------------------------------------------------
// particle terms
// ---------------------------------------------------------------------------

// terminal_vel for liquid particles (Beard 1976), mo_particle_external_properties.f90:61-90
__device__ inline double terminal_vel_liquid(double diam, double rhop, double rhoa, double visc, double beta) {
  if (diam <= 19.e-6) return (diam * diam) * (rhop - rhoa) * kGrav * beta / (18. * visc);
  if (diam < 1.07e-3) {
    double nda = 4. / 3. * cubed(diam) * kGrav * (rhop - rhoa) * rhoa / (visc * visc);
    double X = ::log(nda);
    double X2 = X * X;
    double Y = -0.318657E1 + 0.992696 * X - 0.153193E-2 * X2 - 0.987059E-3 * X2 * X
               - 0.578878E-3 * X2 * X2 + 0.855176E-4 * X2 * X2 * X - 0.327815E-5 * X2 * X2 * X2;
    return visc / rhoa / diam * (beta * ::exp(Y));
  }
  double np_ = (kSurfw0 * kSurfw0 * kSurfw0) * (rhoa * rhoa) / ((visc * visc) * (visc * visc) * (rhop - rhoa) * kGrav);
  double bo = 4. / 3. * (rhop - rhoa) * kGrav / kSurfw0 * (diam * diam);
  double np6 = ::pow(np_, 1. / 6.);
  double X3 = ::log(bo * np6);
  double X2 = X3 * X3;
  double Y3 = -0.500015E1 + 0.523778E1 * X3 - 0.204914E1 * X2 + 0.475294 * X2 * X3
              - 0.542819E-1 * X2 * X2 + 0.238449E-2 * X2 * X2 * X3;
  return fmin(visc / rhoa / diam * (np6 * ::exp(Y3)), 9.2);
}

struct Shape { double alpha, beta, gamma, sigma; };

// getShapeCoefficients, mo_ice_shape.f90:145-191, for the shape parameters
// (alpha, beta) of the mass and (gamma, sigma) of the area power laws
__device__ inline Shape shape_coeffs_p(double ia, double ib, double ig, double isg, double mpri, double mrim,
                                       double numc) {
  Shape s{kPi6 * kRhoic, 3., kPi / 4., 2.};
  if (numc < 1.e-6) return s;
  double mtot = (mpri + mrim) / numc;
  double fr = fmax(fmin(mrim / (mtot * numc), 0.9999), 0.0001);
  double rho_b = fr * kRhori + (1. - fr) * kRhoic;
  double dth = ::pow(kPi6 * kRhoic / ia, 1. / (ib - 3.));
  double dgr = ::pow(ia / (kPi6 * rho_b), 1. / (3. - ib));
  double dcr = ::pow((1. / (1. - fr)) * ia / (kPi6 * rho_b), 1. / (3. - ib));
  double mth = kRhoic * kPi6 * cubed(dth), mgr = rho_b * kPi6 * cubed(dgr), mcr = rho_b * kPi6 * cubed(dcr);
  if (mtot >= mth && mtot < mgr) {
    s = Shape{ia, ib, ig, isg};
  } else if (mtot > mcr) {
    double h = 1. / (1. - fr);
    s = Shape{ia * h, ib, ig * (1. - fr) + kPi / 4. * fr, isg * (1. - fr) + 2. * fr};
  }
  return s;
}
__device__ inline Shape shape_coeffs(const DevLayout &D, double mpri, double mrim, double numc) {
  return shape_coeffs_p(D.ice_alpha, D.ice_beta, D.ice_gamma, D.ice_sigma, mpri, mrim, numc);
}


// terminal_vel for ice (Khvorostyanov & Curry 2002), :91-98 with kcVt..kcbre :103-158
__device__ inline double terminal_vel_ice(double dsph, double rhop, double rhoa, double visc, Shape sh, double dnsp) {
  const double c1 = 0.0902, delta0 = 9.06;
  double vb = kPi6 * cubed(dsph);
  double ap = sh.gamma * ::pow(dnsp, sh.sigma);
  double X = (2. * vb * (rhop - rhoa) * kGrav * dnsp * dnsp) / (ap * rhoa * visc * visc);
  double k1213 = ::sqrt(1. + c1 * ::sqrt(X));
  double bre = 0.5 * c1 * ::sqrt(X) / (k1213 - 1.) / k1213;
  double are = 0.25 * delta0 * delta0 * (k1213 - 1.) * (k1213 - 1.) / ::pow(X, bre);
  double av = are * ::pow(visc, 1. - 2. * bre) * ::pow((2. * sh.alpha * kGrav) / (rhoa * sh.gamma), bre);
  double bv = bre * (sh.beta - sh.sigma + 2.) - 1.;
  return av * ::pow(dnsp, bv);
}

// Per-cell constants of coagc (mo_salsa_coagulation_kernels.f90:264-266, 323-325, 389-391)
struct CellC {
  double T, visc, mfp, zrhoa, vkin, cshear, cinert;
};
__device__ inline CellC cell_consts(const DevLayout &D, double T, double P) {
  CellC c;
  c.T = T;
  c.visc = (7.44523e-3 * ::sqrt(T * T * T)) / (5093. * (T + 110.4));
  c.mfp = (1.656e-10 * T + 1.828e-8) * kPstand / P;
  c.zrhoa = P / (kRd * T);
  c.vkin = c.visc / c.zrhoa;
  c.cshear = ::sqrt(8. * kPi * D.eddy_dis / (15. * c.vkin));
  c.cinert = kPi * D.eddy_075 / (kGrav * ::sqrt(::sqrt(c.vkin)));
  return c;
}

// Particle terms of one colliding particle: diffusion coefficient, mean
// thermal speed, flux-matching omega and terminal velocity (coagc :280-286, 327-339).
struct PTerms { double D, c, w, v; };
__device__ inline PTerms particle_terms(const CellC &cc, double diam, double mass, double rhop, int phase,
                                 double dsph, const Shape &sh) {
  PTerms r;
  double knud = 2. * cc.mfp / diam;
  double beta = 1. + knud * (1.142 + 0.558 * ::exp(-0.999 / knud));
  r.D = beta * kBoltz * cc.T / (3. * kPi * cc.visc * diam);
  r.c = ::sqrt((8. * kBoltz * cc.T) / (kPi * mass));
  r.w = 8. * r.D / (kPi * r.c);
  r.v = phase < 4 ? terminal_vel_liquid(diam, rhop, cc.zrhoa, cc.visc, beta)
                  : terminal_vel_ice(dsph, rhop, cc.zrhoa, cc.visc, sh, diam);
  return r;
}

// The same terms for a liquid particle in the mixed path: diffusion
// coefficient, thermal speed and omega in FP32 (they only feed the FP32
// kernel coefficients); the Cunningham slip correction and the terminal
// velocity in FP64 (vd; the kernel keeps it as its FP32 value v plus an FP32
// residual), since pair_coef takes the difference of two particles' fall
// speeds, which cancels (twin pairs, neighbouring drop bins), and then the
// gravitational efficiency takes stok - 1.214, which cancels again -- FP32
// speeds with an FP32 slip factor left some kernels 7e-6 off.  Only the small
// exponential term of the slip correction is FP32.  The regime is chosen on
// the FP64 diameter, as in the FP64 path.
struct PTermsF { float D, c, w, v; double vd; };
__device__ inline PTermsF particle_terms_liquid_f(const CellC &cc, double diam_d, double mass_d, double rhop_d) {
  using M = Mth<float>;
  const float diam = (float)diam_d, mass = (float)mass_d;
  const float T = (float)cc.T, visc = (float)cc.visc;
  const float pi = (float)kPi, kb = (float)kBoltz;
  PTermsF r;
  const double knud_d = div_nr(2. * cc.mfp, diam_d);
  const float knud = (float)knud_d;
  const double beta_d = 1. + knud_d * (1.142 + 0.558 * (double)M::exp_(M::div_(-0.999f, knud)));
  const float beta = (float)beta_d;
  r.D = M::div_(beta * kb * T, 3.f * pi * visc * diam);
  r.c = M::sqrt_(M::div_(8.f * kb * T, pi * mass));
  r.w = M::div_(8.f * r.D, pi * r.c);
  // terminal_vel, Beard (1976), mo_particle_external_properties.f90:61-90
  if (diam_d <= 19.e-6)
    r.vd = div_nr((diam_d * diam_d) * (rhop_d - cc.zrhoa) * kGrav * beta_d, 18. * cc.visc);
  else
    r.vd = terminal_vel_liquid(diam_d, rhop_d, cc.zrhoa, cc.visc, beta_d);
  r.v = (float)r.vd;
  return r;
}

// Collision kernel of one pair from the two particles' terms (coagc :288-407).
// a_lrg: particle A is the larger (diam(1) >= diam(2), evaluated in FP64).
// vAl, vBl: in the mixed path the residuals of the FP32 fall speeds
// (vX + vXl is the FP64 speed to about 2^-48 relative); unused in FP64.
template <class KT>
__device__ __forceinline__ KT pair_coef(KT dA, KT DA, KT cA, KT wA, KT vA, KT dB, KT DB, KT cB, KT wB,
                                        KT vB, bool a_lrg, KT vkin, KT cshear, KT cinert, KT vAl = KT(0.),
                                        KT vBl = KT(0.)) {
  using M = Mth<KT>;
  const KT pi = KT(kPi);
  KT md = KT(0.5) * (dA + dB);
  KT flux1 = KT(4.) * pi * md * (DA + DB);
  KT flux2 = pi * M::sqrt_(cA * cA + cB * cB) * (md * md);
  KT tA, tB;
  if constexpr (sizeof(KT) == 8) {
    // reference form (:300-310)
    KT a2 = md * md + wA * wA, b2 = md * md + wB * wB;
    tA = ((md + wA) * (md + wA) * (md + wA) - a2 * M::sqrt_(a2)) / (KT(3.) * md * wA) - md;
    tB = ((md + wB) * (md + wB) * (md + wB) - b2 * M::sqrt_(b2)) / (KT(3.) * md * wB) - md;
  } else {
    // the same quantity without the two cancellations of the reference form:
    // tva = w (m + 2s + 4w - m w/(m+s)) / (3 (m + w + s)),  s = sqrt(m^2 + w^2)
    KT sA = M::sqrt_(md * md + wA * wA), sB = M::sqrt_(md * md + wB * wB);
    tA = M::div_(wA * (md + KT(2.) * sA + KT(4.) * wA - M::div_(md * wA, md + sA)), KT(3.) * (md + wA + sA));
    tB = M::div_(wB * (md + KT(2.) * sB + KT(4.) * wB - M::div_(md * wB, md + sB)), KT(3.) * (md + wB + sB));
  }
  KT fmd = M::sqrt_(tA * tA + tB * tB);
  KT zbrown = M::div_(flux1, M::div_(md, md + fmd) + M::div_(flux1, flux2));
  KT d_l = a_lrg ? dA : dB, v_l = a_lrg ? vA : vB, v_s = a_lrg ? vB : vA, D_s = a_lrg ? DB : DA;
  KT reyn = M::div_(d_l * v_l, vkin);
  KT schm = M::div_(vkin, D_s);
  KT dv, stok, stm;
  bool big;
  if constexpr (sizeof(KT) == 8) {
    dv = vA > vB ? vA - vB : vB - vA;
    stok = M::div_(KT(2.) * v_s * dv, d_l * KT(kGrav));
    big = stok > KT(1.214);
    stm = stok - KT(1.214);
  } else {
    // The fall-speed difference cancels (twin pairs, neighbouring drop bins):
    // vA - vB is exact in FP32 when the speeds are within a factor two of
    // each other (Sterbenz), and the residuals restore what rounding the
    // speeds to FP32 lost.  Where the collision efficiency takes stok - 1.214
    // close to the threshold (a second cancellation: below stok = 3 it loses
    // more than a factor 1.7) the Stokes number is formed in FP64 from the
    // FP64 speeds.
    dv = fabsf((vA - vB) + (vAl - vBl));
    stok = M::div_(KT(2.) * v_s * dv, d_l * KT(kGrav));
    big = stok > KT(1.214);
    stm = stok - KT(1.214);
    if (stok > KT(1.2) && stok < KT(3.)) {
      const double vAx = (double)vA + (double)vAl, vBx = (double)vB + (double)vBl;
      const double stokd = div_nr(2. * (a_lrg ? vBx : vAx) * fabs(vAx - vBx), (double)d_l * kGrav);
      stok = (KT)stokd;
      big = stokd > 1.214;
      stm = (KT)(stokd - 1.214);
    }
  }
  KT zbrconv = KT(0.);
  if (reyn <= KT(1.)) zbrconv = KT(0.45) * zbrown * M::cbrt_(reyn * schm);
  else if (reyn > KT(1.)) zbrconv = KT(0.45) * zbrown * M::sqrt_(reyn) * M::cbrt_(schm);
  KT ztshear = cshear * (md * md * md);
  KT zturb = cinert * (md * md) * dv;
  KT s1 = stok + KT(0.5);
  KT zea = M::div_(stok * stok, s1 * s1);
  KT zev = KT(0.);
  if (big) {
    zev = KT(0.75) * M::div_(M::log_(KT(2.) * stok), stm);
    zev = M::div_(KT(1.), (KT(1.) + zev) * (KT(1.) + zev));
  }
  KT zecoll = M::div_(KT(60.) * zev + zea * reyn, KT(60.) + reyn);
  KT zgrav = zecoll * pi * (md * md) * dv;
  KT hyp;
  if constexpr (sizeof(KT) == 8) {
    hyp = M::sqrt_(zgrav * zgrav + ztshear * ztshear + zturb * zturb);
  } else {
    // scaled so that the squares stay clear of the FP32 denormal range
    KT m = fmaxf(zgrav, fmaxf(ztshear, zturb));
    KT r = m > 0.f ? __builtin_amdgcn_rcpf(m) : 0.f;
    KT a = zgrav * r, b = ztshear * r, c = zturb * r;
    hyp = m * M::sqrt_(a * a + b * b + c * c);
  }
  return zbrown + zbrconv + hyp;
}

// ---------------------------------------------------------------------------
// per-cell LDS views
// ---------------------------------------------------------------------------
template <class KT>
struct PT4 { KT D, c, w, v; };      // particle terms of one bin: diffusion, speed, omega, fall speed

// rime fraction, classSection.f90:202-224
__device__ __forceinline__ double rime_frac(const DevLayout &D, double n, const double *v, int b) {
  if (D.phase[b] == 4 && v[D.iwa] > 1.e-23 && v[D.irim] > 1.e-23 && n > D.nlim[b])
    return v[D.irim] * kRhori / (v[D.iwa] * kRhoic + v[D.irim] * kRhori);
  return 0.;
}

// acth2o, mo_salsa_dynamics.f90:1127-1157, with the dry-solute term precomputed
template <bool FAST>
__device__ __forceinline__ double act_h2o(double znw, double zns) {
  double den = fmax(kEpsR8, znw + zns);
  double q;
  if constexpr (FAST) {
    double r = __builtin_amdgcn_rcp(den);      // ~2^-23, refined twice
    r = r * (2. - den * r);
    r = r * (2. - den * r);
    q = znw * r;
  } else {
    q = znw / den;
  }
  return fmax(0.1, q);
}

// Particle terms of an ice bin and of its self-offset twin (FP64), written to
// LDS; updateRhomean (classSection.f90:169-195) and getShapeCoefficients
// (mo_ice_shape.f90:145-191).
template <class KT>
__device__ __forceinline__ void ice_bin_terms(const DevLayout &D, const CellC &cc, double n, double nlim, const double *v,
                                           double dw, PT4<KT> *dst, PT4<KT> *dst_twin, KT *vres = nullptr,
                                           KT *vres_twin = nullptr) {
  double rho;
  if (n > nlim) {
    const double mp = kRhoic * v[D.iwa], mr = kRhori * v[D.irim];
    rho = fmin((mp * kRhoic + mr * kRhori) / fmax(mp + mr, 7.e-25), kRhoic);
  } else {
    rho = kRhoic;
  }
  double sdry = 0.;
  for (int s = 0; s < D.nspec - 1; ++s) sdry += v[s] * kRhoic;
  const Shape sh = shape_coeffs(D, sdry, v[D.nspec - 1] * kRhori, n);
  const double mass = rho * kPi6 * cubed(dw);
  const PTerms p0 = particle_terms(cc, dw, mass, rho, 4, dw, sh);
  const PTerms p1 = particle_terms(cc, 1.1 * dw, 1.33 * mass, rho, 4, dw, sh);
  *dst = PT4<KT>{(KT)p0.D, (KT)p0.c, (KT)p0.w, (KT)p0.v};
  *dst_twin = PT4<KT>{(KT)p1.D, (KT)p1.c, (KT)p1.w, (KT)p1.v};
  if (vres) { *vres = (KT)(p0.v - (double)(KT)p0.v); *vres_twin = (KT)(p1.v - (double)(KT)p1.v); }
}

}  // namespace salsa
