// salsa_dual_physics.h -- the collision-kernel physics in forward-mode dual
// arithmetic (salsa_dual.h): particle terms, terminal velocities, ice shape,
// coagc of one pair, rime fraction, water activity.  Shared by the
// forward-mode derivative kernel (salsa_gradient.hip) and the adjoint kernel
// (salsa_adjoint.hip), which takes the local derivatives of each pair kernel
// with respect to the two wet diameters from it.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include "salsa_dual.h"
#include "salsa_internal.h"

namespace salsa {
namespace {

using DD = Dual<double>;
using DF = Dual<float>;

// Per-cell constants of coagc (mo_salsa_coagulation_kernels.f90:264-266, 323-325, 389-391)
struct GCellC {
  double T, visc, mfp, zrhoa, vkin, cshear, cinert;
};
__device__ GCellC g_cell_consts(const DevLayout &D, double T, double P) {
  GCellC c;
  c.T = T;
  c.visc = (7.44523e-3 * ::sqrt(T * T * T)) / (5093. * (T + 110.4));
  c.mfp = (1.656e-10 * T + 1.828e-8) * kPstand / P;
  c.zrhoa = P / (kRd * T);
  c.vkin = c.visc / c.zrhoa;
  c.cshear = ::sqrt(8. * kPi * D.eddy_dis / (15. * c.vkin));
  c.cinert = kPi * D.eddy_075 / (kGrav * ::sqrt(::sqrt(c.vkin)));
  return c;
}

template <class R>
struct GPT { R D, c, w, v; };      // diffusion coefficient, mean speed, omega, fall speed

// terminal_vel for liquid particles (Beard 1976), mo_particle_external_properties.f90:61-90
__device__ DD g_tv_liquid(DD diam, double rhop, double rhoa, double visc, DD beta) {
  if (diam.v <= 19.e-6) return (diam * diam) * (rhop - rhoa) * kGrav * beta / (18. * visc);
  if (diam.v < 1.07e-3) {
    const DD nda = 4. / 3. * dcube(diam) * kGrav * (rhop - rhoa) * rhoa / (visc * visc);
    const DD X = dlog(nda);
    const DD X2 = X * X;
    const DD Y = -0.318657E1 + 0.992696 * X - 0.153193E-2 * X2 - 0.987059E-3 * X2 * X
                 - 0.578878E-3 * X2 * X2 + 0.855176E-4 * X2 * X2 * X - 0.327815E-5 * X2 * X2 * X2;
    return visc / rhoa / diam * (beta * dexp(Y));
  }
  const double np_ = (kSurfw0 * kSurfw0 * kSurfw0) * (rhoa * rhoa) /
                     ((visc * visc) * (visc * visc) * (rhop - rhoa) * kGrav);
  const DD bo = 4. / 3. * (rhop - rhoa) * kGrav / kSurfw0 * (diam * diam);
  const double np6 = ::pow(np_, 1. / 6.);
  const DD X3 = dlog(bo * np6);
  const DD X2 = X3 * X3;
  const DD Y3 = -0.500015E1 + 0.523778E1 * X3 - 0.204914E1 * X2 + 0.475294 * X2 * X3
                - 0.542819E-1 * X2 * X2 + 0.238449E-2 * X2 * X2 * X3;
  return dmin(visc / rhoa / diam * (np6 * dexp(Y3)), 9.2);
}

struct GShape { DD alpha, beta, gamma, sigma; };

// getShapeCoefficients, mo_ice_shape.f90:145-191
__device__ GShape g_shape(const DevLayout &D, DD mpri, DD mrim, DD numc) {
  GShape s{dconst(kPi6 * kRhoic), dconst(3.), dconst(kPi / 4.), dconst(2.)};
  if (numc.v < 1.e-6) return s;
  const DD mtot = (mpri + mrim) / numc;
  const DD fr = dmax(dmin(mrim / (mtot * numc), 0.9999), 0.0001);
  const DD rho_b = fr * kRhori + (1. - fr) * kRhoic;
  const double dth = ::pow(kPi6 * kRhoic / D.ice_alpha, 1. / (D.ice_beta - 3.));
  const DD dgr = dpow(D.ice_alpha / (kPi6 * rho_b), 1. / (3. - D.ice_beta));
  const DD dcr = dpow((1. / (1. - fr)) * D.ice_alpha / (kPi6 * rho_b), 1. / (3. - D.ice_beta));
  const double mth = kRhoic * kPi6 * (dth * dth * dth);
  const DD mgr = rho_b * kPi6 * dcube(dgr), mcr = rho_b * kPi6 * dcube(dcr);
  if (mtot >= mth && mtot < mgr) {
    s = GShape{dconst(D.ice_alpha), dconst(D.ice_beta), dconst(D.ice_gamma), dconst(D.ice_sigma)};
  } else if (mtot > mcr) {
    const DD h = 1. / (1. - fr);
    s = GShape{D.ice_alpha * h, dconst(D.ice_beta), D.ice_gamma * (1. - fr) + kPi / 4. * fr,
               D.ice_sigma * (1. - fr) + 2. * fr};
  }
  return s;
}

// terminal_vel for ice (Khvorostyanov & Curry 2002), :91-98 with kcVt..kcbre :103-158
__device__ DD g_tv_ice(DD dsph, DD rhop, double rhoa, double visc, const GShape &sh, DD dnsp) {
  const double c1 = 0.0902, delta0 = 9.06;
  const DD vb = kPi6 * dcube(dsph);
  const DD ap = sh.gamma * dpow(dnsp, sh.sigma);
  const DD X = (2. * vb * (rhop - rhoa) * kGrav * dnsp * dnsp) / (ap * rhoa * visc * visc);
  const DD sX = dsqrt(X);
  const DD k1213 = dsqrt(1. + c1 * sX);
  const DD bre = 0.5 * c1 * sX / (k1213 - 1.) / k1213;
  const DD are = 0.25 * delta0 * delta0 * (k1213 - 1.) * (k1213 - 1.) / dpow(X, bre);
  const DD av = are * dpow(visc, 1. - 2. * bre) * dpow((2. * sh.alpha * kGrav) / (rhoa * sh.gamma), bre);
  const DD bv = bre * (sh.beta - sh.sigma + 2.) - 1.;
  return av * dpow(dnsp, bv);
}

// Particle terms of one colliding particle (coagc :280-286, 327-339)
__device__ GPT<DD> g_pterms(const GCellC &cc, DD diam, DD mass, DD rhop, int phase, DD dsph, const GShape &sh) {
  GPT<DD> r;
  const DD knud = 2. * cc.mfp / diam;
  const DD beta = 1. + knud * (1.142 + 0.558 * dexp(-0.999 / knud));
  r.D = beta * kBoltz * cc.T / (3. * kPi * cc.visc * diam);
  r.c = dsqrt((8. * kBoltz * cc.T) / (kPi * mass));
  r.w = 8. * r.D / (kPi * r.c);
  r.v = phase < 4 ? g_tv_liquid(diam, rhop.v, cc.zrhoa, cc.visc, beta)
                  : g_tv_ice(dsph, rhop, cc.zrhoa, cc.visc, sh, diam);
  return r;
}

// The liquid terms in FP32 (mixed path), with the regime of the fall speed
// chosen on the FP64 diameter and its polynomial regimes kept in FP64.
__device__ GPT<DF> g_pterms_liq_f(const GCellC &cc, DD diam_d, DD mass_d, double rhop_d) {
  const DF diam = dcast<float>(diam_d), mass = dcast<float>(mass_d);
  const float rhop = (float)rhop_d;
  const float T = (float)cc.T, mfp = (float)cc.mfp, visc = (float)cc.visc, rhoa = (float)cc.zrhoa;
  const float pi = (float)kPi, g = (float)kGrav, kb = (float)kBoltz;
  GPT<DF> r;
  const DF knud = (2.f * mfp) / diam;
  const DF beta = 1.f + knud * (1.142f + 0.558f * dexp(-0.999f / knud));
  r.D = (beta * kb * T) / (3.f * pi * visc * diam);
  r.c = dsqrt((8.f * kb * T) / (pi * mass));
  r.w = (8.f * r.D) / (pi * r.c);
  if (diam_d.v <= 19.e-6)
    r.v = ((diam * diam) * (rhop - rhoa) * g * beta) / (18.f * visc);
  else
    r.v = dcast<float>(g_tv_liquid(diam_d, rhop_d, cc.zrhoa, cc.visc, dcast<double>(beta)));
  return r;
}

// Collision kernel of one pair (coagc :288-407); FP64 in the reference form,
// FP32 in the cancellation-free form of the step kernel's mixed path.
template <class T>
__device__ Dual<T> g_pair_coef(Dual<T> dA, Dual<T> DA, Dual<T> cA, Dual<T> wA, Dual<T> vA, Dual<T> dB, Dual<T> DB,
                               Dual<T> cB, Dual<T> wB, Dual<T> vB, bool a_lrg, T vkin, T cshear, T cinert) {
  using X = Dual<T>;
  const T pi = T(kPi);
  const X md = T(0.5) * (dA + dB);
  const X flux1 = T(4.) * pi * md * (DA + DB);
  const X flux2 = pi * dsqrt(cA * cA + cB * cB) * (md * md);
  X tA, tB;
  if constexpr (sizeof(T) == 8) {
    const X a2 = md * md + wA * wA, b2 = md * md + wB * wB;
    tA = ((md + wA) * (md + wA) * (md + wA) - a2 * dsqrt(a2)) / (T(3.) * md * wA) - md;
    tB = ((md + wB) * (md + wB) * (md + wB) - b2 * dsqrt(b2)) / (T(3.) * md * wB) - md;
  } else {
    const X sA = dsqrt(md * md + wA * wA), sB = dsqrt(md * md + wB * wB);
    tA = (wA * (md + T(2.) * sA + T(4.) * wA - (md * wA) / (md + sA))) / (T(3.) * (md + wA + sA));
    tB = (wB * (md + T(2.) * sB + T(4.) * wB - (md * wB) / (md + sB))) / (T(3.) * (md + wB + sB));
  }
  const X fmd = dsqrt(tA * tA + tB * tB);
  const X zbrown = flux1 / (md / (md + fmd) + flux1 / flux2);
  const X d_l = a_lrg ? dA : dB, v_l = a_lrg ? vA : vB, v_s = a_lrg ? vB : vA, D_s = a_lrg ? DB : DA;
  const X reyn = (d_l * v_l) / vkin;
  const X schm = vkin / D_s;
  const X dv = vA > vB ? vA - vB : vB - vA;
  const X stok = (T(2.) * v_s * dv) / (d_l * T(kGrav));
  X zbrconv = dconst(T(0.));
  if (reyn <= T(1.)) zbrconv = T(0.45) * zbrown * dcbrt(reyn * schm);
  else if (reyn > T(1.)) zbrconv = T(0.45) * zbrown * dsqrt(reyn) * dcbrt(schm);
  const X ztshear = cshear * (md * md * md);
  const X zturb = cinert * (md * md) * dv;
  const X s1 = stok + T(0.5);
  const X zea = (stok * stok) / (s1 * s1);
  X zev = dconst(T(0.));
  if (stok > T(1.214)) {
    zev = T(0.75) * (dlog(T(2.) * stok) / (stok - T(1.214)));
    zev = T(1.) / ((T(1.) + zev) * (T(1.) + zev));
  }
  const X zecoll = (T(60.) * zev + zea * reyn) / (T(60.) + reyn);
  const X zgrav = zecoll * pi * (md * md) * dv;
  X hyp;
  if constexpr (sizeof(T) == 8) {
    hyp = dsqrt(zgrav * zgrav + ztshear * ztshear + zturb * zturb);
  } else {
    // scaled so that the squares stay clear of the FP32 denormal range; the
    // tangent (g dg + s ds + t dt) / hyp in the same scaled form
    const T m = fmaxf(zgrav.v, fmaxf(ztshear.v, zturb.v));
    const T r = m > 0.f ? DMath<T>::rcp_(m) : 0.f;
    const T a = zgrav.v * r, b = ztshear.v * r, c = zturb.v * r;
    const T q = DMath<T>::sqrt_(a * a + b * b + c * c);
    hyp.v = m * q;
    hyp.d = q > 0.f ? (a * zgrav.d + b * ztshear.d + c * zturb.d) * DMath<T>::rcp_(q) : 0.f;
  }
  return zbrown + zbrconv + hyp;
}

// rime fraction, classSection.f90:202-224
__device__ DD g_rime_frac(const DevLayout &D, DD n, const DD *v, int b) {
  if (D.phase[b] == 4 && v[D.iwa] > 1.e-23 && v[D.irim] > 1.e-23 && n > D.nlim[b])
    return v[D.irim] * kRhori / (v[D.iwa] * kRhoic + v[D.irim] * kRhori);
  return dconst(0.);
}

// acth2o, mo_salsa_dynamics.f90:1127-1157, dry-solute term precomputed
__device__ __forceinline__ DD g_act_h2o(DD znw, DD zns) {
  const DD den = dmax(kEpsR8, znw + zns);
  return dmax(0.1, znw / den);
}

// Particle terms of an ice bin and of its self-offset twin; updateRhomean
// (classSection.f90:169-195) and getShapeCoefficients (mo_ice_shape.f90:145-191).
template <class KT>
__device__ void g_ice_terms(const DevLayout &D, const GCellC &cc, DD n, double nlim, const DD *v, DD dw,
                            GPT<Dual<KT>> &o0, GPT<Dual<KT>> &o1) {
  DD rho;
  if (n > nlim) {
    const DD mp = kRhoic * v[D.iwa], mr = kRhori * v[D.irim];
    rho = dmin((mp * kRhoic + mr * kRhori) / dmax(mp + mr, 7.e-25), kRhoic);
  } else {
    rho = dconst(kRhoic);
  }
  DD sdry = dconst(0.);
  for (int s = 0; s < D.nspec - 1; ++s) sdry += v[s] * kRhoic;
  const GShape sh = g_shape(D, sdry, v[D.nspec - 1] * kRhori, n);
  const DD mass = rho * kPi6 * dcube(dw);
  const GPT<DD> p0 = g_pterms(cc, dw, mass, rho, 4, dw, sh);
  const GPT<DD> p1 = g_pterms(cc, 1.1 * dw, 1.33 * mass, rho, 4, dw, sh);
  o0 = GPT<Dual<KT>>{dcast<KT>(p0.D), dcast<KT>(p0.c), dcast<KT>(p0.w), dcast<KT>(p0.v)};
  o1 = GPT<Dual<KT>>{dcast<KT>(p1.D), dcast<KT>(p1.c), dcast<KT>(p1.w), dcast<KT>(p1.v)};
}

__device__ __forceinline__ double species_rho(const DevLayout &D, int c, int s) {
  return c < 3 ? D.rholiq[s] : (s == D.irim ? kRhori : D.rhoice[s]);
}

}  // namespace
}  // namespace salsa
