// salsa_kernels.hip -- CDNA4 (gfx950) kernels for the SALSA coagulation +
// condensation step.
//
// One fused kernel handles a group of 64/L grid cells per 64-lane wavefront;
// the L lanes of a cell cooperate:
//   1. pack     : LES fields -> SALSA units in LDS        (mo_salsa_driver.f90:230-279)
//   2. bin props: wet diameter, density and the per-particle terms of coagc
//                 for every bin and its "self-offset" twin  (classSection.f90:91-195,
//                 mo_salsa_coagulation_kernels.f90:233-262, mo_particle_external_properties.f90:43-158)
//   3. pairs    : collision kernel K for every active bin pair -> LDS
//                 (mo_salsa_coagulation_kernels.f90:94-180, 264-420)
//   4. sweep    : the sequential semi-implicit target sweep, each target's
//                 sink/source sums reduced across the cell's lanes
//                 (mo_salsa_coagulation_processes.f90:24-1611)
//   5. condense : equilibration + analytic predictor-corrector substeps with
//                 lane-owned bins in registers (mo_salsa_dynamics.f90:691-1157,
//                 mo_salsa_properties.f90:73-380)
//   6. unpack   : tendencies (+=) or in-place state       (mo_salsa_driver.f90:310-388)
//
// KT is the arithmetic type of step 3 (double for the FP64 path, float for the
// FP32-mixed path); everything else -- state, sums, predicates, equilibration
// and the vapour budget -- is FP64 in both paths.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "salsa_activation.h"
#include "salsa_autoconv.h"
#include "salsa_distr.h"
#include "salsa_internal.h"
#include "salsa_physics.h"

namespace salsa {

// LES array index of interior cell `cell` (k fastest)
__device__ __forceinline__ int64_t les_index(const StepArgs &a, int64_t cell) {
  const int k = (int)(cell % a.nz);
  const int64_t r = cell / a.nz;
  const int i = (int)(r % a.nx), j = (int)(r / a.nx);
  return (int64_t)(k + a.hz) + (int64_t)a.nzp * ((int64_t)(i + a.hxy) + (int64_t)a.nxp * (j + a.hxy));
}

// Equilibration (mo_salsa_properties.f90:73-380) of one cell's aerosol bins,
// both regimes in one loop.  Regime 1 (1a bins, :165-269) is one sequential
// chain -- the Kelvin factor carries over from bin to bin -- and runs on lane
// 0 of the cell; the bins of regime 2 (2a/2b, the init branch :275-378, run
// when reg2) are independent and go to the cell's lanes 1, 2, ... (lane 0
// takes what is left after its chain).  Every lane iterates its current
// bin's Kelvin fixed point; the wavefront loops until all lanes are done, so
// the two regimes overlap.  zrh is the relative humidity before the regime
// clamps ([0.05, 0.98] and [0.37, 0.98]); the initial wet diameters come from
// sDW.  The aerosol water volumes are written to sV.
template <int L>
__device__ __forceinline__ void equilibrate(const DevLayout &D, double *sN, double *sV, int nsV, const double *sDW,
                                            double T, double zrh, bool reg2, bool live, int sub, uint32_t &flags) {
#define VV(b, s) sV[(b) * nsV + (s)]
  const int iwa = D.iwa, ndry = D.ndry;
  const int A0 = D.off[0];
  int isu = -1, ioc = -1;
  for (int s = 0; s < ndry; ++s) {
    if (D.dry_id[s] == 0) isu = s;
    if (D.dry_id[s] == 1) ioc = s;
  }
  const int n1 = D.fn1a - D.in1a + 1;
  const int n2 = reg2 ? D.fn2b - D.in2a + 1 : 0;
  // this lane's queue: regime-1 bins (lane 0), then regime-2 bins j with
  // (j + 1) % L == sub
  int q1 = sub == 0 ? 0 : n1;           // next regime-1 position
  int q2 = (sub + L - 1) % L;           // next regime-2 index
  double zke1 = 1.001;                  // regime-1 Kelvin factor, carried over
  int b = -1, reg = 0, count = 0;
  double n = 1., zcore = 0., c1 = 0., c2 = 0., c5 = 0., zr = 0., tol = 0.;
  double zdwet = 0., zdold = 1., zlwc = 0., zke = 1.;
  auto next_task = [&]() {
    reg = 0;
    while (q1 < n1) {                   // regime 1: skip bins with numc <= nlim
      const int bb = A0 + D.in1a + q1++;
      if (sN[bb] > kNlim) {
        b = bb; reg = 1; n = sN[bb];
        const double v1 = isu >= 0 ? VV(bb, isu) : 0., v2 = ioc >= 0 ? VV(bb, ioc) : 0.;
        zcore = v1 / n + v2 / n;
        c1 = v1 * (kRhosu / kMsu);
        c2 = kEpsoc * v2 * (kRhooc / kMoc);
        c5 = 0.;
        zr = fmin(fmax(zrh, 0.05), 0.98);
        tol = 1.e-2;
        zke = zke1;
        zdwet = sDW[bb]; zdold = 1.; zlwc = 0.; count = 0;
        return;
      }
    }
    while (q2 < n2) {                   // regime 2
      const int bb = A0 + D.in2a + q2;
      q2 += L;
      if (sN[bb] > kNlim) {
        b = bb; reg = 2; n = sN[bb];
        double zv[5] = {0., 0., 0., 0., 0.};
        for (int s = 0; s < ndry; ++s)
          if (D.dry_id[s] < 5) zv[D.dry_id[s]] = VV(bb, s);
        zcore = zv[0] / n + zv[1] / n + zv[2] / n + zv[3] / n + zv[4] / n;
        c1 = zv[0] * (kRhosu / kMsu);
        c2 = kEpsoc * (zv[1] * (kRhooc / kMoc));
        c5 = zv[4] * (kRhoss / kMss);
        zr = fmin(fmax(zrh, 0.37), 0.98);
        tol = 1.e-12;
        zke = 1.02;
        zdwet = sDW[bb]; zdold = 1.; zlwc = 0.; count = 0;
        return;
      }
    }
  };
  if (live) next_task();
  // a task is live while the reference's loop condition holds
  auto finish = [&]() {
    VV(b, iwa) = zlwc / kRhowa;
    if (reg == 1) zke1 = zke;
    next_task();
  };
  while (reg && !(fabs(zdwet / zdold - 1.) > tol)) finish();   // condition false on entry
#pragma unroll 1
  while (__any(reg != 0)) {
    if (reg) {
      zdold = fmax(zdwet, 1.e-20);
      const double aw = div_nr(zr, zke);
      const double aw2 = aw * aw;
      const double bm1 = 1.1065495e+2 - 3.6759197e+2 * aw + 5.0462934e+2 * aw2
                         - 3.1543839e+2 * (aw2 * aw) + 6.770824e+1 * (aw2 * aw2);
      const double bm2 = rcp_nr(aw * kMwa) - 1. / kMwa;
      double l = div_nr(c1, bm1) + div_nr(c2, bm2);
      if (reg == 2) {
        const double bm5 = 5.875248e+1 - 1.8781997e+2 * aw + 2.7211377e+2 * aw2
                           - 1.8458287e+2 * (aw2 * aw) + 4.153689e+1 * (aw2 * aw2);
        l = l + div_nr(c5, bm5);
      }
      zlwc = l;
      zdwet = cbrt_fast(div_nr(div_nr(div_nr(zlwc, n), kRhowa), kPi6) + div_nr(zcore, kPi6));
      zke = ::exp(div_nr(2. * kSurfw0 * kMvwa, kBoltz * T * zdwet));
      bool done = !(fabs(zdwet / zdold - 1.) > tol);
      if (++count > 100) { flags |= SALSA_CELL_NOCONV; done = true; }
      if (done) {
        finish();
        while (reg && !(fabs(zdwet / zdold - 1.) > tol)) finish();
      }
    }
  }
#undef VV
}

// Per-cell rate accumulators of the process-rate diagnostics
// (classProcessRates.f90:60-98), as the volumes the LES fields are made from.
enum RateAcc {
  RA_AUTOC = 0, RA_AUTOC50, RA_AUTOC80,   // Autoconversion(50/80)%volc(iwa)
  RA_ACCR, RA_ACCR50, RA_ACCR80,         // Accretion(50/80)%volc(iwa)
  RA_ACCOLL, RA_APCOLL,                  // SUM(ACcoll/APcoll%volc(1:ndry)*rholiq)
  RA_CONDA, RA_CONDC, RA_CONDP,          // Cond_a/c/p%volc(iwa)
  RA_CONDI,                              // Cond_i%volc(iwa)*rhoic + %volc(irim)*rhori
  RA_ACTIV,                              // Activation%numc
  RA_N
};

// Rate contributions of one source entry of a cloud (tcat 1) or precipitation
// (tcat 2) target, accumulateSource with its 50/80 um diagnostics
// (mo_salsa_coagulation_processes.f90:1031-1210).
__device__ __forceinline__ void rate_source_entry(const DevLayout &D, int tcat, int t, int j, double kj,
                                                  const double *Vj, const double *sDW, double &rdry,
                                                  double &racc, double &a80, double &a50, double &u80,
                                                  double &u50) {
  const int jc = j < D.off[1] ? 0 : (j < D.off[2] ? 1 : 2);
  if (jc == 0) {
    double dd = 0.;
    for (int s = 0; s < D.ndry; ++s) dd += Vj[s] * D.rholiq[s];
    rdry += kj * dd;
    return;
  }
  if (tcat != 2) return;
  const double w = kj * Vj[D.iwa];
  if (jc == 1) racc += w;
  const int R = D.off[2];
  const double Dt = sDW[t], Dj = sDW[j];
  const double lims[2] = {80.e-6, 50.e-6};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const double lim = lims[q];
    const bool okj = jc == 1 || D.precpbins[j - R] < lim;      // ex = MIN(iend, cnt-1) for precip collectors
    double &accr = q == 0 ? a80 : a50;
    double &au = q == 0 ? u80 : u50;
    if (okj && D.precpbins[t - R] > lim) accr += w;
    if (okj && Dt > 0. && Dt < lim && cubed(Dt) + cubed(Dj) > cubed(lim)) au += w;
  }
}

template <int L>
__device__ __forceinline__ uint64_t cell_or64(uint64_t x) {
  return (uint64_t)cell_or<L>((uint32_t)x) | ((uint64_t)cell_or<L>((uint32_t)(x >> 32)) << 32);
}

// ---------------------------------------------------------------------------
// the step kernel
//
// HYD = false is the variant for cells whose cloud, precipitation and ice bins
// are all below their activity limit (numc < nlim): their kernel values are
// zero and their sweep targets are skipped in the reference as well, so only
// the aerosol-aerosol pairs and the aerosol targets are evaluated.  The
// condensation stage still visits every bin.
// ---------------------------------------------------------------------------
template <int L, class KT, int MB, int NS, bool HYD, bool RATES>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8)))
k_salsa_step(const DevLayout *__restrict__ Dp, StepArgs a, SmemGeom G) {
  constexpr int CPW = 64 / L, NQ = (MB + L - 1) / L;
  extern __shared__ __align__(16) char smem[];
  const DevLayout &D = *Dp;
  const int lane = threadIdx.x, slot = lane / L, sub = lane % L;
  // Blocks are dealt to the 8 XCDs round-robin; give each XCD a contiguous
  // range of cells so that neighbouring cells (which share cache lines of
  // every field) meet in one L2.
  const uint32_t nblk = gridDim.x, bid = blockIdx.x;
  const uint32_t xcd = bid & 7u, rnk = bid >> 3, per = nblk >> 3, rem = nblk & 7u;
  const uint32_t lb = xcd * per + (xcd < rem ? xcd : rem) + rnk;
  const int64_t cid = (int64_t)lb * CPW + slot;
  const bool live = cid < a.ncells;
  const int nt = D.nt, nsV = G.nsV, ntab = G.ntab;
  const double dt = a.dt;

  uint32_t *tent = reinterpret_cast<uint32_t *>(smem);       // sweep entries of the layout
  char *cb = smem + G.tab_bytes + slot * G.cell_bytes;
  double *sN = reinterpret_cast<double *>(cb + G.oN);
  double *sV = reinterpret_cast<double *>(cb + G.oV);       // [nt][nsV]
  double *sDW = reinterpret_cast<double *>(cb + G.odw);
  PT4<KT> *sPR = reinterpret_cast<PT4<KT> *>(cb + G.opr);   // [2][ntab]: bin, self-offset twin
  KT *sK = reinterpret_cast<KT *>(cb + G.oK);
  // mixed path: residuals of the FP32 fall speeds of every bin and its twin [2][ntab]
  KT *sVL = sizeof(KT) == 4 ? reinterpret_cast<KT *>(cb + G.ovd) : nullptr;
  // precipitation formation per cloud-cloud pair: bits 0-4 target precip bin + 1
  // (0: none), bit 5 D_new > 50 um, bit 6 D_new > 80 um
  uint8_t *sPF = reinterpret_cast<uint8_t *>(cb + G.opf);
  constexpr bool R = RATES;                                    // process-rate diagnostics
  double *rr = RATES ? reinterpret_cast<double *>(cb + G.orr) : nullptr;
  if (R && sub == 0)
    for (int k = 0; k < RA_N; ++k) rr[k] = 0.;
#define VV(b, s) sV[(b) * nsV + (s)]

  int32_t icell = 0;
  const double *rec = nullptr;
  double T = 280., P = 1.e5, rs = 0.01, rsi = 0.01, dn = 1., rv = 0.;
  if (live) {
    icell = a.cell_list ? a.cell_list[cid] : (int32_t)cid;
    rec = a.rec_in + (int64_t)icell * D.nslot;
    T = rec[0]; P = rec[1]; rs = rec[2]; rsi = rec[3]; dn = rec[4]; rv = rec[5];
  }
  if (!a.lsfreeRH) rv = fmin(rv, rs * a.rhlim);     // mo_salsa_driver.f90:193-197
  const double rv0 = rv;                             // the tendency is taken against the limited value
  uint32_t flags = 0;

  // ---- 1. pack (mo_salsa_driver.f90:236-279) from the cell record ----
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int b = sub + q * L;
    if (b < nt) {
      double n = 0.;
      double v[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) v[s] = 0.;
      if (live) {
        const int c = D.cat_of[b];
        const int nsw = c < 3 ? D.nwet : D.nspec;
        const double *r = rec + D.sbase[b];
        n = r[0] * dn;
#pragma unroll
        for (int s = 0; s < NS; ++s)
          if (s < nsw) {
            const double rho = c < 3 ? D.rholiq[s] : (s == D.irim ? kRhori : D.rhoice[s]);
            v[s] = r[1 + s] * dn / rho;
          }
      }
      sN[b] = n;
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (s < nsV) VV(b, s) = v[s];
    }
  }
  for (int e = lane; e < D.nent_total; e += 64) tent[e] = D.ent[e];
  __syncthreads();

  // ---- 0. initialization call (run_SALSA with initialize = .TRUE.,
  //      mo_salsa_driver.f90:282-283): equilibration(init_rh = 0.3,
  //      init = .TRUE.) of the packed cell before the step.  The initial wet
  //      diameters are the cell's own updateDiameter(limit, "all") values
  //      (classSection.f90:91-159); the reference uses whatever the shared
  //      Section array was left holding by the cell processed before. ----
  if (a.init) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int b = sub + q * L;
      if (b < D.nb[0]) {
        const double n = sN[b];
        double vs = 0.;
#pragma unroll
        for (int s = 0; s < NS; ++s)
          if (s < nsV) vs += VV(b, s);
        const double dw = n > D.nlim[b] ? cbrt_fast(vs / n / kPi6) : D.dmid[b];
        sDW[b] = fmin(dw, D.dlim[b]);
      }
    }
    __syncthreads();
    equilibrate<L>(D, sN, sV, nsV, sDW, T, 0.3, true, live, sub, flags);
    __syncthreads();
  }

  // ---- 2./3./4. coagulation (mo_salsa_dynamics.f90:92-202) ----
  if (D.lscoag) {
    const CellC cc = cell_consts(D, T, P);
    uint64_t abits = 0;      // bins taking part in kernels: numc >= nlim (kernels.f90:116-168)
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) {
      const int b = sub + q * L;
      if (b < nt) {
        const double n = sN[b];
        double vs = 0.;
#pragma unroll
        for (int s = 0; s < NS; ++s)
          if (s < nsV) vs += VV(b, s);
        // updateDiameter(limit, "all"): dnsp ends equal to the limited dwet (classSection.f90:147-156)
        double dw = n > D.nlim[b] ? cbrt_fast(vs / n / kPi6) : D.dmid[b];
        dw = fmin(dw, D.dlim[b]);
        sDW[b] = dw;
        if (b < ntab) {
          if (!(n < D.nlim[b])) abits |= 1ull << b;
          const int ph = D.phase[b];
          if (ph == 4) {
            if (sizeof(KT) == 4)
              ice_bin_terms<KT>(D, cc, n, D.nlim[b], sV + b * nsV, dw, sPR + b, sPR + ntab + b, sVL + b,
                                sVL + ntab + b);
            else
              ice_bin_terms<KT>(D, cc, n, D.nlim[b], sV + b * nsV, dw, sPR + b, sPR + ntab + b);
          } else {
            const double rho = kRhowa;                       // updateRhomean, classSection.f90:169-195
            const double mass = rho * kPi6 * cubed(dw);
            // self-offset twin: diameter 1.1x, mass 1.33x (kernels.f90:273-276)
            if (sizeof(KT) == 4) {
              const PTermsF p0 = particle_terms_liquid_f(cc, dw, mass, rho);
              const PTermsF p1 = particle_terms_liquid_f(cc, 1.1 * dw, 1.33 * mass, rho);
              sPR[b] = PT4<KT>{(KT)p0.D, (KT)p0.c, (KT)p0.w, (KT)p0.v};
              sPR[ntab + b] = PT4<KT>{(KT)p1.D, (KT)p1.c, (KT)p1.w, (KT)p1.v};
              sVL[b] = (KT)(p0.vd - (double)p0.v);
              sVL[ntab + b] = (KT)(p1.vd - (double)p1.v);
            } else {
              const Shape sh{0., 0., 0., 0.};
              const PTerms p0 = particle_terms(cc, dw, mass, rho, ph, dw, sh);
              const PTerms p1 = particle_terms(cc, 1.1 * dw, 1.33 * mass, rho, ph, dw, sh);
              sPR[b] = PT4<KT>{(KT)p0.D, (KT)p0.c, (KT)p0.w, (KT)p0.v};
              sPR[ntab + b] = PT4<KT>{(KT)p1.D, (KT)p1.c, (KT)p1.w, (KT)p1.v};
            }
          }
        }
      }
    }
    abits = cell_or64<L>(abits);
    __syncthreads();

    // ---- 3. pairs (buildKernelSelf / buildKernel, kernels.f90:94-180) ----
    {
      const KT vkin = (KT)cc.vkin, cshear = (KT)cc.cshear, cinert = (KT)cc.cinert;
      const KT cfac = (KT)D.charge_factor;
      const int np = G.npairs, pf0 = D.pf0, pf1 = D.pf0 + D.ncc;
      // lscoag mode 2 (mo_salsa_driver.f90:601-697): kernels of this cell as
      // last stored (fetch steps) or stored after this evaluation (update steps)
      KT *ks = a.kmode && live ? reinterpret_cast<KT *>(a.kstore) + (int64_t)icell * D.npair : nullptr;
      const bool fetch = a.kmode == 2;
      for (int p = sub; p < np; p += L) {
        const uint32_t pw = D.pw[p];
        const int b1 = pw & 255u, b2 = (pw >> 8) & 255u, tw = (pw >> 16) & 1u;
        KT k = KT(0.);
        uint32_t pfb = 0;
        if (fetch) {
          if (ks) k = ks[p];
          // every cloud pair: a stored kernel may be non-zero for a bin now below nlim
          if (HYD && p >= pf0 && p < pf1 && live) {
            const double dnew = cbrt_fast(cubed(sDW[b1]) + cubed(sDW[b2]));
            int cnt = 0;
            for (int qq = 0; qq < D.nb[2]; ++qq) cnt += dnew > D.precpbins[qq];
            pfb = dnew > D.precpbins[0] ? (uint32_t)cnt : 0u;
            pfb |= (dnew > 50.e-6 ? 32u : 0u) | (dnew > 80.e-6 ? 64u : 0u);
          }
        } else if (live && ((abits >> b1) & (abits >> b2) & 1ull)) {
          // particle 1 is the twin of particle 2 when the bins' dmid agree within 1%
          const double dB = sDW[b2];
          const double dA = tw ? 1.1 * dB : sDW[b1];
          const int ia = tw ? ntab + b2 : b1;
          const PT4<KT> A = sPR[ia], B = sPR[b2];
          k = pair_coef<KT>((KT)dA, A.D, A.c, A.w, A.v, (KT)dB, B.D, B.c, B.w, B.v, dA >= dB, vkin, cshear,
                            cinert, sVL ? sVL[ia] : KT(0.), sVL ? sVL[b2] : KT(0.));
          k = k * cfac;                                 // charge enhancement (:411-416)
          if (HYD && p >= pf0 && p < pf1) {
            // precip bin reached by coalescence (accumulatePrecipFormation, processes.f90:1461-1467)
            double dnew = cbrt_fast(cubed(sDW[b1]) + cubed(dB));
            int cnt = 0;
            for (int qq = 0; qq < D.nb[2]; ++qq) cnt += dnew > D.precpbins[qq];
            pfb = dnew > D.precpbins[0] ? (uint32_t)cnt : 0u;     // COUNT(D_new > precpbins): bin + 1
            pfb |= (dnew > 50.e-6 ? 32u : 0u) | (dnew > 80.e-6 ? 64u : 0u);
          }
        }
        sK[p] = k;
        if (ks && !fetch) ks[p] = k;
        if (HYD && p >= pf0 && p < pf1) sPF[p - pf0] = (uint8_t)pfb;
      }
      // the aerosol-only variant stores zeros for the pairs it does not evaluate
      if (!HYD && ks && !fetch)
        for (int p = np + sub; p < D.npair; p += L) ks[p] = KT(0.);
    }
    __syncthreads();

    // ---- 4. sweep (coag_ice, coag_precp, coag_aero, coag_cloud) ----
    // Each target visits only its entry list (the bins its masks name, with
    // their roles); the lanes of a cell split the entries and the partial
    // sums are reduced across them.
    const int nspec = D.nspec, ndry = D.ndry, iwa = D.iwa, irim = D.irim;
    const int Roff = D.off[2], pf0 = D.pf0;
    for (int it = G.it0; it < G.it1; ++it) {
      const DevTarget &TG = D.tgt[it];
      const int t = TG.t;
      const uint32_t fl = TG.flags;
      const double Nt = sN[t];
      double Vt[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) Vt[s] = s < nsV ? VV(t, s) : 0.;
      bool act = live && !(Nt < D.nlim[t]);
      if (fl & TF_ICE) act = act && T < 273.15;          // mo_salsa_dynamics.f90:166
      if (!__any(act)) continue;            // no cell of this wavefront updates the target
      const double rft = (fl & TF_ICE) ? rime_frac(D, Nt, Vt, t) : 0.;
      const double emax = D.Eiagg_max, emin = D.Eiagg_min;
      const int nu = (fl & TF_ICE) ? nspec : TG.nupd;
      const int nupd = TG.nupd;
      const uint32_t *ent = tent + TG.ent0;
      const int ne = HYD ? TG.nent : TG.nent_aero;

      double sink = 0., so = 0., sx = 0., src[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) src[s] = 0.;
      uint32_t pmask = 0;
      // rate diagnostics of this target (cloud and precipitation targets)
      const bool RT = R && (TG.cat == 1 || TG.cat == 2);
      double rdry = 0., racc = 0., a80 = 0., a50 = 0., u80 = 0., u50 = 0.;
      if (!HYD || !(fl & (TF_ICE | TF_PF))) {
        // plain collection (accumulateSink / accumulateSource, :707-1210)
        for (int e = sub; e < ne; e += L) {
          const uint32_t w = ent[e];
          const int j = w & 63u, pid = (int)(w >> 16);
          const uint32_t role = (w >> 6) & 1023u;
          const double kj = (double)sK[pid];
          const double ks = (role & R_SINK) ? kj : 0., kv = (role & R_SRC) ? kj : 0.;
          const double *Vj = sV + j * nsV;
          sink += ks * sN[j];
#pragma unroll
          for (int s = 0; s < NS; ++s)
            if (s < nupd) src[s] += kv * Vj[s];
          if (RT && (role & R_SRC)) rate_source_entry(D, TG.cat, t, j, kj, Vj, sDW, rdry, racc, a80, a50, u80, u50);
        }
      } else {
        for (int e = sub; e < ne; e += L) {
          const uint32_t w = ent[e];
          const int j = w & 63u, pid = (int)(w >> 16);
          const uint32_t role = (w >> 6) & 1023u;
          const double kj = (double)sK[pid];
          const double Nj = sN[j];
          const double *Vj = sV + j * nsV;
          if (role & R_SINK) sink += kj * Nj;
          if (role & (R_SINK_EAGG | R_SRC_EAGG)) {      // ice aggregation efficiency (:758-769)
            const double eg = emax - fmax(rime_frac(D, Nj, Vj, j), rft) * (emax - emin);
            if (role & R_SINK_EAGG) sink += eg * kj * Nj;
            if (role & R_SRC_EAGG) {
#pragma unroll
              for (int s = 0; s < NS; ++s)
                if (s < nspec) src[s] += eg * kj * Vj[s];
            }
          }
          if (role & R_SRC) {
#pragma unroll
            for (int s = 0; s < NS; ++s)
              if (s < nupd) src[s] += kj * Vj[s];
            if (RT) rate_source_entry(D, TG.cat, t, j, kj, Vj, sDW, rdry, racc, a80, a50, u80, u50);
          }
          if (role & (R_PC_WA | R_PC_RI)) {              // accumulateSourcePhaseChange (:1236-1266)
#pragma unroll
            for (int s = 0; s < NS; ++s)
              if (s < ndry) src[s] += kj * Vj[s];
            if (role & R_PC_WA) src[iwa] += kj * Vj[iwa] * kRhowa / kRhoic;
            else src[irim] += kj * Vj[iwa] * kRhowa / kRhori;
          }
          if (role & (R_PF_OWN | R_PF_OTH)) {            // accumulatePrecipFormation (:1458-1576)
            const int pb = (int)(sPF[pid - pf0] & 31u) - 1;
            if (pb >= 0) {
              sink += kj * Nj;
              pmask |= 1u << pb;
            } else {
#pragma unroll
              for (int s = 0; s < NS; ++s)
                if (s < nupd) src[s] += kj * Vj[s];
            }
          }
          if (role & R_FIX_OWN) so += kj * Nj;
          if (role & R_FIX_OTH) sx += kj * Nj;
        }
      }
      sink = cell_sum<L>(sink);
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (s < nu) src[s] = cell_sum<L>(src[s]);
      if (RT) {
        rdry = cell_sum<L>(rdry);
        if (TG.cat == 2) {
          racc = cell_sum<L>(racc);
          a80 = cell_sum<L>(a80); a50 = cell_sum<L>(a50);
          u80 = cell_sum<L>(u80); u50 = cell_sum<L>(u50);
        }
        if (act && sub == 0) {                            // rates scale with the target numc (:1170-1196)
          if (TG.cat == 1) {
            rr[RA_ACCOLL] += rdry * Nt;
          } else {
            rr[RA_APCOLL] += rdry * Nt;
            rr[RA_ACCR] += racc * Nt;
            rr[RA_ACCR80] += a80 * Nt; rr[RA_ACCR50] += a50 * Nt;
            rr[RA_AUTOC80] += u80 * Nt; rr[RA_AUTOC50] += u50 * Nt;
          }
        }
      }
      double fix_own = 1., fix_oth = 1.;
      if (HYD && (fl & TF_PF)) {                          // fix_coag (:1440-1456)
        so = cell_sum<L>(so);
        sx = cell_sum<L>(sx);
        fix_own = fmax(1. - dt * so, 0.1);
        fix_oth = fmax(1. - dt * sx, 0.1);
      }

      // self terms (computed identically by every lane of the cell)
      double sink_self = 0.;
      double volsink[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) volsink[s] = 0.;
      int pself = -1;
      double kself = 0., fixs = 1.;
      if (fl & (TF_SELF_HALF | TF_SELF_HALF_EAGG | TF_PF_SELF | TF_PRECIP_SELFCOLL)) {
        const int pid = D.pidx[t][t];
        kself = pid >= 0 ? (double)sK[pid] : 0.;
        if (fl & TF_SELF_HALF) sink_self = 0.5 * kself * Nt;
        if (fl & TF_SELF_HALF_EAGG) sink_self = (emax - rft * (emax - emin)) * 0.5 * kself * Nt;
        if (HYD && (fl & TF_PF_SELF)) {
          fixs = fmax(1. - dt * (0.5 * kself * (Nt * Nt) / Nt), 0.1);
          pself = pid >= 0 ? (int)(sPF[pid - pf0] & 31u) - 1 : -1;
          sink_self = 0.5 * kself * Nt;
          if (pself >= 0) {
#pragma unroll
            for (int s = 0; s < NS; ++s)
              if (s < nupd) volsink[s] = Vt[s] * Nt * kself * fixs;
          }
        }
      }

      // applyCoag / applyCoagIce / applyCoagPrecp (:539-645)
      double Vn[NS];
      const double rden = rcp_nr(1. + dt * sink);
#pragma unroll
      for (int s = 0; s < NS; ++s) Vn[s] = s < nu ? (Vt[s] + dt * src[s] * Nt) * rden : Vt[s];
      double Nn = Nt * rcp_nr(1. + dt * (sink + sink_self));
      if (HYD && (fl & TF_PRECIP_SELFCOLL)) {         // precipSelfCoag (:1588-1611)
        const double nslf = 0.5 * kself * Nt * Nt;
        if (act && sub == 0) {
#pragma unroll
          for (int s = 0; s < NS; ++s)
            if (s < nupd) {
              const double vslf = Vt[s] * Nt * kself;
              VV(t + 1, s) = VV(t + 1, s) + dt * vslf;
              Vn[s] = Vn[s] - dt * vslf;
            }
          sN[t + 1] = sN[t + 1] + dt * nslf;
        }
        Nn = Nn - dt * nslf;
      }
      if (HYD && (fl & TF_PF_SELF)) {
#pragma unroll
        for (int s = 0; s < NS; ++s) Vn[s] = Vn[s] - dt * volsink[s];
      }
      if (act && sub == 0) {
        sN[t] = Nn;
#pragma unroll
        for (int s = 0; s < NS; ++s)
          if (s < nsV) VV(t, s) = Vn[s];
      }

      // transfer of coalesced cloud drops into precipitation bins (applyCoagPrecipFormation :649-703)
      if (HYD && (fl & (TF_PF | TF_PF_SELF))) {
        uint32_t pm = cell_or<L>(pmask);
        if (pself >= 0) pm |= 1u << pself;
        if (!act) pm = 0;
        while (pm) {
          const int pb = __builtin_ctz(pm);
          pm &= pm - 1;
          double vp[NS], npp = 0., p50 = 0., p80 = 0.;
#pragma unroll
          for (int s = 0; s < NS; ++s) vp[s] = 0.;
          for (int e = sub; e < ne; e += L) {
            const uint32_t w = ent[e];
            const uint32_t role = (w >> 6) & 1023u;
            if (!(role & (R_PF_OWN | R_PF_OTH))) continue;
            const int j = w & 63u, pid = (int)(w >> 16);
            const uint32_t pfv = sPF[pid - pf0];
            if ((int)(pfv & 31u) - 1 == pb) {
              const double kj = (double)sK[pid];
              const double fx = (role & R_PF_OWN) ? fix_own : fix_oth;
              const double Nj = sN[j];
#pragma unroll
              for (int s = 0; s < NS; ++s)
                if (s < nupd) vp[s] += (VV(j, s) * Nt + Vt[s] * Nj) * kj * fx;
              npp += kj * Nj * Nt * fx;
              if (R) {                                      // Autoconversion50/80 (:1487-1500)
                const double dvw = (VV(j, iwa) * Nt + Vt[iwa] * Nj) * kj * fx;
                if (pfv & 32u) p50 += dvw;
                if (pfv & 64u) p80 += dvw;
              }
            }
          }
          npp = cell_sum<L>(npp);
#pragma unroll
          for (int s = 0; s < NS; ++s)
            if (s < nupd) vp[s] = cell_sum<L>(vp[s]);
          if (R) {
            p50 = cell_sum<L>(p50);
            p80 = cell_sum<L>(p80);
            if (sub == 0) {
              rr[RA_AUTOC] += vp[iwa];
              rr[RA_AUTOC50] += p50;
              rr[RA_AUTOC80] += p80;
              if (pb == pself) {
                const uint32_t pfs = sPF[D.pidx[t][t] - pf0];
                rr[RA_AUTOC] += volsink[iwa];
                if (pfs & 32u) rr[RA_AUTOC50] += volsink[iwa];
                if (pfs & 64u) rr[RA_AUTOC80] += volsink[iwa];
              }
            }
          }
          if (pb == pself) {
#pragma unroll
            for (int s = 0; s < NS; ++s) vp[s] += volsink[s];
            npp += 0.5 * kself * Nt * Nt * fixs;
          }
          if (sub == 0) {
            const int r = Roff + pb;
#pragma unroll
            for (int s = 0; s < NS; ++s)
              if (s < nupd) VV(r, s) = VV(r, s) + dt * vp[s];
            sN[r] = sN[r] + dt * npp;
          }
        }
      }
      __syncthreads();
    }

    // clamp (mo_salsa_dynamics.f90:189-198)
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int b = sub + q * L;
      if (b < nt) {
        if (sN[b] < -1.e-6) flags |= SALSA_CELL_NEG_NUMC;
        sN[b] = fmax(0., sN[b]);
#pragma unroll
        for (int s = 0; s < NS; ++s)
          if (s < nsV) VV(b, s) = fmax(0., VV(b, s));
      }
    }
  } else {
    // without coagulation the wet diameters are refreshed from the state
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int b = sub + q * L;
      if (b < nt) {
        double vs = 0.;
        for (int s = 0; s < nsV; ++s) vs += VV(b, s);
        double dw = sN[b] > D.nlim[b] ? cbrt_fast(vs / sN[b] / kPi6) : D.dmid[b];
        sDW[b] = fmin(dw, D.dlim[b]);
      }
    }
  }
  __syncthreads();

  // ---- 5. condensation (condensation/gpparth2o, mo_salsa_dynamics.f90:276-342, 691-1125) ----
  if (D.lscnd && (D.lscndh2oae || D.lscndh2ocl || D.lscndh2oic)) {
    const int iwa = D.iwa, irim = D.irim, ndry = D.ndry;
    const int A0 = D.off[0];
    const double zrh = rv / rs;
    double w1 = 0.;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int b = sub + q * L;
      if (b < D.nb[0]) w1 += VV(b, iwa);
    }
    w1 = cell_sum<L>(w1) * kRhowa;

    // Equilibration (mo_salsa_properties.f90:73-380): regime 1 always, regime
    // 2 only when zrh < 0.98 or without aerosol condensation (:769-779)
    equilibrate<L>(D, sN, sV, nsV, sDW, T, zrh, zrh < 0.98 || !D.lscndh2oae, live, sub, flags);
    __syncthreads();
    double w2 = 0.;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int b = sub + q * L;
      if (b < D.nb[0]) w2 += VV(b, iwa);
    }
    w2 = cell_sum<L>(w2) * kRhowa;
    rv = rv - (w2 - w1) / (P * kMair / (kRg * T));

    // cell-skip predicate (:784-788; .AND. binds before .OR.)
    uint32_t any = 0;   // bit0 cloud>nlim, bit1 precip>prlim, bit2 ice>prlim, bit3 aero>nlim
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int b = sub + q * L;
      if (b < nt) {
        const int c = D.cat_of[b];
        const double n = sN[b];
        if (c == 0 && n > kNlim) any |= 8u;
        if (c == 1 && n > kNlim) any |= 1u;
        if (c == 2 && n > kPrlim) any |= 2u;
        if (c == 3 && n > kPrlim) any |= 4u;
      }
    }
    any = cell_or<L>(any);
    const bool pass = live && (((any & 1u) || ((any & 2u) && D.lscndh2ocl)) || ((any & 4u) && D.lscndh2oic) ||
                               ((any & 8u) && zrh > 0.98 && D.lscndh2oae));
    if (pass) {
      const double rhoair = kMair * P / (kRg * T);
      double zdf = (5. / (16. * kAvog * rhoair * 1.e-3 * (3.11e-8 * 3.11e-8))) *
                   ::sqrt(kRg * 1.e7 * T * kMair * 1.e3 * (kMwa + kMair) * 1.e3 / (2. * kPi * kMwa * 1.e3));
      zdf = zdf * 1.e-4;
      const double zmfp = 3. * zdf * ::sqrt(kPi * kMwa / (8. * kRg * T));
      const double zthc = 0.023807 + 7.1128e-5 * (T - 273.16);
      const double hl = ((kAlv * kMwa) / (kRg * T)) - 1.;
      const double hs = ((kAls * kMwa) / (kRg * T)) - 1.;
      // flat-surface saturation concentrations of liquid and ice
      const double csul = rs * rhoair / kMwa, csui = rsi * rhoair / kMwa;
      // Per bin: moles of water, transfer coefficient, Kelvin factor, solute
      // moles and saturation ratio.  The set-up runs as a plain loop into the
      // cell's (now free) kernel region of LDS, so that its long transcendental
      // chains are not interleaved across a lane's bins; the substeps then keep
      // the lane's bins in registers.
      double *cs = reinterpret_cast<double *>(cb + G.opr);   // [5][nt]
      uint32_t upd = 0;        // bit q: bin updated in the substeps
      uint32_t iceq = 0;       // bit q: ice bin
#pragma unroll 1
      for (int q = 0; q < NQ; ++q) {
        const int b = sub + q * L;
        if (b < nt) {
          double zcw0, zmt0 = 0., zkl0 = 1., zws0 = 0.;
          const int c = D.cat_of[b];
          const double n = sN[b];
          double vs = 0.;
#pragma unroll
          for (int s = 0; s < NS; ++s)
            if (s < nsV) vs += VV(b, s);
          double dw = n > D.nlim[b] ? cbrt_fast(vs / n / kPi6) : D.dmid[b];
          dw = fmin(dw, D.dlim[b]);                        // updateDiameter (:803-806); dnsp == dwet
          double zn = 0.;
          for (int s = 0; s < ndry; ++s) zn += D.diss[s] * VV(b, s) * D.rholiq[s] / D.mm[s];
          if (c == 3) {
            iceq |= 1u << q;
            zcw0 = VV(b, iwa) * kRhoic / kMwa + VV(b, irim) * kRhori / kMwa;
          } else {
            zcw0 = VV(b, iwa) * kRhowa / kMwa;
          }
          const bool on = n > D.nlim[b] &&
                          (c == 0 ? (zrh > 0.98 && D.lscndh2oae)
                                  : c == 3 ? (D.lscndh2oic && T < 273.15) : D.lscndh2ocl);
          if (on) {
            const double zknud = 2. * zmfp / dw;
            const double zbeta = (zknud + 1.) / (0.377 * zknud + 1. + 4. / 3. * (zknud + zknud * zknud));
            if (c == 3) {                                   // :892-930 (capacitance 0.5*dnsp, activity 1)
              zkl0 = ::exp(4. * kSurfi0 * kMwa / (kRg * T * kRhowa * dw));
              zws0 = 1.0 * zkl0;
              const double h1 = n * 4. * kPi * (0.5 * dw) * zdf * zbeta;
              const double h2 = kMwa * zdf * kAls * zws0 * csui / (zthc * T);
              zmt0 = h1 / (h2 * hs + 1.);
            } else {                                        // :826-890, :932-955
              const double zact = act_h2o<false>(VV(b, iwa) * D.rholiq[iwa] / D.mm[iwa], zn);
              zkl0 = ::exp(4. * kSurfw0 * kMwa / (kRg * T * kRhowa * dw));
              zws0 = zact * zkl0;
              const double h1 = n * 2. * kPi * dw * zdf * zbeta;
              const double h2 = kMwa * zdf * kAlv * zws0 * csul / (zthc * T);
              zmt0 = h1 / (h2 * hl + 1.);
            }
          }
          cs[b] = zcw0; cs[nt + b] = zmt0; cs[2 * nt + b] = zkl0; cs[3 * nt + b] = zn; cs[4 * nt + b] = zws0;
          // aerosol 1a bins sit outside the predictor and are never updated (nstr = in2a)
          const bool inpred = c != 0 || b >= A0 + D.in2a;
          if (inpred && on) upd |= 1u << q;
        }
      }
      double zcwc = rv * rhoair / kMwa;
      // total and fixed water over every bin (bins not updated keep theirs)
      double stot = 0., sfix = 0.;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int b = sub + q * L;
        if (b < nt) {
          stot += cs[b];
          if (!((upd >> q) & 1u)) sfix += cs[b];
        }
      }
      // The substeps touch only the updated bins (a predictor bin that is not
      // updated has zmt = 0 and adds nothing): deal them round-robin onto the
      // cell's lanes in bin order, so that each lane iterates ceil(k/L) bins
      // instead of every bin it owns.
      double zcw[NQ], zmt[NQ], zkl[NQ], zns[NQ], zws[NQ];
      int cb_[NQ];
      uint32_t iceq_c = 0;
      int nq = 0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) cb_[q] = -1;
      if (live) {
        int rank = 0;
        for (int b = 0; b < nt; ++b) {
          const int c = D.cat_of[b];
          const bool on = sN[b] > D.nlim[b] &&
                          (c == 0 ? (zrh > 0.98 && D.lscndh2oae)
                                  : c == 3 ? (D.lscndh2oic && T < 273.15) : D.lscndh2ocl);
          if (!(on && (c != 0 || b >= A0 + D.in2a))) continue;
          if (rank % L == sub) {
#pragma unroll
            for (int q = 0; q < NQ; ++q)
              if (q == nq) cb_[q] = b;
            ++nq;
          }
          ++rank;
        }
      }
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int b = cb_[q];
        zcw[q] = 0.; zmt[q] = 0.; zkl[q] = 1.; zns[q] = 0.; zws[q] = 0.;
        if (b >= 0) {
          zcw[q] = cs[b]; zmt[q] = cs[nt + b]; zkl[q] = cs[2 * nt + b]; zns[q] = cs[3 * nt + b];
          zws[q] = cs[4 * nt + b];
          if (D.cat_of[b] == 3) iceq_c |= 1u << q;
        }
      }
      // slots in use anywhere in the wavefront (a uniform bound for the loops)
      uint32_t qw = 0;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (__any(q < nq)) qw |= 1u << q;
      double s1 = 0., s2 = 0.;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if ((qw >> q) & 1u) {
          s1 += zmt[q] * zws[q] * (((iceq_c >> q) & 1u) ? csui : csul);
          s2 += zmt[q];
        }
      stot = cell_sum<L>(stot);
      s1 = cell_sum<L>(s1);
      s2 = cell_sum<L>(s2);
      sfix = cell_sum<L>(sfix);
      const double zcwtot = zcwc + stot;

      // substeps of adt = 0.02 s (:986-1053); the count comes from FP64 time accumulation
      const double adt = 2.e-2;
      double zcwint = 0.;
      const int nsub = a.nsub;
      for (int st = 0; st < nsub; ++st) {
        const double zhlp1 = zcwc + adt * s1;
        const double zhlp2 = 1. + adt * s2;
        zcwint = fmin(div_nr(zhlp1, zhlp2), zcwtot);
        double tsum = 0., s1n = 0.;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          if ((qw >> q) & 1u) {
            if (q < nq) {
              const bool ic = (iceq_c >> q) & 1u;
              const double csu = ic ? csui : csul;
              const double d = adt * zmt[q] * (zcwint - zws[q] * csu);
              double nw = zcw[q] + fmin(fmax(d, -1.e-1 * zcw[q]), 1.e-1 * zcw[q]);
              nw = fmax(nw, 0.);
              zcw[q] = nw;
              zws[q] = ic ? zkl[q] : act_h2o<true>(nw, zns[q]) * zkl[q];
              tsum += nw;
              s1n += zmt[q] * zws[q] * csu;
            }
          }
        }
        tsum = cell_sum<L>(tsum);
        s1 = cell_sum<L>(s1n);
        zcwint = zcwtot - (sfix + tsum);
        zcwc = zcwint;
      }
      if (nsub > 100) flags |= SALSA_CELL_SUBSTEPS;
      rv = zcwint * kMwa / rhoair;
      // updated water back to the per-bin table, then every bin is written back
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        if (q < nq) cs[cb_[q]] = zcw[q];
      __syncthreads();
      // Cond_i in the reference's order (:1086-1110): vrate is zeroed once
      // before the ice bins, so a depositing bin also adds the rime entry that
      // the last sublimating bin before it left there (before the write-back
      // below changes the volumes)
      if (R && D.level == 5 && sub == 0) {
        double vw = 0., vri = 0., sw = 0., sr = 0.;
        for (int cc = 0; cc < D.nb[3]; ++cc) {
          const int b = D.off[3] + cc;
          const double org_i = VV(b, iwa) * kRhoic / kMwa, org_r = VV(b, irim) * kRhori / kMwa;
          const double orig = org_i + org_r;
          const double dvt = cs[b] - orig;
          if (dvt > 0.) {
            vw = dvt * kMwa / kRhoic;
          } else if (dvt < 0.) {
            const double fr = (cs[b] / orig) - 1.;
            vw = (fr * org_i) * kMwa / kRhoic;
            vri = (fr * org_r) * kMwa / kRhori;
          } else {
            continue;
          }
          sw += vw;
          sr += vri;
        }
        rr[RA_CONDI] += sw * kRhoic + sr * kRhori;
      }
      __syncthreads();
      // write-back (:1058-1117)
      double dcat[3] = {0., 0., 0.};         // Cond_a/c/p%volc(iwa)
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int b = sub + q * L;
        if (b < nt) {
          const double zw = cs[b];
          if (!((iceq >> q) & 1u)) {
            const double nv = fmax(0., zw * kMwa / kRhowa);
            if (R) {
              const int c = D.cat_of[b];
              const double dv = nv - VV(b, iwa);
              dcat[0] += c == 0 ? dv : 0.;
              dcat[1] += c == 1 ? dv : 0.;
              dcat[2] += c == 2 ? dv : 0.;
            }
            VV(b, iwa) = nv;
          } else {
            const double org_i = VV(b, iwa) * kRhoic / kMwa, org_r = VV(b, irim) * kRhori / kMwa;
            const double orig = org_i + org_r;
            const double dvt = zw - orig;
            if (dvt > 0.) {
              VV(b, iwa) = VV(b, iwa) + dvt * kMwa / kRhoic;
            } else if (dvt < 0.) {
              const double fr = (zw / orig) - 1.;
              VV(b, iwa) = VV(b, iwa) + (fr * org_i) * kMwa / kRhoic;
              VV(b, irim) = VV(b, irim) + (fr * org_r) * kMwa / kRhori;
            }
          }
        }
      }
      if (R) {
#pragma unroll
        for (int k = 0; k < 3; ++k) dcat[k] = cell_sum<L>(dcat[k]);
        if (sub == 0) {
          rr[RA_CONDA] += dcat[0]; rr[RA_CONDC] += dcat[1]; rr[RA_CONDP] += dcat[2];
        }
      }
    }
  }
  __syncthreads();

  // ---- 5a. autoconversion by lognormal fitting (autoconv2, lsauto%mode 2,
  //      mo_salsa.f90:110-112): serial over the cloud bins on one lane ----
  if (D.autoconv2) {
    if (sub == 0 && live) autoconv2_cell<double>(D, sN, sV, nsV, dt);
    __syncthreads();
  }

  // ---- 5a2. cloud activation (cloud_activation, mo_salsa.f90:115-119):
  //      interstitial (lsactiv%mode 1), serial over the 2a/2b bins on one
  //      lane; cloud base (mode 2), which leaves the state as it is and writes
  //      the activated droplets to pa_nactd / pa_vactd, on one lane ----
  if (D.lsactiv == 2 && a.act_state && live) {
    // the aerosol state the activation sees, to k_act_cloud_base
    double *st = a.act_state + (int64_t)icell * D.nb[0] * (1 + D.nwet);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int b = sub + q * L;
      if (b < D.nb[0]) {
        st[b] = sN[b];
        for (int s = 0; s < D.nwet; ++s) st[D.nb[0] + b * D.nwet + s] = VV(b, s);
      }
    }
  }
  if (D.lsactiv == 1) {
    if (sub == 0 && live) {
      const double na = act_interst_cell<double>(D, sN, sV, nsV, rv, rs, T, reinterpret_cast<double *>(cb + G.opr));
      if (R) rr[RA_ACTIV] += na;
    }
    __syncthreads();
  }

  // ---- 5b. size-distribution update (distr_update, mo_salsa.f90:148-149):
  //      serial over the bins, on one lane of the cell ----
  if (D.lsdistupdate) {
    if (sub == 0 && live) flags |= distr_update_cell<double>(D, sN, sV, nsV, reinterpret_cast<double *>(cb + G.opr));
    __syncthreads();
  }

  // ---- 6. unpack (mo_salsa_driver.f90:310-388): per slot of the cell's
  //      output record the new state in LES units (in place) or its change
  //      new - old (tendency mode; the scatter divides by dt and adds) ----
  if (live) {
    double *o = a.rec_out + (int64_t)icell * D.nout;
    const bool inp = a.inplace;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int b = sub + q * L;
      if (b < nt) {
        const int c = D.cat_of[b];
        double *ob = o + (D.sbase[b] - kThermo + 1);
        const double nn = sN[b] / dn;
        if (!isfinite(nn)) flags |= SALSA_CELL_NONFINITE;
        ob[0] = inp ? nn : nn - rec[D.sbase[b]];
        const int nsw = c < 3 ? D.nwet : D.nspec;
#pragma unroll
        for (int s = 0; s < NS; ++s)
          if (s < nsw) {
            const double rho = c < 3 ? D.rholiq[s] : (s == D.irim ? kRhori : D.rhoice[s]);
            const double mm = VV(b, s) * rho / dn;
            if (!isfinite(mm)) flags |= SALSA_CELL_NONFINITE;
            ob[1 + s] = inp ? mm : mm - rec[D.sbase[b] + 1 + s];
          }
      }
    }
    if (sub == 0) {
      if (!isfinite(rv)) flags |= SALSA_CELL_NONFINITE;
      o[0] = inp ? rv : rv - rv0;
    }
  }
  // rate diagnostics into the LES Diag fields (mo_salsa_driver.f90:392-411)
  if (R && live) {
    const int nrf = D.level == 5 ? SALSA_NRATE_L5 : SALSA_NRATE_L4;
    for (int f = sub; f < nrf; f += L) {
      double v = 0.;
      switch (f) {
        case SALSA_RATE_AUTOC: v = rr[RA_AUTOC] * kRhowa / dn; break;
        case SALSA_RATE_AUTOC50: v = rr[RA_AUTOC50] * kRhowa / dn; break;
        case SALSA_RATE_AUTOC80: v = rr[RA_AUTOC80] * kRhowa / dn; break;
        case SALSA_RATE_ACCR: v = rr[RA_ACCR] * kRhowa / dn; break;
        case SALSA_RATE_ACCR50: v = rr[RA_ACCR50] * kRhowa / dn; break;
        case SALSA_RATE_ACCR80: v = rr[RA_ACCR80] * kRhowa / dn; break;
        case SALSA_RATE_ACCOLL: v = rr[RA_ACCOLL] / dn; break;
        case SALSA_RATE_APCOLL: v = rr[RA_APCOLL] / dn; break;
        case SALSA_RATE_CONDA: v = rr[RA_CONDA] / dn; break;
        case SALSA_RATE_CONDC: v = rr[RA_CONDC] / dn; break;
        case SALSA_RATE_CONDP: v = rr[RA_CONDP] / dn; break;
        case SALSA_RATE_CONDI: v = rr[RA_CONDI] / dn; break;
        case SALSA_RATE_ACTIV: v = rr[RA_ACTIV] / dn; break;
        default: v = 0.;      // ice nucleation: not part of this step
      }
      a.rates[(int64_t)f * a.n3 + les_index(a, icell)] = v;
    }
  }
  flags = cell_or<L>(flags);
  if (live && sub == 0) {
    if (a.cell_flags) a.cell_flags[icell] = (uint8_t)flags;
    if (flags && a.fault_any) atomicOr(a.fault_any, (int)flags);
  }
#undef VV
}

// ---------------------------------------------------------------------------
// gather / scatter between the LES field rows and the per-cell records.
// A 256-thread block handles a tile of kTileCells consecutive interior cells:
// every LES row is read (written) as one contiguous run of the tile's cells,
// and the records pass through LDS so that both sides are coalesced.  Each
// row visit thus covers 32 cells, which keeps the address-translation misses
// per cell low (every bin of every field is its own row, 8*nzp*nxp*nyp bytes
// from the next).  The row base addresses are resolved once per block into
// LDS, and every thread issues its row loads in batches of eight independent
// loads before storing them, so that the block keeps many loads in flight.
// ---------------------------------------------------------------------------
constexpr int kTileCells = 32;

// Tile of this block: blocks are dealt to the 8 XCDs round-robin, so give
// each XCD a contiguous range of tiles; neighbouring tiles share the cache
// lines at their ends, which then meet in one L2.
__device__ __forceinline__ int64_t xcd_tile() {
  const uint32_t nblk = gridDim.x, bid = blockIdx.x;
  const uint32_t xcd = bid & 7u, rnk = bid >> 3, per = nblk >> 3, rem = nblk & 7u;
  return (int64_t)(xcd * per + (xcd < rem ? xcd : rem) + rnk);
}
constexpr int kRowBatch = 8;
constexpr int kRowsPerPass = 256 / kTileCells;

// LES rows -> cell records, plus the cost class of every cell: 2 = some cloud /
// precipitation / ice bin at or above its activity limit (numc >= nlim: it
// takes part in the collision kernels, mo_salsa_coagulation_kernels.f90:116-168,
// and in the condensation predicate, mo_salsa_dynamics.f90:784-788); 1 = no
// such bin and RH > 0.98 (aerosol substeps); 0 = no such bin and drier
// (equilibrium water only).  Classes 0 and 1 run the step-kernel variant
// without hydrometeor pairs, which is exact for them.  The class of every cell
// goes to cls, the tile's count per class to tcount[tile][3]; k_class_scan
// and k_class_fill then build one list of cell ids per class, in ascending
// cell order.
__global__ void __launch_bounds__(256) k_gather(const DevLayout *__restrict__ Dp, StepArgs a, double *rec,
                                                uint8_t *cls, int32_t *tcount) {
  extern __shared__ __align__(16) char smem[];
  const DevLayout &D = *Dp;
  const int ns = D.nslot, ld = ns + 1;
  double *tile = reinterpret_cast<double *>(smem);                               // [kTileCells][ld]
  const double **src = reinterpret_cast<const double **>(tile + kTileCells * ld); // [ns] row bases
  int64_t *gi = reinterpret_cast<int64_t *>(src + ns);                            // [kTileCells]
  const int64_t tile_id = xcd_tile();
  const int64_t c0 = tile_id * kTileCells;
  const int nc = (int)(a.ncells - c0 < kTileCells ? a.ncells - c0 : kTileCells);
  const int tid = threadIdx.x;
  if (tid < kTileCells) gi[tid] = tid < nc ? les_index(a, c0 + tid) : 0;
  for (int r = tid; r < ns; r += 256) {
    const int cat = D.slot_cat[r];
    const double *p;
    if (cat >= 0)
      p = (D.slot_mass[r] ? a.m_p[cat] : a.n_p[cat]) + a.n3 * D.slot_row[r];
    else
      p = r == 0 ? a.temp : r == 1 ? a.press : r == 2 ? a.rsl
        : r == 3 ? (D.level == 5 ? a.rsi : nullptr) : r == 4 ? a.dn : a.rv_p;
    src[r] = p;
  }
  __syncthreads();
  {
    const int c = tid % kTileCells, r0 = tid / kTileCells;
    const int64_t g = gi[c];
    for (int j0 = r0; j0 < ns; j0 += kRowsPerPass * kRowBatch) {
      double v[kRowBatch];
#pragma unroll
      for (int u = 0; u < kRowBatch; ++u) {
        const int r = j0 + kRowsPerPass * u;
        const double *p = r < ns ? src[r] : nullptr;
        v[u] = (c < nc && p) ? p[g] : 0.;
      }
#pragma unroll
      for (int u = 0; u < kRowBatch; ++u) {
        const int r = j0 + kRowsPerPass * u;
        if (r < ns) tile[c * ld + r] = v[u];
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < nc * ns; e += 256) {
    const int c = e / ns, r = e % ns;
    rec[(c0 + c) * ns + r] = tile[c * ld + r];
  }
  if (cls && tid < kTileCells) {            // the first wavefront classifies the tile
    int cl = -1;
    if (tid < nc) {
      const double *t = tile + tid * ld;
      const double dn = t[4];
      bool hyd = false;
      if (a.kmode == 2)          // stored kernels may pair any bin with non-zero numc
        for (int b = D.off[1]; b < D.nt; ++b) hyd |= t[D.sbase[b]] * dn != 0.;
      else
        for (int b = D.off[1]; b < D.nt; ++b) hyd |= !(t[D.sbase[b]] * dn < D.nlim[b]);
      double rv = t[5];
      const double rs = t[2];
      if (!a.lsfreeRH) rv = fmin(rv, rs * a.rhlim);
      cl = hyd ? 2 : (rv / rs > 0.98 ? 1 : 0);
      cls[c0 + tid] = (uint8_t)cl;
    }
    const uint64_t m0 = __ballot(cl == 0), m1 = __ballot(cl == 1), m2 = __ballot(cl == 2);
    if (tid == 0) {
      int32_t *tc = tcount + 3 * tile_id;
      tc[0] = __popcll(m0);
      tc[1] = __popcll(m1);
      tc[2] = __popcll(m2);
    }
  }
}

// Exclusive prefix over the tiles of each class's count, in two levels:
// k_class_scan scans 256 tiles per block (wavefront shuffle scans, one tile
// per thread, coalesced loads) and leaves each block's totals in bsum;
// k_class_fill adds the totals of the blocks before its tile (one wavefront
// sums them) and block 0 also forms the class totals.
constexpr int kScanTiles = 256;

__device__ __forceinline__ int wave_incl_scan(int x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  return x;
}

__device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

__global__ void __launch_bounds__(kScanTiles) k_class_scan(const int32_t *__restrict__ tcount, int32_t ntile,
                                                           int32_t *toff, int32_t *bsum) {
  __shared__ int32_t ws[3][kScanTiles / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t i = (int64_t)blockIdx.x * kScanTiles + t;
  int x[3], inc[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    x[k] = i < ntile ? tcount[3 * i + k] : 0;
    inc[k] = wave_incl_scan(x[k]);
    if (lane == 63) ws[k][w] = inc[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int off = 0, tot = 0;
    for (int v = 0; v < kScanTiles / 64; ++v) {
      off += v < w ? ws[k][v] : 0;
      tot += ws[k][v];
    }
    if (i < ntile) toff[3 * i + k] = off + inc[k] - x[k];
    if (t == 0) bsum[3 * blockIdx.x + k] = tot;
  }
}

// Cell ids into the class lists (lists + class * ncells), tile by tile.
__global__ void __launch_bounds__(64) k_class_fill(const uint8_t *__restrict__ cls, const int32_t *__restrict__ toff,
                                                   const int32_t *__restrict__ bsum, int32_t nsb, int64_t ncells,
                                                   int32_t *lists, int32_t *counts) {
  const int tid = threadIdx.x;
  const int64_t tile = blockIdx.x;
  const int B = (int)(tile / kScanTiles);
  const int nread = tile == 0 ? nsb : B;
  int p[3] = {0, 0, 0}, q[3] = {0, 0, 0};
  for (int b = tid; b < nread; b += 64)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int v = bsum[3 * b + k];
      p[k] += b < B ? v : 0;
      q[k] += v;
    }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p[k] = wave_sum(p[k]);
    q[k] = wave_sum(q[k]);
  }
  if (tile == 0 && tid == 0) { counts[0] = q[0]; counts[1] = q[1]; counts[2] = q[2]; }
  const int64_t c0 = tile * kTileCells;
  const int nc = (int)(ncells - c0 < kTileCells ? ncells - c0 : kTileCells);
  const int cl = tid < nc ? (int)cls[c0 + tid] : -1;
  const uint64_t m0 = __ballot(cl == 0), m1 = __ballot(cl == 1), m2 = __ballot(cl == 2);
  if (cl >= 0) {
    const uint64_t below = (1ull << tid) - 1ull;
    const uint64_t m = cl == 0 ? m0 : cl == 1 ? m1 : m2;
    const int base = cl == 0 ? p[0] : cl == 1 ? p[1] : p[2];
    lists[(int64_t)cl * ncells + base + toff[3 * tile + cl] + __popcll(m & below)] = (int32_t)(c0 + tid);
  }
}

// output records -> tendencies (t += change / dt) or, in place, the new state
__global__ void __launch_bounds__(256) k_scatter(const DevLayout *__restrict__ Dp, StepArgs a) {
  extern __shared__ __align__(16) char smem[];
  const DevLayout &D = *Dp;
  const int no = D.nout, ld = no + 1;
  double *tile = reinterpret_cast<double *>(smem);
  double **dst = reinterpret_cast<double **>(tile + kTileCells * ld);   // [no] row bases
  int64_t *gi = reinterpret_cast<int64_t *>(dst + no);
  const int64_t c0 = xcd_tile() * kTileCells;
  const int nc = (int)(a.ncells - c0 < kTileCells ? a.ncells - c0 : kTileCells);
  const int tid = threadIdx.x;
  const bool inp = a.inplace;
  if (tid < kTileCells) gi[tid] = tid < nc ? les_index(a, c0 + tid) : 0;
  for (int r = tid; r < no; r += 256) {
    double *p;
    if (r == 0) {
      p = inp ? a.rv_o : a.rv_t;                      // mo_salsa_driver.f90:387-388
    } else {                                         // :321-323, :354-355
      const int sr = r - 1 + kThermo;
      const int cat = D.slot_cat[sr];
      double *base = D.slot_mass[sr] ? (inp ? a.m_o[cat] : a.m_t[cat]) : (inp ? a.n_o[cat] : a.n_t[cat]);
      p = base + a.n3 * D.slot_row[sr];
    }
    dst[r] = p;
  }
  for (int e = tid; e < nc * no; e += 256) {
    const int c = e / no, r = e % no;
    tile[c * ld + r] = a.rec_out[(c0 + c) * no + r];
  }
  __syncthreads();
  const int c = tid % kTileCells, r0 = tid / kTileCells;
  const int64_t g = gi[c];
  for (int j0 = r0; j0 < no; j0 += kRowsPerPass * kRowBatch) {
    double v[kRowBatch];
#pragma unroll
    for (int u = 0; u < kRowBatch; ++u) {
      const int r = j0 + kRowsPerPass * u;
      v[u] = (!inp && r < no && c < nc) ? dst[r][g] : 0.;
    }
#pragma unroll
    for (int u = 0; u < kRowBatch; ++u) {
      const int r = j0 + kRowsPerPass * u;
      if (r < no && c < nc) {
        const double x = tile[c * ld + r];
        dst[r][g] = inp ? x : v[u] + x / a.dt;
      }
    }
  }
}

size_t record_tile_lds(const DevLayout &Lay) {
  const int n = Lay.nslot > Lay.nout ? Lay.nslot : Lay.nout;
  return (size_t)kTileCells * (n + 1) * sizeof(double) + (size_t)n * sizeof(void *) + kTileCells * sizeof(int64_t);
}

int64_t gather_tiles(int64_t ncells) { return (ncells + kTileCells - 1) / kTileCells; }

// ints of the toff work array: the tiles' offsets, then the scan blocks' totals
int64_t class_work_ints(int64_t ncells) {
  const int64_t nt = gather_tiles(ncells);
  return 3 * (nt + (nt + kScanTiles - 1) / kScanTiles);
}

// Records of every cell; with cls != null also the class lists (tcount and
// toff: [tiles][3] work arrays) and the class counts.
int launch_gather(const DevLayout &Lay, const DevLayout *d, const StepArgs &a, double *rec, uint8_t *cls,
                  int32_t *tcount, int32_t *toff, int32_t *lists, int32_t *counts, hipStream_t s) {
  const unsigned nb = (unsigned)gather_tiles(a.ncells);
  hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), record_tile_lds(Lay), s, d, a, rec, cls, tcount);
  if (cls) {
    const unsigned nsb = (nb + kScanTiles - 1) / kScanTiles;
    int32_t *bsum = toff + 3 * (int64_t)nb;
    hipLaunchKernelGGL(k_class_scan, dim3(nsb), dim3(kScanTiles), 0, s, tcount, (int32_t)nb, toff, bsum);
    hipLaunchKernelGGL(k_class_fill, dim3(nb), dim3(64), 0, s, cls, toff, bsum, (int32_t)nsb, a.ncells, lists,
                       counts);
  }
  return (int)hipGetLastError();
}

int launch_scatter(const DevLayout &Lay, const DevLayout *d, const StepArgs &a, hipStream_t s) {
  const unsigned nb = (unsigned)((a.ncells + kTileCells - 1) / kTileCells);
  hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), record_tile_lds(Lay), s, d, a);
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// LDS geometry and launch wrappers
// ---------------------------------------------------------------------------
static inline int32_t align_up(int32_t x, int32_t a) { return (x + a - 1) / a * a; }

int make_geom(const DevLayout &Lay, bool hyd, int kt_bytes, int lanes, bool rates, SmemGeom &G) {
  if (rates) lanes = kRateLanes;            // the rate-diagnostic variant has one shape
  const int cpw = 64 / lanes;
  G.ntab = hyd ? Lay.nt : Lay.nb[0];
  G.npairs = hyd ? Lay.npair : Lay.npair_aa;
  G.nsV = Lay.level == 5 ? Lay.nspec : Lay.nwet;
  G.it0 = hyd ? 0 : Lay.it_aero0;
  G.it1 = hyd ? Lay.ntarget : Lay.it_aero1;
  G.jmask = hyd ? ~0ull : (Lay.nb[0] >= 64 ? ~0ull : ((1ull << Lay.nb[0]) - 1ull));
  G.tab_bytes = align_up(4 * Lay.nent_total, 16);
  int32_t o = 0;
  G.oN = o; o += 8 * Lay.nt;
  G.oV = o; o += 8 * Lay.nt * G.nsV;
  G.odw = o; o += 8 * Lay.nt;
  o = align_up(o, 16);
  G.opr = o; o += 2 * G.ntab * 4 * kt_bytes;
  G.ovd = -1;
  if (kt_bytes == 4) { G.ovd = o; o += align_up(2 * G.ntab * 4, 16); }   // fall-speed residuals (mixed path)
  G.oK = o; o += G.npairs * kt_bytes;
  // the condensation set-up reuses the particle-term and kernel regions
  if (o - G.opr < 5 * 8 * Lay.nt) o = G.opr + 5 * 8 * Lay.nt;
  G.opf = o; o += hyd ? Lay.ncc : 0;
  o = align_up(o, 16);
  G.orr = -1;
  if (rates) { G.orr = o; o += 8 * RA_N; }
  // keep neighbouring cells of a block off the same LDS banks
  if (o % 256 == 0) o += 16;
  G.cell_bytes = o;
  G.total_bytes = G.tab_bytes + cpw * G.cell_bytes;
  return G.total_bytes <= 65536 ? 0 : -1;
}

template <int L, class KT, int MB, int NS, bool HYD, bool RATES>
static hipError_t launch_t(const DevLayout *d, const StepArgs &a, const SmemGeom &G, hipStream_t s) {
  constexpr int CPW = 64 / L;
  if (a.ncells <= 0) return hipSuccess;
  const int64_t nb = (a.ncells + CPW - 1) / CPW;
  hipLaunchKernelGGL((k_salsa_step<L, KT, MB, NS, HYD, RATES>), dim3((unsigned)nb), dim3(64), (size_t)G.total_bytes,
                     s, d, a, G);
  return hipGetLastError();
}

template <class KT, int MB, int NS, bool HYD>
static int launch_l(int lanes, const DevLayout *d, const StepArgs &a, const SmemGeom &G, hipStream_t s) {
  if (a.rates) return (int)launch_t<kRateLanes, KT, MB, NS, HYD, true>(d, a, G, s);
  switch (lanes) {
    case 8: return (int)launch_t<8, KT, MB, NS, HYD, false>(d, a, G, s);
    case 16: return (int)launch_t<16, KT, MB, NS, HYD, false>(d, a, G, s);
    case 32: return (int)launch_t<32, KT, MB, NS, HYD, false>(d, a, G, s);
    case 64: return (int)launch_t<64, KT, MB, NS, HYD, false>(d, a, G, s);
  }
  return -1;
}

// capacity classes: warm (level 4) and mixed-phase (level 5)
int launch_step(int cap, int lanes, int precision, bool hyd, const DevLayout *d, const StepArgs &a,
                const SmemGeom &G, hipStream_t s) {
#define SALSA_CASE(CAP, MB, NS)                                                                   \
  if (cap == CAP) {                                                                              \
    if (precision == SALSA_FP64)                                                                 \
      return hyd ? launch_l<double, MB, NS, true>(lanes, d, a, G, s)                             \
                 : launch_l<double, MB, NS, false>(lanes, d, a, G, s);                           \
    return hyd ? launch_l<float, MB, NS, true>(lanes, d, a, G, s)                                \
               : launch_l<float, MB, NS, false>(lanes, d, a, G, s);                              \
  }
  SALSA_CASE(4, 32, 3)
  SALSA_CASE(5, 40, 4)
#undef SALSA_CASE
  return -1;
}

// Cloud-base activation (lsactiv 2) of every interior cell from the aerosol
// state the step kernel left in act_state (the state after condensation and
// autoconv2, which cloud_activation sees, mo_salsa.f90:110-119): one thread
// per cell, writing pa_nactd / pa_vactd.
__global__ void __launch_bounds__(256) k_act_cloud_base(const DevLayout *__restrict__ dl, StepArgs a) {
  const DevLayout &D = *dl;
  const int64_t icell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (icell >= a.ncells) return;
  const double *st = a.act_state + icell * D.nb[0] * (1 + D.nwet);
  const double *rec = a.rec_in + icell * D.nslot;
  const int64_t li = les_index(a, icell);
  act_cloud_base_cell(D, st, st + D.nb[0], D.nwet, rec[0], rec[1], a.w[li], rec[4], a.nactd + li, a.vactd + li,
                      a.n3);
}

int launch_act_cloud_base(const DevLayout *d, const StepArgs &a, hipStream_t s) {
  if (a.ncells <= 0) return 0;
  hipLaunchKernelGGL(k_act_cloud_base, dim3((unsigned)((a.ncells + 255) / 256)), dim3(256), 0, s, d, a);
  return (int)hipGetLastError();
}

int cap_bins(int cap) { return cap == 4 ? 32 : 40; }
int cap_spec(int cap) { return cap == 4 ? 3 : 4; }

}  // namespace salsa
