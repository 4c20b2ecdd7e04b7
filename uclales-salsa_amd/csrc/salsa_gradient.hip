// salsa_gradient.hip -- derivatives of a multi-step SALSA coagulation +
// condensation trajectory by forward-mode (tangent-linear) differentiation.
//
// Objective: J = sum over interior cells of the precipitation-bin volume
// concentration (all species, m^3 m^-3) after nsteps in-place steps of the
// box model (fixed T, p, rs, rsi, dn; N, masses and rv evolve).  Cells are
// independent inside a step, so dJ/dN0 of a cell only involves that cell's
// trajectory: with one tangent direction per bin of the cell plus one for
// the collision-kernel enhancement factor E (K = E * coagc, the charge factor
// of mo_salsa_coagulation_kernels.f90:411-416 with chargeTime =
// chargeTimeMax), nt + 1 directions give the complete gradient -- the same
// numbers a reverse sweep would produce, without storing or recomputing the
// trajectory.
//
// Layout: one lane per direction (L = 32 or 64 lanes per cell); the lanes of
// a cell evaluate the same values and take the same branches, so the
// wavefront stays converged; each lane carries the tangent of its direction.
// The per-lane state (numbers, volumes, particle terms, pair kernels, the
// condensation set-up) lives in the lane's private memory.  The physics is
// the step kernel's, in dual arithmetic: pack (mo_salsa_driver.f90:236-279),
// bin properties (classSection.f90:91-224), coagc
// (mo_salsa_coagulation_kernels.f90:185-420), the sequential sweep
// (mo_salsa_coagulation_processes.f90:24-1611), equilibration
// (mo_salsa_properties.f90:73-380) and the predictor-corrector water
// condensation (mo_salsa_dynamics.f90:691-1157); between steps the state
// passes through LES units as in run_SALSA.
//
// KT is the arithmetic type of the collision-kernel coefficients, as in the
// step kernel (double: FP64 path; float: FP32-mixed path).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "salsa_activation.h"
#include "salsa_autoconv.h"
#include "salsa_distr.h"
#include "salsa_dual.h"
#include "salsa_dual_physics.h"
#include "salsa_internal.h"

namespace salsa {

namespace {

using DD = Dual<double>;
using DF = Dual<float>;

__device__ __forceinline__ double cube(double x) { return x * x * x; }

__device__ __forceinline__ int64_t g_les_index(const StepArgs &a, int64_t cell) {
  const int k = (int)(cell % a.nz);
  const int64_t r = cell / a.nz;
  const int i = (int)(r % a.nx), j = (int)(r / a.nx);
  return (int64_t)(k + a.hz) + (int64_t)a.nzp * ((int64_t)(i + a.hxy) + (int64_t)a.nxp * (j + a.hxy));
}

}  // namespace

// ---------------------------------------------------------------------------
// the derivative kernel
// ---------------------------------------------------------------------------
// One cell, one direction (dir < nt: dN0 of bin dir; dir == nt: dE; larger:
// no seed); returns the cell's fault bits.
template <class KT, int MB, int NS, int NP>
__device__ uint32_t gradient_cell(const DevLayout &D, const GradArgs &g, int64_t cell, int dir) {
  using DK = Dual<KT>;
  const StepArgs &a = g.s;
  const int nt = D.nt;
  const int nsV = D.level == 5 ? D.nspec : D.nwet;
  const double dt = a.dt;
  const int nspec = D.nspec, ndry = D.ndry, iwa = D.iwa, irim = D.irim;
  {
    const double *rec = a.rec_in + cell * D.nslot;
    const double T = rec[0], P = rec[1], rs = rec[2], rsi = rec[3], dn = rec[4];
    DD rv = dconst(rec[5]);
    uint32_t flags = 0;

    // state of the cell with the tangent of this lane's direction
    DD N[MB], V[MB][NS], DW[MB];
    DD tmpd[2 * kMaxSpec];              // work space of the activation and the size-distribution update
    for (int b = 0; b < nt; ++b) {
      const int c = D.cat_of[b];
      const int nsw = c < 3 ? D.nwet : D.nspec;
      const double *r = rec + D.sbase[b];
      N[b] = dvar(r[0] * dn, b == dir ? dn : 0.);       // direction b: dN_b = 1 in LES units
#pragma unroll
      for (int s = 0; s < NS; ++s) V[b][s] = dconst(s < nsw ? r[1 + s] * dn / species_rho(D, c, s) : 0.);
    }
    const DK E = dvar((KT)D.charge_factor, dir == nt ? KT(1) : KT(0));
    // this direction's input: a relative change x0 * eps moves a quantity X
    // by |dX/dx| x0 eps
    const double x0 = dir < nt ? rec[D.sbase[dir]] : (dir == nt ? D.charge_factor : 0.);

    for (int step = 0; step < g.nsteps; ++step) {
      if (!a.lsfreeRH) rv = dmin(rv, rs * a.rhlim);       // mo_salsa_driver.f90:193-197

      // ---- coagulation (mo_salsa_dynamics.f90:92-202) ----
      if (D.lscoag) {
        const GCellC cc = g_cell_consts(D, T, P);
        uint64_t abits = 0;                 // bins with numc >= nlim take part in kernels
        GPT<DK> PR0[MB], PR1[MB];           // bin and its self-offset twin
        for (int b = 0; b < nt; ++b) {
          const DD n = N[b];
          DD vs = dconst(0.);
          for (int s = 0; s < nsV; ++s) vs += V[b][s];
          DD dw = n > D.nlim[b] ? dcbrt(vs / n / kPi6) : dconst(D.dmid[b]);
          dw = dmin(dw, D.dlim[b]);
          DW[b] = dw;
          if (!(n < D.nlim[b])) abits |= 1ull << b;
          if (D.phase[b] == 4) {
            g_ice_terms<KT>(D, cc, n, D.nlim[b], V[b], dw, PR0[b], PR1[b]);
          } else {
            const DD mass = kRhowa * kPi6 * dcube(dw);
            if constexpr (sizeof(KT) == 4) {
              PR0[b] = g_pterms_liq_f(cc, dw, mass, kRhowa);
              PR1[b] = g_pterms_liq_f(cc, 1.1 * dw, 1.33 * mass, kRhowa);
            } else {
              const GShape sh{};
              PR0[b] = g_pterms(cc, dw, mass, dconst(kRhowa), D.phase[b], dw, sh);
              PR1[b] = g_pterms(cc, 1.1 * dw, 1.33 * mass, dconst(kRhowa), D.phase[b], dw, sh);
            }
          }
        }

        // pair kernels (buildKernelSelf / buildKernel, kernels.f90:94-180)
        DK K[NP];
        uint8_t PF[NP];
        const KT vkin = (KT)cc.vkin, cshear = (KT)cc.cshear, cinert = (KT)cc.cinert;
        const int np = D.npair, pf0 = D.pf0, pf1 = D.pf0 + D.ncc;
        for (int p = 0; p < np; ++p) {
          const uint32_t pw = D.pw[p];
          const int b1 = pw & 255u, b2 = (pw >> 8) & 255u, tw = (pw >> 16) & 1u;
          DK k = dconst(KT(0));
          uint8_t pfb = 0;
          if ((abits >> b1) & (abits >> b2) & 1ull) {
            const DD dB = DW[b2];
            const DD dA = tw ? 1.1 * dB : DW[b1];
            const GPT<DK> &A = tw ? PR1[b2] : PR0[b1];
            const GPT<DK> &B = PR0[b2];
            k = g_pair_coef<KT>(dcast<KT>(dA), A.D, A.c, A.w, A.v, dcast<KT>(dB), B.D, B.c, B.w, B.v, dA.v >= dB.v,
                                vkin, cshear, cinert);
            k = k * E;                                    // charge enhancement (:411-416)
            if (p >= pf0 && p < pf1) {
              // precip bin reached by coalescence (accumulatePrecipFormation, processes.f90:1461-1467)
              const double dnew = ::cbrt(cube(DW[b1].v) + cube(dB.v));
              int cnt = 0;
              for (int q = 0; q < D.nb[2]; ++q) cnt += dnew > D.precpbins[q];
              pfb = dnew > D.precpbins[0] ? (uint8_t)cnt : (uint8_t)0;
            }
          }
          K[p] = k;
          if (p >= pf0 && p < pf1) PF[p - pf0] = pfb;
        }

        // sweep: coag_ice, coag_precp, coag_aero, coag_cloud in sequence
        const int Roff = D.off[2];
        const double emax = D.Eiagg_max, emin = D.Eiagg_min;
        for (int it = 0; it < D.ntarget; ++it) {
          const DevTarget &TG = D.tgt[it];
          const int t = TG.t;
          const uint32_t fl = TG.flags;
          const DD Nt = N[t];
          DD Vt[NS];
#pragma unroll
          for (int s = 0; s < NS; ++s) Vt[s] = s < nsV ? V[t][s] : dconst(0.);
          bool act = !(Nt < D.nlim[t]);
          if (fl & TF_ICE) act = act && T < 273.15;      // mo_salsa_dynamics.f90:166
          if (!act) continue;
          const DD rft = (fl & TF_ICE) ? g_rime_frac(D, Nt, Vt, t) : dconst(0.);
          const int nu = (fl & TF_ICE) ? nspec : TG.nupd;
          const int nupd = TG.nupd;
          const uint32_t *ent = D.ent + TG.ent0;
          const int ne = TG.nent;

          DD sink = dconst(0.), so = dconst(0.), sx = dconst(0.), src[NS];
#pragma unroll
          for (int s = 0; s < NS; ++s) src[s] = dconst(0.);
          uint32_t pmask = 0;
          for (int e = 0; e < ne; ++e) {
            const uint32_t w = ent[e];
            const int j = w & 63u, pid = (int)(w >> 16);
            const uint32_t role = (w >> 6) & 1023u;
            const DD kj = dcast<double>(K[pid]);
            const DD Nj = N[j];
            const DD *Vj = V[j];
            if (role & R_SINK) sink += kj * Nj;
            if (role & (R_SINK_EAGG | R_SRC_EAGG)) {      // ice aggregation efficiency (:758-769)
              const DD eg = emax - dmax(g_rime_frac(D, Nj, Vj, j), rft) * (emax - emin);
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
            }
            if (role & (R_PC_WA | R_PC_RI)) {              // accumulateSourcePhaseChange (:1236-1266)
#pragma unroll
              for (int s = 0; s < NS; ++s)
                if (s < ndry) src[s] += kj * Vj[s];
              if (role & R_PC_WA) src[iwa] += kj * Vj[iwa] * kRhowa / kRhoic;
              else src[irim] += kj * Vj[iwa] * kRhowa / kRhori;
            }
            if (role & (R_PF_OWN | R_PF_OTH)) {            // accumulatePrecipFormation (:1458-1576)
              const int pb = (int)PF[pid - D.pf0] - 1;
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
          DD fix_own = dconst(1.), fix_oth = dconst(1.);
          if (fl & TF_PF) {                                // fix_coag (:1440-1456)
            fix_own = dmax(1. - dt * so, 0.1);
            fix_oth = dmax(1. - dt * sx, 0.1);
          }

          // self terms
          DD sink_self = dconst(0.), volsink[NS];
#pragma unroll
          for (int s = 0; s < NS; ++s) volsink[s] = dconst(0.);
          int pself = -1;
          DD kself = dconst(0.), fixs = dconst(1.);
          if (fl & (TF_SELF_HALF | TF_SELF_HALF_EAGG | TF_PF_SELF | TF_PRECIP_SELFCOLL)) {
            const int pid = D.pidx[t][t];
            kself = pid >= 0 ? dcast<double>(K[pid]) : dconst(0.);
            if (fl & TF_SELF_HALF) sink_self = 0.5 * kself * Nt;
            if (fl & TF_SELF_HALF_EAGG) sink_self = (emax - rft * (emax - emin)) * 0.5 * kself * Nt;
            if (fl & TF_PF_SELF) {
              fixs = dmax(1. - dt * (0.5 * kself * (Nt * Nt) / Nt), 0.1);
              pself = pid >= 0 ? (int)PF[pid - D.pf0] - 1 : -1;
              sink_self = 0.5 * kself * Nt;
              if (pself >= 0) {
#pragma unroll
                for (int s = 0; s < NS; ++s)
                  if (s < nupd) volsink[s] = Vt[s] * Nt * kself * fixs;
              }
            }
          }

          // applyCoag / applyCoagIce / applyCoagPrecp (:539-645)
          DD Vn[NS];
          const DD rden = 1. / (1. + dt * sink);
#pragma unroll
          for (int s = 0; s < NS; ++s) Vn[s] = s < nu ? (Vt[s] + dt * src[s] * Nt) * rden : Vt[s];
          DD Nn = Nt / (1. + dt * (sink + sink_self));
          if (fl & TF_PRECIP_SELFCOLL) {                   // precipSelfCoag (:1588-1611)
            const DD nslf = 0.5 * kself * Nt * Nt;
#pragma unroll
            for (int s = 0; s < NS; ++s)
              if (s < nupd) {
                const DD vslf = Vt[s] * Nt * kself;
                V[t + 1][s] = V[t + 1][s] + dt * vslf;
                Vn[s] = Vn[s] - dt * vslf;
              }
            N[t + 1] = N[t + 1] + dt * nslf;
            Nn = Nn - dt * nslf;
          }
          if (fl & TF_PF_SELF) {
#pragma unroll
            for (int s = 0; s < NS; ++s) Vn[s] = Vn[s] - dt * volsink[s];
          }
          N[t] = Nn;
#pragma unroll
          for (int s = 0; s < NS; ++s)
            if (s < nsV) V[t][s] = Vn[s];

          // transfer of coalesced cloud drops into precipitation bins (applyCoagPrecipFormation :649-703)
          if (fl & (TF_PF | TF_PF_SELF)) {
            uint32_t pm = pmask;
            if (pself >= 0) pm |= 1u << pself;
            while (pm) {
              const int pb = __builtin_ctz(pm);
              pm &= pm - 1;
              DD vp[NS], npp = dconst(0.);
#pragma unroll
              for (int s = 0; s < NS; ++s) vp[s] = dconst(0.);
              for (int e = 0; e < ne; ++e) {
                const uint32_t w = ent[e];
                const uint32_t role = (w >> 6) & 1023u;
                if (!(role & (R_PF_OWN | R_PF_OTH))) continue;
                const int j = w & 63u, pid = (int)(w >> 16);
                if ((int)PF[pid - D.pf0] - 1 == pb) {
                  const DD kj = dcast<double>(K[pid]);
                  const DD fx = (role & R_PF_OWN) ? fix_own : fix_oth;
                  const DD Nj = N[j];
#pragma unroll
                  for (int s = 0; s < NS; ++s)
                    if (s < nupd) vp[s] += (V[j][s] * Nt + Vt[s] * Nj) * kj * fx;
                  npp += kj * Nj * Nt * fx;
                }
              }
              if (pb == pself) {
#pragma unroll
                for (int s = 0; s < NS; ++s) vp[s] += volsink[s];
                npp += 0.5 * kself * Nt * Nt * fixs;
              }
              const int r = Roff + pb;
#pragma unroll
              for (int s = 0; s < NS; ++s)
                if (s < nupd) V[r][s] = V[r][s] + dt * vp[s];
              N[r] = N[r] + dt * npp;
            }
          }
        }

        // clamp (mo_salsa_dynamics.f90:189-198)
        for (int b = 0; b < nt; ++b) {
          if (N[b].v < -1.e-6) flags |= SALSA_CELL_NEG_NUMC;
          N[b] = dmax(0., N[b]);
          for (int s = 0; s < nsV; ++s) V[b][s] = dmax(0., V[b][s]);
        }
      } else {
        for (int b = 0; b < nt; ++b) {
          DD vs = dconst(0.);
          for (int s = 0; s < nsV; ++s) vs += V[b][s];
          const DD dw = N[b] > D.nlim[b] ? dcbrt(vs / N[b] / kPi6) : dconst(D.dmid[b]);
          DW[b] = dmin(dw, D.dlim[b]);
        }
      }

      // ---- condensation (mo_salsa_dynamics.f90:276-342, 691-1125) ----
      if (D.lscnd && (D.lscndh2oae || D.lscndh2ocl || D.lscndh2oic)) {
        const int A0 = D.off[0];
        const DD zrh = rv / rs;
        DD w1 = dconst(0.);
        for (int b = 0; b < D.nb[0]; ++b) w1 += V[b][iwa];
        w1 = w1 * kRhowa;

        int isu = -1, ioc = -1;
        for (int s = 0; s < ndry; ++s) {
          if (D.dry_id[s] == 0) isu = s;
          if (D.dry_id[s] == 1) ioc = s;
        }
        // equilibration regime 1 (1a), mo_salsa_properties.f90:165-269
        {
          DD zke = dconst(1.001);
          for (int kk = D.in1a; kk <= D.fn1a; ++kk) {
            const int b = A0 + kk;
            const DD n = N[b];
            if (n > kNlim) {
              const DD v1 = isu >= 0 ? V[b][isu] : dconst(0.), v2 = ioc >= 0 ? V[b][ioc] : dconst(0.);
              const DD zcore = v1 / n + v2 / n;
              DD zdwet = DW[b], zdold = dconst(1.), zlwc = dconst(0.);
              const DD zr = dmin(dmax(zrh, 0.05), 0.98);
              int count = 0;
              while (dabs(zdwet / zdold - 1.).v > 1.e-2) {
                zdold = dmax(zdwet, 1.e-20);
                const DD aw = zr / zke;
                const DD bm1 = 1.1065495e+2 - 3.6759197e+2 * aw + 5.0462934e+2 * (aw * aw)
                               - 3.1543839e+2 * (aw * aw * aw) + 6.770824e+1 * ((aw * aw) * (aw * aw));
                const DD bm2 = 1. / (aw * kMwa) - 1. / kMwa;
                zlwc = v1 * (kRhosu / kMsu) / bm1 + kEpsoc * v2 * (kRhooc / kMoc) / bm2;
                zdwet = dcbrt(zlwc / n / kRhowa / kPi6 + zcore / kPi6);
                zke = dexp(2. * kSurfw0 * kMvwa / (kBoltz * T * zdwet));
                if (++count > 100) { flags |= SALSA_CELL_NOCONV; break; }
              }
              V[b][iwa] = zlwc / kRhowa;
            }
          }
        }
        // regime 2 (2a/2b, init call), mo_salsa_properties.f90:275-378
        if (zrh.v < 0.98 || !D.lscndh2oae) {
          for (int b = A0 + D.in2a; b <= A0 + D.fn2b; ++b) {
            const DD n = N[b];
            if (n > kNlim) {
              DD zv[5] = {dconst(0.), dconst(0.), dconst(0.), dconst(0.), dconst(0.)};
              for (int s = 0; s < ndry; ++s)
                if (D.dry_id[s] < 5) zv[D.dry_id[s]] = V[b][s];
              const DD zcore = zv[0] / n + zv[1] / n + zv[2] / n + zv[3] / n + zv[4] / n;
              DD zdwet = DW[b], zdold = dconst(1.), zlwc = dconst(0.), zke = dconst(1.02);
              const DD zr = dmin(dmax(zrh, 0.37), 0.98);
              int count = 0;
              while (dabs(zdwet / zdold - 1.).v > 1.e-12) {
                zdold = dmax(zdwet, 1.e-20);
                const DD aw = zr / zke;
                const DD bm1 = 1.1065495e+2 - 3.6759197e+2 * aw + 5.0462934e+2 * (aw * aw)
                               - 3.1543839e+2 * (aw * aw * aw) + 6.770824e+1 * ((aw * aw) * (aw * aw));
                const DD bm2 = 1. / (aw * kMwa) - 1. / kMwa;
                const DD bm5 = 5.875248e+1 - 1.8781997e+2 * aw + 2.7211377e+2 * (aw * aw)
                               - 1.8458287e+2 * (aw * aw * aw) + 4.153689e+1 * ((aw * aw) * (aw * aw));
                zlwc = zv[0] * (kRhosu / kMsu) / bm1 + kEpsoc * (zv[1] * (kRhooc / kMoc)) / bm2 +
                       zv[4] * (kRhoss / kMss) / bm5;
                zdwet = dcbrt(zlwc / n / kRhowa / kPi6 + zcore / kPi6);
                zke = dexp(2. * kSurfw0 * kMvwa / (kBoltz * T * zdwet));
                if (++count > 100) { flags |= SALSA_CELL_NOCONV; break; }
              }
              V[b][iwa] = zlwc / kRhowa;
            }
          }
        }
        DD w2 = dconst(0.);
        for (int b = 0; b < D.nb[0]; ++b) w2 += V[b][iwa];
        w2 = w2 * kRhowa;
        rv = rv - (w2 - w1) / (P * kMair / (kRg * T));

        // cell-skip predicate (:784-788; .AND. binds before .OR.)
        uint32_t any = 0;
        for (int b = 0; b < nt; ++b) {
          const int c = D.cat_of[b];
          const double n = N[b].v;
          if (c == 0 && n > kNlim) any |= 8u;
          if (c == 1 && n > kNlim) any |= 1u;
          if (c == 2 && n > kPrlim) any |= 2u;
          if (c == 3 && n > kPrlim) any |= 4u;
        }
        const bool pass = ((any & 1u) || ((any & 2u) && D.lscndh2ocl)) || ((any & 4u) && D.lscndh2oic) ||
                          ((any & 8u) && zrh.v > 0.98 && D.lscndh2oae);
        if (pass) {
          const double rhoair = kMair * P / (kRg * T);
          double zdf = (5. / (16. * kAvog * rhoair * 1.e-3 * (3.11e-8 * 3.11e-8))) *
                       ::sqrt(kRg * 1.e7 * T * kMair * 1.e3 * (kMwa + kMair) * 1.e3 / (2. * kPi * kMwa * 1.e3));
          zdf = zdf * 1.e-4;
          const double zmfp = 3. * zdf * ::sqrt(kPi * kMwa / (8. * kRg * T));
          const double zthc = 0.023807 + 7.1128e-5 * (T - 273.16);
          const double hl = ((kAlv * kMwa) / (kRg * T)) - 1.;
          const double hs = ((kAls * kMwa) / (kRg * T)) - 1.;
          const double csul = rs * rhoair / kMwa, csui = rsi * rhoair / kMwa;
          DD zcw[MB], zmt[MB], zkl[MB], zns[MB], zws[MB];
          uint64_t upd = 0, pred = 0, iceq = 0;
          for (int b = 0; b < nt; ++b) {
            const int c = D.cat_of[b];
            const DD n = N[b];
            DD vs = dconst(0.);
            for (int s = 0; s < nsV; ++s) vs += V[b][s];
            DD dw = n > D.nlim[b] ? dcbrt(vs / n / kPi6) : dconst(D.dmid[b]);
            dw = dmin(dw, D.dlim[b]);                     // updateDiameter (:803-806)
            DD zn = dconst(0.);
            for (int s = 0; s < ndry; ++s) zn += D.diss[s] * V[b][s] * D.rholiq[s] / D.mm[s];
            DD zcw0, zmt0 = dconst(0.), zkl0 = dconst(1.), zws0 = dconst(0.);
            if (c == 3) {
              iceq |= 1ull << b;
              zcw0 = V[b][iwa] * kRhoic / kMwa + V[b][irim] * kRhori / kMwa;
            } else {
              zcw0 = V[b][iwa] * kRhowa / kMwa;
            }
            const bool on = n > D.nlim[b] &&
                            (c == 0 ? (zrh.v > 0.98 && D.lscndh2oae)
                                    : c == 3 ? (D.lscndh2oic && T < 273.15) : D.lscndh2ocl);
            if (on) {
              const DD zknud = 2. * zmfp / dw;
              const DD zbeta = (zknud + 1.) / (0.377 * zknud + 1. + 4. / 3. * (zknud + zknud * zknud));
              if (c == 3) {                                   // :892-930
                zkl0 = dexp(4. * kSurfi0 * kMwa / (kRg * T * kRhowa * dw));
                zws0 = 1.0 * zkl0;
                const DD h1 = n * 4. * kPi * (0.5 * dw) * zdf * zbeta;
                const DD h2 = kMwa * zdf * kAls * zws0 * csui / (zthc * T);
                zmt0 = h1 / (h2 * hs + 1.);
              } else {                                        // :826-890, :932-955
                const DD zact = g_act_h2o(V[b][iwa] * D.rholiq[iwa] / D.mm[iwa], zn);
                zkl0 = dexp(4. * kSurfw0 * kMwa / (kRg * T * kRhowa * dw));
                zws0 = zact * zkl0;
                const DD h1 = n * 2. * kPi * dw * zdf * zbeta;
                const DD h2 = kMwa * zdf * kAlv * zws0 * csul / (zthc * T);
                zmt0 = h1 / (h2 * hl + 1.);
              }
            }
            zcw[b] = zcw0; zmt[b] = zmt0; zkl[b] = zkl0; zns[b] = zn; zws[b] = zws0;
            // aerosol 1a bins sit outside the predictor and are never updated (nstr = in2a)
            const bool inpred = c != 0 || b >= A0 + D.in2a;
            if (inpred) pred |= 1ull << b;
            if (inpred && on) upd |= 1ull << b;
          }
          DD zcwc = rv * rhoair / kMwa;
          DD stot = dconst(0.), s1 = dconst(0.), s2 = dconst(0.), sfix = dconst(0.);
          for (int b = 0; b < nt; ++b) {
            stot += zcw[b];
            if ((pred >> b) & 1ull) {
              s1 += zmt[b] * zws[b] * (((iceq >> b) & 1ull) ? csui : csul);
              s2 += zmt[b];
            }
            if (!((upd >> b) & 1ull)) sfix += zcw[b];
          }
          const DD zcwtot = zcwc + stot;
          const double adt = 2.e-2;
          DD zcwint = dconst(0.);
          // Stability of the explicit water update.  Where the +-10 % limiter
          // does not bind, a perturbation of bin b's water is multiplied per
          // substep by m_b = (1 - a_b csu w'_b) (1 - a_b / (1 + adt s2)), with
          // a_b = adt zmt_b and w'_b = d zws_b / d zcw_b (the vapour predictor
          // takes the rest).  A haze bin near activation can have |m_b| >> 1:
          // its water then alternates between the limits and every substep the
          // limiter misses multiplies the tangent, while finite perturbations
          // stay within the limits.  grow is the product over the substeps of
          // the largest such |m_b| (the value path, alike on every lane).
          double grow = 1.;
          for (int st = 0; st < a.nsub; ++st) {          // :986-1053
            const DD zhlp1 = zcwc + adt * s1;
            const DD zhlp2 = 1. + adt * s2;
            zcwint = dmin(zhlp1 / zhlp2, zcwtot);
            DD tsum = dconst(0.), s1n = dconst(0.);
            double mmax = 1.;
            for (int b = 0; b < nt; ++b) {
              const bool ic = (iceq >> b) & 1ull;
              if ((upd >> b) & 1ull) {
                const double csu = ic ? csui : csul;
                const DD d = adt * zmt[b] * (zcwint - zws[b] * csu);
                if (!ic && fabs(d.v) < 1.e-1 * zcw[b].v) {
                  const double znw = zcw[b].v, zs = zns[b].v, den = znw + zs;
                  const double wp = den > kEpsR8 && znw / den > 0.1 ? zkl[b].v * zs / (den * den) : 0.;
                  const double ab = adt * zmt[b].v;
                  mmax = fmax(mmax, fabs((1. - ab * csu * wp) * (1. - ab / zhlp2.v)));
                }
                DD nw = zcw[b] + dmin(dmax(d, -1.e-1 * zcw[b]), 1.e-1 * zcw[b]);
                nw = dmax(nw, 0.);
                zcw[b] = nw;
                zws[b] = ic ? zkl[b] : g_act_h2o(nw, zns[b]) * zkl[b];
                tsum += nw;
              }
              if ((pred >> b) & 1ull) s1n += zmt[b] * zws[b] * (ic ? csui : csul);
            }
            s1 = s1n;
            zcwint = zcwtot - (sfix + tsum);
            zcwc = zcwint;
            grow = fmin(grow * mmax, 1.e300);
          }
          // a growth beyond 1e4 in one step: the linearisation holds only for
          // perturbations far below any that a difference can resolve
          if (grow > 1.e4) flags |= SALSA_CELL_GRAD_UNSTABLE;
          if (a.nsub > 100) flags |= SALSA_CELL_SUBSTEPS;
          rv = zcwint * kMwa / rhoair;
          // write-back (:1058-1117)
          for (int b = 0; b < nt; ++b) {
            if (!((iceq >> b) & 1ull)) {
              V[b][iwa] = dmax(0., zcw[b] * kMwa / kRhowa);
            } else {
              const DD org_i = V[b][iwa] * kRhoic / kMwa, org_r = V[b][irim] * kRhori / kMwa;
              const DD orig = org_i + org_r;
              const DD dvt = zcw[b] - orig;
              if (dvt.v > 0.) {
                V[b][iwa] = V[b][iwa] + dvt * kMwa / kRhoic;
              } else if (dvt.v < 0.) {
                const DD fr = (zcw[b] / orig) - 1.;
                V[b][iwa] = V[b][iwa] + (fr * org_i) * kMwa / kRhoic;
                V[b][irim] = V[b][irim] + (fr * org_r) * kMwa / kRhori;
              }
            }
          }
        }
      }

      // ---- autoconversion (autoconv2), cloud activation and size-distribution
      //      update (distr_update), mo_salsa.f90:110-119, 148-149 ----
      if (D.autoconv2) autoconv2_cell<DD>(D, N, &V[0][0], NS, dt);           // mo_salsa.f90:110-112
      if (D.lsactiv == 1) act_interst_cell<DD>(D, N, &V[0][0], NS, rv, rs, T, tmpd);   // :115-119
      if (D.lsdistupdate) flags |= distr_update_cell<DD>(D, N, &V[0][0], NS, tmpd);

      // ---- the new state through LES units, as the next run_SALSA call sees it ----
      // The cell's linearisation is unstable when the state of an active bin
      // (numc > nlim) or the vapour amplifies a relative input change more
      // than 1e8 times: no perturbation resolvable in FP64 follows it then
      // (the reference's discrete condensation trajectory is unstable there).
      for (int b = 0; b < nt; ++b) {
        const int c = D.cat_of[b];
        const int nsw = c < 3 ? D.nwet : D.nspec;
        const DD nn = N[b] / dn;
        if (!isfinite(nn.v)) flags |= SALSA_CELL_NONFINITE;
        if (!isfinite(nn.d)) flags |= SALSA_CELL_GRAD_UNSTABLE;
        const bool active = N[b].v > D.nlim[b];
        if (active && fabs(N[b].d) * x0 > 1.e8 * N[b].v) flags |= SALSA_CELL_GRAD_UNSTABLE;
        N[b] = nn * dn;
        double vtot = 0.;
        for (int s = 0; s < nsw; ++s) vtot += V[b][s].v;
        for (int s = 0; s < nsw; ++s) {
          const double rho = species_rho(D, c, s);
          if (active && fabs(V[b][s].d) * x0 > 1.e8 * vtot) flags |= SALSA_CELL_GRAD_UNSTABLE;
          const DD mm = V[b][s] * rho / dn;
          if (!isfinite(mm.v)) flags |= SALSA_CELL_NONFINITE;
          if (!isfinite(mm.d)) flags |= SALSA_CELL_GRAD_UNSTABLE;
          V[b][s] = mm * dn / rho;
        }
      }
      if (!isfinite(rv.v)) flags |= SALSA_CELL_NONFINITE;
      if (!isfinite(rv.d)) flags |= SALSA_CELL_GRAD_UNSTABLE;
      if (fabs(rv.d) * x0 > 1.e8 * rv.v) flags |= SALSA_CELL_GRAD_UNSTABLE;
    }

    // objective of this cell: precipitation volume concentration
    DD J = dconst(0.);
    for (int b = D.off[2]; b < D.off[2] + D.nb[2]; ++b)
      for (int s = 0; s < D.nwet; ++s) J += V[b][s];
    // the same bound for the objective
    if (J.v > 0. && fabs(J.d) * x0 > 1.e8 * J.v) flags |= SALSA_CELL_GRAD_UNSTABLE;
    if (!isfinite(J.d)) flags |= SALSA_CELL_GRAD_UNSTABLE;
    if (dir < nt) {
      g.djdn[D.cat_of[dir]][g_les_index(a, cell) + a.n3 * D.loc_of[dir]] = J.d;
    } else if (dir == nt) {
      g.j_cell[cell] = J.v;
      g.djde_cell[cell] = J.d;
    }
    return dir <= nt ? flags : 0u;
  }
}

template <class KT, int MB, int NS, int NP>
__global__ void __launch_bounds__(64) k_salsa_gradient(const DevLayout *__restrict__ Dp, GradArgs g) {
  const DevLayout &D = *Dp;
  const int LPC = g.lanes, cpb = 64 / LPC;
  const int dir = (int)threadIdx.x % LPC;
  for (int64_t grp = blockIdx.x; grp * cpb < g.s.ncells; grp += gridDim.x) {
    const int64_t cell = grp * cpb + (int64_t)threadIdx.x / LPC;
    if (cell >= g.s.ncells) break;
    uint32_t flags = gradient_cell<KT, MB, NS, NP>(D, g, cell, dir);
    for (int m = 1; m < LPC; m <<= 1) flags |= (uint32_t)__shfl_xor((int)flags, m, LPC);   // over the cell's lanes
    if (dir == 0) {
      if (g.s.cell_flags) g.s.cell_flags[cell] = (uint8_t)flags;
      if (flags && g.s.fault_any) atomicOr(g.s.fault_any, (int)flags);
    }
  }
}

// Sums of the per-cell objective and of dJ/dE over the cells with a stable
// linearisation, in a fixed order (one block).
__global__ void __launch_bounds__(1024) k_gradient_sum(const double *__restrict__ x, const double *__restrict__ y,
                                                       const uint8_t *__restrict__ fl, int64_t n, double *out) {
  __shared__ double sx[1024], sy[1024];
  const int t = threadIdx.x;
  double ax = 0., ay = 0.;
  for (int64_t i = t; i < n; i += 1024) {
    ax += x[i];
    if (!(fl[i] & SALSA_CELL_GRAD_UNSTABLE)) ay += y[i];
  }
  sx[t] = ax;
  sy[t] = ay;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (t < s) {
      sx[t] += sx[t + s];
      sy[t] += sy[t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = sx[0];
    out[1] = sy[0];
  }
}

template <class KT, int MB, int NS, int NP>
static int launch_gradient_t(const DevLayout *d, const GradArgs &g, hipStream_t s) {
  const int cpb = 64 / g.lanes;
  const int64_t ngrp = (g.s.ncells + cpb - 1) / cpb;
  // enough wavefronts to cover the machine a few times over; each loops over cell groups
  const unsigned nblk = (unsigned)(ngrp < 4096 ? ngrp : 4096);
  if (nblk == 0) return 0;
  hipLaunchKernelGGL((k_salsa_gradient<KT, MB, NS, NP>), dim3(nblk), dim3(64), 0, s, d, g);
  return (int)hipGetLastError();
}

int launch_gradient(int cap, int precision, const DevLayout *d, const GradArgs &g, double *sum_out,
                    hipStream_t s) {
  int rc = -1;
  if (cap == 4)
    rc = precision == SALSA_FP64 ? launch_gradient_t<double, 32, 3, 528>(d, g, s)
                                 : launch_gradient_t<float, 32, 3, 528>(d, g, s);
  else if (cap == 5)
    rc = precision == SALSA_FP64 ? launch_gradient_t<double, 40, 4, 820>(d, g, s)
                                 : launch_gradient_t<float, 40, 4, 820>(d, g, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_gradient_sum, dim3(1), dim3(1024), 0, s, g.j_cell, g.djde_cell, g.s.cell_flags, g.s.ncells,
                     sum_out);
  return (int)hipGetLastError();
}

}  // namespace salsa
