// salsa_adjoint.hip -- reverse-mode (adjoint) derivatives of a multi-step
// SALSA coagulation + condensation trajectory at level 4.
//
// Objective (the forward-mode kernel's, salsa_gradient.hip): J = sum over
// interior cells of the precipitation-bin volume concentration after nsteps
// in-place steps of the box model (T, p, rs, rsi, dn fixed).  The adjoint of
// one step maps lam' = dJ/dx' (x' the cell state after the step) to
// lam = dJ/dx (the state before it) in one reverse pass, together with the
// step's share of dJ/dE (E the collision-kernel enhancement factor,
// K = E * coagc): one sweep gives the derivative in every direction at once,
// where forward mode carries one tangent per direction.
//
// The state of a cell is x = (rv, N[nt], V[nt][nwet]) in SALSA units.  One
// 64-lane wavefront works on one cell; lane b owns bin b.  Each launch runs
// one step for every cell:
//   mode 0  the plain forward step x_in -> x_out (the forward pass, which
//           keeps a checkpoint every few steps, and the recomputation of the
//           states between checkpoints);
//   mode 1  the forward step again, keeping what its reverse needs, then the
//           reverse: lam (the adjoint of x_out) is replaced in place by the
//           adjoint of x_in, and the step's dJ/dE is added to ebar.
// What the reverse keeps: the local derivatives of every pair kernel with
// respect to the two wet diameters (dual-number coagc in FP64), each sweep
// target's values before its update (in the registers of the target's lane),
// the precipitation bins before a cloud target moves drops into them, the
// state before the clamp, the condensation set-up (registers of the bin's
// lane), the Kelvin-factor iterates of the regime-1 equilibration, and the
// water of every bin and the vapour before each condensation substep.  The
// last three and the pair derivatives live in a per-block tape in HBM.
//
// Forward code: the step kernel's level-4 path (salsa_kernels.hip) with
// coagulation-based precipitation formation -- bin diameters
// (classSection.f90:91-159), coagc (mo_salsa_coagulation_kernels.f90:185-420),
// the target sweep (mo_salsa_coagulation_processes.f90:24-1611), the clamp
// (mo_salsa_dynamics.f90:189-198), equilibration (mo_salsa_properties.f90:
// 73-380) and gpparth2o (mo_salsa_dynamics.f90:691-1125).  Reverse code, the
// same lines backwards:
//   * substeps: through the limiter branches the forward took, from the stored
//     waters; the vapour closure zcwint = zcwtot - (sfix + sum of new water);
//   * equilibration: regime 1 (tolerance 1e-2) through its actual iterations,
//     regime 2 (tolerance 1e-12) at its fixed point (implicit function
//     theorem: the output's adjoint through (1 - dH/dz)^-1);
//   * sweep: targets in reverse order; the bins a target changed are restored
//     to their values before its update first, so every sum is reversed at
//     the state it was formed in;
//   * pair kernels: dK/d(dw) onto both bins' wet diameters, and those onto
//     the bins' number and volumes.
// Between steps the state passes through LES units as run_SALSA leaves it
// (N -> N/dn*dn, V -> V*rho/dn*dn/rho), whose derivative is the identity.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "salsa_dual_physics.h"
#include "salsa_internal.h"
#include "salsa_physics.h"

namespace salsa {

namespace {

constexpr int kEqIt = 104;     // stored regime-1 iterates per bin (at most 101 are used)

// largest value over the L lanes of a cell (every lane of the cell gets it):
// DPP within rows, then the row and half exchanges
template <int L>
__device__ __forceinline__ double cell_max(double x) {
  x = fmax(x, dpp_d<0xB1>(x));
  x = fmax(x, dpp_d<0x4E>(x));
  x = fmax(x, dpp_d<0x141>(x));
  x = fmax(x, dpp_d<0x140>(x));
  if constexpr (L >= 32) {
    auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(x), __double2loint(x), false, false);
    auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(x), __double2hiint(x), false, false);
    x = fmax(__hiloint2double(hi[0], lo[0]), __hiloint2double(hi[1], lo[1]));
  }
  if constexpr (L >= 64) {
    auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(x), __double2loint(x), false, false);
    auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(x), __double2hiint(x), false, false);
    x = fmax(__hiloint2double(hi[0], lo[0]), __hiloint2double(hi[1], lo[1]));
  }
  return x;
}

// Ordering of the cell's LDS traffic between its lanes.  A block is one
// wavefront, so no barrier between waves is needed; the fence keeps the
// compiler from moving LDS accesses across it and waits for the
// outstanding ones.  Unlike a block barrier it is valid where the two cells
// of the wavefront take different branches.
__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// a polynomial of degree 4 in aw and its derivative
__device__ __forceinline__ void poly4(double aw, double c0, double c1, double c2, double c3, double c4, double &p,
                                      double &dp) {
  const double aw2 = aw * aw;
  p = c0 + c1 * aw + c2 * aw2 + c3 * (aw2 * aw) + c4 * (aw2 * aw2);
  dp = c1 + 2. * c2 * aw + 3. * c3 * aw2 + 4. * c4 * (aw2 * aw);
}

// One equilibration task: regime 1 (mo_salsa_properties.f90:185-262) or
// regime 2 (:296-374), with the iteration's parameters.
struct EqTask {
  double zr, c1, c2, c5, n, zcore;
  int reg;
};
struct EqBar {
  double zr, c1, c2, c5, n, zcore;
};

// The task of aerosol bin bb, read from the cell's state.
__device__ __forceinline__ EqTask eq_task(const DevLayout &D, const double *sN, const double *sV, int nsV, int bb,
                                          int reg, double zrh, int isu, int ioc) {
  EqTask q;
  q.reg = reg;
  q.n = sN[bb];
  if (reg == 1) {
    const double v1 = isu >= 0 ? sV[bb * nsV + isu] : 0., v2 = ioc >= 0 ? sV[bb * nsV + ioc] : 0.;
    q.zcore = v1 / q.n + v2 / q.n;
    q.c1 = v1 * (kRhosu / kMsu);
    q.c2 = kEpsoc * v2 * (kRhooc / kMoc);
    q.c5 = 0.;
    q.zr = fmin(fmax(zrh, 0.05), 0.98);
  } else {
    double zv[5] = {0., 0., 0., 0., 0.};
    for (int s = 0; s < D.ndry; ++s)
      if (D.dry_id[s] < 5) zv[D.dry_id[s]] = sV[bb * nsV + s];
    q.zcore = zv[0] / q.n + zv[1] / q.n + zv[2] / q.n + zv[3] / q.n + zv[4] / q.n;
    q.c1 = zv[0] * (kRhosu / kMsu);
    q.c2 = kEpsoc * (zv[1] * (kRhooc / kMoc));
    q.c5 = zv[4] * (kRhoss / kMss);
    q.zr = fmin(fmax(zrh, 0.37), 0.98);
  }
  return q;
}

// One iteration zke -> (zlwc, zdwet, zke'), the step kernel's arithmetic.
__device__ __forceinline__ void eq_iter(const EqTask &q, double T, double zke, double &zlwc, double &zdwet,
                                        double &zke_out) {
  const double aw = div_nr(q.zr, zke);
  const double aw2 = aw * aw;
  const double bm1 = 1.1065495e+2 - 3.6759197e+2 * aw + 5.0462934e+2 * aw2 - 3.1543839e+2 * (aw2 * aw) +
                     6.770824e+1 * (aw2 * aw2);
  const double bm2 = rcp_nr(aw * kMwa) - 1. / kMwa;
  double l = div_nr(q.c1, bm1) + div_nr(q.c2, bm2);
  if (q.reg == 2) {
    const double bm5 = 5.875248e+1 - 1.8781997e+2 * aw + 2.7211377e+2 * aw2 - 1.8458287e+2 * (aw2 * aw) +
                       4.153689e+1 * (aw2 * aw2);
    l = l + div_nr(q.c5, bm5);
  }
  zlwc = l;
  zdwet = cbrt_fast(div_nr(div_nr(div_nr(zlwc, q.n), kRhowa), kPi6) + div_nr(q.zcore, kPi6));
  zke_out = ::exp(div_nr(2. * kSurfw0 * kMvwa, kBoltz * T * zdwet));
}

// Reverse of one iteration at zke: al and az are the adjoints of zlwc and of
// zke'; returns the adjoint of zke and adds those of the parameters to pb.
__device__ __forceinline__ double eq_iter_rev(const EqTask &q, double T, double zke, double al, double az,
                                              EqBar &pb) {
  const double aw = q.zr / zke;
  double bm1, dbm1, bm5 = 1., dbm5 = 0.;
  poly4(aw, 1.1065495e+2, -3.6759197e+2, 5.0462934e+2, -3.1543839e+2, 6.770824e+1, bm1, dbm1);
  const double bm2 = 1. / (aw * kMwa) - 1. / kMwa, dbm2 = -1. / (aw * aw * kMwa);
  double l = q.c1 / bm1 + q.c2 / bm2;
  if (q.reg == 2) {
    poly4(aw, 5.875248e+1, -1.8781997e+2, 2.7211377e+2, -1.8458287e+2, 4.153689e+1, bm5, dbm5);
    l += q.c5 / bm5;
  }
  const double cden = q.n * kRhowa * kPi6;
  const double u = l / cden + q.zcore / kPi6;
  const double zdwet = cbrt_fast(u);
  const double B = 2. * kSurfw0 * kMvwa / (kBoltz * T);
  const double zkeo = ::exp(B / zdwet);
  // zke' = exp(B / zdwet), zdwet = u^(1/3)
  const double ubar = az * (-zkeo * B / (zdwet * zdwet)) * (zdwet / (3. * u));
  const double lbar = al + ubar / cden;
  pb.n += ubar * (-l / (q.n * cden));
  pb.zcore += ubar / kPi6;
  double dl_daw = -q.c1 * dbm1 / (bm1 * bm1) - q.c2 * dbm2 / (bm2 * bm2);
  pb.c1 += lbar / bm1;
  pb.c2 += lbar / bm2;
  if (q.reg == 2) {
    dl_daw += -q.c5 * dbm5 / (bm5 * bm5);
    pb.c5 += lbar / bm5;
  }
  const double awbar = lbar * dl_daw;
  pb.zr += awbar / zke;
  return awbar * (-aw / zke);
}

// The parameters' adjoints onto the bin's number and dry volumes; returns the
// relative humidity's (zero where its clamp binds).
__device__ __forceinline__ double eq_task_rev(const DevLayout &D, const EqTask &q, const EqBar &pb, int nsV, int bb,
                                              double zrh, double *aN, double *aV) {
  const double lo = q.reg == 1 ? 0.05 : 0.37;
  aN[bb] += pb.n - pb.zcore * q.zcore / q.n;       // zcore = sum_s v_s / n
  for (int s = 0; s < D.ndry; ++s) {
    const int id = D.dry_id[s];
    double g = 0.;
    if (q.reg == 1 ? (id == 0 || id == 1) : id < 5) g += pb.zcore / q.n;
    if (id == 0) g += pb.c1 * (kRhosu / kMsu);
    if (id == 1) g += pb.c2 * kEpsoc * (kRhooc / kMoc);
    if (id == 4 && q.reg == 2) g += pb.c5 * (kRhoss / kMss);
    aV[bb * nsV + s] += g;
  }
  return (zrh > lo && zrh < 0.98) ? pb.zr : 0.;
}

// derivatives of acth2o, max(0.1, x / max(eps, x + s)), in x and s; below
// the floor (x + s <= eps, reached by a bin of a few 1e-16 mol m^-3 of
// water) the quotient is x / eps
__device__ __forceinline__ void act_h2o_d(double x, double s, double &dx, double &ds) {
  const double den = x + s;
  dx = ds = 0.;
  if (den > kEpsR8) {
    if (x / den > 0.1) {
      const double r = 1. / (den * den);
      dx = s * r;
      ds = -x * r;
    }
  } else if (x / kEpsR8 > 0.1) {
    dx = 1. / kEpsR8;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// the step / adjoint kernel: one cell per wavefront, cells strided over the grid
// ---------------------------------------------------------------------------
template <class KT, int MB, int NS, int NP, bool REV, int LPC>
__global__ void __launch_bounds__(64, 2) k_salsa_adjoint(const DevLayout *__restrict__ Dp, AdjArgs a) {
  static_assert(NS <= 3, "a target's tape row holds three species sums");
  static_assert(MB <= LPC, "one lane per bin");
  constexpr int CPW = 64 / LPC;        // cells per wavefront
  const DevLayout &D = *Dp;
  const int lane = threadIdx.x % LPC;  // the cell's lane (its bin)
  const int cs = threadIdx.x / LPC;    // the cell's slot in the wavefront
  const int lb = cs * LPC;             // its first lane in the wavefront
  const int nt = D.nt, nsV = D.nwet, np = D.npair, iwa = D.iwa, ndry = D.ndry;
  const int R = D.off[2], A0 = D.off[0], pf0 = D.pf0, pf1 = D.pf0 + D.ncc, nb2 = D.nb[2];
  const double dt = a.dt;
  constexpr bool rev = REV;
  const bool coag = D.lscoag, cond = D.lscnd && (D.lscndh2oae || D.lscndh2ocl);
  const int n1 = D.fn1a - D.in1a + 1;

  // LDS of the wavefront's cells
  __shared__ double sN_[CPW][MB], sV_[CPW][MB * NS], sDW_[CPW][MB];
  __shared__ KT sK_[CPW][NP];
  __shared__ uint8_t sPF_[CPW][NP];
  __shared__ double aN_[CPW][MB], aV_[CPW][MB * NS], aDW_[CPW][MB];
  __shared__ PT4<KT> sPR_[CPW][2 * MB];
  __shared__ KT sVL_[CPW][2 * MB];
  __shared__ int eqK_[CPW][4];         // regime-1 iterations per 1a bin (-1: bin skipped)
  double *sN = sN_[cs], *sV = sV_[cs], *sDW = sDW_[cs], *aN = aN_[cs], *aV = aV_[cs], *aDW = aDW_[cs];
  KT *sK = sK_[cs], *sVL = sVL_[cs];
  uint8_t *sPF = sPF_[cs];
  PT4<KT> *sPR = sPR_[cs];
  int *eqK = eqK_[cs];

  int isu = -1, ioc = -1;
  for (int s = 0; s < ndry; ++s) {
    if (D.dry_id[s] == 0) isu = s;
    if (D.dry_id[s] == 1) ioc = s;
  }
  double *tape = a.tape + ((int64_t)blockIdx.x * CPW + cs) * a.tape_stride;
  double *tP = tape + a.t_p;           // [2][np] dK/d(dw_b1), dK/d(dw_b2)
  double *tPF = tape + a.t_pf;         // [ntarget][nb2][1 + nsV] precipitation bins before a transfer
  double *tEQ = tape + a.t_eq;         // [4][kEqIt] regime-1 iterates
  double *tTG = tape + a.t_tg;         // [ntarget][8] each target's sums: sink, src[3], so, sx, transfer mask
  double (*gT)[8] = reinterpret_cast<double (*)[8]>(tape + a.t_gt);   // [2 MB][8] FP64 particle terms
                                       // (D c w v) and their d/d(dw): bins, then twins

  for (int64_t grp = blockIdx.x; grp * CPW < a.ncells; grp += gridDim.x) {
    const int64_t cell = grp * CPW + cs;
    if (cell >= a.ncells) continue;                      // the wavefront's other cell goes on alone
    const double *rec = a.rec + cell * a.nslot;
    const double T = rec[0], P = rec[1], rs = rec[2], dn = rec[4];
    const double *x = a.x_in + cell * a.nvar;
    double rv = x[0];
    bool rvlim = false;
    if (!a.lsfreeRH) {                                   // mo_salsa_driver.f90:193-197
      rvlim = rs * a.rhlim < rv;
      rv = fmin(rv, rs * a.rhlim);
    }
    uint32_t flags = 0;
    const bool own = lane < nt;                          // this lane owns bin `lane`
    if (own) {
      sN[lane] = x[1 + lane];
      for (int s = 0; s < NS; ++s) if (s < nsV) sV[lane * nsV + s] = x[1 + nt + lane * nsV + s];
    }
    wsync();

    // ---- bin diameters (updateDiameter, classSection.f90:91-159) and particle terms ----
    const CellC cc = cell_consts(D, T, P);
    double dwn = 0., dwv = 0.;           // d(dw)/dN, d(dw)/d(sum V) of this lane's bin
    bool actb = false;
    if (own) {
      const int b = lane;
      const double n = sN[b];
      double vs = 0.;
      for (int s = 0; s < NS; ++s) if (s < nsV) vs += sV[b * nsV + s];
      double dw = D.dmid[b];
      if (n > D.nlim[b]) {
        dw = cbrt_fast(vs / n / kPi6);
        dwn = -dw / (3. * n);
        dwv = vs > 0. ? dw / (3. * vs) : 0.;
      }
      if (D.dlim[b] < dw) { dw = D.dlim[b]; dwn = 0.; dwv = 0.; }
      sDW[b] = dw;
      actb = !(n < D.nlim[b]);
      if (coag) {
        const double mass = kRhowa * kPi6 * cubed(dw);
        if constexpr (sizeof(KT) == 4) {
          const PTermsF p0 = particle_terms_liquid_f(cc, dw, mass, kRhowa);
          const PTermsF p1 = particle_terms_liquid_f(cc, 1.1 * dw, 1.33 * mass, kRhowa);
          sPR[b] = PT4<KT>{(KT)p0.D, (KT)p0.c, (KT)p0.w, (KT)p0.v};
          sPR[MB + b] = PT4<KT>{(KT)p1.D, (KT)p1.c, (KT)p1.w, (KT)p1.v};
          sVL[b] = (KT)(p0.vd - (double)p0.v);
          sVL[MB + b] = (KT)(p1.vd - (double)p1.v);
        } else {
          const Shape sh{0., 0., 0., 0.};
          const PTerms p0 = particle_terms(cc, dw, mass, kRhowa, D.phase[b], dw, sh);
          const PTerms p1 = particle_terms(cc, 1.1 * dw, 1.33 * mass, kRhowa, D.phase[b], dw, sh);
          sPR[b] = PT4<KT>{(KT)p0.D, (KT)p0.c, (KT)p0.w, (KT)p0.v};
          sPR[MB + b] = PT4<KT>{(KT)p1.D, (KT)p1.c, (KT)p1.w, (KT)p1.v};
        }
        if constexpr (rev) {
          // the FP64 terms with their derivative in this bin's wet diameter
          const GCellC gc = g_cell_consts(D, T, P);
          const GShape gsh{};
          const DD dwd = dvar(dw, 1.), massd = kRhowa * kPi6 * dcube(dwd);
          const GPT<DD> q0 = g_pterms(gc, dwd, massd, dconst(kRhowa), D.phase[b], dwd, gsh);
          const GPT<DD> q1 = g_pterms(gc, 1.1 * dwd, 1.33 * massd, dconst(kRhowa), D.phase[b], dwd, gsh);
          double *g0 = gT[b], *g1 = gT[MB + b];
          g0[0] = q0.D.v; g0[1] = q0.c.v; g0[2] = q0.w.v; g0[3] = q0.v.v;
          g0[4] = q0.D.d; g0[5] = q0.c.d; g0[6] = q0.w.d; g0[7] = q0.v.d;
          g1[0] = q1.D.v; g1[1] = q1.c.v; g1[2] = q1.w.v; g1[3] = q1.v.v;
          g1[4] = q1.D.d; g1[5] = q1.c.d; g1[6] = q1.w.d; g1[7] = q1.v.d;
        }
      }
    }
    // bins with numc >= nlim take part in kernels
    const uint64_t abits = (__ballot(actb) >> lb) & (LPC == 64 ? ~0ull : (1ull << (LPC & 63)) - 1);
    wsync();

    // ---- pair kernels (buildKernel*, kernels.f90:94-180) and their diameter derivatives ----
    if (coag) {
      const KT vkin = (KT)cc.vkin, cshear = (KT)cc.cshear, cinert = (KT)cc.cinert;
      const KT cfac = (KT)D.charge_factor;
      GCellC gc{};
      if (rev) gc = g_cell_consts(D, T, P);
      for (int p = lane; p < np; p += LPC) {
        const uint32_t pw = D.pw[p];
        const int b1 = pw & 255u, b2 = (pw >> 8) & 255u, tw = (pw >> 16) & 1u;
        KT k = KT(0.);
        uint32_t pfb = 0;
        double p1 = 0., p2 = 0.;
        if ((abits >> b1) & (abits >> b2) & 1ull) {
          // particle 1 is the twin of particle 2 when the bins' dmid agree within 1 %
          const double dB = sDW[b2];
          const double dA = tw ? 1.1 * dB : sDW[b1];
          const int ia = tw ? MB + b2 : b1;
          const PT4<KT> A = sPR[ia], B = sPR[b2];
          k = pair_coef<KT>((KT)dA, A.D, A.c, A.w, A.v, (KT)dB, B.D, B.c, B.w, B.v, dA >= dB, vkin, cshear, cinert,
                            sizeof(KT) == 4 ? sVL[ia] : KT(0.), sizeof(KT) == 4 ? sVL[b2] : KT(0.));
          k = k * cfac;                                  // charge enhancement (:411-416)
          if (p >= pf0 && p < pf1) {
            // precip bin reached by coalescence (accumulatePrecipFormation, processes.f90:1461-1467)
            const double dnew = cbrt_fast(cubed(sDW[b1]) + cubed(dB));
            int cnt = 0;
            for (int qq = 0; qq < nb2; ++qq) cnt += dnew > D.precpbins[qq];
            pfb = dnew > D.precpbins[0] ? (uint32_t)cnt : 0u;
          }
          if constexpr (rev) {
            // dK/d(dw) of each particle; a twin pair depends on dw_b2 alone
            const double *ga = gT[ia], *gb = gT[b2];
            const double cf = D.charge_factor;
            for (int pass = 0; pass < (tw ? 1 : 2); ++pass) {
              const double ta = (tw || pass == 0) ? 1. : 0., tb = (tw || pass == 1) ? 1. : 0.;
              const DD kd = g_pair_coef<double>(
                  dvar(dA, ta * (tw ? 1.1 : 1.)), dvar(ga[0], ta * ga[4]), dvar(ga[1], ta * ga[5]),
                  dvar(ga[2], ta * ga[6]), dvar(ga[3], ta * ga[7]), dvar(dB, tb), dvar(gb[0], tb * gb[4]),
                  dvar(gb[1], tb * gb[5]), dvar(gb[2], tb * gb[6]), dvar(gb[3], tb * gb[7]), dA >= dB, gc.vkin,
                  gc.cshear, gc.cinert);
              if (tw || pass == 1) p2 = kd.d * cf;
              else p1 = kd.d * cf;
            }
          }
        }
        sK[p] = k;
        if (p >= pf0 && p < pf1) sPF[p - pf0] = (uint8_t)pfb;
        if constexpr (rev) {
          tP[p] = p1;
          tP[np + p] = p2;
        }
      }
    }
    wsync();

    // ---- the sweep: coag_precp, coag_aero, coag_cloud (mo_salsa_dynamics.f90:166-183) ----
    // Lane t keeps target t's values before its update (and bin t+1's, for
    // precipitation self-collection).
    bool tAct = false;
    double tNr = 0., tVr[NS], t1Nr = 0., t1Vr[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) { tVr[s] = 0.; t1Vr[s] = 0.; }
    const int ntg = coag ? D.ntarget : 0;
    for (int it = 0; it < ntg; ++it) {
      const DevTarget &TG = D.tgt[it];
      const int t = TG.t;
      const uint32_t fl = TG.flags;
      const double Nt = sN[t];
      double Vt[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) Vt[s] = s < nsV ? sV[t * nsV + s] : 0.;
      if (Nt < D.nlim[t]) continue;                       // target not active
      if constexpr (rev) {
        if (lane == t) {
          tAct = true;
          tNr = Nt;
#pragma unroll
          for (int s = 0; s < NS; ++s) tVr[s] = Vt[s];
          if (fl & TF_PRECIP_SELFCOLL) {
            t1Nr = sN[t + 1];
#pragma unroll
            for (int s = 0; s < NS; ++s) t1Vr[s] = s < nsV ? sV[(t + 1) * nsV + s] : 0.;
          }
        }
        if ((fl & (TF_PF | TF_PF_SELF)) && lane < nb2) {
          double *row = tPF + ((int64_t)it * nb2 + lane) * (1 + nsV);
          row[0] = sN[R + lane];
          for (int s = 0; s < NS; ++s) if (s < nsV) row[1 + s] = sV[(R + lane) * nsV + s];
        }
      }
      const int nupd = TG.nupd, ne = TG.nent;
      const uint32_t *ent = D.ent + TG.ent0;
      double sink = 0., so = 0., sx = 0., src[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) src[s] = 0.;
      uint32_t pmask = 0;
      if (lane < ne) {                                   // one entry per lane
        const uint32_t w = ent[lane];
        const int j = w & 63u, pid = (int)(w >> 16);
        const uint32_t role = (w >> 6) & 1023u;
        const double kj = sK[pid], Nj = sN[j];
        if (role & R_SINK) sink += kj * Nj;
        if (role & R_SRC) {
#pragma unroll
          for (int s = 0; s < NS; ++s)
            if (s < nupd) src[s] += kj * sV[j * nsV + s];
        }
        if (role & (R_PF_OWN | R_PF_OTH)) {              // accumulatePrecipFormation (:1458-1576)
          const int pb = (int)(sPF[pid - pf0] & 31u) - 1;
          if (pb >= 0) {
            sink += kj * Nj;
            pmask |= 1u << pb;
          } else {
#pragma unroll
            for (int s = 0; s < NS; ++s)
              if (s < nupd) src[s] += kj * sV[j * nsV + s];
          }
        }
        if (role & R_FIX_OWN) so += kj * Nj;
        if (role & R_FIX_OTH) sx += kj * Nj;
      }
      sink = cell_sum<LPC>(sink);
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (s < nupd) src[s] = cell_sum<LPC>(src[s]);
      double fix_own = 1., fix_oth = 1.;
      if (fl & TF_PF) {                                  // fix_coag (:1440-1456)
        so = cell_sum<LPC>(so);
        sx = cell_sum<LPC>(sx);
        fix_own = fmax(1. - dt * so, 0.1);
        fix_oth = fmax(1. - dt * sx, 0.1);
      }
      if constexpr (rev) {
        if (lane == 0) {
          double *row = tTG + it * 8;
          row[0] = sink;
#pragma unroll
          for (int s = 0; s < NS; ++s) row[1 + s] = s < nupd ? src[s] : 0.;
          row[4] = so;
          row[5] = sx;
        }
      }
      double sink_self = 0., kself = 0., fixs = 1., volsink[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) volsink[s] = 0.;
      int pself = -1;
      if (fl & (TF_SELF_HALF | TF_PF_SELF | TF_PRECIP_SELFCOLL)) {
        const int pid = D.pidx[t][t];
        kself = pid >= 0 ? sK[pid] : 0.;
        if (fl & TF_SELF_HALF) sink_self = 0.5 * kself * Nt;
        if (fl & TF_PF_SELF) {
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
      // applyCoag / applyCoagPrecp (:539-645)
      const double rden = rcp_nr(1. + dt * sink);
      double Vn[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) Vn[s] = s < nupd ? (Vt[s] + dt * src[s] * Nt) * rden : Vt[s];
      double Nn = Nt * rcp_nr(1. + dt * (sink + sink_self));
      wsync();                                   // every lane has read the state
      if (fl & TF_PRECIP_SELFCOLL) {                     // precipSelfCoag (:1588-1611)
        const double nslf = 0.5 * kself * Nt * Nt;
#pragma unroll
        for (int s = 0; s < NS; ++s)
          if (s < nupd) {
            const double vslf = Vt[s] * Nt * kself;
            if (lane == 0) sV[(t + 1) * nsV + s] = sV[(t + 1) * nsV + s] + dt * vslf;
            Vn[s] = Vn[s] - dt * vslf;
          }
        if (lane == 0) sN[t + 1] = sN[t + 1] + dt * nslf;
        Nn = Nn - dt * nslf;
      }
      if (fl & TF_PF_SELF) {
#pragma unroll
        for (int s = 0; s < NS; ++s) Vn[s] = Vn[s] - dt * volsink[s];
      }
      if (lane == 0) {
        sN[t] = Nn;
        for (int s = 0; s < NS; ++s) if (s < nsV) sV[t * nsV + s] = Vn[s];
      }
      // coalesced cloud drops into precipitation bins (applyCoagPrecipFormation :649-703);
      // the entries read cloud bins other than t, the writes go to precipitation bins
      if (fl & (TF_PF | TF_PF_SELF)) {
        uint32_t pm = cell_or<LPC>(pmask);
        if (pself >= 0) pm |= 1u << pself;
        if constexpr (rev) {
          if (lane == 0) tTG[it * 8 + 6] = (double)pm;
        }
        while (pm) {
          const int pb = __builtin_ctz(pm);
          pm &= pm - 1;
          double vp[NS], npp = 0.;
#pragma unroll
          for (int s = 0; s < NS; ++s) vp[s] = 0.;
          if (lane < ne) {
            const uint32_t w = ent[lane];
            const uint32_t role = (w >> 6) & 1023u;
            const int j = w & 63u, pid = (int)(w >> 16);
            if ((role & (R_PF_OWN | R_PF_OTH)) && (int)(sPF[pid - pf0] & 31u) - 1 == pb) {
              const double kj = sK[pid];
              const double fx = (role & R_PF_OWN) ? fix_own : fix_oth;
              const double Nj = sN[j];
#pragma unroll
              for (int s = 0; s < NS; ++s)
                if (s < nupd) vp[s] += (sV[j * nsV + s] * Nt + Vt[s] * Nj) * kj * fx;
              npp += kj * Nj * Nt * fx;
            }
          }
          npp = cell_sum<LPC>(npp);
#pragma unroll
          for (int s = 0; s < NS; ++s)
            if (s < nupd) vp[s] = cell_sum<LPC>(vp[s]);
          if (pb == pself) {
#pragma unroll
            for (int s = 0; s < NS; ++s) vp[s] += volsink[s];
            npp += 0.5 * kself * Nt * Nt * fixs;
          }
          if (lane == 0) {
            const int r = R + pb;
#pragma unroll
            for (int s = 0; s < NS; ++s)
              if (s < nupd) sV[r * nsV + s] = sV[r * nsV + s] + dt * vp[s];
            sN[r] = sN[r] + dt * npp;
          }
        }
      }
      wsync();
    }
    // clamp (mo_salsa_dynamics.f90:189-198); lane b keeps its bin before it
    double cNr = 0., cVr[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) cVr[s] = 0.;
    if (own && coag) {
      const int b = lane;
      cNr = sN[b];
      if (sN[b] < -1.e-6) flags |= SALSA_CELL_NEG_NUMC;
      sN[b] = fmax(0., sN[b]);
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (s < nsV) {
          cVr[s] = sV[b * nsV + s];
          sV[b * nsV + s] = fmax(0., sV[b * nsV + s]);
        }
    }
    wsync();

    // ---- condensation (mo_salsa_dynamics.f90:276-342, 691-1125) ----
    const double zrh = rv / rs;
    const double Qv = P * kMair / (kRg * T);
    const double rhoair = kMair * P / (kRg * T);
    const double adt = 2.e-2;
    const int nsub = a.nsub;
    bool pass = false;
    double zdf = 0., zmfp = 0., zthc = 0., hl = 0., csul = 0., s2 = 0., zcwtot = 0.;
    // this lane's bin in the condensation set-up
    double zcw0 = 0., zmt = 0., zkl = 1., zns = 0., zws0 = 0., dwc = 0., cdn = 0., cdv = 0.;
    bool on = false, upd = false, dwvar = false;
    // regime-2 equilibration of this lane's bin: the iterate entering the last iteration
    double eqz = 0.;
    int eq2 = -1;                       // -1: no task; else the iteration count
    if (cond) {
      double w1 = cell_sum<LPC>(lane < D.nb[0] ? sV[lane * nsV + iwa] : 0.) * kRhowa;
      // Equilibration (mo_salsa_properties.f90:73-380): regime 1 on lane 0 (a
      // chain: the Kelvin factor carries over), regime 2 bin j on lane j + 1
      const bool reg2 = zrh < 0.98 || !D.lscndh2oae;
      const int n2 = reg2 ? D.fn2b - D.in2a + 1 : 0;
      if (lane == 0) {
        double zke1 = 1.001;
        for (int i1 = 0; i1 < n1; ++i1) {
          const int bb = A0 + D.in1a + i1;
          int K = -1;
          if (sN[bb] > kNlim) {
            const EqTask q = eq_task(D, sN, sV, nsV, bb, 1, zrh, isu, ioc);
            double zdwet = sDW[bb], zdold = 1., zlwc = 0., zke = zke1;
            K = 0;
            while (fabs(zdwet / zdold - 1.) > 1.e-2) {
              zdold = fmax(zdwet, 1.e-20);
              if (rev) tEQ[i1 * kEqIt + K] = zke;
              eq_iter(q, T, zke, zlwc, zdwet, zke);
              if (++K > 100) { flags |= SALSA_CELL_NOCONV; break; }
            }
            sV[bb * nsV + iwa] = zlwc / kRhowa;
            zke1 = zke;
          }
          eqK[i1] = K;
        }
      } else if (lane - 1 < n2) {
        const int bb = A0 + D.in2a + (lane - 1);
        if (sN[bb] > kNlim) {
          const EqTask q = eq_task(D, sN, sV, nsV, bb, 2, zrh, isu, ioc);
          double zdwet = sDW[bb], zdold = 1., zlwc = 0., zke = 1.02;
          int K = 0;
          while (fabs(zdwet / zdold - 1.) > 1.e-12) {
            zdold = fmax(zdwet, 1.e-20);
            eqz = zke;
            eq_iter(q, T, zke, zlwc, zdwet, zke);
            if (++K > 100) { flags |= SALSA_CELL_NOCONV; break; }
          }
          sV[bb * nsV + iwa] = zlwc / kRhowa;
          eq2 = K;
        }
      }
      wsync();
      const double w2 = cell_sum<LPC>(lane < D.nb[0] ? sV[lane * nsV + iwa] : 0.) * kRhowa;
      rv = rv - (w2 - w1) / Qv;

      // cell-skip predicate (:784-788; .AND. binds before .OR.)
      uint32_t any = 0;
      if (own) {
        const int c = D.cat_of[lane];
        const double n = sN[lane];
        if (c == 0 && n > kNlim) any |= 8u;
        if (c == 1 && n > kNlim) any |= 1u;
        if (c == 2 && n > kPrlim) any |= 2u;
      }
      any = cell_or<LPC>(any);
      pass = ((any & 1u) || ((any & 2u) && D.lscndh2ocl)) || ((any & 8u) && zrh > 0.98 && D.lscndh2oae);
    }
    if (pass) {
      zdf = (5. / (16. * kAvog * rhoair * 1.e-3 * (3.11e-8 * 3.11e-8))) *
            ::sqrt(kRg * 1.e7 * T * kMair * 1.e3 * (kMwa + kMair) * 1.e3 / (2. * kPi * kMwa * 1.e3));
      zdf = zdf * 1.e-4;
      zmfp = 3. * zdf * ::sqrt(kPi * kMwa / (8. * kRg * T));
      zthc = 0.023807 + 7.1128e-5 * (T - 273.16);
      hl = ((kAlv * kMwa) / (kRg * T)) - 1.;
      csul = rs * rhoair / kMwa;
      if (own) {                                         // set-up of this lane's bin (:803-955)
        const int b = lane;
        const int c = D.cat_of[b];
        const double n = sN[b];
        double vs = 0.;
        for (int s = 0; s < NS; ++s) if (s < nsV) vs += sV[b * nsV + s];
        double dw = D.dmid[b];
        if (n > D.nlim[b]) {
          dw = cbrt_fast(vs / n / kPi6);
          dwvar = true;
          cdn = -dw / (3. * n);
          cdv = vs > 0. ? dw / (3. * vs) : 0.;
        }
        if (D.dlim[b] < dw) { dw = D.dlim[b]; dwvar = false; }
        dwc = dw;
        for (int s = 0; s < ndry; ++s) zns += D.diss[s] * sV[b * nsV + s] * D.rholiq[s] / D.mm[s];
        zcw0 = sV[b * nsV + iwa] * kRhowa / kMwa;
        on = n > D.nlim[b] && (c == 0 ? (zrh > 0.98 && D.lscndh2oae) : D.lscndh2ocl);
        if (on) {
          const double zknud = 2. * zmfp / dw;
          const double zbeta = (zknud + 1.) / (0.377 * zknud + 1. + 4. / 3. * (zknud + zknud * zknud));
          const double zact = act_h2o<false>(sV[b * nsV + iwa] * D.rholiq[iwa] / D.mm[iwa], zns);
          zkl = ::exp(4. * kSurfw0 * kMwa / (kRg * T * kRhowa * dw));
          zws0 = zact * zkl;
          const double h1 = n * 2. * kPi * dw * zdf * zbeta;
          const double h2 = kMwa * zdf * kAlv * zws0 * csul / (zthc * T);
          zmt = h1 / (h2 * hl + 1.);
        }
        // aerosol 1a bins sit outside the predictor and are never updated (nstr = in2a)
        upd = on && (c != 0 || b >= A0 + D.in2a);
      }
      double zcw = zcw0, zws = upd ? zws0 : 0.;
      double zcwc = rv * rhoair / kMwa;
      const double stot = cell_sum<LPC>(own ? zcw0 : 0.);
      const double sfix = cell_sum<LPC>((own && !upd) ? zcw0 : 0.);
      double s1 = cell_sum<LPC>(upd ? zmt * zws * csul : 0.);
      s2 = cell_sum<LPC>(upd ? zmt : 0.);
      zcwtot = zcwc + stot;
      double zcwint = 0.;
      // Stability of the linearisation: where the +-10 % limiter does not
      // bind, a water perturbation of bin b is multiplied per substep by
      // (1 - a_b csu w'_b)(1 - a_b / (1 + adt s2)), a_b = adt zmt_b; grow is
      // the product over the substeps of the largest such factor (>= 1).
      double grow = 1.;
      for (int st = 0; st < nsub; ++st) {               // :986-1053
        if constexpr (rev) {
          double *row = tape + (int64_t)st * (nt + 2);
          if (own) row[lane] = zcw;
          if (lane == 0) {
            row[nt] = zcwc;
            row[nt + 1] = s1;
          }
        }
        const double zhlp1 = zcwc + adt * s1;
        const double zhlp2 = 1. + adt * s2;
        zcwint = fmin(div_nr(zhlp1, zhlp2), zcwtot);
        double nw = 0., s1n = 0., m = 1.;
        if (upd) {
          const double d = adt * zmt * (zcwint - zws * csul);
          if (rev && fabs(d) < 1.e-1 * zcw) {
            double dx, ds;
            act_h2o_d(zcw, zns, dx, ds);
            const double ab = adt * zmt;
            m = fmax(1., fabs((1. - ab * csul * dx * zkl) * (1. - ab / zhlp2)));
          }
          nw = zcw + fmin(fmax(d, -1.e-1 * zcw), 1.e-1 * zcw);
          nw = fmax(nw, 0.);
          zcw = nw;
          zws = act_h2o<true>(nw, zns) * zkl;
          s1n = zmt * zws * csul;
        }
        const double tsum = cell_sum<LPC>(nw);
        s1 = cell_sum<LPC>(s1n);
        zcwint = zcwtot - (sfix + tsum);
        zcwc = zcwint;
        if (rev) grow = fmin(grow * cell_max<LPC>(m), 1.e300);
      }
      if (grow > 1.e4) flags |= SALSA_CELL_GRAD_UNSTABLE;
      if (nsub > 100) flags |= SALSA_CELL_SUBSTEPS;
      rv = zcwint * kMwa / rhoair;
      if (own) sV[lane * nsV + iwa] = fmax(0., zcw * kMwa / kRhowa);   // write-back (:1058-1117)
      wsync();
    }

    if constexpr (!rev) {
      // the new state through LES units, as the next run_SALSA call sees it
      double *xo = a.x_out + cell * a.nvar;
      if (own) {
        const int b = lane, c = D.cat_of[b];
        const double nn = sN[b] / dn;
        if (!isfinite(nn)) flags |= SALSA_CELL_NONFINITE;
        xo[1 + b] = nn * dn;
        for (int s = 0; s < NS; ++s) if (s < nsV) {
          const double rho = species_rho(D, c, s);
          const double mm = sV[b * nsV + s] * rho / dn;
          if (!isfinite(mm)) flags |= SALSA_CELL_NONFINITE;
          xo[1 + nt + b * nsV + s] = mm * dn / rho;
        }
      }
      if (!isfinite(rv)) flags |= SALSA_CELL_NONFINITE;
      if (lane == 0) xo[0] = rv;
      flags = cell_or<LPC>(flags);
      if (lane == 0 && flags) {
        if (a.cell_flags) a.cell_flags[cell] |= (uint8_t)flags;
        if (a.fault_any) atomicOr(a.fault_any, (int)flags);
      }
      wsync();
      continue;
    }

    // =====================================================================
    // reverse of the step
    // =====================================================================
    double *lam = a.lam + cell * a.nvar;
    if (own) {
      aN[lane] = lam[1 + lane];
      for (int s = 0; s < NS; ++s) if (s < nsV) aV[lane * nsV + s] = lam[1 + nt + lane * nsV + s];
      aDW[lane] = 0.;
    }
    const double arv = lam[0];
    double aE = 0.;                                      // this lane's share of dJ/dE
    wsync();

    double arv_l = arv;                                  // adjoint of rv after the RH limit
    if (cond) {
      // ---- condensation ----
      double arv_e = arv;
      if (pass) {
        double acw = 0.;                                 // adjoint of this bin's zcw after the substep
        if (own) {
          if (sV[lane * nsV + iwa] > 0.) acw = aV[lane * nsV + iwa] * (kMwa / kRhowa);
          aV[lane * nsV + iwa] = 0.;
        }
        double ac = arv * kMwa / rhoair;                 // adjoint of zcwc after the substep
        double azmt = 0., azkl = 0., azns = 0., azws0 = 0.;
        double azcwtot = 0., asfix = 0., as2 = 0.;
        const double zhlp2 = 1. + adt * s2;
        for (int st = nsub - 1; st >= 0; --st) {
          const double *row = tape + (int64_t)st * (nt + 2);
          const double zcw = own ? row[lane] : 0.;
          const double zcwc = row[nt];
          const double act = upd && st > 0 ? act_h2o<true>(zcw, zns) : 0.;
          const double zws = upd ? (st == 0 ? zws0 : act * zkl) : 0.;
          const double s1 = row[nt + 1];
          const double zhlp1 = zcwc + adt * s1;
          const double A = div_nr(zhlp1, zhlp2);
          const bool tookA = !(zcwtot < A);
          const double zcwi = tookA ? A : zcwtot;
          // zcwc' = zcwtot - (sfix + sum of the updated bins' new water)
          azcwtot += ac;
          asfix -= ac;
          double azcw = 0., azws = 0., azcwi = 0.;
          if (upd) {
            const double anw = acw - ac;
            const double d = adt * zmt * (zcwi - zws * csul);
            const double lo = -1.e-1 * zcw, hi = 1.e-1 * zcw;
            const double cl = fmin(fmax(d, lo), hi);
            double ad = 0.;
            if (zcw + cl > 0.) {                         // nw = max(zcw + cl, 0)
              azcw = anw;
              if (d < lo) azcw += -0.1 * anw;
              else if (hi < d) azcw += 0.1 * anw;
              else ad = anw;
            }
            azmt += ad * adt * (zcwi - zws * csul);
            azcwi = ad * adt * zmt;
            azws = -ad * adt * zmt * csul;
          } else if (own) {
            azcw = acw;                                  // a bin outside the substeps keeps its water
          }
          azcwi = cell_sum<LPC>(azcwi);
          double aA = 0.;
          if (tookA) aA = azcwi;
          else azcwtot += azcwi;
          as2 += adt * (-aA * A / zhlp2);
          const double ahlp1 = aA / zhlp2;
          const double as1 = adt * ahlp1;
          if (upd) {
            azws += as1 * zmt * csul;
            azmt += as1 * zws * csul;
            if (st == 0) {
              azws0 += azws;
            } else {
              double dx, ds;
              act_h2o_d(zcw, zns, dx, ds);
              azcw += azws * zkl * dx;
              azns += azws * zkl * ds;
              azkl += azws * act;
            }
          }
          acw = azcw;
          ac = ahlp1;
        }
        // zcwtot = zcwc0 + sum zcw0, sfix = sum over bins outside the substeps, s2 = sum zmt
        ac += azcwtot;
        double azcw0 = acw + (own ? azcwtot : 0.) + ((own && !upd) ? asfix : 0.);
        if (upd) azmt += as2;
        arv_e = ac * rhoair / kMwa;                      // zcwc0 = rv rhoair / Mwa
        if (own) {
          const int b = lane;
          const double n = sN[b];
          double an = 0., adw = 0.;
          if (on) {
            const double zknud = 2. * zmfp / dwc;
            const double den = 0.377 * zknud + 1. + 4. / 3. * (zknud + zknud * zknud);
            const double zbeta = (zknud + 1.) / den;
            const double h1 = n * 2. * kPi * dwc * zdf * zbeta;
            const double C2 = kMwa * zdf * kAlv * csul / (zthc * T);
            const double g = C2 * zws0 * hl + 1.;
            const double ah1 = azmt / g, ah2 = -azmt * h1 * hl / (g * g);
            const double aws0 = azws0 + ah2 * C2;
            an += ah1 * h1 / n;
            adw += ah1 * n * 2. * kPi * zdf * zbeta;
            const double dbeta = (den - (zknud + 1.) * (0.377 + 4. / 3. * (1. + 2. * zknud))) / (den * den);
            adw += ah1 * n * 2. * kPi * dwc * zdf * dbeta * (-zknud / dwc);
            // zws0 = zact zkl, zkl = exp(Ck / dw)
            const double zact = zws0 / zkl;
            const double akl = azkl + aws0 * zact;
            const double Ck = 4. * kSurfw0 * kMwa / (kRg * T * kRhowa);
            adw += akl * zkl * (-Ck / (dwc * dwc));
            double dx, ds;
            act_h2o_d(zcw0, zns, dx, ds);
            const double aact = aws0 * zkl;
            azns += aact * ds;
            aV[b * nsV + iwa] += aact * dx * D.rholiq[iwa] / D.mm[iwa];
          }
          for (int s = 0; s < ndry; ++s) aV[b * nsV + s] += azns * D.diss[s] * D.rholiq[s] / D.mm[s];
          aV[b * nsV + iwa] += azcw0 * kRhowa / kMwa;
          if (dwvar) {
            an += adw * cdn;
            for (int s = 0; s < NS; ++s) if (s < nsV) aV[b * nsV + s] += adw * cdv;
          }
          aN[b] += an;
        }
        wsync();
      }
      // ---- vapour budget of the equilibration: rv_e = rv_l - (w2 - w1) / Qv ----
      const double aw = arv_e / Qv * kRhowa;
      if (lane < D.nb[0]) aV[lane * nsV + iwa] -= aw;    // w2: the water after the equilibration
      wsync();
      double azrh = 0.;
      if (lane == 0) {
        // regime 1 in reverse bin order, the Kelvin factor carried back
        double azke = 0.;
        for (int i1 = n1 - 1; i1 >= 0; --i1) {
          const int bb = A0 + D.in1a + i1;
          const int K = eqK[i1];
          if (K < 0) continue;
          double al = aV[bb * nsV + iwa] / kRhowa;       // water = zlwc / rhowa
          aV[bb * nsV + iwa] = 0.;
          if (K == 0) continue;
          const EqTask q = eq_task(D, sN, sV, nsV, bb, 1, zrh, isu, ioc);
          EqBar pb{0., 0., 0., 0., 0., 0.};
          double az = azke;
          for (int k = K - 1; k >= 0; --k) {
            az = eq_iter_rev(q, T, tEQ[i1 * kEqIt + k], al, az, pb);
            al = 0.;
          }
          azke = az;
          azrh += eq_task_rev(D, q, pb, nsV, bb, zrh, aN, aV);
        }
      } else if (eq2 >= 0) {
        const int bb = A0 + D.in2a + (lane - 1);
        const double al = aV[bb * nsV + iwa] / kRhowa;
        aV[bb * nsV + iwa] = 0.;
        if (eq2 > 0) {
          // fixed point z = H(z): zlwc = L(z) with dz = (dH/dp) dp / (1 - dH/dz)
          const EqTask q = eq_task(D, sN, sV, nsV, bb, 2, zrh, isu, ioc);
          EqBar pl{0., 0., 0., 0., 0., 0.}, ph{0., 0., 0., 0., 0., 0.};
          const double Lz = eq_iter_rev(q, T, eqz, 1., 0., pl);
          const double Hz = eq_iter_rev(q, T, eqz, 0., 1., ph);
          const double wz = al * Lz / (1. - Hz);
          // the parameters' adjoints are linear in (al, wz)
          const EqBar pb{al * pl.zr + wz * ph.zr, al * pl.c1 + wz * ph.c1, al * pl.c2 + wz * ph.c2,
                         al * pl.c5 + wz * ph.c5, al * pl.n + wz * ph.n, al * pl.zcore + wz * ph.zcore};
          azrh += eq_task_rev(D, q, pb, nsV, bb, zrh, aN, aV);
        }
      }
      wsync();
      if (lane < D.nb[0]) aV[lane * nsV + iwa] += aw;    // w1: the water before it
      arv_l = arv_e + cell_sum<LPC>(azrh) / rs;                   // zrh = rv_l / rs
      wsync();
    }

    if (coag) {
      // ---- clamp: the sweep's reverse starts from the state it left ----
      if (own) {
        if (!(cNr > 0.)) aN[lane] = 0.;
        sN[lane] = cNr;
        for (int s = 0; s < NS; ++s) if (s < nsV) {
          if (!(cVr[s] > 0.)) aV[lane * nsV + s] = 0.;
          sV[lane * nsV + s] = cVr[s];
        }
      }
      wsync();

      // ---- the sweep in reverse ----
      const double cf = D.charge_factor;
      for (int it = ntg - 1; it >= 0; --it) {
        const DevTarget &TG = D.tgt[it];
        const int t = TG.t;
        const uint32_t fl = TG.flags;
        if (!__shfl((int)tAct, lb + t)) continue;       // the target did not change anything
        const bool pft = fl & (TF_PF | TF_PF_SELF);
        const bool scoll = fl & TF_PRECIP_SELFCOLL;
        const int nupd = TG.nupd, ne = TG.nent;
        const uint32_t *ent = D.ent + TG.ent0;
        const double Nt = __shfl(tNr, lb + t);
        double Vt[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) Vt[s] = __shfl(tVr[s], lb + t);
        // restore the bins the target changed to their values before its update
        if (lane == 0) {
          sN[t] = Nt;
          for (int s = 0; s < NS; ++s) if (s < nsV) sV[t * nsV + s] = Vt[s];
        }
        if (scoll) {
          const double N1 = __shfl(t1Nr, lb + t);
          double V1[NS];
#pragma unroll
          for (int s = 0; s < NS; ++s) V1[s] = __shfl(t1Vr[s], lb + t);
          if (lane == 0) {
            sN[t + 1] = N1;
            for (int s = 0; s < NS; ++s) if (s < nsV) sV[(t + 1) * nsV + s] = V1[s];
          }
        }
        if (pft && lane < nb2) {
          const double *row = tPF + ((int64_t)it * nb2 + lane) * (1 + nsV);
          sN[R + lane] = row[0];
          for (int s = 0; s < NS; ++s) if (s < nsV) sV[(R + lane) * nsV + s] = row[1 + s];
        }
        // adjoints of the target's outputs
        const double aNn = aN[t];
        double aVn[NS], aN1 = 0., aV1[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          aVn[s] = s < nsV ? aV[t * nsV + s] : 0.;
          aV1[s] = 0.;
        }
        if (scoll) {
          aN1 = aN[t + 1];
#pragma unroll
          for (int s = 0; s < NS; ++s) aV1[s] = s < nsV ? aV[(t + 1) * nsV + s] : 0.;
        }
        wsync();
        if (lane == 0) {                                 // bin t is rewritten: its adjoint restarts
          aN[t] = 0.;
          for (int s = 0; s < NS; ++s) if (s < nsV) aV[t * nsV + s] = 0.;
        }
        // the entries at the restored state; the target's sums as the forward formed them
        const double *trow = tTG + it * 8;
        const double sink = trow[0], so = trow[4], sx = trow[5];
        double src[NS], Vj[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          src[s] = trow[1 + s];
          Vj[s] = 0.;
        }
        uint32_t role = 0;
        int j = 0, pid = 0, epb = -1;
        double kj = 0., Nj = 0.;
        if (lane < ne) {
          const uint32_t w = ent[lane];
          j = w & 63u;
          pid = (int)(w >> 16);
          role = (w >> 6) & 1023u;
          kj = sK[pid];
          Nj = sN[j];
#pragma unroll
          for (int s = 0; s < NS; ++s) Vj[s] = s < nsV ? sV[j * nsV + s] : 0.;
          if (role & (R_PF_OWN | R_PF_OTH)) epb = (int)(sPF[pid - pf0] & 31u) - 1;
        }
        double fix_own = 1., fix_oth = 1.;
        if (fl & TF_PF) {
          fix_own = fmax(1. - dt * so, 0.1);
          fix_oth = fmax(1. - dt * sx, 0.1);
        }
        double sink_self = 0., kself = 0., fixs = 1.;
        int pself = -1, pid_self = -1;
        if (fl & (TF_SELF_HALF | TF_PF_SELF | TF_PRECIP_SELFCOLL)) {
          pid_self = D.pidx[t][t];
          kself = pid_self >= 0 ? sK[pid_self] : 0.;
          if (fl & TF_SELF_HALF) sink_self = 0.5 * kself * Nt;
          if (fl & TF_PF_SELF) {
            fixs = fmax(1. - dt * (0.5 * kself * (Nt * Nt) / Nt), 0.1);
            pself = pid_self >= 0 ? (int)(sPF[pid_self - pf0] & 31u) - 1 : -1;
            sink_self = 0.5 * kself * Nt;
          }
        }
        const double rden = rcp_nr(1. + dt * sink);
        const double qn = rcp_nr(1. + dt * (sink + sink_self));

        // uniform adjoints (the same on every lane) and the lane's entry adjoints
        double aNt = 0., aVt[NS], akself = 0., afixs = 0., avol[NS];
        double aK = 0., aNj = 0., aVj[NS], aNt_l = 0., aVt_l[NS], afo_l = 0., afx_l = 0.;
#pragma unroll
        for (int s = 0; s < NS; ++s) { aVt[s] = 0.; avol[s] = 0.; aVj[s] = 0.; aVt_l[s] = 0.; }
        // transfer into the precipitation bins: N_r += dt npp, V_r += dt vp
        if (pft) {
          uint32_t pm = (uint32_t)trow[6];
          while (pm) {
            const int pb = __builtin_ctz(pm);
            pm &= pm - 1;
            const int r = R + pb;
            const double gn = dt * aN[r];
            double gv[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) gv[s] = s < nupd ? dt * aV[r * nsV + s] : 0.;
            if (pb == pself) {                           // npp += 0.5 kself Nt^2 fixs, vp += volsink
#pragma unroll
              for (int s = 0; s < NS; ++s) avol[s] += gv[s];
              akself += gn * 0.5 * Nt * Nt * fixs;
              aNt += gn * kself * Nt * fixs;
              afixs += gn * 0.5 * kself * Nt * Nt;
            }
            if (lane < ne && (role & (R_PF_OWN | R_PF_OTH)) && epb == pb) {
              // vp += (V_j Nt + Vt N_j) kj fx, npp += kj N_j Nt fx
              const double fx = (role & R_PF_OWN) ? fix_own : fix_oth;
              double g = gn * Nj * Nt, hj = gn * Nt, ht = gn * Nj;
#pragma unroll
              for (int s = 0; s < NS; ++s)
                if (s < nupd) {
                  g += gv[s] * (Vj[s] * Nt + Vt[s] * Nj);
                  hj += gv[s] * Vt[s];
                  ht += gv[s] * Vj[s];
                  aVj[s] += gv[s] * Nt * kj * fx;
                  aVt_l[s] += gv[s] * Nj * kj * fx;
                }
              aK += g * fx;
              aNj += hj * kj * fx;
              aNt_l += ht * kj * fx;
              if (role & R_PF_OWN) afo_l += g * kj;
              else afx_l += g * kj;
            }
          }
        }
        // precipitation self-collection: V(t+1) += dt Vt Nt kself, N(t+1) += dt 0.5 kself Nt^2,
        // subtracted from bin t
        if (scoll) {
          const double anslf = dt * aN1 - dt * aNn;
#pragma unroll
          for (int s = 0; s < NS; ++s)
            if (s < nupd) {
              const double avs = dt * aV1[s] - dt * aVn[s];
              aVt[s] += avs * Nt * kself;
              aNt += avs * Vt[s] * kself;
              akself += avs * Vt[s] * Nt;
            }
          akself += anslf * 0.5 * Nt * Nt;
          aNt += anslf * kself * Nt;
        }
        if (fl & TF_PF_SELF) {
#pragma unroll
          for (int s = 0; s < NS; ++s) avol[s] -= dt * aVn[s];          // Vn -= dt volsink
          if (pself >= 0) {                             // volsink = Vt Nt kself fixs
#pragma unroll
            for (int s = 0; s < NS; ++s)
              if (s < nupd) {
                aVt[s] += avol[s] * Nt * kself * fixs;
                aNt += avol[s] * Vt[s] * kself * fixs;
                akself += avol[s] * Vt[s] * Nt * fixs;
                afixs += avol[s] * Vt[s] * Nt * kself;
              }
          }
          if (1. - dt * (0.5 * kself * (Nt * Nt) / Nt) > 0.1) {
            akself += afixs * (-dt * 0.5 * Nt);
            aNt += afixs * (-dt * 0.5 * kself);
          }
        }
        // Vn = (Vt + dt src Nt) rden, Nn = Nt qn
        double arden = 0., asrc[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          asrc[s] = 0.;
          if (s < nsV) {
            if (s < nupd) {
              aVt[s] += aVn[s] * rden;
              asrc[s] = aVn[s] * dt * Nt * rden;
              aNt += aVn[s] * dt * src[s] * rden;
              arden += aVn[s] * (Vt[s] + dt * src[s] * Nt);
            } else {
              aVt[s] += aVn[s];
            }
          }
        }
        double asink = -arden * dt * rden * rden;
        aNt += aNn * qn;
        const double ass = -aNn * Nt * dt * qn * qn;    // adjoint of sink + sink_self
        asink += ass;
        if (fl & (TF_SELF_HALF | TF_PF_SELF)) {
          akself += ass * 0.5 * Nt;
          aNt += ass * 0.5 * kself;
        }
        double aso = 0., asx = 0.;
        if (fl & TF_PF) {
          const double afo = cell_sum<LPC>(afo_l), afx = cell_sum<LPC>(afx_l);
          if (1. - dt * so > 0.1) aso = -dt * afo;
          if (1. - dt * sx > 0.1) asx = -dt * afx;
        }
        // the sums, entry by entry
        double adwt = 0.;
        if (lane < ne) {
          if (role & R_SINK) { aK += asink * Nj; aNj += asink * kj; }
          if (role & R_SRC) {
#pragma unroll
            for (int s = 0; s < NS; ++s)
              if (s < nupd) { aK += asrc[s] * Vj[s]; aVj[s] += asrc[s] * kj; }
          }
          if (role & (R_PF_OWN | R_PF_OTH)) {
            if (epb >= 0) {
              aK += asink * Nj;
              aNj += asink * kj;
            } else {
#pragma unroll
              for (int s = 0; s < NS; ++s)
                if (s < nupd) { aK += asrc[s] * Vj[s]; aVj[s] += asrc[s] * kj; }
            }
          }
          if (role & R_FIX_OWN) { aK += aso * Nj; aNj += aso * kj; }
          if (role & R_FIX_OTH) { aK += asx * Nj; aNj += asx * kj; }
          // K of the entry's pair onto the two wet diameters, and onto E
          const int b1 = D.pw[pid] & 255u;
          adwt = aK * (b1 == t ? tP[pid] : tP[np + pid]);
          aE += aK * kj / cf;
        }
        wsync();                                 // bin t's adjoint is cleared
        if (lane < ne) {
          const int b1 = D.pw[pid] & 255u;
          aDW[j] += aK * (b1 == t ? tP[np + pid] : tP[pid]);   // one lane per bin j
          aN[j] += aNj;
          for (int s = 0; s < NS; ++s) if (s < nsV) aV[j * nsV + s] += aVj[s];
        }
        adwt = cell_sum<LPC>(adwt);
        aNt_l = cell_sum<LPC>(aNt_l);
#pragma unroll
        for (int s = 0; s < NS; ++s)
          if (s < nupd) aVt_l[s] = cell_sum<LPC>(aVt_l[s]);
        wsync();
        if (lane == 0) {
          if (pid_self >= 0) {                           // the self pair's kernel
            adwt += akself * (tP[pid_self] + tP[np + pid_self]);
            aE += akself * kself / cf;
          }
          aN[t] += aNt + aNt_l;
          for (int s = 0; s < NS; ++s) if (s < nsV) aV[t * nsV + s] += aVt[s] + aVt_l[s];
          aDW[t] += adwt;
        }
        wsync();
      }
      // ---- wet diameters (of the pair kernels) onto N and V ----
      if (own) {
        aN[lane] += aDW[lane] * dwn;
        for (int s = 0; s < NS; ++s) if (s < nsV) aV[lane * nsV + s] += aDW[lane] * dwv;
      }
      wsync();
    }

    // ---- the relative-humidity limit, and out ----
    if (rvlim) arv_l = 0.;
    aE = cell_sum<LPC>(aE);
    if (own) {
      lam[1 + lane] = aN[lane];
      if (!isfinite(aN[lane])) flags |= SALSA_CELL_GRAD_UNSTABLE;
      for (int s = 0; s < NS; ++s) if (s < nsV) {
        lam[1 + nt + lane * nsV + s] = aV[lane * nsV + s];
        if (!isfinite(aV[lane * nsV + s])) flags |= SALSA_CELL_GRAD_UNSTABLE;
      }
    }
    if (!isfinite(arv_l) || !isfinite(aE)) flags |= SALSA_CELL_GRAD_UNSTABLE;
    // Stability of the linearisation: a relative change of the state before
    // this step, or of E, amplified more than 1e8 times in the cell's J
    // (no perturbation resolvable in FP64 follows the derivative then)
    double amp = 0.;
    if (own) {
      amp = fabs(aN[lane] * x[1 + lane]);
      for (int s = 0; s < NS; ++s)
        if (s < nsV) amp += fabs(aV[lane * nsV + s] * x[1 + nt + lane * nsV + s]);
    }
    amp = cell_sum<LPC>(amp) + fabs(arv_l * x[0]);
    if (lane == 0) {
      lam[0] = arv_l;
      const double eb = a.ebar[cell] + aE;
      a.ebar[cell] = eb;
      const double Jc = a.jcell[cell];
      if (amp > 1.e8 * Jc || fabs(eb) * D.charge_factor > 1.e8 * Jc) flags |= SALSA_CELL_GRAD_UNSTABLE;
    }
    flags = cell_or<LPC>(flags);
    if (lane == 0 && flags) {
      if (a.cell_flags) a.cell_flags[cell] |= (uint8_t)flags;
      if (a.fault_any) atomicOr(a.fault_any, (int)flags);
    }
    wsync();
  }
}

// ---------------------------------------------------------------------------
// pack, seed, output and totals
// ---------------------------------------------------------------------------

// x0 from the cell records: N = numc dn, V = mass dn / rho (mo_salsa_driver.f90:236-279)
__global__ void __launch_bounds__(64) k_adj_pack(const DevLayout *__restrict__ Dp, const double *__restrict__ rec,
                                                 int nslot, int64_t ncells, int nvar, double *__restrict__ x) {
  const DevLayout &D = *Dp;
  const int nt = D.nt, nsV = D.nwet;
  for (int64_t cell = blockIdx.x; cell < ncells; cell += gridDim.x) {
    const double *r0 = rec + cell * nslot;
    const double dn = r0[4];
    double *xc = x + cell * nvar;
    const int b = threadIdx.x;
    if (b < nt) {
      const int c = D.cat_of[b];
      const double *r = r0 + D.sbase[b];
      xc[1 + b] = r[0] * dn;
      for (int s = 0; s < nsV; ++s) xc[1 + nt + b * nsV + s] = r[1 + s] * dn / species_rho(D, c, s);
    }
    if (b == 0) xc[0] = r0[5];
  }
}

// J per cell from the final state, the seed lam = dJ/dx_n, ebar = 0, flags = 0
__global__ void __launch_bounds__(64) k_adj_seed(const DevLayout *__restrict__ Dp, const double *__restrict__ x,
                                                 int64_t ncells, int nvar, double *__restrict__ lam,
                                                 double *__restrict__ ebar, double *__restrict__ jc,
                                                 uint8_t *__restrict__ flags) {
  const DevLayout &D = *Dp;
  const int nt = D.nt, nsV = D.nwet, R = D.off[2], nb2 = D.nb[2];
  for (int64_t cell = blockIdx.x; cell < ncells; cell += gridDim.x) {
    const double *xc = x + cell * nvar;
    double *lc = lam + cell * nvar;
    for (int i = threadIdx.x; i < nvar; i += 64) {
      const int k = i - 1 - nt;                        // volume slot b * nsV + s
      const int b = k >= 0 ? k / nsV : -1;
      lc[i] = (b >= R && b < R + nb2) ? 1. : 0.;
    }
    if (threadIdx.x == 0) {
      double J = 0.;
      for (int b = R; b < R + nb2; ++b)
        for (int s = 0; s < nsV; ++s) J += xc[1 + nt + b * nsV + s];
      jc[cell] = J;
      ebar[cell] = 0.;
    }
  }
}

// dJ/dN0 in LES number units (N = numc dn) scattered to the LES layout
__global__ void __launch_bounds__(64) k_adj_out(const DevLayout *__restrict__ Dp, AdjOut o) {
  const DevLayout &D = *Dp;
  const int nt = D.nt;
  for (int64_t cell = blockIdx.x; cell < o.ncells; cell += gridDim.x) {
    const double dn = o.rec[cell * o.nslot + 4];
    const int k = (int)(cell % o.nz);
    const int64_t r = cell / o.nz;
    const int i = (int)(r % o.nx), jj = (int)(r / o.nx);
    const int64_t li = (int64_t)(k + o.hz) + (int64_t)o.nzp * ((int64_t)(i + o.hxy) + (int64_t)o.nxp * (jj + o.hxy));
    const int b = threadIdx.x;
    if (b < nt) o.djdn[D.cat_of[b]][li + o.n3 * D.loc_of[b]] = o.lam[cell * o.nvar + 1 + b] * dn;
  }
}

// sum of J over every cell and of dJ/dE over the cells without an unstable
// linearisation, in a fixed order (one block)
__global__ void __launch_bounds__(1024) k_adj_sum(const double *__restrict__ x, const double *__restrict__ y,
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

__global__ void k_adj_zero_u8(uint8_t *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

int launch_adjoint_step(int precision, const DevLayout *d, const AdjArgs &a, unsigned nblk, hipStream_t s) {
  if (nblk == 0) return 0;
  const bool fp64 = precision == SALSA_FP64;
  if (a.mode == 1) {
    if (fp64) hipLaunchKernelGGL((k_salsa_adjoint<double, 32, 3, 528, true, 32>), dim3(nblk), dim3(64), 0, s, d, a);
    else hipLaunchKernelGGL((k_salsa_adjoint<float, 32, 3, 528, true, 32>), dim3(nblk), dim3(64), 0, s, d, a);
  } else {
    if (fp64) hipLaunchKernelGGL((k_salsa_adjoint<double, 32, 3, 528, false, 32>), dim3(nblk), dim3(64), 0, s, d, a);
    else hipLaunchKernelGGL((k_salsa_adjoint<float, 32, 3, 528, false, 32>), dim3(nblk), dim3(64), 0, s, d, a);
  }
  return (int)hipGetLastError();
}

int launch_adjoint_pack(const DevLayout *d, const double *rec, int nslot, int64_t ncells, int nvar, double *x,
                        uint8_t *flags, hipStream_t s) {
  const unsigned nb = (unsigned)(ncells < 65536 ? ncells : 65536);
  if (nb) {
    hipLaunchKernelGGL(k_adj_pack, dim3(nb), dim3(64), 0, s, d, rec, nslot, ncells, nvar, x);
    hipLaunchKernelGGL(k_adj_zero_u8, dim3(1024), dim3(256), 0, s, flags, ncells);
  }
  return (int)hipGetLastError();
}

int launch_adjoint_seed(const DevLayout *d, const double *x, int64_t ncells, int nvar, double *lam, double *ebar,
                        double *jc, hipStream_t s) {
  const unsigned nb = (unsigned)(ncells < 65536 ? ncells : 65536);
  if (nb) hipLaunchKernelGGL(k_adj_seed, dim3(nb), dim3(64), 0, s, d, x, ncells, nvar, lam, ebar, jc, nullptr);
  return (int)hipGetLastError();
}

int launch_adjoint_out(const DevLayout *d, const AdjOut &o, const double *jc, const double *ebar,
                       const uint8_t *flags, double *sum_out, hipStream_t s) {
  const unsigned nb = (unsigned)(o.ncells < 65536 ? o.ncells : 65536);
  if (nb) hipLaunchKernelGGL(k_adj_out, dim3(nb), dim3(64), 0, s, d, o);
  hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(1024), 0, s, jc, ebar, flags, o.ncells, sum_out);
  return (int)hipGetLastError();
}

}  // namespace salsa
