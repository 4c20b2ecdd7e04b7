// salsa_internal.h -- device tables and launch arguments shared by the host
// table builder (salsa_layout.cpp), the kernels (salsa_kernels.hip) and the
// C ABI (salsa_capi.cpp).
#pragma once

#include <stdint.h>

#include "salsa_amd.h"

namespace salsa {

constexpr int kMaxBin = 64;           // master bins (aerosol|cloud|precip|ice)
constexpr int kMaxSpec = 9;           // Section%volc(9)
constexpr int kThermo = 6;            // T, p, rs, rsi, dn, rv at the head of a cell record
constexpr int kRateLanes = 16;        // lanes per cell of the step variant that writes rate diagnostics
constexpr int kMaxSlot = kThermo + kMaxBin * (1 + kMaxSpec);

// Physical constants exactly as written in src/src_shared/mo_submctl.f90:253-338.
constexpr double kAvog = 6.0221e+23, kBoltz = 1.3807e-23, kGrav = 9.81;
constexpr double kPstand = 1.01325e+5, kRg = 8.314, kPi = 3.1415927;
constexpr double kPi6 = 0.5235988, kMair = 28.97e-3, kRd = 287.04;
constexpr double kAlv = 2.5e6, kAls = 2.834e6, kSurfw0 = 0.073, kSurfi0 = 0.105;
constexpr double kEpsoc = 0.15, kNlim = 1.0, kPrlim = 1.e-6;
constexpr double kRhowa = 1000., kRhoic = 917., kRhori = 400., kMwa = 18.016e-3;
constexpr double kRhosu = 1830., kMsu = 98.08e-3, kRhooc = 1400., kMoc = 150.e-3;
constexpr double kRhoss = 2165., kMss = 58.44e-3;
constexpr double kEpsR8 = 2.220446049250313080847e-16;   // epsilon(1.0) for real(8)
constexpr double kMvwa = (18.016e-3 / 6.0221e+23) / 1000.;

// Target flags of the coagulation sweep program.
enum TargetFlags : uint32_t {
  TF_SELF_HALF = 1u << 0,        // sink_self += 0.5*K(t,t)*N_t   (accumulateSink multp=0.5)
  TF_SELF_HALF_EAGG = 1u << 1,   // same, ice-ice with aggregation efficiency
  TF_PRECIP_SELFCOLL = 1u << 2,  // precipSelfCoag: products moved to bin t+1
  TF_PF_SELF = 1u << 3,          // accumulatePrecipFormation self collection
  TF_PF = 1u << 4,               // pf_own / pf_oth masks are live
  TF_ICE = 1u << 5               // applyCoagIce update (dry, rime, water)
};

// Roles of a collector/source bin j in the sweep entry of a target t.
enum EntryRole : uint32_t {
  R_SINK = 1u << 0,       // sink += K(t,j) N_j
  R_SINK_EAGG = 1u << 1,  // sink += Eagg K N_j
  R_SRC = 1u << 2,        // src[s] += K V_j[s], s < nupd
  R_SRC_EAGG = 1u << 3,   // src[s] += Eagg K V_j[s], all species
  R_PC_WA = 1u << 4,      // phase change into the ice water slot
  R_PC_RI = 1u << 5,      // phase change into the rime slot
  R_PF_OWN = 1u << 6,     // precipitation formation, own cloud series
  R_PF_OTH = 1u << 7,     // ... other series
  R_FIX_OWN = 1u << 8,    // fix_coag sums
  R_FIX_OTH = 1u << 9
};

// One target bin of the sequential coagulation sweep; the sweep visits the
// targets in the reference order ice -> precip -> aerosol -> cloud
// (mo_salsa_dynamics.f90:166-183), ascending bins within each category.
// Masks select collector / source bins by master index.
struct DevTarget {
  int32_t t;             // master bin index
  int32_t cat;           // 0 aero, 1 cloud, 2 precip, 3 ice
  uint32_t flags;
  int32_t nupd;          // species updated by applyCoag (nwet; ice handled by TF_ICE)
  uint64_t sink;         // sink += K(t,j) N_j
  uint64_t sink_eagg;    // sink += Eagg(t,j) K(t,j) N_j   (ice-ice)
  uint64_t src;          // src[s] += K(j,t) V_j[s],  s < nupd
  uint64_t src_eagg;     // src[s] += Eagg K V_j[s],  all species (ice-ice)
  uint64_t src_pc_wa;    // phase change into ice water slot (aerosol collected by ice)
  uint64_t src_pc_ri;    // phase change into rime slot (cloud/precip collected by ice)
  uint64_t pf_own, fix_own;   // precip formation from the own cloud series, fix_coag range
  uint64_t pf_oth, fix_oth;   // ... from the other series
  // the bins named by the masks above as a list of entries
  // j | role << 6 | pair id << 16, aerosol bins first
  int32_t ent0, nent, nent_aero;
};

struct DevLayout {
  int32_t nt;                         // bins handled (ice dropped at level 4)
  int32_t nb[4], off[4];              // bins and master offsets per category
  int32_t in1a, fn1a, in2a, fn2a, in2b, fn2b;
  int32_t nspec, nwet, ndry, iwa, irim;
  int32_t ncat;                       // 3 at level 4, 4 at level 5
  double rholiq[kMaxSpec], rhoice[kMaxSpec], mm[kMaxSpec], diss[kMaxSpec];
  int32_t dry_id[kMaxSpec];           // species id (0 SO4 .. 6 NH) per dry slot
  int32_t phase[kMaxBin];
  int32_t cat_of[kMaxBin], loc_of[kMaxBin];
  // Cell records.  The gather kernel copies each cell's LES values into one
  // contiguous record of nslot doubles: T, p, rs, rsi, dn, rv, then per bin
  // (master order) its number and its nsw(c) masses.  The step kernel writes
  // nout doubles per cell (rv, then the same per-bin layout) that the scatter
  // kernel adds to the tendencies (or stores as the new state).
  int32_t nslot, nout;
  int32_t sbase[kMaxBin];             // slot of bin b's number in the input record
  int16_t slot_cat[kMaxSlot];         // category of a bin slot (-1: thermodynamic slot)
  int32_t slot_row[kMaxSlot];         // row of the slot inside its category array (units of n3)
  uint8_t slot_mass[kMaxSlot];        // 1: mass slot, 0: number slot
  double dmid[kMaxBin], nlim[kMaxBin], dlim[kMaxBin];
  double vlo[kMaxBin], vhi[kMaxBin];  // bin volume limits (distr_update)
  int32_t cld_a0, cld_a1, cld_b0, cld_b1;   // first/last cloud bin of series a and b (cloud index)
  double precpbins[kMaxBin];
  // pairs: canonical orientation (b1 is "pp1") and the self-offset predicate
  int32_t npair;
  int16_t pb1[kMaxBin * (kMaxBin + 1) / 2], pb2[kMaxBin * (kMaxBin + 1) / 2];
  uint8_t poff[kMaxBin * (kMaxBin + 1) / 2];
  int16_t pidx[kMaxBin][kMaxBin];     // pair id of (i,j), -1 if that block is off
  uint32_t pw[kMaxBin * (kMaxBin + 1) / 2];   // packed pair word: b1 | b2 << 8 | offset << 16
  int32_t npair_aa;                   // aerosol-aerosol pairs (the first block)
  int32_t pf0, ncc;                   // first cloud-cloud pair id and their count
  // sweep program
  int32_t ntarget;
  int32_t it_aero0, it_aero1;         // sweep entries of the aerosol targets [it_aero0, it_aero1)
  DevTarget tgt[kMaxBin];
  uint32_t ent[kMaxBin * kMaxBin];    // sweep entries of all targets
  int32_t nent_total;
  // switches and constants
  int32_t level, lscoag, lscnd, lscndh2oae, lscndh2ocl, lscndh2oic;
  int32_t lsdistupdate;
  int32_t autoconv2;                  // lsauto%state .AND. lsauto%mode == 2 (mo_salsa.f90:110-112)
  int32_t lsactiv;                    // cloud activation (mo_salsa.f90:115-119): 1 interstitial, 2 cloud base
  int32_t any_ice_coag;               // any of lscgii/lscgia/lscgic/lscgip
  double eddy_dis, Eiagg_max, Eiagg_min;
  double ice_alpha, ice_beta, ice_gamma, ice_sigma;
  double charge_factor;               // 1 + chargeCollEnh (lcharge) or 1
  double eddy_075;                    // eddy_dis**0.75 (coagc :391)
};

// LDS geometry of one launch (built on the host for the layout, the kernel
// variant and the cells per wavefront).  The block region starts with the
// pair-id table [ntab][ntab] (int16), followed by one region per cell:
//   N[nt], V[nt][nsV], dw[nt] (double), particle terms [2][ntab] (4 x KT),
//   K[npairs] (KT), precipitation-formation bin per cloud-cloud pair (int8).
struct SmemGeom {
  int32_t tab_bytes, cell_bytes, total_bytes;
  int32_t oN, oV, odw, opr, oK, opf;
  int32_t ntab;        // bins with particle terms and kernels
  int32_t npairs;      // pairs evaluated (the aerosol block only without hydrometeors)
  int32_t nsV;         // volume slots per bin
  int32_t it0, it1;    // sweep entries visited
  uint64_t jmask;      // collector bins that can be active
  int32_t orr;         // byte offset of the rate accumulators, -1 without rates
  int32_t ovd;         // byte offset of the fall-speed residuals [2][ntab] (mixed path), -1 in FP64
};

// Per-launch arguments (passed by value).
struct StepArgs {
  int32_t nzp, nxp, nyp, hz, hxy, nz, nx, ny;
  int64_t n3;                         // nzp*nxp*nyp
  int64_t ncells;                     // cells in this launch
  const int32_t *cell_list;           // interior ids (k fastest), or null = identity
  const double *press, *temp, *rsl, *rsi, *dn, *rv_p;
  double *rv_t;
  const double *n_p[4];
  double *n_t[4];
  const double *m_p[4];
  double *m_t[4];
  double *n_o[4];                     // in-place outputs (== n_p cast) or null
  double *m_o[4];
  double *rv_o;
  double *rates;                      // process-rate fields [nrate][n3], or null
  void *kstore;                       // lscoag mode 2: kernels per cell [ncells_all][npair] (KT)
  int32_t kmode;                      // 0: compute; 1: compute and store; 2: fetch from kstore
  const double *rec_in;               // cell records [ncells_all][nslot]
  double *rec_out;                    // step results [ncells_all][nout]
  double dt;
  int32_t nsub;                       // condensation substeps for dt
  int32_t inplace;
  int32_t init;                       // initialization call: equilibration at RH 0.3 first
  int32_t lsfreeRH;
  double rhlim;
  uint8_t *cell_flags;                // one byte per interior cell, or null
  int32_t *fault_any;                 // OR of all cell flags
  const double *w;                    // cloud-base activation: updraft [n3],
  double *nactd, *vactd;              //   the activated droplets [ncld][n3], [nwet*ncld][n3],
  double *act_state;                  //   the aerosol state it sees [ncells][nbins*(1+nwet)], or null
};

// Arguments of the forward-mode derivative kernel: the trajectory starts from
// the cell records in s.rec_in and runs nsteps in-place steps of s.dt.
struct GradArgs {
  StepArgs s;
  int32_t nsteps;
  int32_t lanes;                      // lanes per cell, one per direction: bins, then E
  double *djdn[4];                    // dJ/dN0 per category, LES layout of the number fields
  double *j_cell, *djde_cell;         // per interior cell: J and dJ/dE
};

// Arguments of the adjoint kernel (salsa_adjoint_device).  The state of a
// cell is kept in SALSA units as x[cell][nvar] = rv, N[nt], V[nt][nsV].
// mode 0: one step x_in -> x_out (the forward pass and the recomputation
// between checkpoints); mode 1: the step's forward internals are recomputed
// from x_in, then its reverse turns lam (adjoint of the state after the
// step) into the adjoint of the state before it, in place, and adds the
// step's share of dJ/dE to ebar.
struct AdjArgs {
  int64_t ncells;
  const double *rec;                  // cell records (T, p, rs, rsi, dn at slots 0..4)
  int32_t nslot, nvar, nsub, mode;
  const double *x_in;
  double *x_out;
  double *lam;
  double *ebar;
  const double *jcell;                // J per cell (mode 1: the amplification bound of the flag)
  // per-block tape of mode 1, tape_stride doubles per block: the condensation
  // substeps [nsub][nt + 2] at 0, pair derivatives [2][npair] at t_p, the
  // precipitation bins before each transfer [ntarget][nprc][1 + nwet] at
  // t_pf, regime-1 equilibration iterates [4][104] at t_eq, each sweep
  // target's sums [ntarget][8] at t_tg, the bins' particle terms with their
  // diameter derivatives [64][8] at t_gt; one tape per cell slot (two cells
  // per block)
  double *tape;
  int64_t tape_stride, t_p, t_pf, t_eq, t_tg, t_gt;
  double dt, rhlim;
  int32_t lsfreeRH;
  uint8_t *cell_flags;                // OR-ed with the step's fault bits
  int32_t *fault_any;
};

// The adjoint's outputs: dJ/dN0 scattered to the LES number fields.
struct AdjOut {
  int64_t ncells, n3;
  int32_t nz, nx, nzp, nxp, hz, hxy, nslot, nvar;
  const double *rec, *lam;
  double *djdn[4];
};

// Arguments of the LES coupling passes after a step (salsa_les_update_device):
// tend_constrain2, update_sclrs, tend0 and SALSA_diagnostics on the LES
// arrays n(c)[b][n3], m(c)[s*nb+b][n3] of a grid with halo 1 in z and 2
// laterally, owned whole by this process (periodic lateral halos).
struct LesArgs {
  int32_t nzp, nxp, nyp, level;
  int64_t n3;
  double dtlt, dzf2;
  int32_t nb[4], nsw[4];              // bins and mass slots per category (ice: + rime)
  int32_t ndry, nwet, iwa, irim;
  int32_t in1a, fn2b, in2a, fn2a;
  int32_t ica_cur, ica_par, fca_cur, icb_cur, icb_par, fcb_cur;
  double rholiq[kMaxSpec], rhoice[kMaxSpec], mm[kMaxSpec], diss[kMaxSpec];
  double aerobins[kMaxBin];           // lower dry-diameter limit of each aerosol bin
  const double *temp, *rh, *rhi, *rp;
  double *n_p[4], *n_t[4], *m_p[4], *m_t[4];
  double *charge_time;                // [nbins+ncld+nprc][n3] or null
  double *diag;                       // [6][n3]: rc, srp, snrp, rtot, ri, riri
  int32_t *list, *count;              // work: [n3] cells the removals visit, [1] their number
  double *gas;                        // [5][n3] a_gaerop (tiny aerosols' SO4, NO, NH, OC go there) or null
  int32_t gas_row[kMaxSpec];          // per dry species slot: its gas row, -1 for none
  double *indef;                      // [nindef][n3] a_indefp (clamped to [0, 1] at level 5) or null
  int32_t nindef;
};

// Arguments of the sedimentation passes (salsa_sedim_device): sedim_SALSA
// with DepositionSlow (aerosol, cloud) and DepositionFast (precipitation;
// ice at level 5) on the LES arrays.
struct SedArgs {
  int32_t nzp, nxp, nyp, level;
  int64_t n3;
  double dtlt;
  int32_t nb[4];                      // aerosol, cloud, precipitation, ice bins
  int32_t ns, ns_ice, iwa, irim;      // mass rows per bin (liquid: wet species; ice: + rime); water, rime
  double rholiq[kMaxSpec], rhoice[kMaxSpec];
  double ice_alpha, ice_beta, ice_gamma, ice_sigma;   // mo_ice_shape.f90:25-30
  double dzt[256];                    // inverse level thicknesses (nzp <= 256)
  const double *tk, *th, *adn, *ustar;
  const double *n_p[4], *m_p[4];
  double *n_t[4], *m_t[4];
  double *tlt, *rrate, *sfcrrate, *vt_prc;
  double *irate, *sfcirate, *vt_ice;  // level 5
};

}  // namespace salsa
