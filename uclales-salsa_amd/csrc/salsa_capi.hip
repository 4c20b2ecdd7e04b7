// salsa_capi.hip -- the extern "C" boundary (include/salsa_amd.h): context
// lifecycle, host- and device-pointer step entry points, cost-class
// compaction of cells, the gradient entry point and the launch-shape
// autotuner.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "salsa_internal.h"

namespace salsa {
extern thread_local char g_err[512];
void set_error(const char *msg);
int build_dev_layout(const salsa_bin_desc &d, const salsa_params &p, DevLayout &L);
int launch_step(int cap, int lanes, int precision, bool hyd, const DevLayout *d, const StepArgs &a,
                const SmemGeom &G, hipStream_t s);
int make_geom(const DevLayout &Lay, bool hyd, int kt_bytes, int lanes, bool rates, SmemGeom &G);
int launch_gather(const DevLayout &Lay, const DevLayout *d, const StepArgs &a, double *rec, uint8_t *cls,
                  int32_t *tcount, int32_t *toff, int32_t *lists, int32_t *counts, hipStream_t s);
int64_t gather_tiles(int64_t ncells);
int64_t class_work_ints(int64_t ncells);
int launch_scatter(const DevLayout &Lay, const DevLayout *d, const StepArgs &a, hipStream_t s);
int launch_act_cloud_base(const DevLayout *d, const StepArgs &a, hipStream_t s);
int cap_bins(int cap);
int cap_spec(int cap);
int launch_gradient(int cap, int precision, const DevLayout *d, const GradArgs &g, double *sum_out,
                    hipStream_t s);
int launch_adjoint_step(int precision, const DevLayout *d, const AdjArgs &a, unsigned nblk, hipStream_t s);
int launch_adjoint_pack(const DevLayout *d, const double *rec, int nslot, int64_t ncells, int nvar, double *x,
                        uint8_t *flags, hipStream_t s);
int launch_adjoint_seed(const DevLayout *d, const double *x, int64_t ncells, int nvar, double *lam, double *ebar,
                        double *jc, hipStream_t s);
int launch_adjoint_out(const DevLayout *d, const AdjOut &o, const double *jc, const double *ebar,
                       const uint8_t *flags, double *sum_out, hipStream_t s);
int launch_les_update(const LesArgs &a, hipStream_t s);
int launch_sedim(const SedArgs &a, int sed_aero, int sed_cloud, int sed_precp, int sed_ice, hipStream_t s);
}  // namespace salsa

using namespace salsa;

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      std::snprintf(g_err, sizeof(g_err), "%s failed: %s", #x, hipGetErrorString(e_));   \
      return SALSA_ERR_HIP;                                                              \
    }                                                                                    \
  } while (0)

struct salsa_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  salsa_bin_desc desc;
  salsa_params params;
  DevLayout hostL;
  DevLayout *devL = nullptr;
  int cap = 4;
  int lanes[3] = {32, 32, 32};
  int32_t nzp = 0, nxp = 0, nyp = 0, nz = 0, nx = 0, ny = 0;
  int64_t n3 = 0, ncells = 0;
  // cell records: LES values gathered per cell, step results per cell
  double *rec_in = nullptr, *rec_out = nullptr;
  // lscoag mode 2: collision kernels of every cell between update steps
  void *kstore = nullptr;
  // per-cell fault flags and the global OR
  uint8_t *flags = nullptr;
  int32_t *fault = nullptr;
  // compaction: class per cell, per-tile class counts and offsets, one list
  // of cell ids per cost class ([3][ncells]), class counts
  uint8_t *cls = nullptr;
  int32_t *tcount = nullptr, *toff = nullptr;
  int32_t *lists = nullptr;
  int32_t *cls_count = nullptr;       // device [3]
  int32_t *cls_count_h = nullptr;     // pinned host [3]
  // salsa_gradient_device: per-cell J and dJ/dE, then the two sums
  double *gbuf = nullptr;
  // salsa_adjoint_device: checkpoints, recomputed states, work states, the
  // adjoint, per-cell J and dJ/dE, totals and the per-block tape
  double *adj_buf = nullptr;
  size_t adj_bytes = 0;
  // cloud-base activation (lsactiv 2): the updraft and the activated
  // droplets, device arrays of the caller (salsa_set_cloud_base_fields)
  const double *act_w = nullptr;
  double *act_n = nullptr, *act_v = nullptr;
  double *act_state = nullptr;        // [ncells][nbins*(1+nwet)]: the aerosol state it sees
  // salsa_les_update_device: the cells the removals visit ([n3]) and their number
  int32_t *les_list = nullptr;
  // its optional fields (salsa_set_les_extra_fields): a_gaerop [5][n3], a_indefp [les_nindef][n3]
  double *les_gas = nullptr, *les_indef = nullptr;
  int32_t les_nindef = 0;
  // staging for salsa_step_host
  double *stage = nullptr;
  size_t stage_bytes = 0;
  // kernel timing: per timed step kEvPerStep events around the launches
  // (gather, the three class kernels, scatter) and the class cell counts
  bool timing = false;
  std::vector<hipEvent_t> ev;
  std::vector<int64_t> ev_cells;      // [3] per timed step
  // this device's copy of the static work space of the allocation-free entry
  double *scratch = nullptr;
  double *tot_h = nullptr;            // pinned host [2]: gradient totals
  int32_t *fault_h = nullptr;         // pinned host [1]: the OR of the cell flags
};
constexpr int kEvPerStep = 10;

// ---------------------------------------------------------------------------
// Work space of salsa_gradient_device_noalloc: per-cell J, per-cell dJ/dE and
// the two totals for up to SALSA_NOALLOC_MAX_CELLS cells, 16 MiB per GPU,
// held until the process ends.  Host code never uses the array's name as an
// address: its device address is resolved once per device (hipGetSymbolAddress
// at context creation, with that device current) and kept in g_scratch_dev.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[2 * (int64_t)SALSA_NOALLOC_MAX_CELLS + 2];
static std::mutex g_scratch_mu;
static double *g_scratch_dev[64];

static int resolve_scratch(int device, double **out) {
  if (device < 0 || device >= 64) { set_error("device index out of range"); return SALSA_ERR_ARG; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      std::snprintf(g_err, sizeof(g_err), "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return SALSA_ERR_HIP;
    }
    g_scratch_dev[device] = static_cast<double *>(ptr);
  }
  *out = g_scratch_dev[device];
  return SALSA_OK;
}

// ---------------------------------------------------------------------------
// small kernels: fault reset, cost classes
// ---------------------------------------------------------------------------
__global__ void k_zero_i32(int32_t *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
static int substeps(double dt) {
  // DO WHILE (ttot < ptstep) with adt = 0.02 accumulated in FP64 (mo_salsa_dynamics.f90:988-1053)
  double ttot = 0.;
  int n = 0;
  while (ttot < dt && n < 100000) { ttot = ttot + 2.e-2; ++n; }
  return n;
}

static int pick_cap(const DevLayout &L) {
  for (int cap : {4, 5})
    if (L.nt <= cap_bins(cap) && L.nspec <= cap_spec(cap)) return cap;
  return 0;
}

static void release(salsa_ctx *c) {
  if (!c) return;
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  c->ev.clear();
  (void)hipFree(c->devL); (void)hipFree(c->flags); (void)hipFree(c->fault);
  (void)hipFree(c->rec_in); (void)hipFree(c->rec_out); (void)hipFree(c->kstore);
  (void)hipFree(c->cls); (void)hipFree(c->tcount); (void)hipFree(c->toff);
  (void)hipFree(c->lists); (void)hipFree(c->cls_count);
  (void)hipFree(c->stage);
  (void)hipFree(c->gbuf);
  (void)hipFree(c->adj_buf);
  (void)hipFree(c->les_list);
  (void)hipFree(c->act_state);
  if (c->cls_count_h) (void)hipHostFree(c->cls_count_h);
  if (c->tot_h) (void)hipHostFree(c->tot_h);
  if (c->fault_h) (void)hipHostFree(c->fault_h);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

extern "C" const char *salsa_last_error(void) { return g_err; }

extern "C" int salsa_ctx_create(const salsa_bin_desc *desc, const salsa_params *params, int32_t nzp,
                                int32_t nxp, int32_t nyp, salsa_ctx **out) {
  if (!desc || !params || !out) { set_error("null argument"); return SALSA_ERR_ARG; }
  *out = nullptr;
  const int hz = params->hz, hxy = params->hxy;
  if (hz < 0 || hxy < 0 || nzp - 2 * hz < 1 || nxp - 2 * hxy < 1 || nyp - 2 * hxy < 1) {
    set_error("grid has no interior cells for the given halo widths");
    return SALSA_ERR_ARG;
  }
  if (params->precision != SALSA_FP64 && params->precision != SALSA_FP32_MIXED) {
    set_error("precision must be SALSA_FP64 or SALSA_FP32_MIXED");
    return SALSA_ERR_ARG;
  }
  if (params->lscoag_mode != 1 && params->lscoag_mode != 2) {
    set_error("lscoag_mode must be 1 or 2");
    return SALSA_ERR_ARG;
  }
  if (params->lsactiv < 0 || params->lsactiv > 2) {
    set_error("lsactiv must be 0, 1 (interstitial activation) or 2 (cloud-base activation)");
    return SALSA_ERR_ARG;
  }
  if (params->lscoag_mode == 2 && !(params->cgintvl > 0.)) {
    set_error("cgintvl must be positive with lscoag_mode 2");
    return SALSA_ERR_ARG;
  }
  salsa_ctx *c = new (std::nothrow) salsa_ctx();
  if (!c) { set_error("out of host memory"); return SALSA_ERR_NOMEM; }
  c->desc = *desc;
  c->params = *params;
  int rc = build_dev_layout(*desc, *params, c->hostL);
  if (rc) { delete c; return rc; }
  c->cap = pick_cap(c->hostL);
  if (!c->cap) { set_error("layout exceeds the compiled bin/pair capacity"); delete c; return SALSA_ERR_ARG; }
  c->nzp = nzp; c->nxp = nxp; c->nyp = nyp;
  c->nz = nzp - 2 * hz; c->nx = nxp - 2 * hxy; c->ny = nyp - 2 * hxy;
  c->n3 = (int64_t)nzp * nxp * nyp;
  c->ncells = (int64_t)c->nz * c->nx * c->ny;
  if (c->ncells > INT32_MAX) { set_error("too many cells"); delete c; return SALSA_ERR_ARG; }
  // default shapes: measured fastest on MI355X for the warm layout (see salsa_autotune)
  c->lanes[0] = 32; c->lanes[1] = 32; c->lanes[2] = 32;
  auto fail = [&](int code) { release(c); delete c; return code; };
  if (hipGetDevice(&c->device) != hipSuccess) { set_error("no HIP device"); return fail(SALSA_ERR_HIP); }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed"); return fail(SALSA_ERR_HIP);
  }
  const size_t nc = (size_t)c->ncells;
  if (hipMalloc(&c->devL, sizeof(DevLayout)) != hipSuccess ||
      hipMalloc(&c->flags, nc) != hipSuccess || hipMalloc(&c->fault, sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->lists, 3 * nc * sizeof(int32_t)) != hipSuccess || hipMalloc(&c->cls, nc) != hipSuccess ||
      hipMalloc(&c->tcount, 3 * (size_t)gather_tiles(c->ncells) * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->toff, (size_t)class_work_ints(c->ncells) * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->cls_count, 3 * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&c->rec_in, nc * c->hostL.nslot * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->rec_out, nc * c->hostL.nout * sizeof(double)) != hipSuccess ||
      hipHostMalloc(&c->cls_count_h, 3 * sizeof(int32_t)) != hipSuccess ||
      hipHostMalloc(&c->tot_h, 2 * sizeof(double)) != hipSuccess ||
      hipHostMalloc(&c->fault_h, sizeof(int32_t)) != hipSuccess) {
    set_error("device allocation failed");
    return fail(SALSA_ERR_NOMEM);
  }
  // no step has run yet: no cell flags, nothing for salsa_ctx_sync to report
  if (hipMemset(c->flags, 0, nc) != hipSuccess || hipMemset(c->fault, 0, sizeof(int32_t)) != hipSuccess) {
    set_error("flag initialisation failed");
    return fail(SALSA_ERR_HIP);
  }
  {
    const int rs = resolve_scratch(c->device, &c->scratch);
    if (rs) return fail(rs);
  }
  if (c->params.lscoag_mode == 2 && c->params.lscoag) {
    const size_t kt = c->params.precision == SALSA_FP64 ? 8 : 4;
    if (hipMalloc(&c->kstore, nc * (size_t)c->hostL.npair * kt) != hipSuccess) {
      set_error("kernel store allocation failed");
      return fail(SALSA_ERR_NOMEM);
    }
    // zero-filled like initialize_coagstorage (mo_salsa_driver.f90:601-648)
    if (hipMemset(c->kstore, 0, nc * (size_t)c->hostL.npair * kt) != hipSuccess) {
      set_error("kernel store initialisation failed");
      return fail(SALSA_ERR_HIP);
    }
  }
  if (hipMemcpy(c->devL, &c->hostL, sizeof(DevLayout), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("copy of device tables failed"); return fail(SALSA_ERR_HIP);
  }
  *out = c;
  return SALSA_OK;
}

extern "C" void salsa_ctx_destroy(salsa_ctx *c) {
  if (!c) return;
  hipStreamSynchronize(c->stream);
  release(c);
  delete c;
}

extern "C" int64_t salsa_ctx_ncells(const salsa_ctx *c) { return c ? c->ncells : -1; }
extern "C" int32_t salsa_ctx_ntotal(const salsa_ctx *c) { return c ? c->hostL.nt : -1; }
extern "C" void *salsa_ctx_stream(salsa_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int salsa_set_launch_shape(salsa_ctx *c, int32_t cls, int32_t lanes) {
  if (!c || cls < 0 || cls > 2) { set_error("bad class"); return SALSA_ERR_ARG; }
  if (lanes == 0) return SALSA_OK;
  if (lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64) {
    set_error("lanes_per_cell must be 8, 16, 32 or 64");
    return SALSA_ERR_ARG;
  }
  SmemGeom G;
  if (make_geom(c->hostL, cls == 2, c->params.precision == SALSA_FP64 ? 8 : 4, lanes, false, G)) {
    set_error("this lanes_per_cell needs more than 64 KiB of LDS per block for the layout");
    return SALSA_ERR_ARG;
  }
  c->lanes[cls] = lanes;
  return SALSA_OK;
}

extern "C" int salsa_get_launch_shape(salsa_ctx *c, int32_t cls) {
  if (!c || cls < 0 || cls > 2) return SALSA_ERR_ARG;
  return c->lanes[cls];
}

static StepArgs make_args(salsa_ctx *c, const double *press, const double *temp, const double *rsl,
                          const double *rsi, const double *dn, const double *rv_p, double *rv_t,
                          const double *const n_p[4], double *const n_t[4], const double *const m_p[4],
                          double *const m_t[4], double *rates, double tstep, int32_t flags,
                          double time = 0., int32_t istp = 1) {
  StepArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nzp = c->nzp; a.nxp = c->nxp; a.nyp = c->nyp; a.hz = c->params.hz; a.hxy = c->params.hxy;
  a.nz = c->nz; a.nx = c->nx; a.ny = c->ny; a.n3 = c->n3; a.ncells = c->ncells;
  a.press = press; a.temp = temp; a.rsl = rsl; a.rsi = rsi; a.dn = dn; a.rv_p = rv_p; a.rv_t = rv_t;
  const int ncat = c->hostL.ncat;
  for (int k = 0; k < 4; ++k) {
    a.n_p[k] = k < ncat ? n_p[k] : nullptr;
    a.m_p[k] = k < ncat ? m_p[k] : nullptr;
    a.n_t[k] = (k < ncat && n_t) ? n_t[k] : nullptr;
    a.m_t[k] = (k < ncat && m_t) ? m_t[k] : nullptr;
    a.n_o[k] = const_cast<double *>(a.n_p[k]);
    a.m_o[k] = const_cast<double *>(a.m_p[k]);
  }
  a.rv_o = const_cast<double *>(rv_p);
  a.dt = tstep;
  a.nsub = substeps(tstep);
  a.inplace = (flags & SALSA_STEP_INPLACE) ? 1 : 0;
  a.init = (flags & SALSA_STEP_INIT) ? 1 : 0;
  a.lsfreeRH = c->params.lsfreeRH;
  a.rhlim = c->params.rhlim;
  a.rates = rates;
  a.kstore = c->kstore;
  a.kmode = 0;
  if (c->kstore) {     // set_SALSA_runtime, mo_salsa_driver.f90:469-472
    const bool upd = std::fmod(time, c->params.cgintvl) < tstep || istp <= 1;
    a.kmode = upd ? 1 : 2;
  }
  a.rec_in = c->rec_in;
  a.rec_out = c->rec_out;
  a.cell_flags = c->flags;
  a.fault_any = c->fault;
  if (c->hostL.lsactiv == 2 && c->act_n) {
    a.w = c->act_w;
    a.nactd = c->act_n;
    a.vactd = c->act_v;
    a.act_state = c->act_state;
  }
  return a;
}

static int check_ptrs(salsa_ctx *c, const double *press, const double *temp, const double *rsl,
                      const double *rsi, const double *dn, const double *rv_p, double *rv_t,
                      const double *const n_p[4], double *const n_t[4], const double *const m_p[4],
                      double *const m_t[4], double *rates, double tstep, int32_t flags) {
  if (!c) { set_error("null context"); return SALSA_ERR_ARG; }
  if (!(tstep > 0.)) { set_error("tstep must be positive"); return SALSA_ERR_ARG; }
  if (!press || !temp || !rsl || !dn || !rv_p || !n_p || !m_p) { set_error("null field pointer"); return SALSA_ERR_ARG; }
  if (c->hostL.level == 5 && !rsi) { set_error("rsi required at level 5"); return SALSA_ERR_ARG; }
  if (c->hostL.lsactiv == 2 && (!c->act_w || !c->act_n || !c->act_v)) {
    set_error("cloud-base activation (lsactiv 2) needs w, nactd and vactd (salsa_set_cloud_base_fields)");
    return SALSA_ERR_ARG;
  }
  const bool inplace = flags & SALSA_STEP_INPLACE;
  if (!inplace && (!rv_t || !n_t || !m_t)) { set_error("null tendency pointer"); return SALSA_ERR_ARG; }
  for (int k = 0; k < c->hostL.ncat; ++k) {
    if (!n_p[k] || !m_p[k]) { set_error("null category field"); return SALSA_ERR_ARG; }
    if (!inplace && (!n_t[k] || !m_t[k])) { set_error("null category tendency"); return SALSA_ERR_ARG; }
  }
  return SALSA_OK;
}

// Enqueue: fault reset, classification + sort (unless disabled), per-class launches.
static int timing_mark(salsa_ctx *c) {
  if (!c->timing) return SALSA_OK;
  hipEvent_t e;
  HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(e, c->stream));
  c->ev.push_back(e);
  return SALSA_OK;
}
#define MARK() do { if (timing_mark(c)) return SALSA_ERR_HIP; } while (0)

static int launch_class(salsa_ctx *c, int k, const StepArgs &a, hipStream_t s) {
  const bool hyd = k == 2;
  SmemGeom G;
  if (make_geom(c->hostL, hyd, c->params.precision == SALSA_FP64 ? 8 : 4, c->lanes[k], a.rates != nullptr, G)) {
    set_error("launch shape exceeds 64 KiB of LDS per block");
    return SALSA_ERR_ARG;
  }
  return launch_step(c->cap, c->lanes[k], c->params.precision, hyd, c->devL, a, G, s);
}

// Gather + classification, then the class counts to the host (they size the
// launches).
static int enqueue_gather(salsa_ctx *c, const StepArgs &a, bool compact) {
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, c->fault, 1);
  int rc = launch_gather(c->hostL, c->devL, a, c->rec_in, compact ? c->cls : nullptr, c->tcount, c->toff, c->lists,
                         c->cls_count, s);
  if (rc) { std::snprintf(g_err, sizeof(g_err), "gather launch failed (%d)", rc); return SALSA_ERR_HIP; }
  MARK();
  if (compact) {
    HIPCHK(hipMemcpyAsync(c->cls_count_h, c->cls_count, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return SALSA_OK;
}

static StepArgs class_args(salsa_ctx *c, const StepArgs &a, int k) {
  StepArgs b = a;
  b.cell_list = c->lists + (int64_t)k * c->ncells;
  b.ncells = c->cls_count_h[k];
  return b;
}

static int enqueue_step(salsa_ctx *c, StepArgs a, int32_t flags) {
  hipStream_t s = c->stream;
  const bool compact = !(flags & SALSA_STEP_NO_COMPACT);
  MARK();
  int rc = enqueue_gather(c, a, compact);
  if (rc) return rc;
  int64_t cells[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k) {
    MARK();
    StepArgs b = a;
    if (!compact) {
      b.cell_list = nullptr;          // every cell through the cloudy variant
      if (k != 2) b.ncells = 0;
    } else {
      b = class_args(c, a, k);
    }
    if (b.ncells > 0) {
      rc = launch_class(c, k, b, s);
      if (rc) { std::snprintf(g_err, sizeof(g_err), "kernel launch failed (%d)", rc); return SALSA_ERR_HIP; }
    }
    cells[k] = b.ncells;
    MARK();
  }
  // cloud-base activation from the aerosol state the class kernels left
  if (a.act_state) {
    rc = launch_act_cloud_base(c->devL, a, s);
    if (rc) { std::snprintf(g_err, sizeof(g_err), "activation launch failed (%d)", rc); return SALSA_ERR_HIP; }
  }
  // cell records -> tendencies or new state
  MARK();
  rc = launch_scatter(c->hostL, c->devL, a, s);
  if (rc) { std::snprintf(g_err, sizeof(g_err), "scatter launch failed (%d)", rc); return SALSA_ERR_HIP; }
  MARK();
  if (c->timing) c->ev_cells.insert(c->ev_cells.end(), cells, cells + 3);
  return SALSA_OK;
}

extern "C" int salsa_ctx_timing(salsa_ctx *c, int32_t enable) {
  if (!c) { set_error("null context"); return SALSA_ERR_ARG; }
  HIPCHK(hipStreamSynchronize(c->stream));
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  c->ev.clear();
  c->ev_cells.clear();
  c->timing = enable != 0;
  return SALSA_OK;
}

extern "C" int salsa_ctx_phase_time(salsa_ctx *c, double ms[5], int64_t cells[3], int64_t *nsteps) {
  if (!c || !ms || !cells || !nsteps) { set_error("null argument"); return SALSA_ERR_ARG; }
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < 5; ++k) ms[k] = 0.;
  for (int k = 0; k < 3; ++k) cells[k] = 0;
  const size_t n = c->ev.size() / kEvPerStep;
  auto el = [&](size_t st, int i, int j, double &acc) -> int {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, c->ev[st * kEvPerStep + i], c->ev[st * kEvPerStep + j]));
    acc += t;
    return SALSA_OK;
  };
  for (size_t st = 0; st < n; ++st) {
    if (el(st, 0, 1, ms[0])) return SALSA_ERR_HIP;                 // gather + classification
    for (int k = 0; k < 3; ++k)
      if (el(st, 2 + 2 * k, 3 + 2 * k, ms[1 + k])) return SALSA_ERR_HIP;   // class k step kernel
    if (el(st, 8, 9, ms[4])) return SALSA_ERR_HIP;                 // scatter
    for (int k = 0; k < 3; ++k) cells[k] += c->ev_cells[3 * st + k];
  }
  *nsteps = (int64_t)n;
  return SALSA_OK;
}

extern "C" int salsa_ctx_kernel_time(salsa_ctx *c, double *ms, int64_t *nsteps) {
  if (!c || !ms || !nsteps) { set_error("null argument"); return SALSA_ERR_ARG; }
  HIPCHK(hipStreamSynchronize(c->stream));
  double tot = 0.;
  const size_t n = c->ev.size() / kEvPerStep;
  for (size_t st = 0; st < n; ++st) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, c->ev[st * kEvPerStep], c->ev[st * kEvPerStep + kEvPerStep - 1]));
    tot += t;
  }
  *ms = tot;
  *nsteps = (int64_t)n;
  return SALSA_OK;
}

extern "C" int salsa_step_device(salsa_ctx *c, const double *press, const double *temp, const double *rsl,
                                 const double *rsi, const double *dn, const double *rv_p, double *rv_t,
                                 const double *const n_p[4], double *const n_t[4],
                                 const double *const m_p[4], double *const m_t[4], double *rates,
                                 double tstep, double time, int32_t istp, int32_t flags) {
  int rc = check_ptrs(c, press, temp, rsl, rsi, dn, rv_p, rv_t, n_p, n_t, m_p, m_t, rates, tstep, flags);
  if (rc) return rc;
  StepArgs a = make_args(c, press, temp, rsl, rsi, dn, rv_p, rv_t, n_p, n_t, m_p, m_t, rates, tstep, flags, time,
                         istp);
  return enqueue_step(c, a, flags);
}

extern "C" int salsa_set_cloud_base_fields(salsa_ctx *c, const double *w, double *nactd, double *vactd) {
  if (!c) { set_error("null context"); return SALSA_ERR_ARG; }
  if ((w == nullptr) != (nactd == nullptr) || (w == nullptr) != (vactd == nullptr)) {
    set_error("w, nactd and vactd are set together (or all null)");
    return SALSA_ERR_ARG;
  }
  if (nactd && !c->act_state &&
      hipMalloc(&c->act_state, (size_t)c->ncells * c->hostL.nb[0] * (1 + c->hostL.nwet) * sizeof(double)) !=
          hipSuccess) {
    set_error("activation work allocation failed");
    return SALSA_ERR_NOMEM;
  }
  c->act_w = w;
  c->act_n = nactd;
  c->act_v = vactd;
  return SALSA_OK;
}

extern "C" int salsa_set_les_extra_fields(salsa_ctx *c, double *gaerop, double *indefp, int32_t nindef) {
  if (!c) { set_error("null context"); return SALSA_ERR_ARG; }
  if ((indefp != nullptr) != (nindef > 0) || nindef < 0 || nindef > 4096) {
    set_error("indefp and nindef (1..4096 rows) are given together");
    return SALSA_ERR_ARG;
  }
  c->les_gas = gaerop;
  c->les_indef = indefp;
  c->les_nindef = indefp ? nindef : 0;
  return SALSA_OK;
}

extern "C" int salsa_ctx_sync(salsa_ctx *c) {
  if (!c) { set_error("null context"); return SALSA_ERR_ARG; }
  HIPCHK(hipMemcpyAsync(c->fault_h, c->fault, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const int32_t f = *c->fault_h;
  if (f & SALSA_CELL_NONFINITE) { set_error("non-finite values produced in some cell"); return SALSA_ERR_NONFINITE; }
  if (f & SALSA_CELL_NOCONV) { set_error("equilibration did not converge in some cell"); return SALSA_ERR_NOCONV; }
  if (f & SALSA_CELL_DISTR) { set_error("size-distribution update did not settle in some cell"); return SALSA_ERR_DISTR; }
  return SALSA_OK;
}

extern "C" int salsa_get_cell_flags(salsa_ctx *c, uint8_t *out, int64_t n) {
  if (!c || !out || n < c->ncells) { set_error("bad arguments"); return SALSA_ERR_ARG; }
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(out, c->flags, (size_t)c->ncells, hipMemcpyDeviceToHost));
  return SALSA_OK;
}

// Host-pointer variant: stage every field through device memory owned by the context.
extern "C" int salsa_step_host(salsa_ctx *c, const double *press, const double *temp, const double *rsl,
                               const double *rsi, const double *dn, const double *rv_p, double *rv_t,
                               const double *const n_p[4], double *const n_t[4],
                               const double *const m_p[4], double *const m_t[4], double *rates,
                               double tstep, double time, int32_t istp, int32_t flags) {
  int rc = check_ptrs(c, press, temp, rsl, rsi, dn, rv_p, rv_t, n_p, n_t, m_p, m_t, rates, tstep, flags);
  if (rc) return rc;
  const DevLayout &L = c->hostL;
  const bool inplace = flags & SALSA_STEP_INPLACE;
  const size_t n3 = (size_t)c->n3;
  const size_t nrate = rates ? (size_t)(L.level == 5 ? SALSA_NRATE_L5 : SALSA_NRATE_L4) : 0;
  size_t nf = 7 + nrate;                               // press temp rsl rsi dn rv_p rv_t (+ rates)
  size_t nmass[4], nnum[4];
  for (int k = 0; k < 4; ++k) {
    nnum[k] = k < L.ncat ? L.nb[k] : 0;
    nmass[k] = k < L.ncat ? (size_t)L.nb[k] * (k < 3 ? L.nwet : L.nspec) : 0;
    nf += 2 * (nnum[k] + nmass[k]);
  }
  const size_t need = nf * n3 * sizeof(double);
  if (need > c->stage_bytes) {
    hipFree(c->stage);
    c->stage = nullptr;
    c->stage_bytes = 0;
    if (hipMalloc(&c->stage, need) != hipSuccess) { set_error("staging allocation failed"); return SALSA_ERR_NOMEM; }
    c->stage_bytes = need;
  }
  hipStream_t s = c->stream;
  double *p = c->stage;
  auto put = [&](const double *h, size_t n) -> double * {
    double *d = p;
    p += n;
    if (h) { if (hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess) return nullptr; }
    return d;
  };
  double *dpress = put(press, n3), *dtemp = put(temp, n3), *drsl = put(rsl, n3), *drsi = put(rsi, n3);
  double *ddn = put(dn, n3), *drvp = put(rv_p, n3), *drvt = put(inplace ? nullptr : rv_t, n3);
  double *drates = rates ? put(rates, nrate * n3) : nullptr;     // halo cells keep their values
  const double *dnp[4] = {0, 0, 0, 0}, *dmp[4] = {0, 0, 0, 0};
  double *dnt[4] = {0, 0, 0, 0}, *dmt[4] = {0, 0, 0, 0};
  for (int k = 0; k < L.ncat; ++k) {
    dnp[k] = put(n_p[k], nnum[k] * n3);
    dmp[k] = put(m_p[k], nmass[k] * n3);
    dnt[k] = put(inplace ? nullptr : n_t[k], nnum[k] * n3);
    dmt[k] = put(inplace ? nullptr : m_t[k], nmass[k] * n3);
  }
  if (!dpress || !drvt) { set_error("host to device copy failed"); return SALSA_ERR_HIP; }
  StepArgs a = make_args(c, dpress, dtemp, drsl, rsi ? drsi : nullptr, ddn, drvp, drvt, dnp, dnt, dmp, dmt,
                         drates, tstep, flags, time, istp);
  rc = enqueue_step(c, a, flags);
  if (rc) return rc;
  if (inplace) {
    HIPCHK(hipMemcpyAsync(const_cast<double *>(rv_p), drvp, n3 * 8, hipMemcpyDeviceToHost, s));
    for (int k = 0; k < L.ncat; ++k) {
      HIPCHK(hipMemcpyAsync(const_cast<double *>(n_p[k]), dnp[k], nnum[k] * n3 * 8, hipMemcpyDeviceToHost, s));
      HIPCHK(hipMemcpyAsync(const_cast<double *>(m_p[k]), dmp[k], nmass[k] * n3 * 8, hipMemcpyDeviceToHost, s));
    }
  } else {
    HIPCHK(hipMemcpyAsync(rv_t, drvt, n3 * 8, hipMemcpyDeviceToHost, s));
    for (int k = 0; k < L.ncat; ++k) {
      HIPCHK(hipMemcpyAsync(n_t[k], dnt[k], nnum[k] * n3 * 8, hipMemcpyDeviceToHost, s));
      HIPCHK(hipMemcpyAsync(m_t[k], dmt[k], nmass[k] * n3 * 8, hipMemcpyDeviceToHost, s));
    }
  }
  if (rates) HIPCHK(hipMemcpyAsync(rates, drates, nrate * n3 * 8, hipMemcpyDeviceToHost, s));
  return salsa_ctx_sync(c);
}

// Forward-mode gradient of the precipitation volume after nsteps in-place
// steps: gather the cell records, run the derivative kernel (one lane per
// direction: the cell's bins, then E) and sum the per-cell results.
static int gradient_run(salsa_ctx *c, const double *press, const double *temp, const double *rsl,
                        const double *rsi, const double *dn, const double *rv, const double *const n_p[4],
                        const double *const m_p[4], double tstep, int32_t nsteps, double *const djdn[4],
                        double *j_cell, double *djde_cell, double totals[2], double *buf);

static int gradient_check(salsa_ctx *c, const double *press, const double *temp, const double *rsl,
                          const double *rsi, const double *dn, const double *rv, const double *const n_p[4],
                          const double *const m_p[4], double tstep, int32_t nsteps, double *const djdn[4],
                          double totals[2]) {
  int rc = check_ptrs(c, press, temp, rsl, rsi, dn, rv, nullptr, n_p, nullptr, m_p, nullptr, nullptr, tstep,
                      SALSA_STEP_INPLACE);
  if (rc) return rc;
  const DevLayout &L = c->hostL;
  if (nsteps < 1) { set_error("nsteps must be at least 1"); return SALSA_ERR_ARG; }
  if (!djdn || !totals) { set_error("null output pointer"); return SALSA_ERR_ARG; }
  for (int k = 0; k < L.ncat; ++k)
    if (!djdn[k]) { set_error("null dJ/dN0 field"); return SALSA_ERR_ARG; }
  if (c->params.lscoag_mode != 1) {
    set_error("the gradient needs lscoag_mode 1 (kernels evaluated every step)");
    return SALSA_ERR_ARG;
  }
  return SALSA_OK;
}

extern "C" int salsa_gradient_device(salsa_ctx *c, const double *press, const double *temp, const double *rsl,
                                     const double *rsi, const double *dn, const double *rv,
                                     const double *const n_p[4], const double *const m_p[4], double tstep,
                                     int32_t nsteps, double *const djdn[4], double *j_cell, double *djde_cell,
                                     double totals[2]) {
  int rc = gradient_check(c, press, temp, rsl, rsi, dn, rv, n_p, m_p, tstep, nsteps, djdn, totals);
  if (rc) return rc;
  const int64_t n = c->ncells;
  if (!c->gbuf && hipMalloc(&c->gbuf, (size_t)(2 * n + 2) * sizeof(double)) != hipSuccess) {
    set_error("gradient buffer allocation failed");
    return SALSA_ERR_NOMEM;
  }
  return gradient_run(c, press, temp, rsl, rsi, dn, rv, n_p, m_p, tstep, nsteps, djdn, j_cell, djde_cell, totals,
                      c->gbuf);
}

// The same with the per-cell results and totals in the library's static
// work space: no device memory is allocated.
extern "C" int salsa_gradient_device_noalloc(salsa_ctx *c, const double *press, const double *temp,
                                             const double *rsl, const double *rsi, const double *dn,
                                             const double *rv, const double *const n_p[4],
                                             const double *const m_p[4], double tstep, int32_t nsteps,
                                             double *const djdn[4], double *j_cell, double *djde_cell,
                                             double totals[2]) {
  int rc = gradient_check(c, press, temp, rsl, rsi, dn, rv, n_p, m_p, tstep, nsteps, djdn, totals);
  if (rc) return rc;
  if (c->ncells > SALSA_NOALLOC_MAX_CELLS) {
    std::snprintf(g_err, sizeof(g_err), "%lld interior cells exceed SALSA_NOALLOC_MAX_CELLS (%d)",
                  (long long)c->ncells, (int)SALSA_NOALLOC_MAX_CELLS);
    return SALSA_ERR_ARG;
  }
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != c->device || !c->scratch) {
    set_error("the context's device is not the current device");
    return SALSA_ERR_ARG;
  }
  return gradient_run(c, press, temp, rsl, rsi, dn, rv, n_p, m_p, tstep, nsteps, djdn, j_cell, djde_cell, totals,
                      c->scratch);
}

// Reverse-mode gradient: gather and pack the initial states, run the steps
// forward keeping a checkpoint every `every` steps, seed the adjoint with
// dJ/dx_n, then for each stretch between checkpoints (last first) recompute
// its states and run the steps' adjoints backwards; scatter dJ/dN0 and sum.
extern "C" int salsa_adjoint_device(salsa_ctx *c, const double *press, const double *temp, const double *rsl,
                                    const double *rsi, const double *dn, const double *rv,
                                    const double *const n_p[4], const double *const m_p[4], double tstep,
                                    int32_t nsteps, int32_t every, double *const djdn[4], double *j_cell,
                                    double *djde_cell, double totals[2]) {
  int rc = gradient_check(c, press, temp, rsl, rsi, dn, rv, n_p, m_p, tstep, nsteps, djdn, totals);
  if (rc) return rc;
  const DevLayout &L = c->hostL;
  if (every < 1) { set_error("checkpoint_every must be at least 1"); return SALSA_ERR_ARG; }
  if (L.level != 4 || L.autoconv2 || L.lsactiv == 1 || L.lsdistupdate) {
    set_error("the adjoint covers level-4 coagulation and condensation (no autoconv2, interstitial activation "
              "or size-distribution update)");
    return SALSA_ERR_ARG;
  }
  bool fits = L.nt <= 32 && L.nwet <= 3 && L.npair <= 528 && L.nb[2] <= 8 && L.ncc <= 528 &&
              L.fn1a - L.in1a + 1 <= 4 && L.fn2b - L.in2a + 1 <= 63;
  for (int it = 0; it < L.ntarget; ++it) fits = fits && L.tgt[it].nent <= 64;
  if (!fits) { set_error("bin layout beyond the adjoint kernel's capacity"); return SALSA_ERR_ARG; }
  const int64_t n = c->ncells;
  const int nvar = 1 + L.nt * (1 + L.nwet);
  const int nsub = substeps(tstep);
  const int nck = (nsteps + every - 1) / every;
  const int nseg = (every < nsteps ? every : nsteps) - 1;
  const int64_t sz = n * nvar;                           // doubles per set of cell states
  const int64_t ngrp = (n + 1) / 2;                      // two cells per wavefront
  const unsigned nblk = (unsigned)(ngrp < 4096 ? ngrp : 4096);
  int64_t stride = (int64_t)nsub * (L.nt + 2);
  const int64_t t_p = stride;
  stride += 2 * (int64_t)L.npair;
  const int64_t t_pf = stride;
  stride += (int64_t)L.ntarget * L.nb[2] * (1 + L.nwet);
  const int64_t t_eq = stride;
  stride += 4 * 104;
  const int64_t t_tg = stride;
  stride += 8 * (int64_t)L.ntarget;
  const int64_t t_gt = stride;
  stride += 2 * 32 * 8;
  stride = (stride + 31) / 32 * 32;
  const size_t need = (size_t)(((int64_t)nck + nseg + 3) * sz + 2 * n + 32 + 2 * (int64_t)nblk * stride) * 8;
  if (need > c->adj_bytes) {
    (void)hipFree(c->adj_buf);
    c->adj_buf = nullptr;
    c->adj_bytes = 0;
    if (hipMalloc(&c->adj_buf, need) != hipSuccess) {
      set_error("adjoint buffer allocation failed");
      return SALSA_ERR_NOMEM;
    }
    c->adj_bytes = need;
  }
  double *ck = c->adj_buf, *seg = ck + nck * sz, *wa = seg + nseg * sz, *wb = wa + sz, *lam = wb + sz;
  double *ebar = lam + sz, *jc = ebar + n, *tot = jc + n, *tape = tot + 32;
  hipStream_t s = c->stream;
  StepArgs a = make_args(c, press, temp, rsl, rsi, dn, rv, nullptr, n_p, nullptr, m_p, nullptr, nullptr, tstep,
                         SALSA_STEP_INPLACE);
  a.kmode = 0;
  a.kstore = nullptr;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, c->fault, 1);
  rc = launch_gather(L, c->devL, a, c->rec_in, nullptr, nullptr, nullptr, nullptr, nullptr, s);
  if (!rc) rc = launch_adjoint_pack(c->devL, c->rec_in, L.nslot, n, nvar, ck, c->flags, s);
  if (rc) { std::snprintf(g_err, sizeof(g_err), "adjoint set-up launch failed (%d)", rc); return SALSA_ERR_HIP; }
  AdjArgs g;
  std::memset(&g, 0, sizeof(g));
  g.ncells = n;
  g.rec = c->rec_in;
  g.nslot = L.nslot;
  g.nvar = nvar;
  g.nsub = nsub;
  g.lam = lam;
  g.ebar = ebar;
  g.jcell = jc;
  g.tape = tape;
  g.tape_stride = stride;
  g.t_p = t_p;
  g.t_pf = t_pf;
  g.t_eq = t_eq;
  g.t_tg = t_tg;
  g.t_gt = t_gt;
  g.dt = tstep;
  g.rhlim = c->params.rhlim;
  g.lsfreeRH = c->params.lsfreeRH;
  g.cell_flags = c->flags;
  g.fault_any = c->fault;
  const int prec = c->params.precision;
  // forward pass
  const double *in = ck;
  for (int k = 0; k < nsteps && !rc; ++k) {
    double *out = ((k + 1) % every == 0 && k + 1 < nsteps) ? ck + (int64_t)((k + 1) / every) * sz
                                                           : (in == wa ? wb : wa);
    g.mode = 0;
    g.x_in = in;
    g.x_out = out;
    rc = launch_adjoint_step(prec, c->devL, g, nblk, s);
    in = out;
  }
  if (!rc) rc = launch_adjoint_seed(c->devL, in, n, nvar, lam, ebar, jc, s);
  // reverse pass, one stretch between checkpoints at a time
  for (int q = nck - 1; q >= 0 && !rc; --q) {
    const int s0 = q * every, s1 = (s0 + every < nsteps) ? s0 + every : nsteps;
    const double *x0 = ck + (int64_t)q * sz;
    const double *prev = x0;
    for (int k = s0 + 1; k < s1 && !rc; ++k) {           // states x_{s0+1} .. x_{s1-1}
      g.mode = 0;
      g.x_in = prev;
      g.x_out = seg + (int64_t)(k - s0 - 1) * sz;
      rc = launch_adjoint_step(prec, c->devL, g, nblk, s);
      prev = g.x_out;
    }
    for (int k = s1 - 1; k >= s0 && !rc; --k) {
      g.mode = 1;
      g.x_in = k == s0 ? x0 : seg + (int64_t)(k - s0 - 1) * sz;
      g.x_out = nullptr;
      rc = launch_adjoint_step(prec, c->devL, g, nblk, s);
    }
  }
  if (!rc) {
    AdjOut o;
    std::memset(&o, 0, sizeof(o));
    o.ncells = n; o.n3 = c->n3;
    o.nz = c->nz; o.nx = c->nx; o.nzp = c->nzp; o.nxp = c->nxp; o.hz = c->params.hz; o.hxy = c->params.hxy;
    o.nslot = L.nslot; o.nvar = nvar;
    o.rec = c->rec_in; o.lam = lam;
    for (int k = 0; k < 4; ++k) o.djdn[k] = k < L.ncat ? djdn[k] : nullptr;
    rc = launch_adjoint_out(c->devL, o, jc, ebar, c->flags, tot, s);
  }
  if (rc) { std::snprintf(g_err, sizeof(g_err), "adjoint launch failed (%d)", rc); return SALSA_ERR_HIP; }
  if (j_cell) HIPCHK(hipMemcpyAsync(j_cell, jc, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
  if (djde_cell) HIPCHK(hipMemcpyAsync(djde_cell, ebar, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(c->tot_h, tot, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  rc = salsa_ctx_sync(c);
  totals[0] = c->tot_h[0];
  totals[1] = c->tot_h[1];
  return rc;
}

// Gather the cell records, run the derivative kernel (one lane per
// direction: the cell's bins, then E) and sum the per-cell results; buf holds
// [J per cell | dJ/dE per cell | the two totals].
static int gradient_run(salsa_ctx *c, const double *press, const double *temp, const double *rsl,
                        const double *rsi, const double *dn, const double *rv, const double *const n_p[4],
                        const double *const m_p[4], double tstep, int32_t nsteps, double *const djdn[4],
                        double *j_cell, double *djde_cell, double totals[2], double *buf) {
  const DevLayout &L = c->hostL;
  const int64_t n = c->ncells;
  int rc;
  hipStream_t s = c->stream;
  StepArgs a = make_args(c, press, temp, rsl, rsi, dn, rv, nullptr, n_p, nullptr, m_p, nullptr, nullptr, tstep,
                         SALSA_STEP_INPLACE);
  a.kmode = 0;
  a.kstore = nullptr;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, c->fault, 1);
  rc = launch_gather(L, c->devL, a, c->rec_in, nullptr, nullptr, nullptr, nullptr, nullptr, s);
  if (rc) { std::snprintf(g_err, sizeof(g_err), "gather launch failed (%d)", rc); return SALSA_ERR_HIP; }
  GradArgs g;
  std::memset(&g, 0, sizeof(g));
  g.s = a;
  g.nsteps = nsteps;
  g.lanes = L.nt + 1 <= 32 ? 32 : 64;
  for (int k = 0; k < 4; ++k) g.djdn[k] = k < L.ncat ? djdn[k] : nullptr;
  g.j_cell = buf;
  g.djde_cell = buf + n;
  rc = launch_gradient(c->cap, c->params.precision, c->devL, g, buf + 2 * n, s);
  if (rc) { std::snprintf(g_err, sizeof(g_err), "gradient launch failed (%d)", rc); return SALSA_ERR_HIP; }
  if (j_cell) HIPCHK(hipMemcpyAsync(j_cell, buf, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
  if (djde_cell) HIPCHK(hipMemcpyAsync(djde_cell, buf + n, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(c->tot_h, buf + 2 * n, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  rc = salsa_ctx_sync(c);
  totals[0] = c->tot_h[0];
  totals[1] = c->tot_h[1];
  return rc;
}

// Time each lanes-per-cell candidate per cost class on the caller's
// device-resident inputs (tendencies go to scratch) and keep the fastest.
extern "C" int salsa_autotune(salsa_ctx *c, const double *press, const double *temp, const double *rsl,
                              const double *rsi, const double *dn, const double *rv_p,
                              const double *const n_p[4], const double *const m_p[4], double tstep) {
  if (!c) { set_error("null context"); return SALSA_ERR_ARG; }
  const DevLayout &L = c->hostL;
  // scratch tendencies
  size_t nf = 1;
  for (int k = 0; k < L.ncat; ++k) nf += (size_t)L.nb[k] * (1 + (k < 3 ? L.nwet : L.nspec));
  double *scr = nullptr;
  HIPCHK(hipMalloc(&scr, nf * c->n3 * sizeof(double)));
  double *rvt = scr, *p = scr + c->n3;
  double *nt[4] = {0, 0, 0, 0}, *mt[4] = {0, 0, 0, 0};
  for (int k = 0; k < L.ncat; ++k) {
    nt[k] = p; p += (size_t)L.nb[k] * c->n3;
    mt[k] = p; p += (size_t)L.nb[k] * (k < 3 ? L.nwet : L.nspec) * c->n3;
  }
  StepArgs a = make_args(c, press, temp, rsl, rsi, dn, rv_p, rvt, n_p, nt, m_p, mt, nullptr, tstep, 0);
  a.kmode = 0;                                           // tuning neither reads nor writes the kernel store
  int rc = enqueue_step(c, a, 0);                        // classification + warm-up
  if (rc) { hipFree(scr); return rc; }
  hipStream_t s = c->stream;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  for (int k = 0; k < 3; ++k) {
    const StepArgs b = class_args(c, a, k);
    if (b.ncells == 0) continue;
    float best = 1e30f;
    const int keep = c->lanes[k];
    int bestl = keep;
    for (int lanes : {8, 16, 32, 64}) {
      SmemGeom G;
      if (make_geom(c->hostL, k == 2, c->params.precision == SALSA_FP64 ? 8 : 4, lanes, false, G)) continue;
      c->lanes[k] = lanes;
      if (launch_class(c, k, b, s)) { c->lanes[k] = keep; continue; }   // warm
      hipEventRecord(e0, s);
      for (int r = 0; r < 3; ++r) launch_class(c, k, b, s);
      hipEventRecord(e1, s);
      hipEventSynchronize(e1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) { best = ms; bestl = lanes; }
    }
    c->lanes[k] = bestl;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamSynchronize(s);
  hipFree(scr);
  return SALSA_OK;
}

// ---------------------------------------------------------------------------
// LES coupling after a step: tend_constrain2, update_sclrs, tend0 and
// SALSA_diagnostics on the device-resident prognostic arrays
// ---------------------------------------------------------------------------
extern "C" int salsa_les_update_device(salsa_ctx *c, double dtlt, const double *dzt, const double *temp,
                                       const double *rh, const double *rhi, const double *rp,
                                       double *const n_p[4], double *const n_t[4], double *const m_p[4],
                                       double *const m_t[4], double *charge_time, double *diag) {
  if (!c) { set_error("null context"); return SALSA_ERR_ARG; }
  const salsa_params &P = c->params;
  const salsa_bin_desc &d = c->desc;
  if (P.hz != 1 || P.hxy != 2) {
    set_error("the LES update needs the LES halo widths (hz 1, hxy 2)");
    return SALSA_ERR_ARG;
  }
  if (c->nzp < 4 || c->nxp < 6 || c->nyp < 6) {      // two interior columns per halo width for the periodic images
    set_error("grid too small for the LES update (nzp >= 4, nxp and nyp >= 6)");
    return SALSA_ERR_ARG;
  }
  if (!(dtlt > 0.) || !dzt || !temp || !rh || !rp || !diag || !n_p || !n_t || !m_p || !m_t) {
    set_error("null argument or non-positive dtlt");
    return SALSA_ERR_ARG;
  }
  const int ncat = P.level == 5 ? 4 : 3;
  if (P.level == 5 && !rhi) { set_error("rhi is needed at level 5"); return SALSA_ERR_ARG; }
  for (int k = 0; k < ncat; ++k)
    if (!n_p[k] || !n_t[k] || !m_p[k] || !m_t[k]) { set_error("null field pointer"); return SALSA_ERR_ARG; }
  const double dz1 = dzt[c->nzp - 2];
  if (!(dz1 != 0.)) { set_error("dzt(nzp-1) is zero"); return SALSA_ERR_ARG; }
  LesArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nzp = c->nzp; a.nxp = c->nxp; a.nyp = c->nyp; a.level = P.level;
  a.n3 = c->n3;
  a.dtlt = dtlt;
  a.dzf2 = dzt[c->nzp - 3] / dz1;                     // sclrset 'mixd' (util.f90:89-91)
  const int nbc[4] = {d.nbins, d.ncld, d.nprc, d.nice};
  const int nwet = d.iwa + 1;
  for (int k = 0; k < 4; ++k) {
    a.nb[k] = k < ncat ? nbc[k] : 0;
    a.nsw[k] = k == 3 ? nwet + 1 : nwet;
  }
  a.ndry = d.iwa; a.nwet = nwet; a.iwa = d.iwa; a.irim = d.irim;
  a.in1a = d.in1a; a.fn2b = d.fn2b; a.in2a = d.in2a; a.fn2a = d.fn2a;
  a.ica_cur = d.ica.cur; a.ica_par = d.ica.par; a.fca_cur = d.fca.cur;
  a.icb_cur = d.icb.cur; a.icb_par = d.icb.par; a.fcb_cur = d.fcb.cur;
  for (int s = 0; s < SALSA_MAXSPEC; ++s) {
    a.rholiq[s] = d.rholiq[s]; a.rhoice[s] = d.rhoice[s]; a.mm[s] = d.mm[s]; a.diss[s] = d.diss[s];
  }
  for (int b = 0; b < d.nbins; ++b) a.aerobins[b] = std::pow(d.vlolim[b] / 0.5235988, 1. / 3.);   // set_aerobins
  a.temp = temp; a.rh = rh; a.rhi = rhi ? rhi : rh; a.rp = rp;
  for (int k = 0; k < 4; ++k) {
    a.n_p[k] = k < ncat ? n_p[k] : nullptr; a.n_t[k] = k < ncat ? n_t[k] : nullptr;
    a.m_p[k] = k < ncat ? m_p[k] : nullptr; a.m_t[k] = k < ncat ? m_t[k] : nullptr;
  }
  a.charge_time = charge_time;
  a.diag = diag;
  // the gas rows of the dry species (constrain_SALSA.f90:332-351): SO4 -> 1,
  // NO -> 2, NH -> 3, OC -> 5 of a_gaerop (1-based)
  a.gas = c->les_gas;
  for (int s = 0; s < SALSA_MAXSPEC; ++s) {
    const int id = s < d.nspec_dry ? d.dry_species[s] : -1;
    a.gas_row[s] = id == 0 ? 0 : id == 5 ? 1 : id == 6 ? 2 : id == 1 ? 4 : -1;
  }
  a.indef = c->les_indef;
  a.nindef = c->les_nindef;
  if (c->n3 >= INT32_MAX) { set_error("grid too large for the LES update"); return SALSA_ERR_ARG; }
  if (!c->les_list && hipMalloc(&c->les_list, ((size_t)c->n3 + 1) * sizeof(int32_t)) != hipSuccess) {
    set_error("LES update work allocation failed");
    return SALSA_ERR_NOMEM;
  }
  a.list = c->les_list;
  a.count = c->les_list + c->n3;
  const int rc = launch_les_update(a, c->stream);
  if (rc) { std::snprintf(g_err, sizeof(g_err), "LES update launch failed (%d)", rc); return SALSA_ERR_HIP; }
  return SALSA_OK;
}

// ---------------------------------------------------------------------------
// sedimentation of the liquid categories (level 4)
// ---------------------------------------------------------------------------
extern "C" int salsa_sedim_device(salsa_ctx *c, double dtlt, const double *dzt, const double *tk, const double *th,
                                  const double *adn, const double *ustar, const double *const n_p[4],
                                  const double *const m_p[4], double *const n_t[4], double *const m_t[4],
                                  double *tlt, double *rrate, double *sfcrrate, double *vt_prc, double *irate,
                                  double *sfcirate, double *vt_ice, int32_t sed_aero, int32_t sed_cloud,
                                  int32_t sed_precp, int32_t sed_ice) {
  if (!c) { set_error("null context"); return SALSA_ERR_ARG; }
  const salsa_params &P = c->params;
  const salsa_bin_desc &d = c->desc;
  const int l5 = P.level == 5;
  if (P.level != 4 && !l5) { set_error("sedimentation is carried at levels 4 and 5"); return SALSA_ERR_ARG; }
  if (P.hz != 1 || P.hxy != 2) { set_error("sedimentation needs the LES halo widths (hz 1, hxy 2)"); return SALSA_ERR_ARG; }
  if (c->nzp < 4 || c->nzp > 256 || c->nxp < 5 || c->nyp < 5) {
    set_error("grid out of range for sedimentation (4 <= nzp <= 256, nxp and nyp >= 5)");
    return SALSA_ERR_ARG;
  }
  if (!(dtlt > 0.) || !dzt || !tk || !th || !adn || !ustar || !n_p || !m_p || !n_t || !m_t || !tlt || !rrate ||
      !sfcrrate || !vt_prc || (l5 && (!irate || !sfcirate || !vt_ice))) {
    set_error("null argument or non-positive dtlt");
    return SALSA_ERR_ARG;
  }
  for (int k = 0; k < 3 + l5; ++k)
    if (!n_p[k] || !m_p[k] || !n_t[k] || !m_t[k]) { set_error("null field pointer"); return SALSA_ERR_ARG; }
  if (l5 && d.nice > 16) { set_error("ice sedimentation carries at most 16 ice bins"); return SALSA_ERR_ARG; }
  SedArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nzp = c->nzp; a.nxp = c->nxp; a.nyp = c->nyp; a.level = P.level; a.n3 = c->n3; a.dtlt = dtlt;
  a.nb[0] = d.nbins; a.nb[1] = d.ncld; a.nb[2] = d.nprc; a.nb[3] = l5 ? d.nice : 0;
  a.ns = d.iwa + 1; a.ns_ice = d.iwa + 2; a.iwa = d.iwa; a.irim = d.irim;
  for (int s = 0; s < SALSA_MAXSPEC; ++s) { a.rholiq[s] = d.rholiq[s]; a.rhoice[s] = d.rhoice[s]; }
  a.ice_alpha = P.ice_alpha; a.ice_beta = P.ice_beta; a.ice_gamma = P.ice_gamma; a.ice_sigma = P.ice_sigma;
  for (int k = 0; k < c->nzp; ++k) {
    if (!(dzt[k] > 0.)) { set_error("dzt must be positive"); return SALSA_ERR_ARG; }
    a.dzt[k] = dzt[k];
  }
  if ((size_t)(d.nbins > d.ncld ? d.nbins : d.ncld) * c->nzp * sizeof(double) > 65536) {
    set_error("too many aerosol or cloud bins for this many levels (bins x nzp <= 8192)");
    return SALSA_ERR_ARG;
  }
  a.tk = tk; a.th = th; a.adn = adn; a.ustar = ustar;
  for (int k = 0; k < 3 + l5; ++k) { a.n_p[k] = n_p[k]; a.m_p[k] = m_p[k]; a.n_t[k] = n_t[k]; a.m_t[k] = m_t[k]; }
  a.tlt = tlt; a.rrate = rrate; a.sfcrrate = sfcrrate; a.vt_prc = vt_prc;
  a.irate = irate; a.sfcirate = sfcirate; a.vt_ice = vt_ice;
  const int rc = launch_sedim(a, sed_aero, sed_cloud, sed_precp, sed_ice, c->stream);
  if (rc) { std::snprintf(g_err, sizeof(g_err), "sedimentation launch failed (%d)", rc); return SALSA_ERR_HIP; }
  return SALSA_OK;
}
