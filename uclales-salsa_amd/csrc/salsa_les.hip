// salsa_les.hip -- the LES side of a SALSA step on device: what cfll does
// after run_SALSA (step.f90:330-333) for the SALSA tracers, so that the
// prognostic fields never leave HBM between steps.
//
//   pass A  tend_constrain2 (constrain_SALSA.f90:18-36; a_indefp clamped
//           when registered), update
//           (step.f90:470-489) on the interior, tend0 (step.f90:379-396):
//           one thread per array element, rows streamed contiguously, one
//           launch over the rows of every category.
//   pass B  sclrset 'mixd' (util.f90:71-107): the bottom row copies k = 2,
//           the top row extrapolates with dzf2 = dzt(nzp-2)/dzt(nzp-1); the
//           lateral halos are the periodic images (cyclics/cyclicc of a
//           process that owns the whole grid).  One launch over every row:
//           one thread per (halo column, level) or per interior column.
//   pass D  SALSA_diagnostics (constrain_SALSA.f90:50-402): clamps on every
//           cell; on interior columns the removal of particles with number
//           but no mass, cloud / precipitation / ice bins that evaporated
//           back to the aerosol bins (findDry4Wet :408-470 for
//           precipitation and ice), aerosols below 0.1 nm dry diameter; then
//           the diagnostic sums rc, srp, snrp, rtot, ri, riri on every cell.
//           One thread per cell; per cell the bins are visited in the
//           reference's order, so the result does not depend on the grid
//           traversal.  A streaming part reads each row once (clamps,
//           sums, and whether any removal test holds, listing those cells);
//           the removals then run on the listed cells only, with each
//           cell's rows staged in LDS where they fit.
//
// The passes are HBM-bound byte movers, coalesced along k (the fastest LES
// index).  Floating-point contraction is off so that every sum
// and product rounds as written.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <type_traits>

#include "salsa_internal.h"

#pragma clang fp contract(off)

namespace salsa {

namespace {

constexpr double kMassTH = 1.e-25;    // constrain_SALSA.f90:65
constexpr int kLesGroup = 2;          // bins whose rows the diagnostics pass loads together
constexpr int kLesCells = 64;         // listed cells per workgroup of the LDS removal pass
constexpr int kLesLdsRows = 128;      // rows per cell that fit its 64 KiB of LDS

__device__ __forceinline__ bool interior_col(const LesArgs &a, int i, int j) {
  return i >= 2 && i <= a.nxp - 3 && j >= 2 && j <= a.nyp - 3;
}

// Row r of the concatenated tracer rows of every category (number rows, then
// mass rows, category by category): its prognostic and tendency rows.  The
// same for every thread of a workgroup (r is blockIdx.y).  Rows past the
// tracers: the a_indefp rows of the update pass (t null).
__device__ __forceinline__ void les_row(const LesArgs &a, int r, double *&p, double *&t) {
  const int ncat = a.level == 5 ? 4 : 3;
  p = t = nullptr;
  for (int c = 0; c < ncat; ++c) {
    if (r < a.nb[c]) {
      p = a.n_p[c] + (int64_t)r * a.n3;
      t = a.n_t[c] + (int64_t)r * a.n3;
      return;
    }
    r -= a.nb[c];
    const int nm = a.nb[c] * a.nsw[c];
    if (r < nm) {
      p = a.m_p[c] + (int64_t)r * a.n3;
      t = a.m_t[c] + (int64_t)r * a.n3;
      return;
    }
    r -= nm;
  }
  if (a.indef && r < a.nindef) p = a.indef + (int64_t)r * a.n3;
}

// pass A, every tracer row in one launch: blockIdx.y the row, x the cell
__global__ void __launch_bounds__(256) k_les_update(LesArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (int)a.n3) return;
  double *p, *t;
  les_row(a, (int)blockIdx.y, p, t);
  if (!p) return;
  const int k = x % a.nzp, col = x / a.nzp;
  const int i = col % a.nxp, j = col / a.nxp;
  if (x == 0 && blockIdx.y == 0) *a.count = 0; // the removal list of this call starts empty
  if (!t) {                                   // tend_constrain2's a_indefp clamp (:33-34)
    p[x] = fmin(fmax(p[x], 0.), 1.);
    return;
  }
  const double rdt = a.dtlt;
  const double pv = p[x], tv = t[x];
  const double tc = fmax((-1. + 1.e-8) * pv / rdt, tv);
  if (k >= 1 && k <= a.nzp - 2 && interior_col(a, i, j)) p[x] = pv + tc * rdt;
  t[x] = 0.;
}

// pass B, every tracer row in one launch, blockIdx.y the row.  Workgroups
// below gh fill the halo columns: one thread per (halo column, k), k fastest,
// so a wavefront reads and writes along one column; the halo columns are the
// 2 rows of columns at each y edge (all x) followed by the 2 columns at each x
// edge of the interior y range.  The others set the bottom and top rows of
// the interior columns (sclrset 'mixd'), one thread per column.  The halo
// reads interior levels 1 .. nzp-2 only, which the bottom / top rows leave
// alone, so the two parts need no order.
__global__ void __launch_bounds__(256) k_les_bound(LesArgs a, unsigned gh) {
  double *p, *t;
  les_row(a, (int)blockIdx.y, p, t);
  if (!t) return;
  const int nz = a.nzp, nxp = a.nxp, nyp = a.nyp, nx = nxp - 4, ny = nyp - 4;
  if (blockIdx.x < gh) {
    const int nhc = 4 * nxp + 4 * ny;
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nhc * nz) return;
    const int hc = h / nz, k = h % nz;
    int i, j;
    if (hc < 4 * nxp) {
      const int jq = hc / nxp;
      j = jq < 2 ? jq : nyp - 4 + jq;
      i = hc % nxp;
    } else {
      const int q = (hc - 4 * nxp) % 4;
      j = 2 + (hc - 4 * nxp) / 4;
      i = q < 2 ? q : nxp - 4 + q;
    }
    const int ii = i < 2 ? i + nx : (i > nxp - 3 ? i - nx : i);
    const int jj = j < 2 ? j + ny : (j > nyp - 3 ? j - ny : j);
    const double *src = p + ((int64_t)ii + (int64_t)nxp * jj) * nz;
    double v;
    if (k == 0) {
      const double s1 = src[1], s2 = src[2];
      v = s1 - 0. * (s2 - s1);
    } else if (k == nz - 1) {
      const double sm = src[nz - 2], sm2 = src[nz - 3];
      v = sm + a.dzf2 * (sm - sm2);
    } else {
      v = src[k];
    }
    p[((int64_t)i + (int64_t)nxp * j) * nz + k] = v;
  } else {
    const int c = (blockIdx.x - gh) * blockDim.x + threadIdx.x;
    if (c >= nx * ny) return;
    const int i = 2 + c % nx, j = 2 + c / nx;
    double *col = p + ((int64_t)i + (int64_t)nxp * j) * nz;
    const double s1 = col[1], s2 = col[2], sm = col[nz - 2], sm2 = col[nz - 3];
    col[0] = s1 - 0. * (s2 - s1);
    col[nz - 1] = sm + a.dzf2 * (sm - sm2);
  }
}

// One cell's tracer rows, either in the LES arrays (row stride n3) or staged
// in LDS (row stride = cells per workgroup): N(c, b) = n[c][b * rs],
// M(c, s, b) = m[c][(s * nb[c] + b) * rs].
struct LesCell {
  double *n[4], *m[4];
  int64_t rs;
  __device__ __forceinline__ double &N(const LesArgs &a, int c, int b) const { return n[c][b * rs]; }
  __device__ __forceinline__ double &M(const LesArgs &a, int c, int s, int b) const {
    return m[c][((int64_t)s * a.nb[c] + b) * rs];
  }
};

// findDry4Wet for an evaporating bin (number ne, dry masses md[s * mstride])
__device__ __forceinline__ int les_dry4wet(const LesArgs &a, const LesCell &C, double ne, const double *md,
                                           int64_t mstride) {
  const int ndry = a.ndry;
  double cd = 0.;
  for (int s = 0; s < ndry; ++s) cd = cd + md[s * mstride] / a.rholiq[s];
  cd = cd / ne;
  cd = ::pow(cd / kPi6, 1. / 3.);
  int cnt = 0;
  for (int b = a.in2a; b <= a.fn2a; ++b) cnt += cd > a.aerobins[b];
  int ba = a.in2a + cnt - 1;
  ba = ba < a.in2a ? a.in2a : ba;
  ba = ba > a.fn2a ? a.fn2a : ba;
  const int bb = a.fn2a + (ba - a.in2a);   // as written in the reference (:443)
  if (C.N(a, 0, bb) <= kNlim) return ba;
  if (C.N(a, 0, ba) <= kNlim) return bb;
  // calc_correlation of the dry masses; sxy keeps the last product only
  // (util.f90:1017)
  double r[2];
  for (int q = 0; q < 2; ++q) {
    const int bx = q ? bb : ba;
    double sx = 0., sy = 0., sx2 = 0., sy2 = 0., sxy = 0.;
    for (int s = 0; s < ndry; ++s) {
      const double xv = C.M(a, 0, s, bx), yv = md[s * mstride];
      sx = sx + xv;
      sy = sy + yv;
      sx2 = sx2 + xv * xv;
      sy2 = sy2 + yv * yv;
      sxy = xv * yv;
    }
    const double n = (double)ndry;
    r[q] = (sx2 * n - sx * sx < kEpsR8 || sy2 * n - sy * sy < kEpsR8)
               ? 0.
               : (sxy * n - sx * sy) / (::sqrt(sx2 * n - sx * sx) * ::sqrt(sy2 * n - sy * sy));
  }
  return r[0] < r[1] ? bb : ba;
}

// SALSA_diagnostics' removals on an interior cell (constrain_SALSA.f90:
// 103-366) in the reference's order; x is the cell's LES index (for the
// charge-time rows).  Returns whether any value changed.
__device__ __forceinline__ bool les_remove(const LesArgs &a, const LesCell &C, int64_t x, double T, double rh,
                                           double rhi) {
  const int nwet = a.nwet, ndry = a.ndry, iwa = a.iwa, irim = a.irim;
  const int64_t n3 = a.n3;
  bool mod = false;
  auto N = [&](int c, int b) -> double & { return C.N(a, c, b); };
  auto M = [&](int c, int s, int b) -> double & { return C.M(a, c, s, b); };
  auto ctime = [&](int row) {
    if (a.charge_time) a.charge_time[(int64_t)row * n3 + x] = 0.;
  };
  // particles with number but no mass (:103-154)
#pragma unroll
  for (int c = 0; c < 2; ++c)
    for (int b = 0; b < a.nb[c]; ++b) {
      double sm = 0.;
      for (int s = 0; s < nwet; ++s) sm = sm + M(c, s, b);
      if (N(c, b) > 0. && sm <= 0.) {
        N(c, b) = 0.;
        for (int s = 0; s < nwet; ++s) M(c, s, b) = 0.;
        ctime(c == 0 ? b : a.nb[0] + b);
        mod = true;
      }
    }
  for (int b = 0; b < a.nb[2]; ++b)
    if (N(2, b) > 0. && M(2, iwa, b) <= 0.) {
      N(2, b) = 0.;
      for (int s = 0; s < nwet; ++s) M(2, s, b) = 0.;
      ctime(a.nb[0] + a.nb[1] + b);
      mod = true;
    }
  if (a.level == 5)
    for (int b = 0; b < a.nb[3]; ++b) {
      const double mt = M(3, iwa, b) + M(3, irim, b);
      if (N(3, b) > 0. && mt < kMassTH) {
        N(3, b) = 0.;
        for (int s = 0; s <= irim; ++s) M(3, s, b) = 0.;
        mod = true;
      }
    }
  // cloud droplets that evaporated (:156-207)
  for (int bc = a.ica_cur; bc <= a.fcb_cur; ++bc) {
    if (!(N(1, bc) > kNlim && rh < 0.999 && M(1, iwa, bc) < 1.e-5)) continue;
    double ns = 0.;
    for (int s = 0; s < ndry; ++s) ns = ns + a.diss[s] * M(1, s, bc) / a.mm[s];
    ns = ns / N(1, bc);
    const double zbb = 6. * kMwa * ns / (kPi * kRhowa);
    const double zaa = 4. * kMwa * kSurfw0 / (kRg * kRhowa * T);
    const double cdcld = ::sqrt(3. * zbb / zaa);
    double zvol = 0.;
    for (int s = 0; s < nwet; ++s) zvol = zvol + M(1, s, bc) / a.rholiq[s];
    zvol = zvol / N(1, bc);
    const double zdh2o = ::pow(zvol / kPi6, 1. / 3.);
    if (zdh2o < fmin(0.7 * cdcld, 1.3e-5) || M(1, iwa, bc) < kMassTH * N(1, bc)) {
      const int ba = bc <= a.fca_cur ? a.ica_par + (bc - a.ica_cur) : a.icb_par + (bc - a.icb_cur);
      N(0, ba) = N(0, ba) + N(1, bc);
      N(1, bc) = 0.;
      for (int s = 0; s < nwet; ++s) {
        M(0, s, ba) = M(0, s, ba) + M(1, s, bc);
        M(1, s, bc) = 0.;
      }
      ctime(a.nb[0] + bc);
      mod = true;
    }
  }
  // precipitation that evaporated (:209-258)
  for (int bc = 0; bc < a.nb[2]; ++bc) {
    if (!(N(2, bc) > kPrlim && rh < 0.999 && M(2, iwa, bc) < 1.e-6)) continue;
    double ns = 0.;
    for (int s = 0; s < ndry; ++s) ns = ns + a.diss[s] * M(2, s, bc) / a.mm[s];
    ns = ns / N(2, bc);
    const double zbb = 6. * kMwa * ns / (kPi * kRhowa);
    const double zaa = 4. * kMwa * kSurfw0 / (kRg * kRhowa * T);
    const double cdprc = ::sqrt(3. * zbb / zaa);
    double zvol = 0.;
    for (int s = 0; s < nwet; ++s) zvol = zvol + M(2, s, bc) / a.rholiq[s];
    zvol = zvol / N(2, bc);
    const double zdh2o = ::pow(zvol / kPi6, 1. / 3.);
    if (zdh2o < fmin(0.2 * cdprc, 10.e-6) || M(2, iwa, bc) < kMassTH * N(2, bc)) {
      const int ba = les_dry4wet(a, C, N(2, bc), &M(2, 0, bc), (int64_t)a.nb[2] * C.rs);
      N(0, ba) = N(0, ba) + N(2, bc);
      N(2, bc) = 0.;
      for (int s = 0; s < nwet; ++s) {
        M(0, s, ba) = M(0, s, ba) + M(2, s, bc);
        M(2, s, bc) = 0.;
      }
      ctime(a.nb[0] + a.nb[1] + bc);
      mod = true;
    }
  }
  // ice that sublimated (:260-315)
  if (a.level == 5) {
    for (int bc = 0; bc < a.nb[3]; ++bc) {
      const double mt = M(3, iwa, bc) + M(3, irim, bc);
      double vt = 0.;
      for (int s = 0; s <= irim; ++s) vt = vt + M(3, s, bc) / a.rhoice[s];
      if (!(N(3, bc) > kPrlim && rhi < 0.999 && mt < 1.e-15)) continue;
      double cdice = vt / N(3, bc);
      cdice = ::pow(cdice / kPi6, 1. / 3.);
      double zdrms = 0., zwams = 0.;
      for (int s = 0; s < ndry; ++s) zdrms = zdrms + M(3, s, bc);
      for (int s = 0; s <= irim; ++s) zwams = zwams + M(3, s, bc);
      const double zvol = zdrms / zwams;
      if (zvol > 0.5 || cdice < 2.e-6) {
        const int ba = les_dry4wet(a, C, N(3, bc), &M(3, 0, bc), (int64_t)a.nb[3] * C.rs);
        N(0, ba) = N(0, ba) + N(3, bc);
        N(3, bc) = 0.;
        for (int s = 0; s < nwet; ++s) {
          M(0, s, ba) = M(0, s, ba) + M(3, s, bc);
          M(3, s, bc) = 0.;
        }
        M(0, iwa, ba) = M(0, iwa, ba) + M(3, irim, bc);
        M(3, irim, bc) = 0.;
        mod = true;
      }
    }
  }
  // aerosols below 0.1 nm dry diameter (:317-366); with the gas field
  // (lscndgas) their SO4, OC, NO and NH go to the gas phase as molecules
  for (int b = 0; b < a.nb[0]; ++b) {
    if (!(N(0, b) > kNlim)) continue;
    double zvol = 0.;
    for (int s = 0; s < ndry; ++s) zvol = zvol + M(0, s, b) / a.rholiq[s];
    zvol = zvol / N(0, b);
    if (zvol < kPi6 * 1.e-30) {
      if (a.gas)
        for (int s = 0; s < ndry; ++s) {
          const int g = a.gas_row[s];
          if (g >= 0) a.gas[(int64_t)g * n3 + x] = a.gas[(int64_t)g * n3 + x] + M(0, s, b) / a.mm[s] * kAvog;
        }
      for (int s = 0; s < nwet; ++s) M(0, s, b) = 0.;
      N(0, b) = 0.;
      mod = true;
    }
  }
  return mod;
}

// diagnostic tracers (:370-400) from the cell's (final) rows
__device__ __forceinline__ void les_sums(const LesArgs &a, const LesCell &C, int64_t x) {
  const int iwa = a.iwa, irim = a.irim;
  const int64_t n3 = a.n3;
  double sa = 0., sc = 0., sp = 0., sn = 0.;
  for (int b = a.in1a; b <= a.fn2b; ++b) sa = sa + C.M(a, 0, iwa, b);
  for (int b = a.ica_cur; b <= a.fcb_cur; ++b) sc = sc + C.M(a, 1, iwa, b);
  for (int b = 0; b < a.nb[2]; ++b) sp = sp + C.M(a, 2, iwa, b);
  for (int b = 0; b < a.nb[2]; ++b) sn = sn + C.N(a, 2, b);
  const double rc = sa + sc;
  double rtot = rc + sp + a.rp[x];
  a.diag[x] = rc;
  a.diag[n3 + x] = sp;
  a.diag[2 * n3 + x] = sn;
  if (a.level == 5) {
    double si = 0., sr = 0.;
    for (int b = 0; b < a.nb[3]; ++b) si = si + C.M(a, 3, iwa, b);
    for (int b = 0; b < a.nb[3]; ++b) sr = sr + C.M(a, 3, irim, b);
    rtot = rtot + si + sr;
    a.diag[4 * n3 + x] = si;
    a.diag[5 * n3 + x] = sr;
  }
  a.diag[3 * n3 + x] = rtot;
}

__device__ __forceinline__ bool les_interior_cell(const LesArgs &a, int64_t x) {
  const int64_t col = x / a.nzp;
  return interior_col(a, (int)(col % a.nxp), (int)(col / a.nxp));
}

// pass D, streaming part: one thread per cell, the bins of each category in
// groups of kLesGroup whose rows (number and every mass species) are loaded
// together.  Clamps (written where they change a value), the diagnostic
// sums in bin order, and for interior cells whether any removal test of
// les_remove holds on the clamped values: its outer conditions, which are
// what it tests first on values nothing has changed yet (false means
// les_remove would leave the cell as it is).  Each row is read once.  MS:
// the most mass rows per bin of any category (registers per bin).
template <int MS>
__global__ void __launch_bounds__(256) k_les_diag(LesArgs a) {
  const int ncat = a.level == 5 ? 4 : 3;
  const int iwa = a.iwa, irim = a.irim, nwet = a.nwet, ndry = a.ndry;
  const int64_t n3 = a.n3;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n3; x += (int64_t)gridDim.x * blockDim.x) {
    const bool inter = les_interior_cell(a, x);
    const double rh = a.rh[x], rhi = a.level == 5 ? a.rhi[x] : 0.;
    bool any = false;
    double sa = 0., sc = 0., sp = 0., sn = 0., si = 0., sr = 0.;
    // groups of kLesGroup bins, categories in order; the next group's rows
    // are loaded before the current group's clamps are stored (different
    // rows), so two groups of loads are in flight
    double vn[2][kLesGroup], vm[2][kLesGroup][MS];
    auto load = [&](int c, int b0, double (&n)[kLesGroup], double (&m)[kLesGroup][MS]) {
      const int nb = a.nb[c], nsw = a.nsw[c];
      const double *np = a.n_p[c] + x, *mp = a.m_p[c] + x;
#pragma unroll
      for (int u = 0; u < kLesGroup; ++u) {
        const int b = b0 + u;
        n[u] = b < nb ? np[(int64_t)b * n3] : 0.;
#pragma unroll
        for (int s = 0; s < MS; ++s) m[u][s] = b < nb && s < nsw ? mp[((int64_t)s * nb + b) * n3] : 0.;
      }
    };
    int c = 0, b0 = 0;
    load(0, 0, vn[0], vm[0]);
    // one group; the buffer index is a template constant (no dynamic
    // register index), the loop below alternates the two
    auto group = [&](auto cur_c) -> bool {
      constexpr int cur = decltype(cur_c)::value;
      int c1 = c, b1 = b0 + kLesGroup;
      if (b1 >= a.nb[c]) { c1 = c + 1; b1 = 0; }
      const bool more = c1 < ncat;
      if (more) load(c1, b1, vn[cur ^ 1], vm[cur ^ 1]);
      const int nb = a.nb[c], nsw = a.nsw[c];
      double *np = a.n_p[c] + x, *mp = a.m_p[c] + x;
      double (&n_)[kLesGroup] = vn[cur];
      double (&m_)[kLesGroup][MS] = vm[cur];
      // remove negative values (:88-100); only values that change are written
#pragma unroll
      for (int u = 0; u < kLesGroup; ++u) {
        const int b = b0 + u;
        if (b >= nb) continue;
        if (!(n_[u] >= 0.)) {
          n_[u] = fmax(0., n_[u]);
          np[(int64_t)b * n3] = n_[u];
        }
#pragma unroll
        for (int s = 0; s < MS; ++s)
          if (s < nsw && !(m_[u][s] >= 0.)) {
            m_[u][s] = fmax(0., m_[u][s]);
            mp[((int64_t)s * nb + b) * n3] = m_[u][s];
          }
      }
#pragma unroll
      for (int u = 0; u < kLesGroup; ++u) {
        const int b = b0 + u;
        if (b >= nb) continue;
        const double n = n_[u];
        double wv = 0., rv = 0.;                // water and rime masses (no dynamic register index)
#pragma unroll
        for (int s = 0; s < MS; ++s) {
          if (s == iwa) wv = m_[u][s];
          if (s == irim) rv = m_[u][s];
        }
        if (c < 2) {
          double sm = 0.;
#pragma unroll
          for (int s = 0; s < MS; ++s)
            if (s < nwet) sm = sm + m_[u][s];
          any |= n > 0. && sm <= 0.;
        }
        if (c == 0) {
          double zvol = 0.;
#pragma unroll
          for (int s = 0; s < MS; ++s)
            if (s < ndry) zvol = zvol + m_[u][s] / a.rholiq[s];
          any |= n > kNlim && zvol / n < kPi6 * 1.e-30;
          if (b >= a.in1a && b <= a.fn2b) sa = sa + wv;
        } else if (c == 1) {
          any |= b >= a.ica_cur && b <= a.fcb_cur && rh < 0.999 && n > kNlim && wv < 1.e-5;
          if (b >= a.ica_cur && b <= a.fcb_cur) sc = sc + wv;
        } else if (c == 2) {
          any |= n > 0. && wv <= 0.;
          any |= rh < 0.999 && n > kPrlim && wv < 1.e-6;
          sp = sp + wv;
          sn = sn + n;
        } else {
          const double mt = wv + rv;
          any |= n > 0. && mt < kMassTH;
          any |= n > kPrlim && rhi < 0.999 && mt < 1.e-15;
          si = si + wv;
          sr = sr + rv;
        }
      }
      c = c1;
      b0 = b1;
      return more;
    };
    while (group(std::integral_constant<int, 0>{}) && group(std::integral_constant<int, 1>{})) {
    }
    // diagnostic tracers (:370-400), as les_sums forms them
    const double rc = sa + sc;
    double rtot = rc + sp + a.rp[x];
    a.diag[x] = rc;
    a.diag[n3 + x] = sp;
    a.diag[2 * n3 + x] = sn;
    if (a.level == 5) {
      rtot = rtot + si + sr;
      a.diag[4 * n3 + x] = si;
      a.diag[5 * n3 + x] = sr;
    }
    a.diag[3 * n3 + x] = rtot;
    // append the cell to the list of cells the removals visit (one atomic
    // per wavefront; the list order does not matter, cells are independent)
    const bool fl = inter && any;
    const unsigned long long bal = __ballot(fl);
    if (bal) {
      const int lane = (int)__lane_id(), leader = __ffsll((long long)bal) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(a.count, (int)__popcll(bal));
      base = __shfl(base, leader);
      if (fl) a.list[base + (int)__popcll(bal & ((1ull << lane) - 1ull))] = (int32_t)x;
    }
  }
}

// pass D, removals on the listed cells, on the LES arrays (layouts whose
// rows do not fit kLesLdsRows); their sums are formed again from the final
// values
__global__ void __launch_bounds__(256) k_les_remove(LesArgs a) {
  const int ncat = a.level == 5 ? 4 : 3;
  const int cnt = *a.count;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < cnt; idx += gridDim.x * blockDim.x) {
    const int64_t x = a.list[idx];
    LesCell G;
    G.rs = a.n3;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      G.n[c] = c < ncat ? a.n_p[c] + x : nullptr;
      G.m[c] = c < ncat ? a.m_p[c] + x : nullptr;
    }
    les_remove(a, G, x, a.temp[x], a.rh[x], a.level == 5 ? a.rhi[x] : 0.);
    les_sums(a, G, x);
  }
}

// pass D, removals on the listed cells with each cell's rows staged in LDS
// (one lane per listed cell, its own column lds[row * kLesCells + lane]):
// the sequence of dependent tests runs at LDS latency.  A cell whose
// removals changed something writes all its rows back and its sums again.
__global__ void __launch_bounds__(kLesCells) k_les_remove_lds(LesArgs a) {
  extern __shared__ double lds[];
  const int ncat = a.level == 5 ? 4 : 3;
  const int t = threadIdx.x, cnt = *a.count;
  const int64_t n3 = a.n3;
  for (int idx = blockIdx.x * kLesCells + t; idx - t < cnt; idx += gridDim.x * kLesCells) {
    if (idx >= cnt) continue;                 // lanes work on their own LDS column: no barrier
    const int64_t x = a.list[idx];
    LesCell G, S;
    G.rs = n3;
    S.rs = kLesCells;
    int row = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      G.n[c] = c < ncat ? a.n_p[c] + x : nullptr;
      G.m[c] = c < ncat ? a.m_p[c] + x : nullptr;
      S.n[c] = lds + (int64_t)row * kLesCells + t;
      row += c < ncat ? a.nb[c] : 0;
      S.m[c] = lds + (int64_t)row * kLesCells + t;
      row += c < ncat ? a.nb[c] * a.nsw[c] : 0;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (c >= ncat) continue;
        const double *g = q ? G.m[c] : G.n[c];
        double *l = q ? S.m[c] : S.n[c];
        const int nr = q ? a.nb[c] * a.nsw[c] : a.nb[c];
        for (int r0 = 0; r0 < nr; r0 += 16) {
          double v[16];
#pragma unroll
          for (int u = 0; u < 16; ++u) v[u] = r0 + u < nr ? g[(int64_t)(r0 + u) * n3] : 0.;
#pragma unroll
          for (int u = 0; u < 16; ++u)
            if (r0 + u < nr) l[(r0 + u) * kLesCells] = v[u];
        }
      }
    if (les_remove(a, S, x, a.temp[x], a.rh[x], a.level == 5 ? a.rhi[x] : 0.)) {
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          if (c >= ncat) continue;
          double *g = q ? G.m[c] : G.n[c];
          const double *l = q ? S.m[c] : S.n[c];
          const int nr = q ? a.nb[c] * a.nsw[c] : a.nb[c];
          for (int r = 0; r < nr; ++r) g[(int64_t)r * n3] = l[r * kLesCells];
        }
      les_sums(a, S, x);
    }
  }
}

unsigned grid_for(int64_t n, int block) {
  const int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 65536 ? (g > 0 ? g : 1) : 65536);
}

}  // namespace

int launch_les_update(const LesArgs &a, hipStream_t s) {
  const int ncat = a.level == 5 ? 4 : 3;
  int nrow = 0;
  for (int c = 0; c < ncat; ++c) nrow += a.nb[c] * (1 + a.nsw[c]);
  const int nind = a.indef && a.level == 5 ? a.nindef : 0;
  if (nrow < 1 || nrow + nind > 65535) return (int)hipErrorInvalidValue;
  const unsigned gx = (unsigned)((a.n3 + 255) / 256);
  hipLaunchKernelGGL(k_les_update, dim3(gx, nrow + nind), dim3(256), 0, s, a);
  const unsigned gh = (unsigned)(((4 * a.nxp + 4 * (a.nyp - 4)) * a.nzp + 255) / 256);
  const unsigned gr = (unsigned)(((a.nxp - 4) * (a.nyp - 4) + 255) / 256);
  hipLaunchKernelGGL(k_les_bound, dim3(gh + gr, nrow), dim3(256), 0, s, a, gh);
  int msw = 0;
  for (int c = 0; c < ncat; ++c) msw = a.nsw[c] > msw ? a.nsw[c] : msw;
  if (msw <= 3)
    hipLaunchKernelGGL(k_les_diag<3>, dim3(grid_for(a.n3, 256)), dim3(256), 0, s, a);
  else if (msw <= 5)
    hipLaunchKernelGGL(k_les_diag<5>, dim3(grid_for(a.n3, 256)), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_les_diag<kMaxSpec>, dim3(grid_for(a.n3, 256)), dim3(256), 0, s, a);
  if (nrow <= kLesLdsRows) {
    const int64_t g = (a.n3 + kLesCells - 1) / kLesCells;
    hipLaunchKernelGGL(k_les_remove_lds, dim3((unsigned)(g < 4096 ? g : 4096)), dim3(kLesCells),
                       (size_t)nrow * kLesCells * sizeof(double), s, a);
  } else {
    hipLaunchKernelGGL(k_les_remove, dim3(grid_for(a.n3, 256)), dim3(256), 0, s, a);
  }
  return (int)hipGetLastError();
}

}  // namespace salsa
