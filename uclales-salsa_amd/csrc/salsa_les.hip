// salsa_les.hip -- the LES side of a SALSA step on device: what cfll does
// after run_SALSA (step.f90:330-333) for the SALSA tracers, so that the
// prognostic fields never leave HBM between steps.
//
//   pass A  tend_constrain2 (constrain_SALSA.f90:18-36), update
//           (step.f90:470-489) on the interior, tend0 (step.f90:379-396):
//           one thread per array element, rows streamed contiguously.
//   pass B  sclrset 'mixd' (util.f90:71-107): the bottom row copies k = 2,
//           the top row extrapolates with dzf2 = dzt(nzp-2)/dzt(nzp-1); the
//           lateral halos are the periodic images (cyclics/cyclicc of a
//           process that owns the whole grid).  One thread per (row, column).
//   pass D  SALSA_diagnostics (constrain_SALSA.f90:50-402): clamps on every
//           cell; on interior columns the removal of particles with number
//           but no mass, cloud / precipitation / ice bins that evaporated
//           back to the aerosol bins (findDry4Wet :408-470 for
//           precipitation and ice), aerosols below 0.1 nm dry diameter; then
//           the diagnostic sums rc, srp, snrp, rtot, ri, riri on every cell.
//           One thread per cell; per cell the bins are visited in the
//           reference's order, so the result does not depend on the grid
//           traversal.
//
// The passes are HBM-bound byte movers: no tiling, coalesced along k (the
// fastest LES index).  Floating-point contraction is off so that every sum
// and product rounds as written.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "salsa_internal.h"

#pragma clang fp contract(off)

namespace salsa {

namespace {

constexpr double kMassTH = 1.e-25;    // constrain_SALSA.f90:65

__device__ __forceinline__ bool interior_col(const LesArgs &a, int i, int j) {
  return i >= 2 && i <= a.nxp - 3 && j >= 2 && j <= a.nyp - 3;
}

// pass A: one (p, t) array pair; blockIdx.y is the row, x the cell
__global__ void __launch_bounds__(256) k_les_update(LesArgs a, double *__restrict__ p, double *__restrict__ t) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (int)a.n3) return;
  const int k = x % a.nzp, col = x / a.nzp;
  const int i = col % a.nxp, j = col / a.nxp;
  const int64_t e = (int64_t)blockIdx.y * a.n3 + x;
  const double rdt = a.dtlt;
  const double pv = p[e], tv = t[e];
  const double tc = fmax((-1. + 1.e-8) * pv / rdt, tv);
  if (k >= 1 && k <= a.nzp - 2 && interior_col(a, i, j)) p[e] = pv + tc * rdt;
  t[e] = 0.;
}

// pass B, halo columns: one thread per (halo column, k), blockIdx.y the
// row; k fastest, so a wavefront reads and writes along one column.  The
// halo columns are the 2 rows of columns at each y edge (all x) followed by
// the 2 columns at each x edge of the interior y range.
__global__ void __launch_bounds__(256) k_les_halo(LesArgs a, double *__restrict__ p) {
  const int nz = a.nzp, nxp = a.nxp, nyp = a.nyp, nx = nxp - 4, ny = nyp - 4;
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
  double *row = p + (int64_t)blockIdx.y * a.n3;
  const double *src = row + ((int64_t)ii + (int64_t)nxp * jj) * nz;
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
  row[((int64_t)i + (int64_t)nxp * j) * nz + k] = v;
}

// pass B, interior columns: bottom and top rows (sclrset 'mixd'), one thread
// per column, blockIdx.y the row
__global__ void __launch_bounds__(256) k_les_rows(LesArgs a, double *__restrict__ p) {
  const int nz = a.nzp, nx = a.nxp - 4, ny = a.nyp - 4;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = 2 + c % nx, j = 2 + c / nx;
  double *col = p + (int64_t)blockIdx.y * a.n3 + ((int64_t)i + (int64_t)a.nxp * j) * nz;
  const double s1 = col[1], s2 = col[2], sm = col[nz - 2], sm2 = col[nz - 3];
  col[0] = s1 - 0. * (s2 - s1);
  col[nz - 1] = sm + a.dzf2 * (sm - sm2);
}

// findDry4Wet for an evaporating bin (number ne, dry masses md[s * nbev])
__device__ int les_dry4wet(const LesArgs &a, int64_t x, double ne, const double *md, int64_t mstride) {
  const int ndry = a.ndry, nba = a.nb[0];
  const double *na = a.n_p[0] + x;
  const double *ma = a.m_p[0] + x;
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
  if (na[(int64_t)bb * a.n3] <= kNlim) return ba;
  if (na[(int64_t)ba * a.n3] <= kNlim) return bb;
  // calc_correlation of the dry masses; sxy keeps the last product only
  // (util.f90:1017)
  double r[2];
  for (int q = 0; q < 2; ++q) {
    const int bx = q ? bb : ba;
    double sx = 0., sy = 0., sx2 = 0., sy2 = 0., sxy = 0.;
    for (int s = 0; s < ndry; ++s) {
      const double xv = ma[((int64_t)s * nba + bx) * a.n3], yv = md[s * mstride];
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

// pass D: one thread per cell
__global__ void __launch_bounds__(256) k_les_diag(LesArgs a) {
  const int ncat = a.level == 5 ? 4 : 3;
  const int nwet = a.nwet, ndry = a.ndry, iwa = a.iwa, irim = a.irim;
  const int64_t n3 = a.n3;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n3; x += (int64_t)gridDim.x * blockDim.x) {
    // remove negative values (:88-100); only values that change are written.
    // Rows are loaded 8 at a time before any store so that the loads are in
    // flight together.
    for (int c = 0; c < ncat; ++c)
      for (int q = 0; q < 2; ++q) {
        double *base = (q ? a.m_p[c] : a.n_p[c]) + x;
        const int nrow = q ? a.nb[c] * a.nsw[c] : a.nb[c];
        for (int r0 = 0; r0 < nrow; r0 += 8) {
          double v[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) v[u] = r0 + u < nrow ? base[(int64_t)(r0 + u) * n3] : 0.;
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (r0 + u < nrow && !(v[u] >= 0.)) base[(int64_t)(r0 + u) * n3] = fmax(0., v[u]);
        }
      }
    const int64_t col = x / a.nzp;
    const int i = (int)(col % a.nxp), j = (int)(col / a.nxp);
    if (interior_col(a, i, j)) {
      auto N = [&](int c, int b) -> double & { return a.n_p[c][(int64_t)b * n3 + x]; };
      auto M = [&](int c, int s, int b) -> double & { return a.m_p[c][((int64_t)s * a.nb[c] + b) * n3 + x]; };
      auto ctime = [&](int row) {
        if (a.charge_time) a.charge_time[(int64_t)row * n3 + x] = 0.;
      };
      // particles with number but no mass (:103-154)
      for (int c = 0; c < 2; ++c)
        for (int b = 0; b < a.nb[c]; ++b) {
          double sm = 0.;
          for (int s = 0; s < nwet; ++s) sm = sm + M(c, s, b);
          if (N(c, b) > 0. && sm <= 0.) {
            N(c, b) = 0.;
            for (int s = 0; s < nwet; ++s) M(c, s, b) = 0.;
            ctime(c == 0 ? b : a.nb[0] + b);
          }
        }
      for (int b = 0; b < a.nb[2]; ++b)
        if (N(2, b) > 0. && M(2, iwa, b) <= 0.) {
          N(2, b) = 0.;
          for (int s = 0; s < nwet; ++s) M(2, s, b) = 0.;
          ctime(a.nb[0] + a.nb[1] + b);
        }
      if (a.level == 5)
        for (int b = 0; b < a.nb[3]; ++b) {
          const double mt = M(3, iwa, b) + M(3, irim, b);
          if (N(3, b) > 0. && mt < kMassTH) {
            N(3, b) = 0.;
            for (int s = 0; s <= irim; ++s) M(3, s, b) = 0.;
          }
        }
      const double T = a.temp[x], rh = a.rh[x];
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
          const int ba = les_dry4wet(a, x, N(2, bc), &M(2, 0, bc), (int64_t)a.nb[2] * n3);
          N(0, ba) = N(0, ba) + N(2, bc);
          N(2, bc) = 0.;
          for (int s = 0; s < nwet; ++s) {
            M(0, s, ba) = M(0, s, ba) + M(2, s, bc);
            M(2, s, bc) = 0.;
          }
          ctime(a.nb[0] + a.nb[1] + bc);
        }
      }
      // ice that sublimated (:260-315)
      if (a.level == 5) {
        const double rhi = a.rhi[x];
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
            const int ba = les_dry4wet(a, x, N(3, bc), &M(3, 0, bc), (int64_t)a.nb[3] * n3);
            N(0, ba) = N(0, ba) + N(3, bc);
            N(3, bc) = 0.;
            for (int s = 0; s < nwet; ++s) {
              M(0, s, ba) = M(0, s, ba) + M(3, s, bc);
              M(3, s, bc) = 0.;
            }
            M(0, iwa, ba) = M(0, iwa, ba) + M(3, irim, bc);
            M(3, irim, bc) = 0.;
          }
        }
      }
      // aerosols below 0.1 nm dry diameter (:317-366); no gas-phase transfer
      for (int b = 0; b < a.nb[0]; ++b) {
        if (!(N(0, b) > kNlim)) continue;
        double zvol = 0.;
        for (int s = 0; s < ndry; ++s) zvol = zvol + M(0, s, b) / a.rholiq[s];
        zvol = zvol / N(0, b);
        if (zvol < kPi6 * 1.e-30) {
          for (int s = 0; s < nwet; ++s) M(0, s, b) = 0.;
          N(0, b) = 0.;
        }
      }
    }
    // diagnostic tracers (:370-400)
    const double *ma = a.m_p[0] + (int64_t)iwa * a.nb[0] * n3 + x;
    const double *mc = a.m_p[1] + (int64_t)iwa * a.nb[1] * n3 + x;
    const double *mp = a.m_p[2] + (int64_t)iwa * a.nb[2] * n3 + x;
    double sa = 0., sc = 0., sp = 0., sn = 0.;
    for (int b = a.in1a; b <= a.fn2b; ++b) sa = sa + ma[(int64_t)b * n3];
    for (int b = a.ica_cur; b <= a.fcb_cur; ++b) sc = sc + mc[(int64_t)b * n3];
    for (int b = 0; b < a.nb[2]; ++b) sp = sp + mp[(int64_t)b * n3];
    for (int b = 0; b < a.nb[2]; ++b) sn = sn + a.n_p[2][(int64_t)b * n3 + x];
    const double rc = sa + sc;
    double rtot = rc + sp + a.rp[x];
    a.diag[x] = rc;
    a.diag[n3 + x] = sp;
    a.diag[2 * n3 + x] = sn;
    if (a.level == 5) {
      const double *mi = a.m_p[3] + (int64_t)iwa * a.nb[3] * n3 + x;
      const double *mr = a.m_p[3] + (int64_t)irim * a.nb[3] * n3 + x;
      double si = 0., sr = 0.;
      for (int b = 0; b < a.nb[3]; ++b) si = si + mi[(int64_t)b * n3];
      for (int b = 0; b < a.nb[3]; ++b) sr = sr + mr[(int64_t)b * n3];
      rtot = rtot + si + sr;
      a.diag[4 * n3 + x] = si;
      a.diag[5 * n3 + x] = sr;
    }
    a.diag[3 * n3 + x] = rtot;
  }
}

unsigned grid_for(int64_t n, int block) {
  const int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 65536 ? (g > 0 ? g : 1) : 65536);
}

}  // namespace

int launch_les_update(const LesArgs &a, hipStream_t s) {
  const int ncat = a.level == 5 ? 4 : 3;
  const unsigned gx = (unsigned)((a.n3 + 255) / 256);
  for (int c = 0; c < ncat; ++c) {
    hipLaunchKernelGGL(k_les_update, dim3(gx, a.nb[c]), dim3(256), 0, s, a, a.n_p[c], a.n_t[c]);
    hipLaunchKernelGGL(k_les_update, dim3(gx, a.nb[c] * a.nsw[c]), dim3(256), 0, s, a, a.m_p[c], a.m_t[c]);
  }
  const unsigned gh = (unsigned)(((4 * a.nxp + 4 * (a.nyp - 4)) * a.nzp + 255) / 256);
  const unsigned gr = (unsigned)(((a.nxp - 4) * (a.nyp - 4) + 255) / 256);
  for (int c = 0; c < ncat; ++c) {
    for (int q = 0; q < 2; ++q) {
      double *arr = q ? a.m_p[c] : a.n_p[c];
      const unsigned nrow = (unsigned)(q ? a.nb[c] * a.nsw[c] : a.nb[c]);
      hipLaunchKernelGGL(k_les_halo, dim3(gh, nrow), dim3(256), 0, s, a, arr);
      hipLaunchKernelGGL(k_les_rows, dim3(gr, nrow), dim3(256), 0, s, a, arr);
    }
  }
  hipLaunchKernelGGL(k_les_diag, dim3(grid_for(a.n3, 256)), dim3(256), 0, s, a);
  return (int)hipGetLastError();
}

}  // namespace salsa
