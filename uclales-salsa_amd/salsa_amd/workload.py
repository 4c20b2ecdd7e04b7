"""Seeded synthetic LES fields for the SALSA coagulation + condensation step.

Every field is a pure function of (seed, stream, linear cell index) through a
splitmix64 counter hash, so any consumer can regenerate identical inputs.
Arrays use the LES FieldArray layout: 3-D fields are flat (nzp*nxp*nyp,)
vectors in Fortran order (k fastest); per-bin fields are [nbin, ncell3d]
(bin-major), mass fields [nspec*nbin, ncell3d] with row s*nb + b
(src/src_LES/util.f90:955-961).

The aerosol spectrum follows size_distribution (src/src_salsa/
mo_salsa_sizedist.f90:49-68); saturation mixing ratios follow rslf/rsif
(src/src_LES/thrm.f90:345-387); air density follows thrm.f90:165.
"""
from dataclasses import dataclass, field

import numpy as np

PI6 = 0.5235988          # mo_submctl.f90:266
PI = 3.1415927           # mo_submctl.f90:265
RHOSU, MSU = 1830., 98.08e-3
RHOWA, MWA = 1000., 18.016e-3
RHOIC, RHORI = 917., 400.
RHODU = 2650.
SURFW0, BOLTZ, AVOG = 0.073, 1.3807e-23, 6.0221e+23
MVWA = (MWA / AVOG) / RHOWA

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = x.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        x += np.uint64(0x9E3779B97F4A7C15)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, stream, idx):
    """U[0,1) per (seed, stream, idx): splitmix64(seed*2^40 + idx) xor-mixed by stream."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = np.uint64(seed) * np.uint64(1 << 40) + idx
        key = key ^ (np.uint64(stream) * np.uint64(0xD1B54A32D192ED03))
    z = _splitmix64(key)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, stream, idx):
    u1 = uniform(seed, 2 * stream + 1000, idx)
    u2 = uniform(seed, 2 * stream + 1001, idx)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def rslf(p, t):
    """thrm.f90:345-365"""
    c = [0.6105851e+03, 0.4440316e+02, 0.1430341e+01, 0.2641412e-01, 0.2995057e-03,
         0.2031998e-05, 0.6936113e-08, 0.2564861e-11, -.3704404e-13]
    x = np.minimum(np.maximum(-80., t - 273.16), 50.)
    esl = c[8]
    for k in range(7, -1, -1):
        esl = c[k] + x * esl
    return .622 * esl / (p - esl)


def rsif(p, t):
    """thrm.f90:369-387"""
    c = [0.6114327e+03, 0.5027041e+02, 0.1875982e+01, 0.4158303e-01, 0.5992408e-03,
         0.5743775e-05, 0.3566847e-07, 0.1306802e-09, 0.2152144e-12]
    x = np.minimum(np.maximum(-80., t - 273.16), 0.)
    esi = c[8]
    for k in range(7, -1, -1):
        esi = c[k] + x * esi
    return .622 * esi / (p - esi)


@dataclass
class Grid:
    nz: int
    nx: int
    ny: int
    hz: int = 1
    hxy: int = 2

    @property
    def nzp(self):
        return self.nz + 2 * self.hz

    @property
    def nxp(self):
        return self.nx + 2 * self.hxy

    @property
    def nyp(self):
        return self.ny + 2 * self.hxy

    @property
    def ncell3d(self):
        return self.nzp * self.nxp * self.nyp

    @property
    def ncells(self):
        return self.nz * self.nx * self.ny

    def interior_index(self):
        """Flat (Fortran-order) indices of interior cells, interior order k fastest."""
        k = np.arange(self.nz) + self.hz
        i = np.arange(self.nx) + self.hxy
        j = np.arange(self.ny) + self.hxy
        J, I, K = np.meshgrid(j, i, k, indexing="ij")
        return (K + self.nzp * (I + self.nxp * J)).ravel()

    def interior_k(self):
        return np.tile(np.arange(self.nz), self.nx * self.ny)


@dataclass
class Fields:
    grid: Grid
    level: int
    press: np.ndarray
    temp: np.ndarray
    rsl: np.ndarray
    rsi: np.ndarray
    dn: np.ndarray
    rv: np.ndarray
    n: list = field(default_factory=list)     # [cat] -> [nb, ncell3d]
    m: list = field(default_factory=list)     # [cat] -> [ns*nb, ncell3d]
    cloudy: np.ndarray = None                 # [ncells] bool, interior order


def _layout_counts(L):
    nice = L.nice
    return L.nbins, L.ncld, L.nprc, nice


def _aerosol_spectrum(L, modes):
    """Bin number fractions from lognormal modes, as size_distribution does
    (10 sub-intervals per bin, mo_salsa_sizedist.f90:49-68).  Returns
    [nbins] number concentration per unit total, for bins in1a..fn2a."""
    out = np.zeros(L.nbins)
    for kk in range(L.in1a, L.fn2a + 1):
        dlo = (L.vlolim[kk] / PI6) ** (1. / 3.)
        dhi = (L.vhilim[kk] / PI6) ** (1. / 3.)
        dd = (dhi - dlo) / 10
        acc = 0.0
        for ib in range(10):
            d1 = dlo + ib * dd
            d2 = d1 + dd
            dm = (d1 + d2) / 2
            dlp = np.log(d2 / d1)
            for (n, dpg, sg) in modes:
                acc += n * dlp / (np.sqrt(2. * np.pi) * np.log(sg)) * \
                    np.exp(-np.log(dm / dpg) ** 2 / (2. * np.log(sg) ** 2))
        out[kk] = acc
    return out


def _equilibrium_water(vso4_per, numc_like, T, rh):
    """Aerosol water volume per particle from the ZSR binary molality of
    sulphate with a Kelvin fixed point (the regime-2 relation of
    mo_salsa_properties.f90:305-360, iterated a fixed 40 times)."""
    rh = np.clip(rh, 0.37, 0.98)
    zke = np.full_like(vso4_per, 1.02)
    d = np.cbrt(vso4_per / PI6)
    for _ in range(40):
        aw = rh / zke
        bm1 = 1.1065495e+2 - 3.6759197e+2 * aw + 5.0462934e+2 * aw ** 2 \
            - 3.1543839e+2 * aw ** 3 + 6.770824e+1 * aw ** 4
        lwc_per = vso4_per * (RHOSU / MSU) / bm1       # kg water per particle
        d = np.cbrt(lwc_per / RHOWA / PI6 + vso4_per / PI6)
        zke = np.exp(2. * SURFW0 * MVWA / (BOLTZ * T * d))
    return lwc_per / RHOWA


def generate(L, grid, seed, level=4, t_range=(275., 295.), p_range=(900e2, 1000e2),
             rh_range=(0.95, 1.01), cloud_mode="random", cloud_frac=0.3,
             cloud_slab=(60, 90), precip_in_cloud=False, ice=False):
    """Build seeded LES-layout fields for layout L (bin descriptor with the
    attribute names of salsa_bin_desc).  cloud_mode: 'random' (seeded
    fraction), 'slab' (interior k in [cloud_slab)), or 'none'."""
    n3 = grid.ncell3d
    nb, nc, npr, ni = _layout_counts(L)
    idx = grid.interior_index()
    cid = np.arange(grid.ncells, dtype=np.uint64)

    def fill3(val_int, halo):
        a = np.full(n3, halo, dtype=np.float64)
        a[idx] = val_int
        return a

    T = t_range[0] + (t_range[1] - t_range[0]) * uniform(seed, 1, cid)
    P = p_range[0] + (p_range[1] - p_range[0]) * uniform(seed, 2, cid)
    RH = rh_range[0] + (rh_range[1] - rh_range[0]) * uniform(seed, 3, cid)
    rs = rslf(P, T)
    rsi_ = rsif(P, T)
    rv = RH * rs
    dn = P / (287.04 * T * (1. + 0.61 * rv))     # thrm.f90:165, defs.f90:25

    if cloud_mode == "slab":
        kk = grid.interior_k()
        cloudy = (kk >= cloud_slab[0]) & (kk < cloud_slab[1])
    elif cloud_mode == "random":
        cloudy = uniform(seed, 4, cid) < cloud_frac
    else:
        cloudy = np.zeros(grid.ncells, dtype=bool)

    # ---- aerosol (#/m3, m3/m3) ----
    modes_a = [(100.e6, 40.e-9, 1.5), (300.e6, 150.e-9, 1.6)]
    frac_a = _aerosol_spectrum(L, modes_a)
    naer = np.zeros((nb, grid.ncells))
    for kk in range(L.in1a, L.fn2a + 1):
        naer[kk] = frac_a[kk]
    for k in range(L.in2b, L.fn2b + 1):
        naer[k] = 0.25 * frac_a[L.in2a + (k - L.in2b)]
    # 5 % seeded per-cell variability of the spectrum amplitude
    naer *= (1. + 0.05 * (2. * uniform(seed, 5, cid) - 1.))[None, :]
    vso4_per = np.array([PI6 * L.dmid[k] ** 3 for k in range(nb)])
    vaer_so4 = naer * vso4_per[:, None]
    vaer_h2o = np.zeros_like(vaer_so4)
    for k in range(nb):
        w = _equilibrium_water(np.full(grid.ncells, vso4_per[k]), None, T, RH)
        vaer_h2o[k] = naer[k] * w
    vaer_h2o[naer <= 1.0] = 0.0

    nspec = L.nspec
    iwa = L.iwa
    dry = list(L.dry_species[:L.nspec_dry])
    iso4 = dry.index(0) if 0 in dry else None
    idu = dry.index(3) if 3 in dry else None

    # ---- cloud ----
    ncl = np.zeros((nc, grid.ncells))
    vcl_dry = np.zeros((nc, grid.ncells))
    vcl_h2o = np.zeros((nc, grid.ncells))
    if cloudy.any():
        Nc = (50.e6 + 150.e6 * uniform(seed, 6, cid))
        par = [L.ica.par + q for q in range(L.fca.cur - L.ica.cur + 1)] + \
              [L.icb.par + q for q in range(L.fcb.cur - L.icb.cur + 1)]
        w = np.array([naer[p] for p in par])          # [nc, ncells]
        wsum = w.sum(axis=0)
        for c in range(nc):
            p = par[c]
            ncell_c = np.where(cloudy, Nc * w[c] / wsum, 0.)
            dwet = 12.e-6 * np.exp(np.log(1.3) * normal(seed, 20 + c, cid))
            core = vso4_per[p]
            wv = np.maximum(PI6 * dwet ** 3 - core, 0.)
            ncl[c] = ncell_c
            vcl_dry[c] = ncell_c * core
            vcl_h2o[c] = ncell_c * wv

    # ---- precipitation ----
    npr_ = np.zeros((npr, grid.ncells))
    vpr_dry = np.zeros((npr, grid.ncells))
    vpr_h2o = np.zeros((npr, grid.ncells))
    if precip_in_cloud and cloudy.any():
        for q in range(npr):
            b = L.nbins + L.ncld + q
            nq = np.where(cloudy, 1.e3 + 9.e3 * uniform(seed, 40 + q, cid), 0.)
            d = L.dmid[b]
            npr_[q] = nq
            vpr_h2o[q] = nq * PI6 * d ** 3
            vpr_dry[q] = vpr_h2o[q] * 1.e-6

    # ---- ice (level 5) ----
    nic = np.zeros((ni, grid.ncells))
    vic_pri = np.zeros((ni, grid.ncells))
    vic_rim = np.zeros((ni, grid.ncells))
    vic_dry = np.zeros((ni, grid.ncells))
    if ice and ni > 0:
        for q in range(ni):
            b = L.nbins + L.ncld + L.nprc + q
            nq = (1.e3 + 9.e3 * uniform(seed, 60 + q, cid)) * np.exp(-0.5 * q)
            fr = 0.5 * uniform(seed, 70 + q, cid)            # rime mass fraction
            d = L.dmid[b]
            mtot = nq * RHOIC * PI6 * d ** 3
            vic_rim[q] = fr * mtot / RHORI
            vic_pri[q] = (1. - fr) * mtot / RHOIC
            vic_dry[q] = vic_pri[q] * 1.e-6
            nic[q] = nq

    # ---- dust (level 5 configs): lognormal 10 cm-3, Dg 500 nm, sigma 1.6 ----
    vaer_du = np.zeros_like(vaer_so4)
    if idu is not None:
        frac_d = _aerosol_spectrum(L, [(10.e6, 500.e-9, 1.6)])
        for kk in range(L.in2b, L.fn2b + 1):
            nd = frac_d[L.in2a + (kk - L.in2b)]
            naer[kk] += nd
            vaer_du[kk] = nd * vso4_per[kk]

    # ---- convert to LES units and layout ----
    dn3 = fill3(dn, 1.0)
    dn_i = dn
    f = Fields(grid=grid, level=level,
               press=fill3(P, 1.e5), temp=fill3(T, 280.), rsl=fill3(rs, rslf(1.e5, 280.)),
               rsi=fill3(rsi_, rsif(1.e5, 280.)), dn=dn3, rv=fill3(rv, 0.),
               cloudy=cloudy)

    def put_n(arr):
        out = np.zeros((arr.shape[0], n3))
        out[:, idx] = arr / dn_i[None, :]
        return out

    def put_m(vols, rhos, ns):
        nbc = vols[0].shape[0]
        out = np.zeros((ns * nbc, n3))
        for s, (v, rho) in enumerate(zip(vols, rhos)):
            if v is None:
                continue
            out[s * nbc:(s + 1) * nbc, idx] = v * rho / dn_i[None, :]
        return out

    def species_vols(vdry_so4, vdry_du, vh2o, nbc, rime=None, icecat=False):
        vols, rhos = [], []
        for s in range(L.nspec_dry):
            sid = dry[s]
            if sid == 0:
                vols.append(vdry_so4)
            elif sid == 3 and vdry_du is not None:
                vols.append(vdry_du)
            else:
                vols.append(np.zeros((nbc, grid.ncells)))
            rhos.append(L.rhoice[s] if icecat else L.rholiq[s])
        vols.append(vh2o)
        rhos.append(RHOIC if icecat else RHOWA)
        if icecat:
            vols.append(rime)
            rhos.append(RHORI)
        return vols, rhos

    nwet = nspec - 1
    va, ra = species_vols(vaer_so4, vaer_du, vaer_h2o, nb)
    vc, rc = species_vols(vcl_dry, None, vcl_h2o, nc)
    vp, rp = species_vols(vpr_dry, None, vpr_h2o, npr)
    f.n = [put_n(naer), put_n(ncl), put_n(npr_)]
    f.m = [put_m(va, ra, nwet), put_m(vc, rc, nwet), put_m(vp, rp, nwet)]
    if level == 5:
        vi, ri = species_vols(vic_dry, None, vic_pri, ni, rime=vic_rim, icecat=True)
        f.n.append(put_n(nic))
        f.m.append(put_m(vi, ri, nspec))
    else:
        f.n.append(None)
        f.m.append(None)
    return f


CONFIGS = {
    # benchmark workloads 1-5 at the reference-legal layout nbin=(3,4): 11/8/7 (+ice 7)
    1: dict(grid=(40, 8, 8), seed=1, level=4, cloud_mode="random", cloud_frac=0.3),
    2: dict(grid=(120, 64, 64), seed=2, level=4, cloud_mode="slab", cloud_slab=(60, 90)),
    3: dict(grid=(80, 32, 32), seed=3, level=4, cloud_mode="random", cloud_frac=0.3),
    4: dict(grid=(100, 128, 128), seed=4, level=4, cloud_mode="random", cloud_frac=0.15,
            precip_in_cloud=True),
    5: dict(grid=(100, 128, 128), seed=5, level=5, cloud_mode="random", cloud_frac=0.3,
            t_range=(240., 275.), ice=True),
}


def config_fields(L, cfg, grid_override=None):
    c = dict(CONFIGS[cfg])
    nz, nx, ny = grid_override or c.pop("grid")
    c.pop("grid", None)
    g = Grid(nz, nx, ny)
    return generate(L, g, **c)


def to_salsa_units(L, f, level=None):
    """Pack interior cells into SALSA units exactly as run_SALSA does
    (mo_salsa_driver.f90:230-279): numc = N*dn; volc = m*dn/rho.  Returns
    (T, p, rs, rsi, rv, numc[ncells, ntotal], volc[ncells, ntotal, 9]) where
    ntotal covers the categories present in f."""
    level = f.level if level is None else level
    idx = f.grid.interior_index()
    dn = f.dn[idx]
    cats = 4 if level == 5 else 3
    nbs = [L.nbins, L.ncld, L.nprc, L.nice][:cats]
    nt = sum(nbs)
    n = len(idx)
    numc = np.zeros((n, nt))
    volc = np.zeros((n, nt, 9))
    off = 0
    nwet = L.nspec - 1
    for c in range(cats):
        nbc = nbs[c]
        for b in range(nbc):
            numc[:, off + b] = f.n[c][b, idx] * dn
            for s in range(nwet):
                rho = L.rhoice[s] if c == 3 else L.rholiq[s]
                volc[:, off + b, s] = f.m[c][s * nbc + b, idx] * dn / rho
            if c == 3:
                volc[:, off + b, L.irim] = f.m[c][L.irim * nbc + b, idx] * dn / RHORI
        off += nbc
    return (f.temp[idx].copy(), f.press[idx].copy(), f.rsl[idx].copy(), f.rsi[idx].copy(),
            f.rv[idx].copy(), numc, volc)
