"""Shared helpers of the test suite: paths, golden loaders, error metrics,
and a flat-grid packer that feeds arbitrary SALSA-unit cells to the C ABI."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (os.path.join(ROOT, "uclales-salsa_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

# error floors: below these magnitudes a value is physically empty
FLOOR_N = 1.e-3        # #/m3
FLOOR_V = 1.e-24       # m3/m3
FLOOR_RV = 1.e-12      # kg/kg


def rel(a, b, floor):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def golden(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def oracle_layout_for(g):
    import oracle as O
    species = tuple(int(s) for s in g["species"])
    return O.make_layout(nbin=tuple(int(x) for x in g["nbin"]), nprc=int(g["nprc"]), nice=int(g["nice"]),
                         species=species)


def oracle_params_for(g):
    import oracle as O
    kw = {}
    if int(g["lcharge"]):
        kw.update(lcharge=1, charge_coll_enh=float(g["charge_enh"]))
    if not int(g["lsauto"]):
        kw.update(lsauto_state=0)
    if "lsauto_mode" in g:
        kw.update(lsauto_mode=int(g["lsauto_mode"]))
    if "distr" in g and int(g["distr"]):
        kw.update(lsdistupdate=1)
    if "lsactiv" in g and int(g["lsactiv"]):
        kw.update(lsactiv=1)
    return O.make_params(level=int(g["level"]), **kw)


def full_volc(volc, ns=None):
    n, nt, k = volc.shape
    out = np.zeros((n, nt, 9))
    out[:, :, :k] = volc
    return out


class FlatCells:
    """SALSA-unit cells laid out as a flat (nzp = n, nxp = nyp = 1, no halo)
    LES grid with dn = 1 kg/m3 per cell scaled by `dn` so the kernel's pack
    reproduces the given numc / volc to rounding."""

    def __init__(self, desc, level, T, p, rs, rsi, rv, numc, volc, dn=None):
        n = len(T)
        self.n = n
        self.level = level
        ncat = 4 if level == 5 else 3
        nbs = [desc.nbins, desc.ncld, desc.nprc, desc.nice][:ncat]
        self.nbs = nbs
        dn = np.ones(n) if dn is None else np.asarray(dn, dtype=np.float64)
        self.f = dict(press=np.ascontiguousarray(p, dtype=np.float64),
                      temp=np.ascontiguousarray(T, dtype=np.float64),
                      rsl=np.ascontiguousarray(rs, dtype=np.float64),
                      rsi=np.ascontiguousarray(rsi, dtype=np.float64),
                      dn=np.ascontiguousarray(dn), rv=np.ascontiguousarray(rv, dtype=np.float64).copy())
        self.n_p, self.m_p = [], []
        off = 0
        nwet = desc.nspec - 1
        for c in range(ncat):
            nb = nbs[c]
            N = np.ascontiguousarray((numc[:, off:off + nb] / dn[:, None]).T)
            nsp = nwet if c < 3 else desc.nspec
            M = np.zeros((nsp * nb, n))
            for s in range(nsp):
                if c < 3:
                    rho = desc.rholiq[s]
                else:
                    rho = 400. if s == desc.irim else desc.rhoice[s]
                M[s * nb:(s + 1) * nb] = (volc[:, off:off + nb, s] * rho / dn[:, None]).T
            self.n_p.append(N)
            self.m_p.append(np.ascontiguousarray(M))
            off += nb
        self.nt = off
        self.desc = desc

    def unpack(self, n_p, m_p, rv):
        """LES-unit arrays back to SALSA units with the kernel's pack arithmetic."""
        desc = self.desc
        dn = self.f["dn"]
        numc = np.zeros((self.n, self.nt))
        volc = np.zeros((self.n, self.nt, 9))
        off = 0
        nwet = desc.nspec - 1
        for c, nb in enumerate(self.nbs):
            numc[:, off:off + nb] = (n_p[c] * dn[None, :]).T
            nsp = nwet if c < 3 else desc.nspec
            for s in range(nsp):
                if c < 3:
                    rho = desc.rholiq[s]
                else:
                    rho = 400. if s == desc.irim else desc.rhoice[s]
                volc[:, off:off + nb, s] = (m_p[c][s * nb:(s + 1) * nb] * dn[None, :] / rho).T
            off += nb
        return np.asarray(rv), numc, volc

    def params_grid(self):
        return self.n, 1, 1


def oracle_sensitivity(oracle_mod, L, P, T, p, rs, rsi, rv, numc, volc, eps=1e-13, seed=0, kernel_eps=0.0,
                       nsteps=1, init=False, kseeds=(1, 2)):
    """Spread of the oracle's own outputs under a relative input perturbation
    of size eps: the part of the reference's answer that its own rounding
    does not determine.  With kernel_eps > 0 the spread under a random
    per-pair relative kernel perturbation of that size (one run per seed in
    kseeds) is folded in as well: what FP32-rounded collision kernels cannot
    determine.
    Returns (d_rv, d_numc, d_volc) relative spreads."""
    rng = np.random.default_rng(seed)
    base = oracle_mod.step_cells(L, P, T, p, rs, rsi, rv, numc, volc, nsteps=nsteps, init=init)
    runs = [oracle_mod.step_cells(L, P, T, p, rs, rsi, rv * (1 + eps * rng.standard_normal(len(rv))),
                                  numc * (1 + eps * rng.standard_normal(numc.shape)),
                                  volc * (1 + eps * rng.standard_normal(volc.shape)), nsteps=nsteps, init=init)]
    if kernel_eps > 0:
        for ks in kseeds:
            Pk = type(P).from_buffer_copy(P)
            Pk.kpert = kernel_eps
            Pk.kseed = ks
            runs.append(oracle_mod.step_cells(L, Pk, T, p, rs, rsi, rv, numc, volc, nsteps=nsteps, init=init))
    out = []
    for i, fl in ((0, FLOOR_RV), (1, FLOOR_N), (2, FLOOR_V)):
        out.append(np.max([rel(r[i], base[i], fl) for r in runs], axis=0))
    return tuple(out)


# Kernel perturbation that stands for FP32 rounding of the collision kernels
# (about two units in the last place of a float).
KERNEL_EPS_FP32 = 2.0 ** -22
# The largest relative difference between the FP32 and FP64 collision
# kernels measured on workload states (8.2e-7 at level 5, pairs of large
# drops and ice; 5.2e-7 at level 4) is about 2^-20.2.  Quantities that are
# linear in the kernels -- the volume the semi-implicit coagulation update
# fails to conserve across categories -- are held to the oracle's spread at
# that size.
KERNEL_EPS_FP32_MAX = 2.0 ** -20


def content_scale(inputs, ref, cats):
    """Per element, the amount its bin could have held during the step: the
    larger of its input value and of its category's total (per species for
    volumes) in the inputs and in the reference output.  Arguments are
    (rv, numc, volc) triples; cats the [first, last+1) bins of each category."""
    irv, inc, ivc = inputs
    _, onc, ovc = ref
    sn = np.abs(np.array(inc, dtype=np.float64))
    sv = np.abs(np.array(ivc, dtype=np.float64))
    for c0, c1 in cats:
        tn = np.maximum(np.abs(inc[:, c0:c1]).sum(axis=1), np.abs(onc[:, c0:c1]).sum(axis=1))
        tv = np.maximum(np.abs(ivc[:, c0:c1]).sum(axis=1), np.abs(ovc[:, c0:c1]).sum(axis=1))
        sn[:, c0:c1] = np.maximum(sn[:, c0:c1], tn[:, None])
        sv[:, c0:c1] = np.maximum(sv[:, c0:c1], tv[:, None, :])
    return np.abs(irv), sn, sv


def check_condition_aware(got, ref, sens, tol, ns, label="", inputs=None):
    """Per element: |got - ref| <= max(tol, 10 * sens) relative.  With
    inputs = content_scale(...) an element also passes when |got - ref| <= 16
    ulp of that scale: what is left of a bin that a process emptied
    (complete activation, N - N_act) is the rounding of the amount it held,
    not a relative quantity.  Returns the raw maxima and the count of
    elements that needed the conditioning allowance."""
    grv, gnc, gvc = got
    orv, onc, ovc = ref
    srv, snc, svc = sens
    ulp16 = 16 * 2.0 ** -52
    ins = inputs if inputs is not None else (np.zeros_like(orv), np.zeros_like(onc), np.zeros_like(ovc))
    out = []
    nflag = 0

    def one(g_, r_, s_, i_, floor, tag):
        e = rel(g_, r_, floor)
        small = np.abs(np.asarray(g_) - np.asarray(r_)) <= ulp16 * np.abs(i_)
        e = np.where(small, 0., e)
        lim = np.maximum(tol, 10. * s_)
        assert np.all(e <= lim), (label, tag, float(e.max()))
        out.append(float(e.max()))
        return int(np.sum(e > tol))

    nflag += one(grv, orv, srv, ins[0], FLOOR_RV, "rv")
    nflag += one(gnc, onc, snc, ins[1], FLOOR_N, "N")
    for s in range(ns):
        nflag += one(gvc[:, :, s], ovc[:, :, s], svc[:, :, s], ins[2][:, :, s], FLOOR_V, "V%d" % s)
    return out, nflag


def les_case(desc, level, grid, seed, dt=1.0):
    """LES fields for the coupling after a step (tend_constrain2,
    update_sclrs, SALSA_diagnostics): a workload state (config 3 at level 4,
    config 5 at level 5) with random tendencies, some pushing below zero,
    and cells planted with every condition SALSA_diagnostics acts on --
    subsaturation with cloud / precipitation / ice bins that lost their water,
    bins with number but no mass, negative values and aerosols below 0.1 nm.
    Returns (shape, dzt, f, n_p, n_t, m_p, m_t) with f = temp, rh, rhi, rp."""
    from salsa_amd import workload as W
    rng = np.random.default_rng(seed)
    f = W.config_fields(desc, 5 if level == 5 else 3, grid_override=grid)
    g = f.grid
    shape = (g.nzp, g.nxp, g.nyp)
    n3 = g.nzp * g.nxp * g.nyp
    ncat = 4 if level == 5 else 3
    nb = [desc.nbins, desc.ncld, desc.nprc, desc.nice][:ncat]
    nwet = desc.iwa + 1
    n_p = [np.array(f.n[c], dtype=np.float64) for c in range(ncat)]
    m_p = [np.array(f.m[c], dtype=np.float64) for c in range(ncat)]
    # precipitation (empty in the workload): drops in a third of the cells
    cells = rng.random(n3) < 0.35
    n_p[2][:, cells] = 10 ** rng.uniform(-7, 3, size=(nb[2], cells.sum()))
    for s in range(nwet):
        scale = 1e-3 if s == desc.iwa else 1e-12
        m_p[2][s * nb[2]:(s + 1) * nb[2], cells] = scale * 10 ** rng.uniform(-8, 0, size=(nb[2], cells.sum()))
    # cloud, precipitation and ice bins that lost most of their water
    for c in range(1, ncat):
        sel = rng.random((nb[c], n3)) < 0.3
        w = m_p[c][desc.iwa * nb[c]:(desc.iwa + 1) * nb[c]]
        w[sel] *= 10 ** rng.uniform(-16, -1, size=sel.sum())
        if c == 3:
            r = m_p[c][desc.irim * nb[c]:(desc.irim + 1) * nb[c]]
            r[sel] *= 10 ** rng.uniform(-16, -1, size=sel.sum())
    # number but no mass
    for c in range(ncat):
        nsw = nwet + (1 if c == 3 else 0)
        sel = rng.random((nb[c], n3)) < 0.02
        n_p[c][sel] = np.maximum(n_p[c][sel], 5.)
        for s in range(nsw):
            m_p[c][s * nb[c]:(s + 1) * nb[c]][sel] = 0.
    # aerosols below 0.1 nm dry diameter
    sel = rng.random((nb[0], n3)) < 0.02
    n_p[0][sel] = np.maximum(n_p[0][sel], 1e3)
    for s in range(desc.iwa):
        m_p[0][s * nb[0]:(s + 1) * nb[0]][sel] *= 1e-30
    # negative values
    for c in range(ncat):
        for a in (n_p[c], m_p[c]):
            neg = rng.random(a.shape) < 0.01
            a[neg] *= -1e-3
    n_t = [a * rng.uniform(-1.5, 0.5, size=a.shape) / dt for a in n_p]
    m_t = [a * rng.uniform(-1.5, 0.5, size=a.shape) / dt for a in m_p]
    rh = np.array(f.rv / f.rsl)
    rhi = np.array(f.rv / f.rsi)
    sub = rng.random(n3) < 0.5
    rh[sub] = rng.uniform(0.9, 0.998, size=sub.sum())
    rhi[sub] = rng.uniform(0.9, 0.998, size=sub.sum())
    fields = dict(temp=np.array(f.temp), rh=rh, rhi=rhi, rp=np.array(f.rv))
    dzt = 1. / np.linspace(20., 35., g.nzp)
    return shape, dzt, fields, n_p, n_t, m_p, m_t


def cell_totals(T, p, rv, volc, ndry, level, nt):
    """Per-cell totals that coagulation and condensation move between bins:
    the volume of every dry species summed over the bins [m3/m3] and, at
    level 4, the water mass (vapour plus condensed, kg/m3).  Returns a list
    of arrays, one per total."""
    out = [volc[:, :nt, s].sum(axis=1) for s in range(ndry)]
    if level == 4:
        rhoair = 28.97e-3 * p / (8.314 * T)
        out.append(rv * rhoair + volc[:, :nt, ndry].sum(axis=1) * 1000.)
    return out


def oracle_total_spread(oracle_mod, L, P, T, p, rs, rsi, rv, numc, volc, level, nt, kernel_eps=0.0, eps=1e-13,
                        seed=0, nsteps=1, kseeds=(1, 2)):
    """Relative spread of the oracle's per-cell totals (cell_totals) under a
    relative input perturbation of size eps and, with kernel_eps > 0, under
    random per-pair kernel perturbations of that size (two seeds): how much
    of the reference's mass bookkeeping its own rounding, and FP32-rounded
    kernels, leave undetermined.  (The semi-implicit coagulation update does
    not conserve volume across categories exactly, so its per-cell loss
    depends on the kernel values.)"""
    rng = np.random.default_rng(seed)
    ndry = L.ndry
    base = oracle_mod.step_cells(L, P, T, p, rs, rsi, rv, numc, volc, nsteps=nsteps)
    runs = [oracle_mod.step_cells(L, P, T, p, rs, rsi, rv * (1 + eps * rng.standard_normal(len(rv))),
                                  numc * (1 + eps * rng.standard_normal(numc.shape)),
                                  volc * (1 + eps * rng.standard_normal(volc.shape)), nsteps=nsteps)]
    if kernel_eps > 0:
        for ks in kseeds:
            Pk = type(P).from_buffer_copy(P)
            Pk.kpert = kernel_eps
            Pk.kseed = ks
            runs.append(oracle_mod.step_cells(L, Pk, T, p, rs, rsi, rv, numc, volc, nsteps=nsteps))
    tb = cell_totals(T, p, base[0], base[2], ndry, level, nt)
    spread = [np.zeros_like(x) for x in tb]
    for r in runs:
        tr = cell_totals(T, p, r[0], r[2], ndry, level, nt)
        for k in range(len(tb)):
            spread[k] = np.maximum(spread[k], np.abs(tr[k] - tb[k]) / np.maximum(np.abs(tb[k]), 1e-300))
    return tb, spread
