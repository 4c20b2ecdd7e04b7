"""ctypes binding of the FP64 CPU oracle (libsalsa_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the thing that
is measured or shipped.  See salsa_oracle.h for the reference citations.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "_build", "libsalsa_oracle.so")

MAXBIN = 64
MAXSPEC = 9
NRATE = 20


class ParBin(C.Structure):
    _fields_ = [("cur", C.c_int), ("par", C.c_int)]


class Layout(C.Structure):
    _fields_ = [
        ("nbin", C.c_int * 2), ("reglim", C.c_double * 4),
        ("nprc", C.c_int), ("prc_dlo", C.c_double), ("prc_vratio", C.c_double),
        ("nice", C.c_int), ("ice_dlo", C.c_double), ("ice_vratio", C.c_double),
        ("nspec_dry", C.c_int), ("dry_species", C.c_int * 7),
        ("in1a", C.c_int), ("fn1a", C.c_int), ("in2a", C.c_int), ("fn2a", C.c_int),
        ("in2b", C.c_int), ("fn2b", C.c_int), ("nbins", C.c_int),
        ("ica", ParBin), ("fca", ParBin), ("icb", ParBin), ("fcb", ParBin),
        ("ncld", C.c_int),
        ("iaero", C.c_int), ("icloud", C.c_int), ("iprecp", C.c_int), ("iice", C.c_int),
        ("ntotal", C.c_int),
        ("nspec", C.c_int), ("nwet", C.c_int), ("ndry", C.c_int), ("iwa", C.c_int), ("irim", C.c_int),
        ("spec_id", C.c_int * MAXSPEC),
        ("rholiq", C.c_double * MAXSPEC), ("rhoice", C.c_double * MAXSPEC),
        ("mm", C.c_double * MAXSPEC), ("diss", C.c_double * MAXSPEC),
        ("phase", C.c_int * MAXBIN),
        ("vlolim", C.c_double * MAXBIN), ("vhilim", C.c_double * MAXBIN), ("dmid", C.c_double * MAXBIN),
        ("nlim", C.c_double * MAXBIN), ("dlim", C.c_double * MAXBIN),
        ("precpbins", C.c_double * MAXBIN),
    ]


class Params(C.Structure):
    _fields_ = [
        ("lscoag", C.c_int), ("lscnd", C.c_int),
        ("lscgaa", C.c_int), ("lscgcc", C.c_int), ("lscgpp", C.c_int), ("lscgii", C.c_int),
        ("lscgca", C.c_int), ("lscgpa", C.c_int), ("lscgia", C.c_int), ("lscgpc", C.c_int),
        ("lscgic", C.c_int), ("lscgip", C.c_int),
        ("lscndh2oae", C.c_int), ("lscndh2ocl", C.c_int), ("lscndh2oic", C.c_int),
        ("lsauto_state", C.c_int), ("lsauto_mode", C.c_int),
        ("eddy_dis", C.c_double), ("Eiagg_max", C.c_double), ("Eiagg_min", C.c_double),
        ("ice_alpha", C.c_double), ("ice_beta", C.c_double), ("ice_gamma", C.c_double),
        ("ice_sigma", C.c_double),
        ("lcharge", C.c_int), ("charge_coll_enh", C.c_double),
        ("kpert", C.c_double), ("kseed", C.c_ulonglong),
        ("lsdistupdate", C.c_int), ("level", C.c_int), ("lsactiv", C.c_int),
        ("diag_growth", C.c_int), ("init", C.c_int),
    ]


class Rate(C.Structure):
    _fields_ = [("volc", C.c_double * MAXSPEC), ("numc", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = C.CDLL(LIBPATH)
        dp = C.POINTER(C.c_double)
        L.so_layout_init.argtypes = [C.POINTER(Layout), C.POINTER(C.c_int), dp, C.c_int, C.c_double,
                                     C.c_double, C.c_int, C.c_double, C.c_double, C.c_int,
                                     C.POINTER(C.c_int)]
        L.so_params_default.argtypes = [C.POINTER(Params), C.c_int]
        L.so_step_cell.argtypes = [C.POINTER(Layout), C.POINTER(Params), C.c_double, C.c_double,
                                   C.c_double, C.c_double, dp, C.c_double, dp, dp, C.POINTER(Rate)]
        L.so_step_cell.restype = C.c_int
        L.so_step_cell_ks.argtypes = L.so_step_cell.argtypes + [C.POINTER(C.c_double), C.c_int]
        L.so_step_cell_ks.restype = C.c_int
        L.so_step_cell_act.argtypes = L.so_step_cell.argtypes + [C.c_double, C.POINTER(C.c_double),
                                                                C.POINTER(C.c_double)]
        L.so_step_cell_act.restype = C.c_int
        L.so_step_batch.argtypes = [C.POINTER(Layout), C.POINTER(Params), C.c_long, dp, dp, dp, dp, dp,
                                    C.c_double, C.c_int, dp, dp]
        L.so_step_batch.restype = C.c_int
        L.so_coagc.argtypes = [C.POINTER(Layout), C.POINTER(Params), dp, dp, C.c_int, C.c_int,
                               C.c_double, C.c_double]
        L.so_coagc.restype = C.c_double
        L.so_terminal_vel_liquid.argtypes = [C.c_double] * 5
        L.so_terminal_vel_liquid.restype = C.c_double
        L.so_acth2o.argtypes = [C.POINTER(Layout), dp, dp]
        L.so_acth2o.restype = C.c_double
        L.so_rslf.argtypes = [C.c_double, C.c_double]
        L.so_rslf.restype = C.c_double
        L.so_rsif.argtypes = [C.c_double, C.c_double]
        L.so_rsif.restype = C.c_double
        P4 = C.c_void_p * 4
        L.so_les_update.argtypes = [C.POINTER(Layout), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    dp, dp, dp, dp, P4, P4, P4, P4, C.c_void_p, dp]
        L.so_les_update.restype = C.c_int
        L.so_les_update_ext.argtypes = L.so_les_update.argtypes + [C.c_void_p, C.c_void_p, C.c_int]
        L.so_les_update_ext.restype = C.c_int
        L.so_les_corr.argtypes = [dp, dp, C.c_int]
        L.so_les_corr.restype = C.c_double
        P3 = C.c_void_p * 3
        L.so_sedim.argtypes = [C.POINTER(Layout), C.c_int, C.c_int, C.c_int, C.c_double, dp, dp, dp, dp, dp,
                               P3, P3, P3, P3, dp, dp, dp, dp, C.c_int, C.c_int, C.c_int]
        L.so_sedim.restype = C.c_int
        L.so_terminal_vel.argtypes = [C.c_double] * 5
        L.so_terminal_vel.restype = C.c_double
        L.so_sedim5.argtypes = [C.POINTER(Layout), C.POINTER(Params), C.c_int, C.c_int, C.c_int, C.c_double,
                                dp, dp, dp, dp, dp, P4, P4, P4, P4, dp, dp, dp, dp, dp, dp, dp,
                                C.c_int, C.c_int, C.c_int, C.c_int]
        L.so_sedim5.restype = C.c_int
        L.so_sed_ice_vc.argtypes = [C.POINTER(Layout), C.POINTER(Params), C.c_double, C.c_double, C.c_double, dp,
                                    dp]
        L.so_sed_ice_vc.restype = None
        L.so_act_cloud_base.argtypes = [C.POINTER(Layout), C.c_double, C.c_double, C.c_double, dp, dp, dp, dp]
        L.so_act_cloud_base.restype = None
        _lib = L
    return _lib


def make_layout(nbin=(3, 4), reglim=(3.e-9, 9.703e-8, 1.e-5, 1.e-5), nprc=7, prc_dlo=20.e-6,
                prc_vratio=4., nice=7, ice_dlo=2.e-6, ice_vratio=20., species=(0,)):
    L = Layout()
    nb = (C.c_int * 2)(*nbin)
    rg = (C.c_double * 4)(*reglim)
    sp = (C.c_int * 7)(*(list(species) + [0] * (7 - len(species))))
    rc = lib().so_layout_init(C.byref(L), nb, rg, nprc, prc_dlo, prc_vratio, nice, ice_dlo,
                              ice_vratio, len(species), sp)
    if rc != 0:
        raise ValueError("so_layout_init failed: %d" % rc)
    return L


def make_params(level=4, **kw):
    P = Params()
    lib().so_params_default(C.byref(P), level)
    for k, v in kw.items():
        setattr(P, k, v)
    return P


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def step_cells_activ(L, P, T, p, rs, rsi, rv, numc, volc, w, dt=1.0):
    """One step per cell (SALSA units) that also returns the cloud-base
    activation of the step (P.lsactiv = 2, updraft w per cell): (rv, numc,
    volc, status, nact[ncell, ncld], vact[ncell, ncld, 9])."""
    lb = lib()
    n = len(T)
    rv = np.array(rv, dtype=np.float64, copy=True)
    numc = np.ascontiguousarray(numc, dtype=np.float64).copy()
    volc = np.ascontiguousarray(volc, dtype=np.float64).copy()
    status = np.zeros(n, dtype=np.int32)
    nact, vact = np.zeros((n, L.ncld)), np.zeros((n, L.ncld, MAXSPEC))
    for i in range(n):
        x = C.c_double(rv[i])
        na, va = nact[i], vact[i]
        status[i] = lb.so_step_cell_act(C.byref(L), C.byref(P), float(T[i]), float(p[i]), float(rs[i]),
                                        float(rsi[i]), C.byref(x), dt, _dp(numc[i]), _dp(volc[i]), None,
                                        float(w[i]), _dp(na), _dp(va))
        rv[i] = x.value
    return rv, numc, volc, status, nact, vact


def step_cells(L, P, T, p, rs, rsi, rv, numc, volc, dt=1.0, nsteps=1, rates=False, kstore=None, kmodes=None,
               init=False):
    """Run `nsteps` box-model steps for each cell (SALSA units).  With init the
    first step is run_SALSA's initialization call (equilibration at RH 0.3
    before the step, mo_salsa_driver.f90:282-283).

    numc: [ncell, ntotal]; volc: [ncell, ntotal, 9].  Returns
    (rv, numc, volc, status[, rates[ncell, 20, 10]]) as new arrays.
    kstore: optional [ncell, ntotal, ntotal] kernel store (lscoag%mode 2),
    updated in place; kmodes: per-step kernel mode (1 update, 2 fetch).
    """
    lb = lib()
    n = len(T)
    rv = np.array(rv, dtype=np.float64, copy=True)
    numc = np.ascontiguousarray(numc, dtype=np.float64).copy()
    volc = np.ascontiguousarray(volc, dtype=np.float64).copy()
    status = np.zeros(n, dtype=np.int32)
    rr = np.zeros((n, NRATE, MAXSPEC + 1)) if rates else None
    rbuf = (Rate * NRATE)()
    Pi = None
    if init:
        Pi = Params()
        C.memmove(C.byref(Pi), C.byref(P), C.sizeof(Params))
        Pi.init = 1
    for i in range(n):
        x = C.c_double(rv[i])
        nc = numc[i]
        vc = volc[i]
        st = 0
        for k in range(nsteps):
            Pk = Pi if (init and k == 0) else P
            if kstore is None:
                st |= lb.so_step_cell(C.byref(L), C.byref(Pk), float(T[i]), float(p[i]), float(rs[i]),
                                      float(rsi[i]), C.byref(x), dt, _dp(nc), _dp(vc),
                                      rbuf if rates else None)
            else:
                ks = kstore[i]
                assert ks.flags["C_CONTIGUOUS"] and ks.shape == (L.ntotal, L.ntotal)
                st |= lb.so_step_cell_ks(C.byref(L), C.byref(Pk), float(T[i]), float(p[i]), float(rs[i]),
                                         float(rsi[i]), C.byref(x), dt, _dp(nc), _dp(vc),
                                         rbuf if rates else None, _dp(ks), int(kmodes[k]))
        rv[i] = x.value
        status[i] = st
        if rates:
            for k in range(NRATE):
                rr[i, k, :MAXSPEC] = rbuf[k].volc[:]
                rr[i, k, MAXSPEC] = rbuf[k].numc
    if rates:
        return rv, numc, volc, status, rr
    return rv, numc, volc, status


def rslf(p, t):
    return lib().so_rslf(p, t)


def rsif(p, t):
    return lib().so_rsif(p, t)


def step_batch(L, P, T, p, rs, rsi, rv, numc, volc, dt=1.0, nsteps=1):
    """C-loop batch step (in place on rv, numc [n, ntotal], volc [n, ntotal, 9])."""
    for a in (T, p, rs, rsi, rv, numc, volc):
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return lib().so_step_batch(C.byref(L), C.byref(P), len(T), _dp(T), _dp(p), _dp(rs), _dp(rsi), _dp(rv),
                               dt, nsteps, _dp(numc), _dp(volc))


def les_update(L, level, shape, dtlt, dzf2, temp, rh, rhi, rp, n_p, n_t, m_p, m_t, charge_time=None,
               gaerop=None, indefp=None):
    """tend_constrain2 + update_sclrs + tend0 + SALSA_diagnostics on LES
    arrays (so_les_update_ext).  n_p/n_t/m_p/m_t: lists of 4 float64 arrays in
    the LES layout ([rows, n3], n3 in Fortran order), None for the ice
    category at level 4; updated in place.  gaerop [5, n3] (the gas phase
    that removed tiny aerosols feed, lscndgas) and indefp [rows, n3]
    (clamped to [0, 1] at level 5, ice_theta_dist), when given, are updated
    in place too.  Returns diag[6, n3]."""
    nzp, nxp, nyp = shape
    n3 = nzp * nxp * nyp
    P4 = C.c_void_p * 4

    def ptrs(lst):
        out = []
        for a in lst:
            if a is None:
                out.append(None)
            else:
                assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
                out.append(a.ctypes.data)
        return P4(*out)
    diag = np.zeros((6, n3))
    ct = None
    if charge_time is not None:
        assert charge_time.dtype == np.float64 and charge_time.flags["C_CONTIGUOUS"]
        ct = charge_time.ctypes.data
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (temp, rh, rhi, rp)]
    for a in (gaerop, indefp):
        assert a is None or (a.dtype == np.float64 and a.flags["C_CONTIGUOUS"] and a.size % n3 == 0)
    assert gaerop is None or gaerop.size == 5 * n3
    gp = gaerop.ctypes.data if gaerop is not None else None
    ip = indefp.ctypes.data if indefp is not None else None
    nind = indefp.size // n3 if indefp is not None else 0
    rc = lib().so_les_update_ext(C.byref(L), level, nzp, nxp, nyp, dtlt, dzf2, *(_dp(a) for a in arrs),
                                 ptrs(n_p), ptrs(n_t), ptrs(m_p), ptrs(m_t), ct, _dp(diag), gp, ip, nind)
    if rc != 0:
        raise ValueError("so_les_update failed: %d" % rc)
    return diag


def les_corr(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().so_les_corr(_dp(x), _dp(y), len(x))


def sedim(L, shape, dtlt, dzt, tk, th, adn, ustar, n_p, m_p, n_t, m_t, tlt, sed=(1, 1, 1)):
    """so_sedim on LES arrays (level 4: aerosol, cloud, precipitation).
    n_t, m_t, tlt are added to in place.  Returns (rrate[n3],
    sfcrrate[nxp*nyp], vt_prc[nprc, n3])."""
    nzp, nxp, nyp = shape
    n3 = nzp * nxp * nyp
    P3 = C.c_void_p * 3

    def ptrs(lst):
        for a in lst:
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        return P3(*[a.ctypes.data for a in lst])
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (dzt, tk, th, adn, ustar)]
    assert tlt.dtype == np.float64 and tlt.flags["C_CONTIGUOUS"]
    rr, sr, vt = np.zeros(n3), np.zeros(nxp * nyp), np.zeros((L.nprc, n3))
    rc = lib().so_sedim(C.byref(L), nzp, nxp, nyp, dtlt, *(_dp(a) for a in arrs), ptrs(n_p[:3]), ptrs(m_p[:3]),
                        ptrs(n_t[:3]), ptrs(m_t[:3]), _dp(tlt), _dp(rr), _dp(sr), _dp(vt), *sed)
    if rc != 0:
        raise ValueError("so_sedim failed: %d" % rc)
    return rr, sr, vt


def sedim5(L, P, shape, dtlt, dzt, tk, th, adn, ustar, n_p, m_p, n_t, m_t, tlt, sed=(1, 1, 1, 1)):
    """so_sedim5 on LES arrays (level 5: the three liquid categories, then
    ice).  n_t, m_t (four categories), tlt are added to in place.  Returns
    (rrate[n3], sfcrrate[nxp*nyp], irate[n3], sfcirate[nxp*nyp],
    vt_prc[nprc, n3], vt_ice[nice, n3])."""
    nzp, nxp, nyp = shape
    n3 = nzp * nxp * nyp
    P4 = C.c_void_p * 4

    def ptrs(lst):
        for a in lst:
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        return P4(*[a.ctypes.data for a in lst])
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (dzt, tk, th, adn, ustar)]
    assert tlt.dtype == np.float64 and tlt.flags["C_CONTIGUOUS"]
    rr, sr, ir, si = np.zeros(n3), np.zeros(nxp * nyp), np.zeros(n3), np.zeros(nxp * nyp)
    vt, vi = np.zeros((L.nprc, n3)), np.zeros((L.nice, n3))
    rc = lib().so_sedim5(C.byref(L), C.byref(P), nzp, nxp, nyp, dtlt, *(_dp(a) for a in arrs), ptrs(n_p[:4]),
                         ptrs(m_p[:4]), ptrs(n_t[:4]), ptrs(m_t[:4]), _dp(tlt), _dp(rr), _dp(sr), _dp(ir), _dp(si),
                         _dp(vt), _dp(vi), *sed)
    if rc != 0:
        raise ValueError("so_sedim5 failed: %d" % rc)
    return rr, sr, ir, si, vt, vi


def act_cloud_base(L, T, p, w, numc, volc):
    """Cloud-base activation of one cell (ActCloudBase + activate3) on a
    state in SALSA units (numc[ntotal], volc[ntotal, 9]): returns (nact[ncld],
    vact[ncld, 9]) of the activated droplets."""
    nc = np.ascontiguousarray(numc, dtype=np.float64)
    vc = np.ascontiguousarray(volc, dtype=np.float64).reshape(-1)
    nact, vact = np.zeros(L.ncld), np.zeros(L.ncld * 9)
    lib().so_act_cloud_base(C.byref(L), T, p, w, _dp(nc), _dp(vc), _dp(nact), _dp(vact))
    return nact, vact.reshape(L.ncld, 9)


def sed_ice_vc(L, P, tk, adn, numc, masses):
    """DepositionFast's fall-speed chain of one ice bin (flag 4): returns
    (dwet, dnsp, alpha, beta, gamma, sigma, zdn, GG, vc).  masses: the bin's
    LES mass mixing ratios of every species, rime last."""
    m = np.ascontiguousarray(masses, dtype=np.float64)
    assert m.size == L.nspec
    out = np.zeros(9)
    lib().so_sed_ice_vc(C.byref(L), C.byref(P), tk, adn, numc, _dp(m), _dp(out))
    return out


def terminal_vel(diam, rhop, rhoa, visc, beta):
    return lib().so_terminal_vel(diam, rhop, rhoa, visc, beta)
