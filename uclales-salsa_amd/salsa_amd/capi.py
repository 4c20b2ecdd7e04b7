"""ctypes binding of libsalsa_amd.so (the C ABI declared in include/salsa_amd.h).

The library is the product path: every call lands in hand-written HIP
kernels.  There is no CPU fallback -- if the shared library is missing the
import of this module raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libsalsa_amd.so")

MAXBIN = 64
MAXSPEC = 9

SALSA_OK = 0
SALSA_ERR_ARG = -1
SALSA_ERR_HIP = -2
SALSA_ERR_NOCONV = -3
SALSA_ERR_NONFINITE = -4
SALSA_ERR_NOMEM = -5

SALSA_FP64 = 0
SALSA_FP32_MIXED = 1

STEP_TENDENCY = 0
STEP_INPLACE = 1
STEP_NO_COMPACT = 2
STEP_INIT = 4

CELL_NEG_NUMC = 1
CELL_NOCONV = 2
CELL_SUBSTEPS = 4
CELL_NONFINITE = 8
CELL_GRAD_UNSTABLE = 16
CELL_DISTR = 32

SPECIES = {"SO4": 0, "OC": 1, "BC": 2, "DU": 3, "SS": 4, "NO": 5, "NH": 6}

EXPORTS = [
    "salsa_bin_desc_init", "salsa_params_default", "salsa_ctx_create", "salsa_ctx_destroy",
    "salsa_step_host", "salsa_step_device", "salsa_ctx_sync", "salsa_get_cell_flags",
    "salsa_set_launch_shape", "salsa_get_launch_shape", "salsa_autotune", "salsa_ctx_ncells",
    "salsa_ctx_ntotal", "salsa_ctx_stream", "salsa_last_error", "salsa_ctx_timing",
    "salsa_ctx_kernel_time", "salsa_ctx_phase_time", "salsa_gradient_device_noalloc", "salsa_adjoint_device",
    "salsa_set_cloud_base_fields",
    "salsa_set_les_extra_fields",
]

NOALLOC_MAX_CELLS = 1 << 20


# process-rate fields (include/salsa_amd.h SALSA_RATE_*)
RATE_NAMES = ["s_m_autoc", "s_m_autoc50", "s_m_autoc80", "s_m_accr", "s_m_accr50", "s_m_accr80",
              "s_m_ACcoll_dry", "s_m_APcoll_dry", "s_n_activ", "s_m_conda", "s_m_condc", "s_m_condp",
              "s_m_AIcoll_dry", "s_n_icehom", "s_n_icedep", "s_n_iceimm", "s_m_condi"]
NRATE_L4, NRATE_L5 = 12, 17


class ParBin(C.Structure):
    _fields_ = [("cur", C.c_int32), ("par", C.c_int32)]


class BinDesc(C.Structure):
    """salsa_bin_desc"""
    _fields_ = [
        ("nbin", C.c_int32 * 2), ("reglim", C.c_double * 4),
        ("nprc", C.c_int32), ("prc_dlo", C.c_double), ("prc_vratio", C.c_double),
        ("nice", C.c_int32), ("ice_dlo", C.c_double), ("ice_vratio", C.c_double),
        ("nspec_dry", C.c_int32), ("dry_species", C.c_int32 * 7),
        ("in1a", C.c_int32), ("fn1a", C.c_int32), ("in2a", C.c_int32), ("fn2a", C.c_int32),
        ("in2b", C.c_int32), ("fn2b", C.c_int32), ("nbins", C.c_int32),
        ("ica", ParBin), ("fca", ParBin), ("icb", ParBin), ("fcb", ParBin),
        ("ncld", C.c_int32), ("ntotal", C.c_int32),
        ("nspec", C.c_int32), ("iwa", C.c_int32), ("irim", C.c_int32),
        ("rholiq", C.c_double * MAXSPEC), ("rhoice", C.c_double * MAXSPEC),
        ("mm", C.c_double * MAXSPEC), ("diss", C.c_double * MAXSPEC),
        ("phase", C.c_int32 * MAXBIN),
        ("vlolim", C.c_double * MAXBIN), ("vhilim", C.c_double * MAXBIN), ("dmid", C.c_double * MAXBIN),
        ("nlim", C.c_double * MAXBIN), ("dlim", C.c_double * MAXBIN),
        ("precpbins", C.c_double * MAXBIN),
    ]

    @property
    def nspec_wet(self):
        return self.nspec - 1


class Params(C.Structure):
    """salsa_params"""
    _fields_ = [
        ("level", C.c_int32), ("lscoag", C.c_int32), ("lscnd", C.c_int32),
        ("lsauto", C.c_int32), ("lsauto_mode", C.c_int32),
        ("lscgaa", C.c_int32), ("lscgcc", C.c_int32), ("lscgpp", C.c_int32), ("lscgii", C.c_int32),
        ("lscgca", C.c_int32), ("lscgpa", C.c_int32), ("lscgia", C.c_int32), ("lscgpc", C.c_int32),
        ("lscgic", C.c_int32), ("lscgip", C.c_int32),
        ("lscndh2oae", C.c_int32), ("lscndh2ocl", C.c_int32), ("lscndh2oic", C.c_int32),
        ("lsfreeRH", C.c_int32), ("rhlim", C.c_double),
        ("eddy_dis", C.c_double), ("Eiagg_max", C.c_double), ("Eiagg_min", C.c_double),
        ("ice_alpha", C.c_double), ("ice_beta", C.c_double), ("ice_gamma", C.c_double),
        ("ice_sigma", C.c_double),
        ("lcharge", C.c_int32), ("charge_coll_enh", C.c_double),
        ("precision", C.c_int32), ("hz", C.c_int32), ("hxy", C.c_int32),
        ("lscoag_mode", C.c_int32), ("cgintvl", C.c_double), ("lsdistupdate", C.c_int32), ("lsactiv", C.c_int32),
    ]


_lib = None
_dp = C.POINTER(C.c_double)
_P4 = C.c_void_p * 4


def lib():
    """Load libsalsa_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError("libsalsa_amd.so is not built (run __graft_entry__.build()); "
                           "there is no CPU fallback for the SALSA step")
    # PyTorch-ROCm wheels carry their own HIP runtime (libamdhip64.so) next to
    # ROCm's libamdhip64.so.7 that this library links.  When both live in one
    # process, torch's runtime has to be loaded first or the second runtime
    # finds no device; importing torch here fixes the order.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIBPATH)
    L.salsa_bin_desc_init.argtypes = [C.POINTER(BinDesc), C.POINTER(C.c_int32), _dp, C.c_int32,
                                      C.c_double, C.c_double, C.c_int32, C.c_double, C.c_double,
                                      C.c_int32, C.POINTER(C.c_int32)]
    L.salsa_bin_desc_init.restype = C.c_int
    L.salsa_params_default.argtypes = [C.POINTER(Params), C.c_int32]
    L.salsa_params_default.restype = C.c_int
    L.salsa_ctx_create.argtypes = [C.POINTER(BinDesc), C.POINTER(Params), C.c_int32, C.c_int32,
                                   C.c_int32, C.POINTER(C.c_void_p)]
    L.salsa_ctx_create.restype = C.c_int
    L.salsa_ctx_destroy.argtypes = [C.c_void_p]
    L.salsa_ctx_destroy.restype = None
    step_args = [C.c_void_p] + [C.c_void_p] * 7 + [_P4] * 4 + [C.c_void_p, C.c_double, C.c_double,
                                                                C.c_int32, C.c_int32]
    L.salsa_step_host.argtypes = step_args
    L.salsa_step_host.restype = C.c_int
    L.salsa_step_device.argtypes = step_args
    L.salsa_step_device.restype = C.c_int
    L.salsa_ctx_sync.argtypes = [C.c_void_p]
    L.salsa_ctx_sync.restype = C.c_int
    L.salsa_get_cell_flags.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    L.salsa_get_cell_flags.restype = C.c_int
    L.salsa_set_launch_shape.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    L.salsa_set_launch_shape.restype = C.c_int
    L.salsa_get_launch_shape.argtypes = [C.c_void_p, C.c_int32]
    L.salsa_get_launch_shape.restype = C.c_int
    L.salsa_autotune.argtypes = [C.c_void_p] + [C.c_void_p] * 6 + [_P4, _P4, C.c_double]
    L.salsa_autotune.restype = C.c_int
    L.salsa_ctx_ncells.argtypes = [C.c_void_p]
    L.salsa_ctx_ncells.restype = C.c_int64
    L.salsa_ctx_ntotal.argtypes = [C.c_void_p]
    L.salsa_ctx_ntotal.restype = C.c_int32
    L.salsa_ctx_stream.argtypes = [C.c_void_p]
    L.salsa_ctx_stream.restype = C.c_void_p
    L.salsa_ctx_timing.argtypes = [C.c_void_p, C.c_int32]
    L.salsa_ctx_timing.restype = C.c_int
    L.salsa_ctx_kernel_time.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.salsa_ctx_kernel_time.restype = C.c_int
    L.salsa_ctx_phase_time.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64)]
    L.salsa_ctx_phase_time.restype = C.c_int
    L.salsa_gradient_device.argtypes = [C.c_void_p] + [C.c_void_p] * 6 + [_P4, _P4, C.c_double, C.c_int32,
                                                                          _P4, C.c_void_p, C.c_void_p, _dp]
    L.salsa_gradient_device.restype = C.c_int
    L.salsa_gradient_device_noalloc.argtypes = L.salsa_gradient_device.argtypes
    L.salsa_gradient_device_noalloc.restype = C.c_int
    L.salsa_adjoint_device.argtypes = [C.c_void_p] + [C.c_void_p] * 6 + [_P4, _P4, C.c_double, C.c_int32,
                                                                         C.c_int32, _P4, C.c_void_p, C.c_void_p, _dp]
    L.salsa_adjoint_device.restype = C.c_int
    L.salsa_les_update_device.argtypes = [C.c_void_p, C.c_double, _dp] + [C.c_void_p] * 4 + [_P4] * 4 + \
        [C.c_void_p, C.c_void_p]
    L.salsa_les_update_device.restype = C.c_int
    L.salsa_sedim_device.argtypes = [C.c_void_p, C.c_double, _dp] + [C.c_void_p] * 4 + [_P4] * 4 + \
        [C.c_void_p] * 7 + [C.c_int32] * 4
    L.salsa_sedim_device.restype = C.c_int
    L.salsa_set_cloud_base_fields.argtypes = [C.c_void_p] * 4
    L.salsa_set_cloud_base_fields.restype = C.c_int
    L.salsa_set_les_extra_fields.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
    L.salsa_set_les_extra_fields.restype = C.c_int
    L.salsa_last_error.argtypes = []
    L.salsa_last_error.restype = C.c_char_p
    _lib = L
    return L


class SalsaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("salsa error %d: %s" % (code, msg))
        self.code = code


def _check(rc):
    if rc != SALSA_OK:
        raise SalsaError(rc, lib().salsa_last_error().decode())
    return rc


def bin_desc(nbin=(3, 4), reglim=(3.e-9, 9.703e-8, 1.e-5, 1.e-5), nprc=7, prc_dlo=20.e-6,
             prc_vratio=4., nice=7, ice_dlo=2.e-6, ice_vratio=20., species=("SO4",)):
    """The bin/species descriptor of NAMELIST &salsa (mo_salsa_init.f90:148-450)."""
    d = BinDesc()
    ids = [SPECIES[s] if isinstance(s, str) else int(s) for s in species]
    sp = (C.c_int32 * 7)(*(ids + [0] * (7 - len(ids))))
    _check(lib().salsa_bin_desc_init(C.byref(d), (C.c_int32 * 2)(*nbin), (C.c_double * 4)(*reglim),
                                     nprc, prc_dlo, prc_vratio, nice, ice_dlo, ice_vratio, len(ids), sp))
    return d


def params(level=4, precision=SALSA_FP32_MIXED, **kw):
    p = Params()
    _check(lib().salsa_params_default(C.byref(p), level))
    p.precision = precision
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def _ptr(x):
    """Address of a numpy array (host) or a torch tensor (device or host)."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        if not x.flags["C_CONTIGUOUS"] or x.dtype != np.float64:
            raise ValueError("expected C-contiguous float64 arrays")
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        if x.dtype.__repr__() != "torch.float64" or not x.is_contiguous():
            raise ValueError("expected contiguous float64 tensors")
        return x.data_ptr()
    raise TypeError("unsupported array type %r" % type(x))


class Context:
    """One SALSA context for a (nzp, nxp, nyp) grid on the current HIP device."""

    def __init__(self, desc, prm, nzp, nxp, nyp):
        self.desc = desc
        self.params = prm
        self._h = C.c_void_p()
        _check(lib().salsa_ctx_create(C.byref(desc), C.byref(prm), nzp, nxp, nyp, C.byref(self._h)))
        self.shape = (nzp, nxp, nyp)

    def close(self):
        if self._h:
            lib().salsa_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ncells(self):
        return lib().salsa_ctx_ncells(self._h)

    @property
    def ntotal(self):
        return lib().salsa_ctx_ntotal(self._h)

    @property
    def stream(self):
        return lib().salsa_ctx_stream(self._h)

    def _call(self, fn, f, n_p, m_p, n_t, m_t, rv_t, tstep, time, istp, flags, rates=None):
        P = lambda xs: _P4(*[_ptr(x) for x in (list(xs) + [None] * (4 - len(xs)))])
        return fn(self._h, _ptr(f["press"]), _ptr(f["temp"]), _ptr(f["rsl"]), _ptr(f.get("rsi")),
                  _ptr(f["dn"]), _ptr(f["rv"]), _ptr(rv_t), P(n_p), P(n_t or []), P(m_p), P(m_t or []),
                  _ptr(rates), tstep, time, istp, flags)

    def step_host(self, f, n_p, m_p, n_t=None, m_t=None, rv_t=None, tstep=1.0, time=10.0, istp=2,
                  flags=STEP_TENDENCY, rates=None):
        """run_SALSA on host numpy arrays (staged through device memory).
        rates: optional float64 array [nrate, nyp, nxp, nzp] (Fortran field
        order, C-contiguous) receiving the process-rate fields (RATE_*)."""
        return _check(self._call(lib().salsa_step_host, f, n_p, m_p, n_t, m_t, rv_t, tstep, time, istp, flags,
                                 rates))

    def _ordered(self, enqueue):
        """Run `enqueue` with the context stream ordered after the work already
        queued on torch's current stream, and torch's current stream ordered
        after the enqueued step (torch and the library share one HIP runtime,
        so the two streams can wait on each other's events)."""
        import torch
        cur = torch.cuda.current_stream()
        ext = torch.cuda.ExternalStream(self.stream)
        ext.wait_stream(cur)
        rc = enqueue()
        cur.wait_stream(ext)
        return rc

    def step_device(self, f, n_p, m_p, n_t=None, m_t=None, rv_t=None, tstep=1.0, time=10.0, istp=2,
                    flags=STEP_TENDENCY, rates=None):
        """run_SALSA on device-resident torch tensors; enqueues only (see sync())."""
        return _check(self._ordered(lambda: self._call(lib().salsa_step_device, f, n_p, m_p, n_t, m_t, rv_t,
                                                       tstep, time, istp, flags, rates)))

    def sync(self):
        return _check(lib().salsa_ctx_sync(self._h))

    def gradient_device(self, f, n_p, m_p, djdn, tstep=1.0, nsteps=1, j_cell=None, djde_cell=None,
                        noalloc=False):
        """Forward-mode gradient of J = precipitation volume concentration
        summed over interior cells after `nsteps` in-place steps (inputs are
        not modified).  djdn: per category a float64 tensor shaped like
        n_p[c] receiving dJ/dN0 (interior cells); j_cell / djde_cell:
        optional per-cell outputs.  Returns (J, dJ/dE) after waiting for the
        context stream.  noalloc: the entry point that allocates no device
        memory (salsa_gradient_device_noalloc)."""
        P = lambda xs: _P4(*[_ptr(x) for x in (list(xs) + [None] * (4 - len(xs)))])
        tot = (C.c_double * 2)()
        fn = lib().salsa_gradient_device_noalloc if noalloc else lib().salsa_gradient_device
        rc = self._ordered(lambda: fn(
            self._h, _ptr(f["press"]), _ptr(f["temp"]), _ptr(f["rsl"]), _ptr(f.get("rsi")), _ptr(f["dn"]),
            _ptr(f["rv"]), P(n_p), P(m_p), tstep, nsteps, P(djdn), _ptr(j_cell), _ptr(djde_cell),
            C.cast(tot, _dp)))
        _check(rc)
        return tot[0], tot[1]

    def adjoint_device(self, f, n_p, m_p, djdn, tstep=1.0, nsteps=1, checkpoint_every=5, j_cell=None,
                       djde_cell=None):
        """Reverse-mode gradient of the objective of gradient_device
        (salsa_adjoint_device): one adjoint sweep gives dJ/dN0 of every bin
        and dJ/dE; the forward pass keeps the state every `checkpoint_every`
        steps.  Arguments and return value as gradient_device."""
        P = lambda xs: _P4(*[_ptr(x) for x in (list(xs) + [None] * (4 - len(xs)))])
        tot = (C.c_double * 2)()
        rc = self._ordered(lambda: lib().salsa_adjoint_device(
            self._h, _ptr(f["press"]), _ptr(f["temp"]), _ptr(f["rsl"]), _ptr(f.get("rsi")), _ptr(f["dn"]),
            _ptr(f["rv"]), P(n_p), P(m_p), tstep, nsteps, checkpoint_every, P(djdn), _ptr(j_cell),
            _ptr(djde_cell), C.cast(tot, _dp)))
        _check(rc)
        return tot[0], tot[1]

    def les_update_device(self, dtlt, dzt, f, n_p, n_t, m_p, m_t, diag, charge_time=None):
        """tend_constrain2 + update_sclrs + tend0 + SALSA_diagnostics on the
        device fields (salsa_les_update_device).  dzt: host float64[nzp];
        f: dict with "temp", "rh", "rp" and (level 5) "rhi" tensors; n_p,
        m_p updated in place, n_t, m_t zeroed; diag: float64 tensor [6, n3]."""
        P = lambda xs: _P4(*[_ptr(x) for x in (list(xs) + [None] * (4 - len(xs)))])
        dz = np.ascontiguousarray(dzt, dtype=np.float64)
        if dz.size != self.shape[0]:
            raise ValueError("dzt must have nzp values")
        rc = self._ordered(lambda: lib().salsa_les_update_device(
            self._h, dtlt, C.cast(dz.ctypes.data, _dp), _ptr(f["temp"]), _ptr(f["rh"]), _ptr(f.get("rhi")),
            _ptr(f["rp"]), P(n_p), P(n_t), P(m_p), P(m_t), _ptr(charge_time), _ptr(diag)))
        _check(rc)

    def sedim_device(self, dtlt, dzt, f, n_p, m_p, n_t, m_t, tlt, rrate, sfcrrate, vt_prc, sed=(1, 1, 1),
                     irate=None, sfcirate=None, vt_ice=None):
        """Sedimentation (salsa_sedim_device).  dzt: host float64[nzp]; f:
        dict of "tk", "th", "adn", "ustar" tensors; n_t, m_t, tlt added to;
        rrate, sfcrrate, vt_prc (level 5 also irate, sfcirate, vt_ice)
        overwritten.  sed: (aerosol, cloud, precipitation[, ice])."""
        P = lambda xs: _P4(*[_ptr(x) for x in (list(xs) + [None] * (4 - len(xs)))])
        dz = np.ascontiguousarray(dzt, dtype=np.float64)
        if dz.size != self.shape[0]:
            raise ValueError("dzt must have nzp values")
        sw = [int(x) for x in sed] + [0] * (4 - len(sed))
        rc = self._ordered(lambda: lib().salsa_sedim_device(
            self._h, dtlt, C.cast(dz.ctypes.data, _dp), _ptr(f["tk"]), _ptr(f["th"]), _ptr(f["adn"]),
            _ptr(f["ustar"]), P(n_p[:4]), P(m_p[:4]), P(n_t[:4]), P(m_t[:4]), _ptr(tlt), _ptr(rrate),
            _ptr(sfcrrate), _ptr(vt_prc), _ptr(irate), _ptr(sfcirate), _ptr(vt_ice), *sw))
        _check(rc)

    def set_cloud_base_fields(self, w, nactd, vactd):
        """Cloud-base activation (lsactiv = 2): the updraft w [n3] and the
        output fields nactd [ncld, n3] and vactd [nwet*ncld, n3] (device
        tensors) used by the following steps (salsa_set_cloud_base_fields).
        None for all three clears them."""
        self._keep_cb = (w, nactd, vactd)
        _check(lib().salsa_set_cloud_base_fields(self._h, _ptr(w), _ptr(nactd), _ptr(vactd)))

    def set_les_extra_fields(self, gaerop=None, indefp=None):
        """The LES update's optional fields (salsa_set_les_extra_fields):
        gaerop [5, n3] (a_gaerop, lscndgas) and indefp [rows, n3] (a_indefp,
        ice_theta_dist), device tensors or None."""
        self._keep_les = (gaerop, indefp)
        n3 = self.shape[0] * self.shape[1] * self.shape[2]
        nind = 0 if indefp is None else int(indefp.numel() // n3)
        _check(lib().salsa_set_les_extra_fields(self._h, _ptr(gaerop), _ptr(indefp), nind))

    def cell_flags(self):
        out = np.zeros(self.ncells, dtype=np.uint8)
        _check(lib().salsa_get_cell_flags(self._h, out.ctypes.data, out.size))
        return out

    def timing(self, enable=True):
        """Reset and enable (or disable) HIP-event timing of the step kernels."""
        _check(lib().salsa_ctx_timing(self._h, 1 if enable else 0))

    def kernel_time(self):
        """(milliseconds, steps) of step-kernel time since timing(True)."""
        ms = C.c_double()
        n = C.c_int64()
        _check(lib().salsa_ctx_kernel_time(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def phase_time(self):
        """(ms[5], cells[3], steps): GPU time of gather + classification, of
        the class 0/1/2 step kernels and of the scatter, and the cells of each
        class, summed over the steps timed since timing(True)."""
        ms = (C.c_double * 5)()
        cells = (C.c_int64 * 3)()
        n = C.c_int64()
        _check(lib().salsa_ctx_phase_time(self._h, ms, cells, C.byref(n)))
        return list(ms), list(cells), n.value

    def set_shape(self, cls, lanes):
        _check(lib().salsa_set_launch_shape(self._h, cls, lanes))

    def shape_of(self, cls):
        return lib().salsa_get_launch_shape(self._h, cls)

    def autotune(self, f, n_p, m_p, tstep=1.0):
        P = lambda xs: _P4(*[_ptr(x) for x in (list(xs) + [None] * (4 - len(xs)))])
        return _check(self._ordered(lambda: lib().salsa_autotune(
            self._h, _ptr(f["press"]), _ptr(f["temp"]), _ptr(f["rsl"]), _ptr(f.get("rsi")), _ptr(f["dn"]),
            _ptr(f["rv"]), P(n_p), P(m_p), tstep)))
