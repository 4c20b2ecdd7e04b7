"""The C ABI library on the CPU: it loads, exports every declared symbol, and
its host-side bin descriptor equals the reference's bin table.  No kernel is
launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import ROOT, golden


def declared_functions():
    src = open(os.path.join(ROOT, "include", "salsa_amd.h")).read()
    return sorted(set(re.findall(r"\b(salsa_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from salsa_amd import capi
    lib = capi.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(capi.EXPORTS) <= set(names)


@pytest.mark.parametrize("name,species", [("bins_legal.npz", ("SO4",)), ("bins_ice.npz", ("SO4", "DU"))])
def test_bin_desc_matches_reference_bins(name, species):
    from salsa_amd import capi
    g = golden(name)
    d = capi.bin_desc(species=species)
    nt = d.ntotal
    assert np.array_equal(np.array(d.vlolim[:nt]), g["vlolim"])
    assert np.array_equal(np.array(d.vhilim[:nt]), g["vhilim"])
    assert np.array_equal(np.array(d.dmid[:nt]), g["dmid"])
    assert np.array_equal(np.array(d.precpbins[:d.nprc]), g["precpbins"])
    assert (d.ica.cur, d.ica.par, d.fca.cur, d.icb.cur, d.icb.par, d.fcb.cur) == (0, 3, 3, 4, 7, 7)


def test_params_defaults_and_level4_overrides():
    from salsa_amd import capi
    p4 = capi.params(level=4)
    p5 = capi.params(level=5)
    assert p4.lscgii == 0 and p4.lscgia == 0 and p4.lscndh2oic == 0
    assert p5.lscgii == 1 and p5.lscndh2oic == 1
    assert p4.eddy_dis == 0.001 and p4.rhlim == 1.2 and p4.hz == 1 and p4.hxy == 2


def test_invalid_arguments_are_rejected_without_touching_the_device():
    from salsa_amd import capi
    with pytest.raises(capi.SalsaError):
        capi.params(level=3)
    with pytest.raises(capi.SalsaError):
        capi.bin_desc(nbin=(3, 2))
    d = capi.bin_desc()
    p = capi.params(level=4)
    h = C.c_void_p()
    # no interior cells for the default halo widths
    rc = capi.lib().salsa_ctx_create(C.byref(d), C.byref(p), 2, 4, 4, C.byref(h))
    assert rc == capi.SALSA_ERR_ARG
    p.precision = 7
    rc = capi.lib().salsa_ctx_create(C.byref(d), C.byref(p), 10, 10, 10, C.byref(h))
    assert rc == capi.SALSA_ERR_ARG
    assert b"precision" in capi.lib().salsa_last_error()
