"""The Fortran ISO_C_BINDING side of the boundary (uclales-salsa_amd/fortran).

* layout: a Fortran program fills salsa_bin_desc / salsa_params through the
  BIND(C) twins and the C library; its fields must equal what the Python
  ctypes binding reads from the same calls (no GPU needed).
* drop-in: mo_salsa_amd_driver (run_SALSA with the reference's dummy list)
  compiles against the reference's own module files where those were built.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "uclales-salsa_amd", "fortran")
LIBDIR = os.path.join(ROOT, "uclales-salsa_amd", "salsa_amd")
FC = "/opt/rocm/bin/amdflang"

needs_fc = pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not available")


def _build_cabi(tmp):
    subprocess.run([FC, "-fdefault-real-8", "-O1", "-J", tmp, "-c", os.path.join(FDIR, "salsa_amd_cabi.f90"),
                    "-o", os.path.join(tmp, "cabi.o")], check=True)


@needs_fc
def test_bindc_layout_matches_c(tmp_path):
    from salsa_amd import capi
    tmp = str(tmp_path)
    _build_cabi(tmp)
    exe = os.path.join(tmp, "layout_check")
    subprocess.run([FC, "-fdefault-real-8", "-O1", "-I", tmp, os.path.join(ROOT, "tests", "fortran", "layout_check.f90"),
                    os.path.join(tmp, "cabi.o"), "-L" + LIBDIR, "-lsalsa_amd", "-Wl,-rpath," + LIBDIR,
                    "-Wl,-rpath,/opt/rocm/lib", "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    rows = {l.split()[0]: l.split()[1:] for l in out if l.strip()}
    d = capi.bin_desc(species=("SO4", "DU"))
    p = capi.params(level=5)
    assert rows["rc"] == ["0"] and rows["prc"] == ["0"]
    assert [int(x) for x in rows["idx"]] == [d.in1a, d.fn1a, d.in2a, d.fn2a, d.in2b, d.fn2b, d.nbins]
    assert [int(x) for x in rows["par"]] == [d.ica.cur, d.ica.par, d.fcb.cur, d.fcb.par]
    assert [int(x) for x in rows["cnt"]] == [d.ncld, d.ntotal, d.nspec, d.iwa, d.irim]
    assert np.allclose([float(x) for x in rows["dbl"]], [d.dmid[4], d.precpbins[2], d.rholiq[1]], rtol=1e-15)
    assert [int(x) for x in rows["ph"]] == [d.phase[11], d.phase[d.ntotal - 1]]
    assert [int(x) for x in rows["psw"]] == [p.level, p.lscgii, p.lscndh2oic, p.lsfreeRH, p.precision, p.hxy]
    assert np.allclose([float(x) for x in rows["pdb"]], [p.rhlim, p.ice_sigma, p.charge_coll_enh, p.cgintvl],
                       rtol=1e-15)
    assert [int(x) for x in rows["pmd"]] == [p.lscoag_mode, 7, 9]


REF_MODS = os.path.join(ROOT, "oracle", "_ref", "obj")


@needs_fc
@pytest.mark.skipif(not os.path.exists(os.path.join(REF_MODS, "classfieldarray.mod")),
                    reason="the reference's module files are built only where the reference tree exists")
def test_dropin_run_salsa_compiles_against_reference_modules(tmp_path):
    tmp = str(tmp_path)
    _build_cabi(tmp)
    subprocess.run([FC, "-fdefault-real-8", "-O1", "-I", REF_MODS, "-I", tmp, "-J", tmp, "-c",
                    os.path.join(FDIR, "mo_salsa_amd_driver.f90"), "-o", os.path.join(tmp, "drv.o")], check=True)
    assert os.path.getsize(os.path.join(tmp, "drv.o")) > 0
