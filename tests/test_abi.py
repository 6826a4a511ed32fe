"""The C ABI: every entry point declared in include/*.h is exported by
libespec_hip.so (loaded, nothing called on a device), and host-side
helpers agree with the oracle."""
import ctypes
import glob
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+[\s\*]+(espec_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("espec_pspofft", "espec_psodfft", "espec_plnz", "espec_fft", "espec_au",
                 "espec_lnz", "espec_dstevx", "espec_fcorrel", "espec_hwhm",
                 "espec_spectrumtd", "espec_lanczs", "espec_spo_plan_create",
                 "espec_sil_plan_create"):
        assert must in names


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    lib = ctypes.CDLL(os.path.join(ROOT, "espec_amd", "libespec_hip.so"))
    assert hasattr(lib, name), f"{name} declared but not exported"


def test_strerror_and_version_need_no_device():
    import espec_amd as E
    L = E.lib()
    assert L.espec_strerror(0).decode()
    assert L.espec_strerror(5).decode()
    assert "gfx950" in L.espec_version().decode()


@pytest.mark.parametrize("ti,tf,dt", [(0.0, 250.0, 1e-4), (0.0, 0.2, 1e-3), (1.0, 2.0, 0.3),
                                      (0.0, 249.856, 5e-5), (0.0, 0.0, 0.1)])
def test_step_count_rule(orc, ti, tf, dt):
    import espec_amd as E
    assert E.nsteps(ti, tf, dt) == orc.nsteps(ti, tf, dt)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_fortran_dropin_exports_reference_names():
    """libespec_f.so exports the reference's subroutine symbols."""
    lib = ctypes.CDLL(os.path.join(ROOT, "espec_amd", "libespec_f.so"))
    for name in ("pspofft_", "psodfft_", "plnz_", "lanczs_", "fft_", "digt_", "au_", "au_1d_",
                 "au_2d_", "au_3d_", "fcorrel_", "ecnorm_", "efft4_", "eigenerg_", "ran_",
                 "hwhm_", "spectrumtd_"):
        assert hasattr(lib, name), name


def test_fortran_ran_is_the_reference_lcg(orc):
    """RAN through the Fortran ABI (by-reference INTEGER), no device needed."""
    lib = ctypes.CDLL(os.path.join(ROOT, "espec_amd", "libespec_f.so"))
    lib.ran_.restype = ctypes.c_double
    s = ctypes.c_int(1)
    got = [lib.ran_(ctypes.byref(s)) for _ in range(50)]
    want, seed = orc.ran(1, 50)
    assert np.array_equal(got, want) and s.value == seed


def test_fortran_driver_reads_case_and_reports_without_device():
    """The host program parses a full test input; with no GPU visible the
    drop-in returns INFO = 3 (HIP error) instead of stopping."""
    import espec_amd as E
    if E.device_count() > 0:
        pytest.skip("a device is visible")
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import setups
    from test_dropin_gpu import DRIVER, driver_input
    c = setups.test_case(12)
    r = subprocess.run([DRIVER], input=driver_input(c), capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr[-500:]
    assert "INFO    3" in r.stdout
