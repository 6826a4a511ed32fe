"""The reference's own main program (espec.f) linked against the drop-in.

oracle/Makefile target `dropin` links the reference's objects, minus every
object whose routines libespec_f.so provides (INTEGRATION.md lists them),
with libespec_f.so and libespec_hip.so into oracle/_ref/espec_dropin.x.
Here, without a GPU: the link resolves, none of the replaced routines is
defined by the executable itself (so the calls reach the library), and a
run of the reference's input.test12 gets as far as the initial
eigenstates (MTRXDIAG, on the device in the drop-in), which report INFO = 3
(no HIP device) through the main program's own error branch
(espec.f:906-909, 928-951).  tests/test_dropin_main_gpu.py runs the same
executable on the GPU against the published tables."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "espec_dropin.x")
OBJ = os.path.join(ROOT, "oracle", "_ref", "obj", "espec.o")
# symbols of the objects the drop-in replaces (flang's names)
REPLACED = ["pspofft_", "pspofftaux_", "psodfft_", "psodfftaux_", "plnz_", "plnzaux_", "lnz_",
            "reort_", "lanczs_", "au_", "au_1d_", "au_2d_", "au_3d_", "fft_", "fftnt_", "digt_",
            "fcorrel_", "ecnorm_", "efft4_", "eigenerg_", "ran_", "hwhm_", "spectrumtd_",
            "pplnz_", "absorbingbc_", "ppsod_", "lanczsg_", "mtrxdiag_", "s2ppsod_"]


@pytest.fixture(scope="module")
def exe():
    if not os.path.isdir("/root/reference/src") or not os.path.exists(OBJ):
        pytest.skip("reference objects not built here (oracle/Makefile ref)")
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "dropin"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return EXE


def test_library_exports_every_replaced_routine():
    r = subprocess.run(["nm", "-D", "--defined-only",
                        os.path.join(ROOT, "espec_amd", "libespec_f.so")],
                       capture_output=True, text=True, check=True)
    names = {ln.split()[-1] for ln in r.stdout.splitlines() if ln.strip()}
    missing = [s for s in REPLACED if s not in names]
    assert not missing, missing


def test_replaced_routines_resolve_to_the_library(exe):
    r = subprocess.run(["nm", exe], capture_output=True, text=True, check=True)
    defined = {ln.split()[-1] for ln in r.stdout.splitlines()
               if len(ln.split()) == 3 and ln.split()[1] in "TtWw"}
    undefined = {ln.split()[-1] for ln in r.stdout.splitlines()
                 if len(ln.split()) == 2 and ln.split()[0] == "U"}
    assert not (defined & set(REPLACED)), sorted(defined & set(REPLACED))
    # the main program's own calls (espec.f) and the kept objects' calls
    assert {"pspofft_", "psodfft_", "plnz_", "lanczs_", "spectrumtd_"} <= undefined
    r = subprocess.run(["ldd", exe], capture_output=True, text=True, check=True)
    assert "libespec_f.so" in r.stdout and "libespec_hip.so" in r.stdout


def test_input_test12_reaches_the_drop_in(exe, tmp_path):
    shutil.copy(os.path.join(ROOT, "tests", "golden", "inputs", "input.test12"),
                tmp_path / "input.spc")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout
    assert "Using complete matrix diagonalization procedure" in out
    assert "possible obtain eignvalues and eignvectors correctly" in out, out[-2000:]
