"""Printed blocks of the reference's own v0.7 program on its regression
inputs, for the drop-in's end-to-end test.

The published results (ttests/results.test*) come from eSPec v0.4; the
v0.7 main program in the reference's sources differs from them (for
instance it shifts the potential by its grid minimum, SHIFTP, espec.f:118,
gvpot.f:64-73), by 3.3e-2 in test12's C(t).  tests/test_dropin_main_gpu.py
runs the v0.7 main program with the hot path linked to libespec_f.so, so
its reference is the same v0.7 program built entirely from the reference's
sources (oracle/_ref/espec_ref.x, `make -C oracle ref`).  Run here, where
that build exists:

    python tests/golden/make_ref_fixtures.py

It writes tests/golden/v07_test<N>_<block>.txt.gz (numbers only, the
blocks of make_fixtures.py).
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import make_fixtures as MF  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "espec_ref.x")
INPUTS = os.path.join(HERE, "inputs")
STOPS = ["Correlation Function", "Spectrum:", "Partial auto", "Total correlation",
         "Calculating", "Propagating", "Eigenvector propagated"]
# (test, blocks): (name, title, columns)
CASES = {12: [("table", "Partial auto-correlation function", 6),
              ("corr", "Correlation Function", 3), ("spec", "Spectrum:", 4)],
         4: [("table", "Partial auto-correlation function", 6),
             ("corr", "Correlation Function", 3), ("spec", "Spectrum:", 4)],
         14: [("table", "Auto-correlation function", 7)]}


def run(t):
    with tempfile.TemporaryDirectory() as d:
        shutil.copy(os.path.join(INPUTS, f"input.test{t}"), os.path.join(d, "input.spc"))
        for f in os.listdir(INPUTS):
            if not f.startswith("input."):
                shutil.copy(os.path.join(INPUTS, f), os.path.join(d, f))
        r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=3600)
        return r.stdout.splitlines()


def main():
    for t, blocks in CASES.items():
        lines = run(t)
        for name, title, nc in blocks:
            stops = [s for s in STOPS if not title.startswith(s)]
            if t == 14:
                stops = ["Eigenvector propagated"]
            MF.write(f"v07_test{t}_{name}", MF.block(lines, title, nc, stops))


if __name__ == "__main__":
    main()
