"""The oracle's LANCZSG against the reference build, and the dense FD matrix
of MTRXDIAG against the mathematics.

LANCZSG (lanczsc.f:1-153) STOPs the process when a state does not converge,
so the reference runs in a child process and its log is read from what it
prints (FORMAT 1031-1033, 8 significant digits).  Converging cases put grid
point 41, whose amplitude the reference's change measure reads
(lanczsc.f:104-109), in the far tail of the ground state.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import EXACT, REF

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def oscillator(n, half=8.0, k=1.0):
    x = np.linspace(-half, half, n)
    h = x[1] - x[0]
    return [1.0 / (24 * h * h)], 0.5 * k * x * x


def fd_matrix_1d(shm, v):
    n = len(v)
    s = shm[0]
    return (np.diag(30 * s + v) + np.diag(-16 * s * np.ones(n - 1), 1)
            + np.diag(-16 * s * np.ones(n - 1), -1) + np.diag(s * np.ones(n - 2), 2)
            + np.diag(s * np.ones(n - 2), -2))


CHILD = r"""
import sys, numpy as np
sys.path.insert(0, {oracle!r})
from oracle import Reference
n, maxint, abstol = {n}, {maxint}, {abstol}
x = np.linspace(-8.0, 8.0, n); h = x[1] - x[0]
info, ev, vc = Reference().lanczsg([n], 1, 1, maxint, abstol, [1.0 / (24 * h * h)], 0.5 * x * x)
sys.stdout.flush()
np.save({out!r}, np.concatenate([[info, ev[0]], vc[0]]))
"""


def run_reference_lanczsg(tmp_path, n, maxint, abstol):
    out = str(tmp_path / "ref.npy")
    code = CHILD.format(oracle=os.path.join(ROOT, "oracle"), n=n, maxint=maxint,
                        abstol=abstol, out=out)
    # the reference prints one Latin-1 byte in its STOP message
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=300, encoding="latin-1")
    log = [(int(m[0]), float(m[1]), float(m[2])) for m in re.findall(
        r"Interaction\s+(\d+);\s*Energy =\s*(\S+)/a\.u\.;\s*WF converg\. =\s*(\S+);", r.stdout)]
    res = np.load(out) if os.path.exists(out) else None
    return r.returncode, log, res, r.stdout


@pytest.mark.parametrize("n", [400, 600])
def test_lanczsg_oracle_matches_reference(ref, orc, tmp_path, n):
    rc, rlog, res, _ = run_reference_lanczsg(tmp_path, n, 60, 1e-6)
    assert rc == 0 and res is not None
    shm, v = oscillator(n)
    info, w, vecs, log = orc.lanczsg([n], 1, 1, 60, 1e-6, shm, v, REF)
    assert info == 0 and int(res[0]) == 0
    # the same restarts, the same change measure and energies within the
    # bisection tolerance.  The measure reads two tail components of size
    # ~1e-9, which carry the Lanczos vectors' absolute rounding (~1e-16)
    assert [r[1] for r in log] == [r[0] for r in rlog]
    for (_, kk, e, s), (_, re_, rs) in zip(log, rlog):
        assert abs(s - rs) <= 1e-6 * abs(rs) + 1e-13
        assert abs(e - re_) <= 1e-6
    assert abs(w[0] - res[1]) <= 1e-6
    assert np.abs(vecs[0] - res[2:]).max() <= 1e-12


def test_lanczsg_reference_stops_without_convergence(ref, orc, tmp_path):
    # maxint = 2: three restarts, then the reference's STOP; the oracle
    # reports it as INFO = 1 after the same three log lines
    rc, rlog, res, out = run_reference_lanczsg(tmp_path, 400, 2, 1e-6)
    assert res is None and "didn" in out and "convergence" in out
    shm, v = oscillator(400)
    info, _, _, log = orc.lanczsg([400], 1, 1, 2, 1e-6, shm, v, REF)
    assert info == 1 and len(log) == len(rlog) == 3


def test_lanczsg_exact_mode_converges_to_the_eigenvector(orc):
    # EXACT measures the whole change |VI - s VIAUX|_1 < ABSTOL, so the
    # vector is the FD ground state to that tolerance
    n = 300
    shm, v = oscillator(n, 7.0)
    info, w, vecs, log = orc.lanczsg([n], 1, 1, 400, 1e-7, shm, v, EXACT)
    assert info == 0 and log[-1][3] <= 1e-7
    e, z = np.linalg.eigh(fd_matrix_1d(shm, v))
    assert abs(w[0] - e[0]) <= 1e-7
    z0 = z[:, 0] * np.sign(z[:, 0] @ vecs[0])
    assert np.abs(vecs[0] - z0).max() <= 1e-7


@pytest.mark.parametrize("shape", [(300,), (18, 14)])
def test_mtrxdiag_reference_matches_dense_eigh(ref, shape):
    rng = np.random.default_rng(3)
    n = int(np.prod(shape))
    shm = [0.7, 0.45][:len(shape)]
    v = rng.uniform(0, 4, n)
    info, ev, vc = ref.mtrxdiag(list(shape), 1, 5, 1e-13, shm, v)
    assert info == 0
    # the matrix of ap1d.f / ap2d.f: 30, -16, 1 times SHM along x inside a
    # grid row, along y between rows, V on the diagonal
    h = np.zeros((n, n))
    nx = shape[0]
    for j in range(n):
        h[j, j] = 30 * sum(shm) + v[j]
        for d, c in ((1, -16.0), (2, 1.0)):
            if j % nx >= d:
                h[j, j - d] = h[j - d, j] = c * shm[0]
            if len(shape) == 2 and j >= d * nx:
                h[j, j - d * nx] = h[j - d * nx, j] = c * shm[1]
    e, z = np.linalg.eigh(h)
    assert np.abs(ev - e[:5]).max() <= 1e-11
    for k in range(5):
        assert np.abs(np.abs(vc[k] @ z[:, k]) - 1) <= 1e-10
