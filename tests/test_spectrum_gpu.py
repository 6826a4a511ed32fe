"""Spectra of the benchmark surfaces on the device against the FP64 oracle:
the whole chain (propagation, C(t) assembly, Gaussian window, FFTNT) for
the 1-D Morse workload (512 points, 4096 steps of dt = 0.5 a.u., window
exp(-(t/500 a.u.)^2)) and the 256x256 coupled Morse surface (1024 steps
of dt = 1 a.u.), on the FP64 and mixed paths.  The peak positions must be
the oracle's (the same energy-grid points), the spectrum the oracle's to
the C(t) bar carried through the transform: |dW| <= sum_t |dC(t)| <= N
max|dC| relative to the peak.  tests/test_oracle_spectrum.py pins the
oracle's peaks to the grid Hamiltonian's eigenvalues."""
import numpy as np
import pytest

import espec_amd as E
from espec_amd import configs

pytestmark = pytest.mark.gpu


def peaks(w, frac=0.01):
    return [i for i in range(1, len(w) - 1)
            if w[i] > w[i - 1] and w[i] >= w[i + 1] and w[i] > frac * w.max()]


CASES = [("config1", 13, 0.5), ("config2", 11, 1.0)]
# (flags, |dC| bar of the path, relative |dW| bar): W is a windowed sum of
# C(t), so |dW| / W_peak <= max|dC| / (population of the peak's state)
PATHS = [(E.FP64 | E.EXACT, 1e-10, 1e-9), (E.MIXED | E.EXACT, 1e-6, 1e-5)]


@pytest.mark.parametrize("name,mf,dt_au", CASES)
def test_spectrum_peaks_match_oracle(gpu, orc, name, mf, dt_au):
    import oracle
    n2 = 1 << (mf - 1)
    c = getattr(configs, name)(nsteps=n2)
    lc, tf = orc.lc_rule(0.0, n2 * c.dt_fs, c.dt_fs, mf)
    assert E.lc_rule(0.0, n2 * c.dt_fs, c.dt_fs, mf) == (lc, tf)
    wp = np.exp(-(1 << mf) * dt_au / 500.0)  # window exp(-(t / 500 a.u.)^2)
    print(f"case: spectrum {name} oracle", flush=True)
    ro = orc.pspofft(c.np, c.dt_fs, 0.0, tf, lc, c.sh, c.sm, c.u, c.v, c.vpot, oracle.EXACT)
    uo, vo = orc.assemble_corr(mf, ro["cr"][:n2], ro["ci"][:n2])
    eo, wo, wio = orc.spectrumtd(1, mf, tf / n2, tf, wp, 0.0, uo, vo, oracle.EXACT)
    po = peaks(wo)
    assert len(po) >= 2
    for flags, cbar, wbar in PATHS:
        print(f"case: spectrum {name} flags={flags}", flush=True)
        plan = E.SpoPlan(c.np, c.dt_fs, lc, c.sh, c.sm, c.vpot[None], c.u[None], c.v[None],
                         maxsamp=n2 + 2, flags=flags)
        plan.run(n2 * lc)
        g = plan.get(0, psi=False)
        plan.close()
        dc = max(np.abs(g["cr"][:n2] - ro["cr"][:n2]).max(),
                 np.abs(g["ci"][:n2] - ro["ci"][:n2]).max())
        u, v = E.assemble_corr(mf, g["cr"][:n2], g["ci"][:n2])
        e, w, wi = E.spectrumtd(1, mf, tf / n2, tf, wp, 0.0, u, v, E.EXACT)
        rel = max(np.abs(w - wo).max(), np.abs(wi - wio).max()) / wo.max()
        print(f"  max|dC| {dc:.2e}  max|dW|/W_peak {rel:.2e}  peaks {len(po)}", flush=True)
        assert np.array_equal(e, eo)
        assert peaks(w) == po
        assert dc <= cbar
        assert rel <= wbar
