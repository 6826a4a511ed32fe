"""The oracle's spectrum of the 1-D Morse workload (512 points on
[0.5, 6.0] bohr, m = 918, psi0 at 1.8 bohr with sigma 0.15, dt = 0.5 a.u.,
4096 steps, C(t) every step) against the mathematics.

C(t) = sum_n |c_n|^2 exp(-i E_n t) for the grid Hamiltonian's eigenpairs
(E_n, phi_n), so the Gaussian-windowed spectrum (SPECTRUMTD .SG,
spectrumtd.f:58-109, window exp(-(t/500 a.u.)^2)) peaks at the E_n the
packet populates.  The peaks sit on the transform's energy grid (spacing
ENC = 2 pi / (N dt) in eV), so each must lie within ENC/2 of an
eigenvalue of H = T + V, T the Fourier-grid kinetic operator the
split-operator step uses; the split-operator error at this dt moves the
eigenphases by far less than ENC.  A sign error in the transform or in
the propagation would put the peaks at -E_n; a wrong kinetic factor would
move them off the eigenvalues.
"""
import numpy as np

from oracle import EXACT

FATE = 27.2113834  # eV per hartree (spectrumtd.f:38)


def spectrum_config1(orc, mf=13):
    from espec_amd import configs
    c = configs.config1(nsteps=1 << (mf - 1))
    lc, tf = orc.lc_rule(0.0, (1 << (mf - 1)) * c.dt_fs, c.dt_fs, mf)
    o = orc.pspofft(c.np, c.dt_fs, 0.0, tf, lc, c.sh, c.sm, c.u, c.v, c.vpot, EXACT)
    n2 = 1 << (mf - 1)
    u, v = orc.assemble_corr(mf, o["cr"][:n2], o["ci"][:n2])
    wp = np.exp(-(1 << mf) * 0.5 / 500.0)  # exp(-(t / 500 a.u.)^2)
    e, w, wi = orc.spectrumtd(1, mf, tf / n2, tf, wp, 0.0, u, v, EXACT)
    return c, e, w


def peaks(e, w, frac=0.01):
    return [i for i in range(1, len(w) - 1)
            if w[i] > w[i - 1] and w[i] >= w[i + 1] and w[i] > frac * w.max()]


def test_config1_spectrum_peaks_at_grid_eigenvalues(orc):
    c, e, w = spectrum_config1(orc)
    n, h, m = c.np[0], c.sh[0], c.sm[0]
    k = 2 * np.pi * np.fft.fftfreq(n, d=h)
    t = np.real(np.fft.ifft((k ** 2 / (2 * m))[:, None] * np.fft.fft(np.eye(n), axis=0), axis=0))
    ev = np.linalg.eigvalsh((t + t.T) / 2 + np.diag(c.vpot)) * FATE
    enc = e[1] - e[0]
    assert abs(enc - 2 * np.pi / (len(e) * 0.5) * FATE) <= 1e-9 * enc
    pk = peaks(e, w)
    assert len(pk) >= 5
    used = set()
    for i in pk:
        j = int(np.argmin(np.abs(ev - e[i])))
        assert abs(ev[j] - e[i]) <= enc / 2, (e[i], ev[j])
        assert j not in used
        used.add(j)
    # the packet (centred 0.4 bohr from the minimum) populates the lowest
    # states most; the strongest peak is one of the two lowest
    assert int(np.argmin(np.abs(ev - e[int(np.argmax(w))]))) <= 1
