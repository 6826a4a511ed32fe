"""Absorber shapes of ABSORBINGBC (absorb.f) in the oracle, against the
reference build (oracle/_ref; skipped where absent) for every TPABSOR the
reference implements, and against what the shapes fix without it: the
masks are 1 and the imaginary potentials 0 inside [AL, AR]; the masks
stay in (0, 1] and fall towards the edges for an odd exponent; the
.VOPTIC_POLIN potential with an even exponent is mirror-symmetric about
the midpoint of [AL, AR] on a symmetric grid; .VOPTIC_KOSLOFF does not
depend on x on either side; the 2-D .SMOOTHW mask along a grid line that
stays inside one axis's window is the 1-D mask of the other axis with
exponent 1."""
import numpy as np
import pytest

from oracle import (ABG_SMOOTHW, ABG_SMRS, ABG_VOPTIC_KOSLOFF, ABG_VOPTIC_LINEAR,
                    ABG_VOPTIC_POLIN)

KINDS = [ABG_SMOOTHW, ABG_SMRS, ABG_VOPTIC_LINEAR, ABG_VOPTIC_KOSLOFF, ABG_VOPTIC_POLIN]
N, XI, SH = 200, 0.0, 0.025
XF = XI + (N - 1) * SH
AL, AR = 1.0, 3.975


def grid():
    return XI + SH * np.arange(N)


def args1(kind, voi):
    return [N], kind, [AL], [AR], [SH], [XI], [XF], voi


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("voi", [[0.3, 1.0], [0.05, 2.0], [1.2, 3.0]])
def test_absorbingbc_matches_reference(orc, ref, kind, voi):
    a = orc.absorbingbc(*args1(kind, voi))
    b = ref.absorbingbc(*args1(kind, voi))
    assert np.array_equal(a[0], b[0])
    if kind == ABG_SMOOTHW:
        assert a[1] == b[1]


def test_absorbingbc_2d_matches_reference(orc, ref):
    args = ([40, 30], ABG_SMOOTHW, [0.4, 0.3], [1.5, 1.2], [0.05, 0.06], [0.0, -0.2],
            [1.95, 1.54], [0.5, 0.8])
    a, b = orc.absorbingbc(*args), ref.absorbingbc(*args)
    assert np.array_equal(a[0], b[0])


@pytest.mark.parametrize("kind", KINDS)
def test_absorbingbc_interior(orc, kind):
    va, _ = orc.absorbingbc(*args1(kind, [0.4, 2.0]))
    x = grid()
    inside = (x >= AL) & (x <= AR)
    want = 1.0 if kind in (ABG_SMOOTHW, ABG_SMRS) else 0.0
    assert np.all(va[inside] == want)
    assert np.all(va[~inside] != want)


def test_absorbingbc_mask_shapes(orc):
    x = grid()
    out = (x < AL) | (x > AR)
    m, alpha = orc.absorbingbc(*args1(ABG_SMOOTHW, [0.4, 1.0]))
    assert np.all((m > 0) & (m <= 1)) and np.all(m[out] < 1)
    left, right = x < AL, x > AR
    assert np.all(np.diff(m[left]) > 0) and np.all(np.diff(m[right]) < 0)
    assert alpha == pytest.approx(1 - 0.1 * (XF - AR), rel=1e-7)
    s, _ = orc.absorbingbc(*args1(ABG_SMRS, [0.4, 1.0]))
    d = np.where(left, AL - x, np.where(right, x - AR, 0.0))
    assert np.allclose(1 - s, 0.4 * d * d, rtol=0, atol=2e-16)


def test_absorbingbc_potential_shapes(orc):
    x = grid()
    left, right = x < AL, x > AR
    # mirror symmetry: AL - XI = 1.0 = XF - AR on this grid
    assert AL - XI == pytest.approx(XF - AR, abs=1e-12)
    w, _ = orc.absorbingbc(*args1(ABG_VOPTIC_POLIN, [0.7, 2.0]))
    assert np.allclose(w[left], w[right][::-1], rtol=1e-12, atol=0)
    assert np.all(w[left | right] > 0)
    k, _ = orc.absorbingbc(*args1(ABG_VOPTIC_KOSLOFF, [0.7, 2.0]))
    assert np.ptp(k[left]) == 0 and np.ptp(k[right]) == 0
    lin, _ = orc.absorbingbc(*args1(ABG_VOPTIC_LINEAR, [0.7, 1.0]))
    assert np.all(lin[left | right] > 0)
    assert np.all(np.diff(lin[left]) < 0) and np.all(np.diff(lin[right]) > 0)


def test_absorbingbc_2d_lines_are_1d_masks(orc):
    np_, al, ar = [48, 36], [0.5, 0.4], [1.8, 1.3]
    sh, xi = [0.05, 0.05], [0.0, 0.0]
    xf = [xi[a] + (np_[a] - 1) * sh[a] for a in range(2)]
    voi = [0.5, 0.8]
    m2, _ = orc.absorbingbc(np_, ABG_SMOOTHW, al, ar, sh, xi, xf, voi)
    m2 = m2.reshape(np_[1], np_[0])
    j = 16   # x2 = 0.8 inside [0.4, 1.3]
    m1, _ = orc.absorbingbc([np_[0]], ABG_SMOOTHW, al[:1], ar[:1], sh[:1], xi[:1], xf[:1],
                            [voi[0], 1.0])
    assert np.array_equal(m2[j], m1)
    i = 20   # x1 = 1.0 inside [0.5, 1.8]
    m1, _ = orc.absorbingbc([np_[1]], ABG_SMOOTHW, al[1:], ar[1:], sh[1:], xi[1:], xf[1:],
                            [voi[1], 1.0])
    assert np.array_equal(m2[:, i], m1)


def test_absorbingbc_unsupported(orc):
    with pytest.raises(ValueError):
        orc.absorbingbc([8, 8], ABG_VOPTIC_LINEAR, [1, 1], [2, 2], [0.1, 0.1], [0, 0],
                        [0.7, 0.7], [1, 1])
