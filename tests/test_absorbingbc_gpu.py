"""Absorber shapes of ABSORBINGBC (absorb.f) on the GPU against the oracle,
which tests/test_oracle_absorbingbc.py pins to the reference build bit for
bit.  Bar: 1e-14 relative (device exp/pow/cosh against the host's)."""
import numpy as np
import pytest

import espec_amd as E

pytestmark = pytest.mark.gpu

KINDS = [E.ABG_SMOOTHW, E.ABG_SMRS, E.ABG_VOPTIC_LINEAR, E.ABG_VOPTIC_KOSLOFF,
         E.ABG_VOPTIC_POLIN]


def close(g, o):
    # the 1-D .SMOOTHW with an even VOI(2) overflows to +inf near XI on fine
    # grids, in the reference as here: the overflowing points must agree
    fin = np.isfinite(o)
    if not np.array_equal(fin, np.isfinite(g)) or not np.array_equal(g[~fin], o[~fin]):
        return False
    return np.all(np.abs(g[fin] - o[fin]) <= 1e-14 * np.maximum(1.0, np.abs(o[fin])))


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n", [1, 200, 70000])
def test_absorbingbc_1d(gpu, orc, kind, n):
    sh, xi = 5.0 / n, 0.0
    xf = xi + (n - 1) * sh
    args = ([n], kind, [1.0], [xf - 1.0], [sh], [xi], [xf], [0.4, 2.0])
    print(f"case: ABSORBINGBC kind={kind} n={n}", flush=True)
    g, o = E.absorbingbc(*args), orc.absorbingbc(*args)
    assert close(g[0], o[0])
    assert g[1] == o[1]


def test_absorbingbc_2d(gpu, orc):
    args = ([512, 300], E.ABG_SMOOTHW, [0.4, 0.3], [1.5, 1.2], [0.004, 0.005], [0.0, -0.2],
            [2.044, 1.295], [0.5, 0.8])
    print("case: ABSORBINGBC 2-D 512x300", flush=True)
    g, o = E.absorbingbc(*args), orc.absorbingbc(*args)
    assert close(g[0], o[0])


def test_absorbingbc_rejects_unsupported(gpu):
    with pytest.raises(E.EspecError):
        E.absorbingbc([8, 8], E.ABG_VOPTIC_POLIN, [1, 1], [2, 2], [0.1, 0.1], [0, 0],
                      [0.7, 0.7], [1, 1])
    with pytest.raises(E.EspecError):
        E.absorbingbc([8], 9, [1], [2], [0.1], [0], [0.7], [1, 1])


def test_absorbingbc_fortran_dropin(gpu, orc):
    """ABSORBINGBC(DIM, TPABSOR, N, NP, AL, AR, SH, XI, XF, VOI, VABC, ALPHA)
    of libespec_f.so with the reference's argument list: by reference, the
    two CHARACTER lengths appended by value (gfortran's convention)."""
    import ctypes as C
    from espec_amd import build as B
    E._load()
    lib = C.CDLL(B.FLIB)
    n = 300
    sh, xi = 0.02, 0.0
    xf = xi + (n - 1) * sh
    d = lambda *v: np.array(v, dtype=np.float64)
    va, alpha = np.zeros(n), C.c_double(0.0)
    dim, tp = b".1D", b".VOPTIC_LINEAR"
    f = lib.absorbingbc_
    f.restype = None
    f(C.c_char_p(dim), C.c_char_p(tp), C.byref(C.c_int(n)),
      np.array([n, 1, 1], dtype=np.int32).ctypes.data_as(C.c_void_p),
      *[a.ctypes.data_as(C.c_void_p) for a in (d(1.0), d(xf - 1.0), d(sh), d(xi), d(xf),
                                              d(0.4, 1.0), va)],
      C.byref(alpha), C.c_size_t(len(dim)), C.c_size_t(len(tp)))
    o, _ = orc.absorbingbc([n], E.ABG_VOPTIC_LINEAR, [1.0], [xf - 1.0], [sh], [xi], [xf],
                           [0.4, 1.0])
    assert close(va, o) and np.any(o != 0)


@pytest.mark.parametrize("tp,kind", [(b".SMOOTHW_au     ", E.ABG_SMOOTHW),
                                     (b".SMRS_au", E.ABG_SMRS),
                                     (b".VOPTIC_POLIN ! comment", E.ABG_VOPTIC_POLIN)])
def test_absorbingbc_fortran_dropin_prefix(gpu, orc, tp, kind):
    """TPABSOR is matched on its leading characters as absorb.f:63-131 does:
    espec.f:652's '.SMOOTHW_au' and '.SMRS_au', and a CHARACTER*16 value
    read with '(A)' that still holds a trailing comment, select the shape
    of their prefix."""
    import ctypes as C
    from espec_amd import build as B
    E._load()
    lib = C.CDLL(B.FLIB)
    n = 257
    sh, xi = 0.02, 0.0
    xf = xi + (n - 1) * sh
    d = lambda *v: np.array(v, dtype=np.float64)
    va, alpha = np.full(n, -7.0), C.c_double(0.0)
    dim = b".1D"
    print(f"case: ABSORBINGBC drop-in TPABSOR={tp!r}", flush=True)
    f = lib.absorbingbc_
    f.restype = None
    f(C.c_char_p(dim), C.c_char_p(tp), C.byref(C.c_int(n)),
      np.array([n, 1, 1], dtype=np.int32).ctypes.data_as(C.c_void_p),
      *[a.ctypes.data_as(C.c_void_p) for a in (d(1.0), d(xf - 1.0), d(sh), d(xi), d(xf),
                                              d(0.4, 3.0), va)],
      C.byref(alpha), C.c_size_t(len(dim)), C.c_size_t(len(tp)))
    o, al = orc.absorbingbc([n], kind, [1.0], [xf - 1.0], [sh], [xi], [xf], [0.4, 3.0])
    assert close(va, o)
    if kind == E.ABG_SMOOTHW:
        assert alpha.value == al
