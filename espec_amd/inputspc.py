"""input.spc for the propagation path: reader and runner.

parse() reads the keywords of eSPec's input groups that the propagation
and spectrum path uses (rdinput.f:120-640):
  **MAIN      *TITLE *TPCALC *DIMENSION *POTENTIAL *GRID_RANGES *MASS
              *INIEIGVC *CHANGE *PRTCRL (other switches are kept as text)
  **TI        *TPDIAG *NIST *ABSTOL *NREORT
  **TD        *PROPAG *PRPGSTATE *TPTRANS *PRPTOL
  **SPECTRUM  *FOURIER *WINDOWING *HWHM
Values follow Fortran list-directed input: blanks or commas separate
them, '/' ends a record, '!' starts a comment and D marks an exponent.

run() follows the main program (espec.f) for .SPECTRUM/.TD and .ONLYSPEC
on 1-D grids: potentials from .OHS/.MORSE (makecst.f:82-175,
potfuncs.f:63-137) or a .FILE (rdpte.f/rdpt.f), the initial state from
LANCZS on the device (.CALC; .MTRXDIAG is served by the same full-
dimension LANCZS) or read from the file (.GETR), the LC/TF rule
(espec.f:1324-1353), the propagation (.PFSPO, .PFSOD, .PSIL, .PPSIL on the
device), the C(t) assembly (espec.f:1546-1566), HWHM and SPECTRUMTD.
The initial state, the propagation, the assembly and the spectrum run
through libespec_hip.so; the potentials' closed forms are evaluated on
the host, as the reference's generators evaluate them once per run.
"""
import os
from dataclasses import dataclass, field

import numpy as np

from . import (A0A, AMU, EXACT, FP64, PULSE_NONE, REF, assemble_corr, hwhm, lanczs, lc_rule,
               pplnz, plnz, psodfft, pspofft, spectrumtd)

# makecst.f:36-50
CCGS = 2.99792458e10
TAU = 2.4188843243e-17
TWOPI = 6.2831853071795864769252867663
LMTREORT = 300  # param.h: LANCZS's Gram-Schmidt window


def _num(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def _values(line):
    """List-directed values of one record (up to '/' or '!')."""
    line = line.split("!")[0]
    line = line.split("/")[0]
    return [t for t in line.replace(",", " ").split() if t]


def _word(line):
    return line.split("!")[0].strip()


@dataclass
class Spc:
    """The parsed input; defaults are the main program's (espec.f:110-180)."""
    title: str = ""
    tpcalc: str = ".SPECTRUM"
    tdi: str = ".NC"
    corr_file: str = ""
    dim: str = ".1D"
    np: list = field(default_factory=lambda: [1])
    potential: str = ".NULL"
    pot_file: str = ""
    cst_i: list = field(default_factory=list)
    cst_f: list = field(default_factory=list)
    xi: list = field(default_factory=list)
    xf: list = field(default_factory=list)
    grid_au: bool = False
    mass: list = field(default_factory=list)
    inieigvc: str = ".CALC"
    change: bool = True
    prtcrl: str = ".NO"
    tpdiag: str = ".MTRXDIAG"
    nist: list = field(default_factory=lambda: [1, 0])
    abstol: float = 1e-6
    nreort: int = LMTREORT
    propag: str = ".NULL"
    ti: float = 0.0
    tf: float = 1e-4
    dt: float = 1e-4
    mp: int = 10
    prpgstate: int = 0
    tptrans: str = ".ONE"
    prptol: float = 3.0
    mf: int = 13
    window: str = ".NONE"
    wp: float = 0.0
    hwhm: tuple = None  # (CHWHM, DE, E1I, ADE) in eV
    tpabsor: str = ".NULL"  # *ABSORBING shape (ABSORB false by default, espec.f)
    efc: str = ".NONE"  # *PULSE_AND_DIPOLE pulse shape
    other: dict = field(default_factory=dict)


def parse(text, base=None):
    """Spc from the text of an input.spc; base: directory of named files."""
    s = Spc()
    lines = [ln.rstrip("\n") for ln in text.splitlines()]
    i = 0
    group = ""

    def nxt():
        nonlocal i
        i += 1
        if i >= len(lines):
            raise ValueError("input.spc ends inside a keyword")
        return lines[i]

    while i < len(lines):
        ln = lines[i]
        w = _word(ln)
        if w.startswith("**"):
            group = w[2:].split()[0] if len(w) > 2 else ""
            if group == "END":
                break
        elif w.startswith("*"):
            key = w.split()[0].upper()
            if key == "*TITLE":
                s.title = nxt().strip()
            elif key == "*TPCALC":
                s.tpcalc = _word(nxt())
                if s.tpcalc.startswith(".SPECTRUM"):
                    s.tdi = _word(nxt())
                elif s.tpcalc.startswith(".ONLYSPEC"):
                    s.corr_file = _word(nxt())
                elif s.tpcalc.startswith((".PROPAGATION", ".CORRELATION", ".COLLISION")):
                    s.tdi = ".TD"
                elif s.tpcalc.startswith(".ENERGY"):
                    s.tdi = ".TI"
            elif key == "*DIMENSION":
                s.dim = _word(nxt())
                s.np = [int(_num(t)) for t in _values(nxt())]
            elif key in ("*POTENTIAL", "*INITIAL_POTENTIAL") and group in ("MAIN", "TI"):
                s.potential = _word(nxt())
                if s.potential.startswith(".FILE"):
                    s.pot_file = _word(nxt())
                else:
                    s.cst_i = [_num(t) for t in _values(nxt())]
                    s.cst_f = [_num(t) for t in _values(nxt())]
            elif key == "*GRID_RANGES":
                v = [_num(t) for t in _values(nxt())]
                s.xi, s.xf = v[0::2], v[1::2]
                s.grid_au = w[13:15] == "au"
            elif key == "*MASS":
                s.mass = [_num(t) for t in _values(nxt())]
            elif key == "*INIEIGVC":
                s.inieigvc = _word(nxt())
            elif key == "*CHANGE":
                s.change = not _word(nxt()).startswith((".NO", ".OFF"))
            elif key == "*PRTCRL":
                p = _word(nxt())
                if p.startswith((".FULL", ".ALL")):
                    p = ".YES"
                elif p.startswith((".NONE", ".NULL", ".EMPTY")):
                    p = ".NO"
                s.prtcrl = p
            elif key == "*TPDIAG":
                s.tpdiag = _word(nxt())
            elif key == "*NIST":
                s.nist = [int(_num(t)) for t in _values(nxt())]
            elif key == "*ABSTOL":
                s.abstol = _num(_values(nxt())[0])
            elif key == "*NREORT":
                s.nreort = int(_num(_values(nxt())[0]))
            elif key == "*PROPAG":
                s.propag = _word(nxt())
                v = [_num(t) for t in _values(nxt())]
                s.ti, s.tf, s.dt = v[0], v[1], v[2]
                if len(v) > 3 and s.propag.startswith((".PSIL", ".PPSIL", ".PPLNZ")):
                    s.mp = int(v[3])
            elif key == "*PRPGSTATE":
                s.prpgstate = int(_num(_values(nxt())[0]))
            elif key == "*TPTRANS":
                s.tptrans = _word(nxt())
            elif key == "*PRPTOL":
                s.prptol = _num(_values(nxt())[0])
            elif key == "*FOURIER":
                s.mf = int(_num(_values(nxt())[0]))
            elif key == "*WINDOWING":
                s.window = _word(nxt())
                s.wp = _num(_values(nxt())[0])
            elif key == "*HWHM":
                s.hwhm = tuple(_num(t) for t in _values(nxt())[:4])
            elif key.startswith("*ABSORBING"):  # rdinput.f:526-548
                s.tpabsor = _word(nxt())
            elif key.startswith("*PULSE_AND_DIPOLE"):  # rdinput.f:549-575
                s.efc = _word(nxt())
            else:
                s.other[key] = _word(nxt()) if i + 1 < len(lines) else ""
        i += 1
    s.other["base"] = base or ""
    return s


def read(path):
    with open(path) as f:
        return parse(f.read(), os.path.dirname(os.path.abspath(path)))


# ----------------------------------------------------------- potentials

def grid(s):
    """Points of axis 1 in bohr and the step (espec.f:345, 619-622)."""
    n = s.np[0]
    xi, xf = s.xi[0], s.xf[0]
    if s.grid_au:  # *GRID_RANGES au: converted to angstrom first (rdinput.f:209-214)
        xi, xf = xi * A0A, xf * A0A
    sh_a = (xf - xi) / (n - 1.0)
    return xi / A0A + np.arange(n) * (sh_a / A0A), sh_a / A0A


def potential(kind, cst, x, m):
    """.OHS: k/2 (x - re)^2 with k = (w c tau 2 pi)^2 m (makecst.f:84-86,
    potfuncs.f:63-64); .MORSE: D (1 - exp(-a (x - re)))^2 with
    D = 2 pi c tau w^2 / (4 wx), a = sqrt(k / (2 D)) (makecst.f:141-148,
    potfuncs.f:104-105).  re in angstrom, w and wx in cm^-1."""
    k = (cst[1] * CCGS * TAU * TWOPI) ** 2 * m
    re = cst[0] / A0A
    if kind.startswith(".OHS"):
        return 0.5 * k * (x - re) ** 2
    if kind.startswith(".MORSE"):
        d = TWOPI * CCGS * TAU * cst[1] ** 2 / (4.0 * cst[2])
        a = np.sqrt(k / (2.0 * d))
        return d * (1.0 - np.exp(-a * (x - re))) ** 2
    raise ValueError(f"potential {kind} is not supported here")


def _sections(path):
    """The numeric sections of a two-section .FILE (eigenvector, then the
    potential; rdpte.f:40-62, rdpt.f:40-98), split at '#' lines."""
    secs, cur = [], None
    with open(path) as f:
        for ln in f:
            if ln.startswith("#"):
                cur = []
                secs.append(cur)
                continue
            p = ln.split()
            if cur is None or len(p) != 2:
                continue
            try:
                cur.append([_num(p[0]), _num(p[1])])
            except ValueError:
                continue
    return [np.array(c) for c in secs if c]


# ---------------------------------------------------------------- run

PROPAGATORS = (".PFSPO", ".PFSOD", ".PSIL", ".PPSIL", ".PPLNZ")


def run(s, flags=FP64 | REF):
    """Run a parsed input on the device; returns dict(lc, tf, cr, ci, t, u, v,
    e, w, wi): the C(t) samples, the assembled correlation function on
    t = -TF + (i-1) DTF and the printed spectrum columns (E, W, Im)."""
    base = s.other.get("base", "")
    iwin = 1 if s.window.startswith(".SG") else 0
    if s.tpcalc.startswith(".ONLYSPEC"):  # getcorr.f: the printed correlation function
        rows = []
        with open(os.path.join(base, s.corr_file)) as f:
            for ln in f:
                q = ln.split()
                if len(q) == 3:
                    try:
                        rows.append([_num(a) for a in q])
                    except ValueError:
                        pass
        c = np.array(rows)
        tf = -c[0, 0]
        u, v = c[:, 1].copy(), c[:, 2].copy()
        out = dict(lc=0, tf=tf, cr=None, ci=None)
    else:
        if not s.tdi.startswith(".TD") or s.dim[:3] != ".1D":
            raise ValueError("run() covers .SPECTRUM .TD and .ONLYSPEC on 1-D grids")
        if not s.propag.startswith(PROPAGATORS):
            raise ValueError(f"propagator {s.propag} is not supported here")
        # An absorber or a laser changes the propagation (absorb.f, ef.f);
        # run() propagates without either, so it refuses inputs that ask
        # for them rather than return a different C(t).
        if not s.tpabsor.startswith(".NULL"):
            raise ValueError(f"*ABSORBING {s.tpabsor}: absorbers are not run by run()")
        if not s.efc.startswith((".NONE", ".NULL")):
            raise ValueError(f"*PULSE_AND_DIPOLE {s.efc}: laser pulses are not run by run()")
        m = s.mass[0] * AMU
        n = s.np[0]
        if s.potential.startswith(".FILE"):
            secs = _sections(os.path.join(base, s.pot_file))
            psi_f, pot = secs[0], secs[1]
            x = pot[:, 0]
            sh = (x[-1] - x[0]) / (n - 1.0)  # the file's grid is in bohr
            v1 = pot[:, 1].copy()
            v0 = v1
        else:
            x, sh = grid(s)
            v0 = potential(s.potential, s.cst_i, x, m)
            v1 = potential(s.potential, s.cst_f, x, m)
            psi_f = None
        shm = 1.0 / (24.0 * m * sh * sh)  # espec.f:630
        if s.inieigvc.startswith(".GETR"):
            if psi_f is None:
                raise ValueError(".GETR needs a .FILE potential")
            u0 = psi_f[:, 1].copy()
        else:
            k = s.prpgstate
            _, vc = lanczs([n], 1, k + 1, min(n, s.nreort), [shm], v0, s.abstol)
            u0 = vc[k] / np.linalg.norm(vc[k])
        v00 = np.zeros(n)
        lc, tf = lc_rule(s.ti, s.tf, s.dt, s.mf, s.change)
        p = s.propag
        if p.startswith(".PFSPO"):
            r = pspofft([n], s.dt, s.ti, tf, lc, [sh], [m], u0, v00, v1, flags)
        elif p.startswith(".PFSOD"):
            r = psodfft([n], s.dt, s.ti, tf, lc, [sh], [m], u0, v00, v1, flags)
        elif p.startswith(".PSIL"):
            r = plnz([n], s.mp, s.abstol, s.dt, s.ti, tf, lc, [shm], u0, v00, v1, flags)
        else:  # .PPSIL / .PPLNZ without a laser section: EFC .NONE
            r = pplnz([n], s.mp, s.abstol, s.dt, s.ti, tf, lc, [shm], u0, v00, v1,
                      np.zeros(n), (PULSE_NONE, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0), flags)
        mc = 1 << (s.mf - 1)
        u, v = assemble_corr(s.mf, r["cr"][:mc], r["ci"][:mc])
        out = dict(lc=lc, tf=tf, cr=r["cr"], ci=r["ci"], psi0=(u0, v00))
    mc = 1 << (s.mf - 1)
    dtf = tf / mc  # espec.f:1600-1601
    out["t"] = -tf + np.arange(2 * mc) * dtf
    out["u"], out["v"] = u.copy(), v.copy()
    if s.hwhm is not None:
        chwhm, de, e1i, ade = s.hwhm
        u, v = hwhm(s.mf, dtf, tf, chwhm, de - e1i, u, v)
    else:
        ade = 0.0
    e, w, wi = spectrumtd(iwin, s.mf, dtf, tf, s.wp, ade, u, v, REF if flags & REF else EXACT)
    out.update(e=e, w=w, wi=wi)
    return out
