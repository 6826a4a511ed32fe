"""Extract the numeric golden blocks of the reference's regression tests.

Run once, in a container where the reference checkout is present:

    python tests/golden/make_fixtures.py [/root/reference]

For each regression output ttests/results.testN it writes
tests/golden/testN_<block>.txt.gz with the numbers of one printed block:

  table  'Partial auto-correlation function' rows: t/fs, E, Re C, Im C, |C|, norm
         (printed by pspofft.f:148 / plnz.f:155, format 1021)
  corr   'Correlation Function' rows: t/fs, Re C, Im C (espec.f:1570-1576)
  spec   'Spectrum' rows: E/eV, W, Re, Im (spectrumtd.f:98-109)

and copies the numeric columns of the test inputs that the propagations
read (ttests/eigpot_test3.inp: x, psi0 and x, V; ttests/correl8.inp: t,
Re C, Im C).  These are data (inputs and expected outputs), no source.
"""
import gzip
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def numeric_rows(lines, ncols):
    out = []
    for ln in lines:
        p = ln.split()
        if len(p) != ncols:
            continue
        try:
            [float(x.replace("D", "E")) for x in p]
        except ValueError:
            continue
        out.append(" ".join(p))
    return out


def block(lines, title, ncols, stop_titles, limit=None):
    start = None
    for i, ln in enumerate(lines):
        if ln.strip().startswith(title):
            start = i + 1
            break
    if start is None:
        return []
    end = len(lines)
    for j in range(start, len(lines)):
        if any(lines[j].strip().startswith(t) for t in stop_titles):
            end = j
            break
    rows = numeric_rows(lines[start:end], ncols)
    return rows[:limit] if limit else rows


def write(name, rows):
    path = os.path.join(HERE, name + ".txt.gz")
    with gzip.open(path, "wt") as f:
        f.write("\n".join(rows) + "\n")
    print(f"{path}: {len(rows)} rows")


def pplnz_table(tt):
    """Rows t/fs, ET, E0, Pr, Pi, |P0|^2, norm of results.test14."""
    lines = open(os.path.join(tt, "results.test14")).read().splitlines()
    return block(lines, "Auto-correlation function", 7, ["Eigenvector propagated"])


def main(ref):
    tt = os.path.join(ref, "ttests")
    stops = ["Correlation Function", "Spectrum:", "Partial auto", "Total correlation",
             "Calculating", "Propagating", "Eigenvector propagated"]
    limits = {"table": {6: 1200}, "corr": {6: 0}, "spec": {6: 0}}
    for t in (4, 5, 6, 8, 10, 12, 13):
        lines = open(os.path.join(tt, f"results.test{t}")).read().splitlines()
        for name, title, nc in (("table", "Partial auto-correlation function", 6),
                                ("corr", "Correlation Function", 3),
                                ("spec", "Spectrum:", 4)):
            lim = limits[name].get(t)
            if lim == 0:
                continue
            rows = block(lines, title, nc, [s for s in stops if not title.startswith(s)], lim)
            if rows:
                write(f"test{t}_{name}", rows)
    # test14 (.PPSIL): PPLNZ's table, 7 columns (pplnz.f:1011/1021)
    write("test14_table", pplnz_table(tt))
    ep = open(os.path.join(tt, "eigpot_test3.inp")).read().splitlines()
    sec = [i for i, ln in enumerate(ep) if ln.startswith("#")]
    write("eigpot_test3_psi", numeric_rows(ep[sec[0] + 1:sec[1]], 2))
    write("eigpot_test3_pot", numeric_rows(ep[sec[1] + 1:], 2))
    c8 = open(os.path.join(tt, "correl8.inp")).read().splitlines()
    write("correl8_corr", block(c8, "Correlation Function", 3, ["Spectrum:"]))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
