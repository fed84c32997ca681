"""Result files in the reference's output format, and their comparison.

* :func:`format_frequency_block` writes one OUTPUTEC block (Outputec.for:9-68):
  the frequency banner, boundary displacements and tractions, interior
  displacements and the nine interior stresses, each complex number as two
  ``D11.4`` fields (FORMAT ``I5,3(2X,D11.4,',',D11.4)`` and
  ``I4,9(2X,D11.4,',',D11.4)``).
* :func:`parse_dynamic_results` reads those blocks back (the reference's
  ``dtester.c`` locates the same sections).
* :func:`compare_results` is the reference's acceptance rule
  (dtester.c:204-267): the largest row sum of |difference| for boundary
  displacements, tractions and interior displacements; for the stresses,
  each point's largest row sum over its 3x3 tensor, summed over the points;
  each against ``eps = 1e-3``, plus a NaN check of the first result set.
"""
from __future__ import annotations

import math
import re
from typing import Dict, List

import numpy as np

_NUM = re.compile(r"[-+]?\d*\.\d+(?:[DdEe][-+]?\d+|[-+]\d{3})")


def fortran_d(x: float, w: int = 11, d: int = 4) -> str:
    """Fortran ``Dw.d`` edit descriptor (0.dddd mantissa, D exponent)."""
    if x != x:
        return "NaN".rjust(w)
    if x == 0.0:
        s = "0." + "0" * d + "D+00"
        return s.rjust(w)
    sign = "-" if x < 0 else ""
    ax = abs(x)
    e = math.floor(math.log10(ax)) + 1
    m = round(ax / 10.0 ** e, d)
    if m >= 1.0:
        m /= 10.0
        e += 1
    if m < 0.1:
        m *= 10.0
        e -= 1
    mant = ("%." + str(d) + "f") % m
    if abs(e) < 100:
        exp = "D%+03d" % e
    else:
        exp = "%+04d" % e
    s = sign + mant + exp
    return s.rjust(w)


def _z(v: complex) -> str:
    return "  " + fortran_d(v.real) + "," + fortran_d(v.imag)


def format_frequency_block(fr: float, zfi, zdfi, zdsol, zssol, nbe: int, l: int) -> str:
    """One OUTPUTEC block (Outputec.for:17-65), ISO-8859-1 text as in the
    reference output files."""
    out: List[str] = []
    star79 = " " + "*" * 79
    out.append("")
    out.append(star79)
    out.append("")
    out.append(" " * 30 + "FREQÜÊNCIA = " + fortran_d(fr, 14, 7))
    out.append("")
    out.append(star79)
    out.append("")
    out.append(" RESULTADOS")
    out.append("")
    out.append("  NÓS DO CONTORNO")
    out.append("")
    out.append(" ELEM      DESLOCAMENTO X           DESLOCAMENTOT Y           DESLOCAMENTO Z")
    out.append("")
    for i in range(nbe):
        out.append("%5d" % (i + 1) + "".join(_z(complex(zfi[3 * i + j])) for j in range(3)))
    out.append("")
    out.append(" ELEM        TRACTION X               TRACTION Y               TRACTION Z")
    out.append("")
    for i in range(nbe):
        out.append("%5d" % (i + 1) + "".join(_z(complex(zdfi[3 * i + j])) for j in range(3)))
    if l > 0:
        out.append("")
        out.append("")
        out.append("  PONTOS INTERNOS")
        out.append("")
        out.append(" PONTO     DESLOCAMENTO X           DESLOCAMENTO Y           DESLOCAMENTO Z")
        for i in range(l):
            out.append("%5d" % (i + 1) + "".join(_z(complex(zdsol[3 * i + j])) for j in range(3)))
        out.append("")
        out.append("")
        out.append("  PONTOS INTERNOS")
        out.append("")
        out.append(" PONTO        SIGMAXX                 SIGMAXY                  SIGMAXZ"
                   "                  SIGMAYX                  SIGMAYY                  SIGMAYZ"
                   "                  SIGMAZX                  SIGMAZY                  SIGMAZZ")
        for i in range(l):
            out.append("%4d" % (i + 1) + "".join(_z(complex(zssol[9 * i + j])) for j in range(9)))
    out.append(" " + "*" * 230)
    return "\n".join(out) + "\n"


def _nums(line: str) -> List[float]:
    vals = []
    for tok in _NUM.findall(line):
        t = tok.replace("D", "E").replace("d", "e")
        if "E" not in t and "e" not in t:  # Fortran drops the D for |exp| >= 100
            k = max(t.rfind("+"), t.rfind("-"))
            t = t[:k] + "E" + t[k:]
        vals.append(float(t))
    return vals


def _rows(lines: List[str], start: int, count: int, ncplx: int):
    out = np.zeros((count, ncplx), dtype=np.complex128)
    got = 0
    k = start
    while got < count and k < len(lines):
        s = lines[k].strip()
        k += 1
        if not s or not s.split()[0].isdigit():
            continue
        v = _nums(lines[k - 1][5:] if ncplx == 3 else lines[k - 1][4:])
        if len(v) < 2 * ncplx:
            continue
        out[got] = np.array(v[0:2 * ncplx:2]) + 1j * np.array(v[1:2 * ncplx:2])
        got += 1
    if got != count:
        raise ValueError("expected %d result rows, found %d" % (count, got))
    return out, k


def parse_dynamic_results(text: str, nbe: int, l: int, block: int = 0) -> Dict[str, np.ndarray]:
    """Boundary u, t (NBE x 3), interior u (L x 3) and sigma (L x 9) of the
    ``block``-th frequency block of a D output file."""
    lines = text.replace("\r\n", "\n").split("\n")
    starts = [i for i, s in enumerate(lines) if s.strip() == "RESULTADOS"]
    k = starts[block]
    u, k = _rows(lines, k, nbe, 3)
    while "TRACTION X" not in lines[k]:
        k += 1
    t, k = _rows(lines, k, nbe, 3)
    res = {"u": u, "t": t}
    if l > 0:
        while "PONTOS INTERNOS" not in lines[k]:
            k += 1
        du, k = _rows(lines, k + 1, l, 3)
        while "PONTOS INTERNOS" not in lines[k]:
            k += 1
        sg, k = _rows(lines, k + 1, l, 9)
        res["du"] = du
        res["sigma"] = sg
    return res


def compare_results(a: Dict[str, np.ndarray], b: Dict[str, np.ndarray], eps: float = 1e-3):
    """dtester.c's acceptance rule.  Returns (ok, errors dict)."""
    def rowmax(x, y):
        return float(np.max(np.sum(np.abs(x - y), axis=1))) if x.size else 0.0

    err = {"u": rowmax(a["u"], b["u"]), "t": rowmax(a["t"], b["t"])}
    if "du" in a and "du" in b:
        err["du"] = rowmax(a["du"], b["du"])
        d = np.abs(a["sigma"] - b["sigma"]).reshape(-1, 3, 3)  # point, row, column
        err["sigma"] = float(np.sum(np.max(np.sum(d, axis=2), axis=1))) if d.size else 0.0
    nan = any(np.isnan(v).any() for v in a.values())
    ok = (not nan) and all(v <= eps for v in err.values())
    return ok, err
