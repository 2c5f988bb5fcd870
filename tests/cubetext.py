"""Cube files for the reader tests: text and binary writers following
critic2's writegrid_cube (crystalmod@write.f90:3476-3570) and an
independent expected-value path (Python's float(), correctly rounded)."""
import re
import struct

import numpy as np

_BARE_EXP = re.compile(r"^([+-]?(?:\d+\.?\d*|\.\d+))([+-]\d+)$")


def fortran_float(tok: str) -> float:
    """A Fortran real constant to binary64 with Python's float (D/Q exponent
    letters, bare signed exponents)."""
    t = tok
    for a in "DdQq":
        t = t.replace(a, "E")
    m = _BARE_EXP.match(t)
    if m:
        t = m.group(1) + "E" + m.group(2)
    return float(t)


def expected_grid(text: bytes, n, data_offset: int) -> np.ndarray:
    """f(n1,n2,n3) x fastest from the value tokens after the header."""
    toks = text[data_offset:].split()
    n1, n2, n3 = n
    npts = n1 * n2 * n3
    v = np.array([fortran_float(t.decode()) for t in toks[:npts]], dtype=np.float64)
    # file order (((f(i,j,k),k),j),i): k fastest
    return v.reshape(n1, n2, n3).transpose(2, 1, 0).reshape(-1)


def header(n, step, x0=(0.0, 0.0, 0.0), atoms=(), ismo=False, precise=False):
    nat = len(atoms)
    lines = ["critic2-cube", "critic2-cube"]
    if precise:
        f = lambda i, v: f"{i:5d} " + " ".join(f"{x:22.14E}" for x in v) + " "
    else:
        f = lambda i, v: f"{i:5d}" + "".join(f"{x:12.6f}" for x in v)
    lines.append(f(-nat if ismo else nat, x0))
    for i in range(3):
        lines.append(f(n[i], step[i]))
    for z, x in atoms:
        lines.append(f"{z:4d}{0.0:5.1f}{x[0]:11.6f}{x[1]:11.6f}{x[2]:11.6f}")
    if ismo:
        lines.append("    1    7")
    return "\n".join(lines) + "\n"


def body(tokens_file_order, per_line=6, n3=None, sep=" ", eol="\n"):
    """Value records: per (i,j) row of n3 values, per_line per line."""
    out = []
    toks = list(tokens_file_order)
    rows = [toks[r:r + n3] for r in range(0, len(toks), n3)] if n3 else [toks]
    for row in rows:
        for s in range(0, len(row), per_line):
            out.append(sep + sep.join(row[s:s + per_line]) + eol)
    return "".join(out)


def fortran_records(*records: bytes, split=None) -> bytes:
    """gfortran unformatted sequential records; split=(record index, parts)
    writes that record as subrecords (negative markers mark continuation)."""
    out = b""
    for ir, rec in enumerate(records):
        if split and split[0] == ir:
            parts = split[1]
            cuts = np.linspace(0, len(rec), parts + 1).astype(int)
            for p in range(parts):
                chunk = rec[cuts[p]:cuts[p + 1]]
                lead = -len(chunk) if p < parts - 1 else len(chunk)
                trail = -len(chunk) if p > 0 else len(chunk)
                out += struct.pack("<i", lead) + chunk + struct.pack("<i", trail)
        else:
            out += struct.pack("<i", len(rec)) + rec + struct.pack("<i", len(rec))
    return out


def bincube(f_xfast, n, step, x0=(0.0, 0.0, 0.0), atoms=(), ismo=False, split=None) -> bytes:
    """step: 3x3, column i = axis-i vector (written as Fortran xd(3,3))."""
    nat = len(atoms)
    recs = [struct.pack("<i3d", -nat if ismo else nat, *x0),
            struct.pack("<3i9d", *n, *np.asarray(step, dtype=np.float64).reshape(-1, order="F"))]
    for z, x in atoms:
        recs.append(struct.pack("<i4d", z, 0.0, *x))
    recs.append(np.ascontiguousarray(f_xfast, dtype="<f8").tobytes())
    if split is not None:
        split = (len(recs) - 1, split)
    return fortran_records(*recs, split=split)


def fortran_e(x: float, style: int) -> str:
    """One field of the cube body formats with Python's correctly rounded
    formatting: style 0 E22.14E3, 1 1p,E12.5E3, 2 1p,e13.5e3, 3 1p,E15.7
    (two-digit exponent; past 99 the letter is dropped)."""
    import math
    W = {0: 22, 1: 12, 2: 13, 3: 15}[style]
    P = {0: 14, 3: 8}.get(style, 6)
    if math.isnan(x):
        t = "NaN"
    elif math.isinf(x):
        t = "-Infinity" if x < 0 else "Infinity"
    else:
        s = f"{x:.{P - 1}e}"
        neg = s.startswith("-")
        mant, ex = s.lstrip("-").split("e")
        digits = mant.replace(".", "")
        e = 0 if x == 0 else int(ex) + (1 if style == 0 else 0)
        es = f"E{'-' if e < 0 else '+'}{abs(e):03d}"
        if style == 3:
            es = f"E{'-' if e < 0 else '+'}{abs(e):02d}" if abs(e) <= 99 else f"{'-' if e < 0 else '+'}{abs(e):03d}"
        body = ("0." + digits if style == 0 else digits[0] + "." + digits[1:]) + es
        t = ("-" if neg else "") + body
    return t.rjust(W) if len(t) <= W else "*" * W


def expected_body(f_xfast, n, style: int) -> str:
    """Body records of writegrid_cube / write_cube_body from Python."""
    n1, n2, n3 = n
    z = np.asarray(f_xfast, dtype=np.float64).reshape(n3, n2, n1).transpose(2, 1, 0)  # [i, j, k]
    out = []
    for i in range(n1):
        for j in range(n2):
            row = [fortran_e(float(v), style) for v in z[i, j]]
            for s in range(0, n3, 6):
                chunk = row[s:s + 6]
                out.append("".join(" " + c for c in chunk) + (" " if len(chunk) < 6 else "") + "\n")
    return "".join(out)


def atoms_of(text: bytes, nat: int, ismo: bool = False):
    """(z, Cartesian positions) from the atom records of a text cube."""
    lines = text.split(b"\n")[6:6 + nat]
    z = [int(l.split()[0]) for l in lines]
    x = [[fortran_float(t.decode()) for t in l.split()[2:5]] for l in lines]
    return np.array(z, dtype=np.int32), np.array(x, dtype=np.float64).reshape(-1, 3)
