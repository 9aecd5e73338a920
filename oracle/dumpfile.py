"""TEST INFRASTRUCTURE ONLY: numpy restatement of the reference's 3-D field dump format, the
checker for the library's lmgx_file_3_device / lmgx_read_3_device.  Only ``tests/`` and
``tests/golden/`` import it.

code/file_3.F:5-158 (one rank): a Fortran sequential file of the level's blocks in order.
* filefrmt 0 (unformatted): records framed by 4-byte little-endian lengths (the record markers
  of gfortran/flang), numbl as int32 (file_3.F:94-97), then per block (nip1, njp1, nkp1) as
  3 x int32 (:99-115), then per block, k outer, j inner, one record of the nip1 values of row
  (j, k) (:130-138) - REAL*4 unless built with -DWRITE_DOUBLE (:28-32), the value converted with
  round-to-nearest.
* filefrmt 1 (formatted): list-directed numbl and dims (:96, :114), then each row written with
  '(8(1x,g15.8))' (:141-146): 16 characters a value, 8 values a line.
code/read_3.F:5-188 reads level 1 back: the numbl and dims checks (ierr 3 / 4, :94-121), one
record per row, the first nip1 values (REAL*4 unless -DREAD_DOUBLE), list-directed for text; a
missing file gives zeros (:155-164).  The list-directed integers are written here the way the
reference compiled with flang writes them (one blank, then the digits); that spacing is the
compiler's choice and read_3 does not depend on it.
"""
from __future__ import annotations

import math
import struct
from typing import List, Sequence, Tuple

import numpy as np


def _blocks(nijkp1: Sequence[Sequence[int]]) -> List[Tuple[int, int, int]]:
    return [tuple(int(v) for v in n) for n in nijkp1]


# ---- binary ---------------------------------------------------------------------------------
def _rec(payload: bytes) -> bytes:
    m = struct.pack("<i", len(payload))
    return m + payload + m


def write_bin(f: np.ndarray, offs: Sequence[int], nijkp1, real_kind: int = 4) -> bytes:
    """file_3 with filefrmt 0 of the field f (blocks at offs, natural layout)."""
    dt = "<f4" if real_kind == 4 else "<f8"
    bl = _blocks(nijkp1)
    out = [_rec(struct.pack("<i", len(bl)))]
    out += [_rec(struct.pack("<3i", *n)) for n in bl]
    for o, (ni, nj, nk) in zip(offs, bl):
        a = np.asarray(f[o:o + ni * nj * nk], np.float64).reshape(nk * nj, ni).astype(dt)
        out += [_rec(row.tobytes()) for row in a]
    return b"".join(out)


def read_bin(data: bytes, offs, nijkp1, n: int, real_kind: int = 4):
    """read_3 with filefrmt 0 -> (f, ierr) with ierr 0, 3 (numbl), 4 (dims) or 5 (read)."""
    f = np.zeros(n)
    bl = _blocks(nijkp1)
    pos = 0

    def rec():
        nonlocal pos
        if pos + 4 > len(data):
            raise EOFError
        (ln,) = struct.unpack_from("<i", data, pos)
        if ln < 0 or pos + 8 + ln > len(data) or struct.unpack_from("<i", data, pos + 4 + ln)[0] != ln:
            raise EOFError
        p = data[pos + 4: pos + 4 + ln]
        pos += 8 + ln
        return p
    try:
        r = rec()
        if len(r) < 4:
            return f, 5
        if struct.unpack_from("<i", r)[0] != len(bl):
            return f, 3
        for n3 in bl:
            r = rec()
            if len(r) < 12:
                return f, 5
            if tuple(struct.unpack_from("<3i", r)) != n3:
                return f, 4
        w = 4 if real_kind == 4 else 8
        dt = "<f4" if real_kind == 4 else "<f8"
        for o, (ni, nj, nk) in zip(offs, bl):
            for q in range(nj * nk):
                r = rec()
                if len(r) < ni * w:
                    return f, 5
                f[o + q * ni: o + (q + 1) * ni] = np.frombuffer(r, dt, ni)
    except EOFError:
        return f, 5
    return f, 0


# ---- text -----------------------------------------------------------------------------------
def g15_8(x: float) -> str:
    """Fortran G15.8 output editing of a double (F2008 10.7.5.2.2, round to nearest)."""
    w, d = 15, 8
    if not math.isfinite(x):                     # flang writes these right-justified
        return ("NaN" if math.isnan(x) else ("-Inf" if x < 0 else "Inf")).rjust(w)
    if x == 0.0:
        s = "0." + "0" * (d - 1)                 # F(w-4).(d-1), zero
        s = ("-" + s) if math.copysign(1.0, x) < 0 else s
        return s.rjust(w - 4) + " " * 4
    ax = abs(x)
    # decimal exponent after rounding to d significant digits
    m = "%.*e" % (d - 1, ax)
    k = int(m[m.index("e") + 1:]) + 1           # 10**(k-1) <= rounded |x| < 10**k
    if 0 <= k <= d:
        s = "%.*f" % (d - k, ax) + ("." if k == d else "")   # F11.0 keeps its point
        s = ("-" + s) if x < 0 else s
        return s.rjust(w - 4) + " " * 4
    # Ew.d: 0.dddddddd E+ee (E dropped for a 3-digit exponent)
    digits = m[0] + m[2:m.index("e")]
    e = k
    ex = ("E%+03d" % e) if abs(e) <= 99 else ("%+04d" % e)
    s = "0." + digits + ex
    s = ("-" + s) if x < 0 else s
    return s.rjust(w)


def write_txt(f: np.ndarray, offs, nijkp1) -> bytes:
    """file_3 with filefrmt 1."""
    bl = _blocks(nijkp1)
    lines = [" %d" % len(bl)]
    lines += [" %d %d %d" % n for n in bl]
    for o, (ni, nj, nk) in zip(offs, bl):
        a = np.asarray(f[o:o + ni * nj * nk], np.float64).reshape(nk * nj, ni)
        for row in a:
            for c in range(0, ni, 8):
                lines.append("".join(" " + g15_8(float(v)) for v in row[c:c + 8]))
    return ("\n".join(lines) + "\n").encode()


def _fortran_real(tok: str) -> float:
    t = tok.replace("D", "E").replace("d", "e")
    if "e" not in t.lower():
        # an exponent without its letter: 0.10000000-099
        for i in range(1, len(t)):
            if t[i] in "+-" and t[i - 1] not in "eE":
                t = t[:i] + "E" + t[i:]
                break
    return float(t)


def read_txt(data: bytes, offs, nijkp1, n: int):
    """read_3 with filefrmt 1 -> (f, ierr): list-directed reads, one read per row, each
    starting on a new line and taking values across lines as needed."""
    f = np.zeros(n)
    bl = _blocks(nijkp1)
    lines = data.decode().split("\n")
    ln = 0

    def read(count):
        nonlocal ln
        vals = []
        while len(vals) < count:
            if ln >= len(lines):
                raise EOFError
            vals += lines[ln].replace(",", " ").split()
            ln += 1
        return vals[:count]
    try:
        if int(read(1)[0]) != len(bl):
            return f, 3
        for n3 in bl:
            if tuple(int(v) for v in read(3)) != n3:
                return f, 4
        for o, (ni, nj, nk) in zip(offs, bl):
            for q in range(nj * nk):
                f[o + q * ni: o + (q + 1) * ni] = [_fortran_real(v) for v in read(ni)]
    except (EOFError, ValueError):
        return f, 5
    return f, 0
