"""Writes new-joe-kuo-6.1023: the first 1023 rows of Joe & Kuo's Sobol direction numbers.

Source of the numbers: the copy of S. Joe and F. Y. Kuo's ``new-joe-kuo-6.21201`` table that
SciPy ships (``scipy/stats/_sobol_direction_numbers.npz``: primitive polynomials and initial
direction numbers). Rows are written in Joe & Kuo's text format ``d s a m_1 ... m_s``, where the
polynomial x^s + a_1 x^(s-1) + ... + a_(s-1) x + 1 has a = (a_1 ... a_(s-1)) in binary.
Dimension 0 of the sampler (van der Corput) needs no row, so rows d = 2..1024 cover sampler
dimensions 1..1023.

Joe & Kuo's licence (BSD-style) permits redistribution with attribution; see
https://web.maths.unsw.edu.au/~fkuo/sobol/ . Reference: S. Joe and F. Y. Kuo, "Constructing Sobol
sequences with better two-dimensional projections", SIAM J. Sci. Comput. 30 (2008) 2635-2654.
"""
from __future__ import annotations

import os
from pathlib import Path

import numpy as np


def rows(count: int = 1023):
    import scipy.stats._sobol as sob

    z = np.load(os.path.join(os.path.dirname(sob.__file__), "_sobol_direction_numbers.npz"))
    poly, vinit = z["poly"], z["vinit"]
    out = []
    for k in range(1, count + 1):  # scipy row 0 is van der Corput
        p = int(poly[k])
        s = p.bit_length() - 1
        a = (p >> 1) & ((1 << (s - 1)) - 1) if s > 1 else 0
        ms = [int(x) for x in vinit[k, :s]]
        out.append((k + 1, s, a, ms))
    return out


def main() -> None:
    dst = Path(__file__).resolve().parent / "new-joe-kuo-6.1023"
    lines = ["d       s       a       m_i"]
    for d, s, a, ms in rows():
        lines.append(f"{d}\t{s}\t{a}\t" + " ".join(str(m) for m in ms) + " ")
    dst.write_text("\n".join(lines) + "\n")
    print(f"wrote {dst}")


if __name__ == "__main__":
    main()
