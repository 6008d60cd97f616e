"""Seismogram files in the reference's format: <fn>_rcv_<i>.out with one
"x y" line per sample (Receivers::write, Receivers.cpp:32-46), each value as a
default-formatted std::ostream writes it (6 significant digits, %g).  The C++
layer writes the same bytes (tspeed::Receivers::write, tsh_write_receivers)."""
from __future__ import annotations

import numpy as np


def _fmt(v: float, digits: int) -> str:
    # C's %g is what a default std::ostream prints for a double
    return "%.*g" % (digits if digits > 0 else 6, v)


def write_receivers(fn: str, seis, digits: int = 0) -> None:
    """seis[sample, receiver, component] -> fn_rcv_<r>.out.  digits <= 0: the
    reference's 6-digit format; digits = 17 keeps every double exactly."""
    seis = np.asarray(seis, dtype=np.float64)
    for r in range(seis.shape[1]):
        with open(f"{fn}_rcv_{r}.out", "w") as f:
            for s in range(seis.shape[0]):
                f.write(f"{_fmt(float(seis[s, r, 0]), digits)} {_fmt(float(seis[s, r, 1]), digits)}\n")


def read_receivers(fn: str, nrec: int) -> np.ndarray:
    cols = [np.loadtxt(f"{fn}_rcv_{r}.out", ndmin=2) for r in range(nrec)]
    return np.stack(cols, axis=1)
