"""Seismogram files in the reference's format: <fn>_rcv_<i>.out with one
"x y" line per sample (Receivers::write, Receivers.cpp:32-46).  The C++ layer
writes the same files (tspeed::Receivers::write)."""
from __future__ import annotations

import numpy as np


def write_receivers(fn: str, seis) -> None:
    """seis[sample, receiver, component] -> fn_rcv_<r>.out"""
    seis = np.asarray(seis, dtype=np.float64)
    for r in range(seis.shape[1]):
        with open(f"{fn}_rcv_{r}.out", "w") as f:
            for s in range(seis.shape[0]):
                f.write(f"{float(seis[s, r, 0])!r} {float(seis[s, r, 1])!r}\n")


def read_receivers(fn: str, nrec: int) -> np.ndarray:
    cols = [np.loadtxt(f"{fn}_rcv_{r}.out", ndmin=2) for r in range(nrec)]
    return np.stack(cols, axis=1)
