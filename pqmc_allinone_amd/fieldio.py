"""Phonon-field files and checkpoints.

Text field format (src/Main_PQMC.f90:355-368, read there when
import_ph_field is set, from ph_field_core1.dat in the MPI build): one line
per slice t = 1..ntime holding boson_field(:, t), the Nb values in F18.6.
The format keeps six decimals, so a field written and read back moves by at
most 5e-7 per value.

Checkpoint (this library's own format; the reference has no resume): the
state the sweep loop needs to continue as if it had not stopped, written at a
loop boundary before a backward sweep, where the reference itself re-enters
through init_g(1) (src/Main_PQMC.f90:87-95).  Resuming sets the fields, runs
init_g(1), restores the counters and the ran2 state and continues with the
next loop.  The ran2 state is the five words ran_seed(get=...) returns
(src/mod_ranstate.f90:81-104).  Layout, little-endian stream:

    char[8]    "PQMCCKP1"
    int32      L1, L2, Lt, ngroup, ne, Nb
    int32      loop (last completed), forward (1 if that loop went forward)
    int32      ran2 state i, j, k, m, n
    int64      positive_accept, negative_accept, global_accept, global_reject
    float64    err_fast_max
    float64    boson_field(Nb, Lt)
    int32      total_width, n_meas
    complex128 meas_sum(total_width)
    int32      bins_on, nbins, nmeas, meas_number, cur_bin, cur_meas
    if bins_on:
      complex128 acc(total_width)         the measurement in progress
      float64    bin_sum(total_width)     real parts of the bin so far
      float64    mean(total_width, nbins) finished bins
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

MAGIC = b"PQMCCKP1"


def write_ph_field(path: str, x: np.ndarray):
    """x: (Lt, Nb), row t-1 = boson_field(:, t)."""
    x = np.asarray(x, np.float64)
    with open(path, "w") as f:
        for row in x:
            f.write("".join(_f18_6(v) for v in row) + "\n")


def read_ph_field(path: str, Lt: int, Nb: int) -> np.ndarray:
    """Fixed-width F18.6 fields, Nb per line, Lt lines."""
    x = np.empty((Lt, Nb))
    with open(path) as f:
        for t in range(Lt):
            line = f.readline().rstrip("\n")
            if len(line) < 18 * Nb:
                raise ValueError(f"{path}: line {t + 1} holds fewer than {Nb} fields of width 18")
            x[t] = [float(line[18 * i:18 * (i + 1)]) for i in range(Nb)]
    return x


def _f18_6(v: float) -> str:
    s = f"{v:18.6f}"
    return s if len(s) <= 18 else "*" * 18


@dataclass
class Checkpoint:
    dims: tuple                 # L1, L2, Lt, ngroup, ne, Nb
    loop: int
    forward: bool
    ran2: tuple                 # i, j, k, m, n
    positive_accept: int
    negative_accept: int
    global_accept: int
    global_reject: int
    err_fast_max: float
    fields: np.ndarray          # (Lt, Nb)
    total_width: int
    n_meas: int = 0
    meas_sum: np.ndarray | None = None
    bins: dict | None = field(default=None)   # nbins, nmeas, meas_number, cur_bin, cur_meas, acc, bin_sum, mean


def save_checkpoint(path: str, c: Checkpoint):
    L1, L2, Lt, ng, ne, Nb = (int(v) for v in c.dims)
    tw = int(c.total_width)
    with open(path, "wb") as f:
        f.write(MAGIC)
        np.array([L1, L2, Lt, ng, ne, Nb, c.loop, int(bool(c.forward))], "<i4").tofile(f)
        np.array([int(np.int32(v)) for v in c.ran2], "<i4").tofile(f)
        np.array([c.positive_accept, c.negative_accept, c.global_accept, c.global_reject], "<i8").tofile(f)
        np.array([c.err_fast_max], "<f8").tofile(f)
        fx = np.ascontiguousarray(c.fields, "<f8")
        if fx.shape != (Lt, Nb):
            raise ValueError("fields must be (Lt, Nb)")
        fx.tofile(f)
        np.array([tw, c.n_meas], "<i4").tofile(f)
        ms = np.zeros(tw, np.complex128) if c.meas_sum is None else np.asarray(c.meas_sum, np.complex128)
        ms.astype("<c16").tofile(f)
        b = c.bins
        if b is None:
            np.zeros(6, "<i4").tofile(f)
            return
        np.array([1, b["nbins"], b["nmeas"], b["meas_number"], b["cur_bin"], b["cur_meas"]], "<i4").tofile(f)
        np.asarray(b["acc"], "<c16").tofile(f)
        np.asarray(b["bin_sum"], "<f8").tofile(f)
        np.ascontiguousarray(b["mean"], "<f8").tofile(f)   # (nbins, total_width) = mean(total_width, nbins)


def load_checkpoint(path: str) -> Checkpoint:
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path}: not a checkpoint of this library")
        h = np.fromfile(f, "<i4", 8)
        L1, L2, Lt, ng, ne, Nb, loop, fwd = (int(v) for v in h)
        ran2 = tuple(int(v) for v in np.fromfile(f, "<i4", 5))
        cnt = np.fromfile(f, "<i8", 4)
        efm = float(np.fromfile(f, "<f8", 1)[0])
        x = np.fromfile(f, "<f8", Nb * Lt).reshape(Lt, Nb)
        tw, n_meas = (int(v) for v in np.fromfile(f, "<i4", 2))
        ms = np.fromfile(f, "<c16", tw).astype(np.complex128)
        bh = np.fromfile(f, "<i4", 6)
        bins = None
        if int(bh[0]):
            nbins = int(bh[1])
            bins = {"nbins": nbins, "nmeas": int(bh[2]), "meas_number": int(bh[3]), "cur_bin": int(bh[4]),
                    "cur_meas": int(bh[5]), "acc": np.fromfile(f, "<c16", tw).astype(np.complex128),
                    "bin_sum": np.fromfile(f, "<f8", tw).astype(np.float64),
                    "mean": np.fromfile(f, "<f8", tw * nbins).reshape(nbins, tw).astype(np.float64)}
        if f.read(1):
            raise ValueError(f"{path}: trailing bytes")
    return Checkpoint((L1, L2, Lt, ng, ne, Nb), loop, bool(fwd), ran2, int(cnt[0]), int(cnt[1]), int(cnt[2]),
                      int(cnt[3]), efm, x, tw, n_meas, ms, bins)


def bins_state(ms) -> dict:
    """The part of a measio.MeasurementSystem a checkpoint keeps."""
    return {"nbins": ms.nbins, "nmeas": ms.nmeas, "meas_number": ms.meas_number, "cur_bin": ms.cur_bin,
            "cur_meas": ms.cur_meas, "acc": ms.acc.copy(), "bin_sum": ms.bin_sum.copy(), "mean": ms.mean.copy()}


def restore_bins(ms, b: dict):
    """Put a checkpoint's bin state back into a measio.MeasurementSystem of the same shape."""
    if (b["nbins"], b["nmeas"], b["meas_number"]) != (ms.nbins, ms.nmeas, ms.meas_number):
        raise ValueError("checkpoint bins do not match this measurement system")
    ms.cur_bin, ms.cur_meas = b["cur_bin"], b["cur_meas"]
    ms.acc = b["acc"].copy()
    ms.bin_sum = b["bin_sum"].copy()
    ms.mean = b["mean"].copy()
