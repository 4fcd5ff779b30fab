"""Measurement bins and the reference's output files.

``MeasurementSystem`` restates the run control of the reference's
measurement module for the observable layout of input/model/EPSOCZ.nml:262-283:

- begin_measure (src/mod_meas.f90:729-763): a window of 2*meas_number + 1
  slices around ntime/2 is summed into one measurement, then divided;
- bin means (:856-884): the real parts of a bin's measurements, summed in
  measurement order, divided by nmeas;
- output_bin_data / output_bin_count (:886-942): each finished bin's mean is
  appended to temp/tempout_core<id>.bin (raw float64, total_width values) and
  a line "C<id>   |    B<bin>" to temp/tempout_core_count.dat;
- output_all_data (:944-971): data/out_core<id>.bin holds mean(total_width,
  nbins), column-major;
- output_reading_guide (:973-1017): data/reading_guide.nml describes the layout.

File names are those of the reference's MPI build (one file per core id).
``postpro`` restates the statistics of the reference's post-processing
program (src/outputnew.f90:94-232): per MPI block, the mean over cores and
bins, the error bar of independent bins and the error bar of core means.

The device returns each slice's measurement with its k-space companions
already transformed.  For meas_number > 0 the window average of those
transforms equals the transform of the window average (the transform is
linear), to rounding.
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

from .host import SCALARS, observable_layout

KIND_SCALAR, KIND_FIELD = 1, 2


class MeasurementError(RuntimeError):
    pass


def observable_table(ncell: int):
    """(name, kind, width, offset_lo, offset_hi) with 1-based offsets, as the
    reading guide lists them."""
    out = []
    for name, lo, hi in observable_layout(ncell):
        kind = KIND_SCALAR if name in SCALARS else KIND_FIELD
        out.append((name, kind, hi - lo, lo + 1, hi))
    return out


class MeasurementSystem:
    """Bins of measurements for one chain (one core id of the reference)."""

    def __init__(self, ncell: int, ntime: int, nbins: int, nmeas: int, meas_number: int = 0,
                 output_addr: str | None = None, myid: int = 0, n_cores: int = 1, mpi_nblock: int = 1,
                 output_format: str = "f18.8"):
        if nbins <= 0 or nmeas <= 0:
            raise MeasurementError("init_from_namelist: nbins/nmeas must be >0")
        self.table = observable_table(ncell)
        self.ncell, self.ntime = ncell, ntime
        self.nbins, self.nmeas, self.meas_number = nbins, nmeas, meas_number
        self.total_width = self.table[-1][4]
        self.field_width = ncell
        self.output_addr, self.myid = output_addr, myid
        self.n_cores, self.mpi_nblock = n_cores, mpi_nblock
        self.output_format = output_format
        self.cur_bin, self.cur_meas = 1, 1
        self.acc = np.zeros(self.total_width, np.complex128)   # data(:, cur_meas, cur_bin)
        self.bin_sum = np.zeros(self.total_width)             # sum of real parts of the bin so far
        self.mean = np.zeros((nbins, self.total_width))        # row b-1 = mean(:, b)
        if output_addr is not None:
            os.makedirs(os.path.join(output_addr, "temp"), exist_ok=True)
            os.makedirs(os.path.join(output_addr, "data"), exist_ok=True)
            if myid == 0:
                self.write_reading_guide()

    # ------------------------------------------------------------ run control
    def is_measure_time(self, loop: int, time: int, warmup: int, meas_interval: int) -> bool:
        """The condition of src/Main_PQMC.f90:143-148."""
        if loop <= warmup or (loop - warmup) % meas_interval != 0:
            return False
        return self.ntime // 2 - self.meas_number <= time <= self.ntime // 2 + self.meas_number

    def begin_measure(self, forward: bool, time: int, sample: np.ndarray):
        """One slice's measurement (total_width complex) into the current bin."""
        if self.cur_bin > self.nbins:
            raise MeasurementError("begin_measure: no more bins left")
        self.acc += np.asarray(sample, np.complex128)
        m = self.meas_number
        last = (forward and time == self.ntime // 2 + m) or ((not forward) and time == self.ntime // 2 - m)
        if not last:
            return
        if m:
            # componentwise division, as Fortran's complex / integer (numpy's
            # complex division would multiply by the reciprocal)
            d = float(2 * m + 1)
            self.acc = (self.acc.real / d) + 1j * (self.acc.imag / d)
        self.bin_sum += self.acc.real
        self.acc = np.zeros(self.total_width, np.complex128)
        self.cur_meas += 1
        if self.cur_meas > self.nmeas:
            self.mean[self.cur_bin - 1] = self.bin_sum / float(self.nmeas)
            self.bin_sum = np.zeros(self.total_width)
            if self.output_addr is not None:
                self.output_bin_data(self.cur_bin)
            self.cur_meas = 1
            self.cur_bin += 1

    def data_analyse(self):
        """Means of every bin (unfinished bins stay zero) to data/out_core<id>.bin."""
        if self.output_addr is not None:
            self.output_all_data()
        return self.mean

    # ------------------------------------------------------------ files
    def _path(self, sub: str, name: str) -> str:
        return os.path.join(self.output_addr, sub, name)

    def output_bin_data(self, b: int):
        with open(self._path("temp", f"tempout_core{self.myid}.bin"), "ab") as f:
            np.ascontiguousarray(self.mean[b - 1], np.float64).tofile(f)
        with open(self._path("temp", "tempout_core_count.dat"), "a") as f:
            f.write(f"C{self.myid:4d}   |    B{b:4d}\n")

    def output_all_data(self):
        with open(self._path("data", f"out_core{self.myid}.bin"), "wb") as f:
            np.ascontiguousarray(self.mean, np.float64).tofile(f)

    def write_reading_guide(self):
        def i12(v):
            return f"{v:12d}"
        lines = [" &basics", " n_obs=" + i12(len(self.table)), " total_width=" + i12(self.total_width),
                 " n_bins=" + i12(self.nbins), " n_meas_per_bin=" + i12(self.nmeas),
                 " field_width=" + i12(self.field_width), " size_of_dp=" + i12(8),
                 " n_cores=" + i12(self.n_cores), " MPI_nblock=" + i12(self.mpi_nblock),
                 " MPI_one_block=" + i12(self.n_cores // self.mpi_nblock),
                 f' output_format="{self.output_format}"', " /", " "]
        for name, kind, width, lo, hi in self.table:
            lines += [" &obs", f' name="{name}"', " kind=" + i12(kind), " width=" + i12(width),
                      " offset_lo=" + i12(lo), " offset_hi=" + i12(hi), " /", " "]
        with open(self._path("data", "reading_guide.nml"), "w") as f:
            f.write("\n".join(lines) + "\n")


# ---------------------------------------------------------------- readers

_NML_ITEM = re.compile(r"(\w+)\s*=\s*(\"[^\"]*\"|'[^']*'|[^\s,/]+)")


def read_reading_guide(path: str) -> tuple[dict, list[dict]]:
    """The &basics group and the list of &obs groups of a reading guide."""
    text = open(path).read()
    groups = re.findall(r"&(\w+)(.*?)/", text, flags=re.S)
    basics, obs = {}, []
    for gname, body in groups:
        d = {}
        for k, v in _NML_ITEM.findall(body):
            v = v.strip("\"'")
            d[k.lower()] = int(v) if re.fullmatch(r"[+-]?\d+", v) else v
        if gname.lower() == "basics":
            basics = d
        elif gname.lower() == "obs":
            obs.append(d)
    return basics, obs


def read_bin_means(path: str, total_width: int) -> np.ndarray:
    """tempout_core<id>.bin or out_core<id>.bin as (nbins, total_width)."""
    a = np.fromfile(path, np.float64)
    if a.size % total_width:
        raise MeasurementError(f"{path}: {a.size} values is not a multiple of total_width {total_width}")
    return a.reshape(-1, total_width)


# ---------------------------------------------------------------- statistics

def postpro(data_dir: str, guide: str = "reading_guide.nml", data_file: str = "out_core") -> dict:
    """compute_out_mean_and_err of src/outputnew.f90:141-185 over the
    out_core<id>.bin files of ``data_dir``.  Returns name -> dict of arrays
    out_mean, out_err, large_err shaped (MPI_nblock, width).  Sums run in the
    reference's order (bins, then cores)."""
    basics, obs = read_reading_guide(os.path.join(data_dir, guide))
    tw, nb = basics["total_width"], basics["n_bins"]
    ncores = basics.get("n_cores", 1)
    nblock = basics.get("mpi_nblock", 1)
    one = basics.get("mpi_one_block", ncores // nblock)
    # bin_data(width, n_bins, one_block, nblock) per observable, as (nblock, one, n_bins, width)
    raw = np.zeros((nblock, one, nb, tw))
    for myid in range(ncores):
        a = np.fromfile(os.path.join(data_dir, f"{data_file}{myid}.bin"), np.float64)
        if a.size != tw * nb:
            raise MeasurementError(f"{data_file}{myid}.bin: expected {tw * nb} values, found {a.size}")
        raw[myid // one, myid % one] = a.reshape(nb, tw)
    out = {}
    n_all = one * nb
    for o in obs:
        lo, hi = o["offset_lo"] - 1, o["offset_hi"]
        d = raw[..., lo:hi]
        w = hi - lo
        core_mean = np.zeros((nblock, one, w))
        mean = np.zeros((nblock, w))
        err = np.zeros((nblock, w))
        lerr = np.zeros((nblock, w))
        for ib in range(nblock):
            tsum = np.zeros((one, w))
            for ic in range(one):
                s = np.zeros(w)
                for b in range(nb):
                    s = s + d[ib, ic, b]
                tsum[ic] = s
                core_mean[ib, ic] = s / float(nb)
            mean[ib] = _colsum(tsum) / float(n_all)
            tsum = np.zeros((one, w))
            for ic in range(one):
                s = np.zeros(w)
                for b in range(nb):
                    s = s + (d[ib, ic, b] - mean[ib]) ** 2
                tsum[ic] = s
            with np.errstate(divide="ignore", invalid="ignore"):
                err[ib] = np.sqrt(_colsum(tsum) / float(n_all - 1) / float(n_all))
                lerr[ib] = np.sqrt(_colsum((core_mean[ib] - mean[ib]) ** 2) / float(one - 1) / float(one))
        out[o["name"]] = {"kind": o["kind"], "width": w, "core_mean": core_mean, "out_mean": mean,
                          "out_err": err, "large_err": lerr}
    return out


def _colsum(t: np.ndarray) -> np.ndarray:
    s = np.zeros(t.shape[1:])
    for row in t:
        s = s + row
    return s


def _ffmt(v: float, fmt: str) -> str:
    m = re.fullmatch(r"[fF](\d+)\.(\d+)", fmt)
    if not m:
        raise MeasurementError(f"unsupported output_format {fmt!r}")
    w, d = int(m.group(1)), int(m.group(2))
    if math.isnan(v):
        return "NaN".rjust(w)
    s = f"{v:{w}.{d}f}"
    return s if len(s) <= w else "*" * w


def write_results(target: str, results: dict, block: int = 1, output_format: str = "f18.8"):
    """output_results of src/outputnew.f90:190-230: keep ``target`` up to its
    '#--observables--#' line, then append the table of block ``block``."""
    keep = []
    with open(target) as f:
        for line in f:
            keep.append(line.rstrip("\n"))
            if line.strip() == "#--observables--#":
                break
        else:
            raise MeasurementError(f"{target}: no '#--observables--#' line")
    lines = keep + ["  ", f"{'Name':15s}{'Index':>6s}{'Average':>18s}{'Error':>18s}{'LError':>18s}", " "]
    for name, r in results.items():
        if r["width"] > 1:
            lines.append("#-----------------------------------------------------------")
        for i in range(r["width"]):
            lines.append(f"{name[:15]:15s}{i + 1:6d}"
                         + _ffmt(r["out_mean"][block - 1, i], output_format)
                         + _ffmt(r["out_err"][block - 1, i], output_format)
                         + _ffmt(r["large_err"][block - 1, i], output_format))
        lines.append("  ")
    with open(target, "w") as f:
        f.write("\n".join(lines) + "\n")
