"""Restart and interoperability files of the SCF caller.

openLOWDIN writes, after an SCF (MultiSCF.f90:1319-1467, MultiSCF_saveWfn):
  * ``lowdin.wfn`` -- Fortran unformatted sequential records; per species the
    labelled matrices and vectors (REMOVED-ORBITALS, TWOPARTICLES, COUPLING,
    EXCHANGE-CORRELATION, EXCHANGE-CORRELATION-ENERGY, COEFFICIENTS, DENSITY,
    KINETIC, ATTRACTION, [EXTERNAL], HCORE, ORBITALS, FOCK, OVERLAP,
    TRANSFORMATION), then TOTALENERGY, COSMO3ENERGY, COUPLINGENERGY,
    PUNTUALINTERACTIONENERGY and VIRIAL;
  * ``<input>.vec`` (binary) or ``<input>.plainvec`` (formatted) -- per species
    COEFFICIENTS and ORBITALS,
and reads the coefficients back as the SCF guess (DensityMatrixSCFGuess.f90:
77-107: plainvec first, else vec), or from a Gaussian formatted checkpoint
file (MolecularSystem.f90:1979-2107, ``Alpha MO coefficients`` with the
Cartesian d functions reordered from xx, yy, zz, xy, xz, yz to xx, xy, xz,
yy, yz, zz, MolecularSystem.f90:1907-1928).

Record layouts, as Matrix_writeToFile / Vector_writeToFile produce them
(Matrix.f90:561-657, Vector.f90:593-707):
  binary:    one record per label (the label variable's full length, blank
             padded: 30 characters for the species labels), one record with
             the element count (integer(8)), one record with the values
             (FP64, column-major); gfortran frames each record with 4-byte
             little-endian lengths.
  formatted: list-directed label lines (a leading blank, the padded label),
             the element count, then the matrix row by row in
             (n(1X,ES15.8)) or the vector list-directed.
A reader looks for the first record sequence whose labels match (trailing
blanks ignored), then checks the element count (Matrix.f90:697-792,
873-950).  The binary reader fills column-major, the formatted one row-major.

The guess density from read coefficients is eta * C_occ C_occ^T
(DensityMatrixSCFGuess.f90:144-152).
"""
from __future__ import annotations

import os
import struct
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

LABEL_LEN = 30   # character(30) labels(2) in MultiSCF_saveWfn


class WfnFormatError(ValueError):
    pass


def _pad(label: str, length: int = LABEL_LEN) -> str:
    return label[:length].ljust(length)


# ------------------------------------------------------------ binary records
def _write_record(fh, payload: bytes):
    fh.write(struct.pack("<i", len(payload)))
    fh.write(payload)
    fh.write(struct.pack("<i", len(payload)))


def read_records(path: str) -> List[bytes]:
    """All records of a gfortran unformatted sequential file."""
    out = []
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0
    while pos < len(data):
        if pos + 4 > len(data):
            raise WfnFormatError("%s: truncated record header at byte %d" % (path, pos))
        (n,) = struct.unpack_from("<i", data, pos)
        end = pos + 4 + n
        if n < 0 or end + 4 > len(data):
            raise WfnFormatError("%s: bad record length %d at byte %d" % (path, n, pos))
        (n2,) = struct.unpack_from("<i", data, end)
        if n2 != n:
            raise WfnFormatError("%s: record trailer %d != header %d at byte %d" % (path, n2, n, pos))
        out.append(data[pos + 4:end])
        pos = end + 4
    return out


def _write_block(fh, labels: Sequence[str], values: np.ndarray, label_len: Optional[int] = LABEL_LEN):
    for lab in labels:
        text = _pad(lab, label_len) if label_len else lab
        _write_record(fh, text.encode("ascii"))
    flat = np.asarray(values, dtype=np.float64).reshape(-1, order="F")
    _write_record(fh, struct.pack("<q", flat.size))
    _write_record(fh, flat.astype("<f8").tobytes())


def find_binary(records: List[bytes], labels: Sequence[str]) -> Optional[np.ndarray]:
    """Values following the first run of label records equal to ``labels``
    (trailing blanks ignored), or None."""
    want = [lab.rstrip() for lab in labels]
    k = len(want)
    for i in range(len(records) - k - 1):
        if all(records[i + j].decode("ascii", "replace").rstrip() == want[j] for j in range(k)):
            size_rec = records[i + k]
            if len(size_rec) != 8:
                continue
            (size,) = struct.unpack("<q", size_rec)
            vals = records[i + k + 1]
            if len(vals) != 8 * size:
                raise WfnFormatError("record of %s holds %d bytes for %d values" % (want, len(vals), size))
            return np.frombuffer(vals, dtype="<f8").astype(np.float64)
    return None


def matrix_from_binary(flat: np.ndarray, rows: int, cols: int) -> np.ndarray:
    if flat.size != rows * cols:
        raise WfnFormatError("matrix of %d values, expected %d x %d" % (flat.size, rows, cols))
    return flat.reshape((rows, cols), order="F").copy()


# ------------------------------------------------------------ formatted
def _fortran_real(x: float) -> str:
    return "%.17E" % x


def _write_formatted_block(fh, labels: Sequence[str], values: np.ndarray, matrix: bool):
    for lab in labels:
        fh.write(" %s\n" % _pad(lab))
    a = np.asarray(values, dtype=np.float64)
    fh.write("%21d\n" % a.size)
    if matrix:
        for row in a:
            fh.write("".join(" %15.8E" % v for v in row) + "\n")
    else:
        fh.write(" " + "  ".join(_fortran_real(v) for v in a.reshape(-1)) + "\n")


def _fortran_float(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def find_formatted(path: str, labels: Sequence[str]) -> Optional[np.ndarray]:
    """Values after the first run of label lines equal to ``labels``: the
    element count line, then that many list-directed reals over any number of
    lines (Matrix.f90:873-950)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    want = [lab.strip() for lab in labels]
    k = len(want)
    for i in range(len(lines) - k):
        if all(lines[i + j].strip() == want[j] for j in range(k)):
            try:
                size = int(lines[i + k].split()[0])
            except (ValueError, IndexError):
                continue
            vals: List[float] = []
            j = i + k + 1
            while len(vals) < size and j < len(lines):
                vals += [_fortran_float(t) for t in lines[j].split()]
                j += 1
            if len(vals) < size:
                raise WfnFormatError("%s: %s holds %d of %d values" % (path, want, len(vals), size))
            return np.array(vals[:size])
    return None


# ------------------------------------------------------------ vec / plainvec
def write_vec(path: str, species: Sequence[Tuple[str, np.ndarray, np.ndarray]], binary: bool):
    """MultiSCF_saveWfn's coefficient file: per species (name, C, orbital
    energies), COEFFICIENTS then ORBITALS."""
    if binary:
        with open(path, "wb") as fh:
            for name, C, eps in species:
                _write_block(fh, ["COEFFICIENTS", name], C)
                _write_block(fh, ["ORBITALS", name], eps)
    else:
        with open(path, "w") as fh:
            for name, C, eps in species:
                _write_formatted_block(fh, ["COEFFICIENTS", name], C, matrix=True)
                _write_formatted_block(fh, ["ORBITALS", name], eps, matrix=False)


def read_coefficients(prefix: str, name: str, n: int) -> Optional[Tuple[np.ndarray, str]]:
    """DensityMatrixSCFGuess's read: ``prefix + "plainvec"`` if it exists
    (row-major), else ``prefix + "vec"`` (column-major); the n x n
    COEFFICIENTS block of species ``name`` and the file it came from, or
    None when neither file holds it (the reference then falls back to its
    guess, failContinue)."""
    plain, binf = prefix + "plainvec", prefix + "vec"
    if os.path.exists(plain):
        flat = find_formatted(plain, ["COEFFICIENTS", name])
        if flat is None or flat.size != n * n:
            return None
        return flat.reshape(n, n).copy(), plain
    if os.path.exists(binf):
        flat = find_binary(read_records(binf), ["COEFFICIENTS", name])
        if flat is None or flat.size != n * n:
            return None
        return matrix_from_binary(flat, n, n), binf
    return None


def read_orbital_energies(path: str, name: str) -> Optional[np.ndarray]:
    if path.endswith("plainvec"):
        return find_formatted(path, ["ORBITALS", name])
    return find_binary(read_records(path), ["ORBITALS", name])


def guess_density(C: np.ndarray, n_occupied: int, eta: float) -> np.ndarray:
    """eta * C_occ C_occ^T (DensityMatrixSCFGuess.f90:144-152)."""
    occ = C[:, :n_occupied]
    return eta * occ @ occ.T


# ------------------------------------------------------------ lowdin.wfn
WFN_SPECIES_ORDER = ["REMOVED-ORBITALS", "TWOPARTICLES", "COUPLING", "EXCHANGE-CORRELATION",
                     "EXCHANGE-CORRELATION-ENERGY", "COEFFICIENTS", "DENSITY", "KINETIC", "ATTRACTION",
                     "EXTERNAL", "HCORE", "ORBITALS", "FOCK", "OVERLAP", "TRANSFORMATION"]
WFN_SCALARS = ["TOTALENERGY", "COSMO3ENERGY", "COUPLINGENERGY", "PUNTUALINTERACTIONENERGY", "VIRIAL"]


def write_wfn(path: str, species: Sequence[Tuple[str, Dict[str, np.ndarray]]], scalars: Dict[str, float]):
    """lowdin.wfn in MultiSCF_saveWfn's record order.  ``species``: (name,
    blocks) with blocks keyed by the labels above (matrices n x n, ORBITALS
    and REMOVED-ORBITALS vectors, EXCHANGE-CORRELATION-ENERGY a scalar);
    EXTERNAL is written only when present, as the reference writes it only
    with an external potential.  ``scalars``: the five energies, each a
    one-label record of the label's own length."""
    with open(path, "wb") as fh:
        for name, blocks in species:
            for lab in WFN_SPECIES_ORDER:
                if lab not in blocks:
                    if lab == "EXTERNAL":
                        continue
                    raise WfnFormatError("species %s lacks %s" % (name, lab))
                _write_block(fh, [lab, name], np.atleast_1d(np.asarray(blocks[lab], dtype=np.float64)))
        for lab in WFN_SCALARS:
            _write_block(fh, [lab], np.atleast_1d(float(scalars.get(lab, 0.0))), label_len=None)


def read_wfn(path: str) -> Dict[Tuple[str, ...], np.ndarray]:
    """Every labelled block of a lowdin.wfn: {(label, species) or (label,):
    flat values}."""
    recs = read_records(path)
    out: Dict[Tuple[str, ...], np.ndarray] = {}
    i = 0
    while i < len(recs):
        labs = []
        while i < len(recs) and len(recs[i]) != 8 and recs[i][:1].isalpha():
            labs.append(recs[i].decode("ascii").rstrip())
            i += 1
        if i + 1 >= len(recs) or not labs:
            raise WfnFormatError("%s: unexpected record %d" % (path, i))
        (size,) = struct.unpack("<q", recs[i])
        vals = np.frombuffer(recs[i + 1], dtype="<f8").astype(np.float64)
        if vals.size != size:
            raise WfnFormatError("%s: %s holds %d of %d values" % (path, labs, vals.size, size))
        out[tuple(labs)] = vals
        i += 2
    return out


# ------------------------------------------------------------ fchk
# Gaussian Cartesian d order xx, yy, zz, xy, xz, yz -> xx, xy, xz, yy, yz, zz
FCHK_D_TO_LOWDIN = [0, 3, 4, 1, 5, 2]


def read_fchk_arrays(path: str) -> Dict[str, np.ndarray]:
    """Integer and real arrays of a formatted checkpoint file by name
    (lines ``(A40,A4,A6,I14)`` with ``N=``, integers 6 per line, reals 5 per
    line; MolecularSystem.f90:2023-2084)."""
    arrays: Dict[str, np.ndarray] = {}
    with open(path) as fh:
        lines = fh.read().splitlines()
    i = 2
    while i < len(lines):
        line = lines[i]
        obj, typ, arr = line[:40].strip(), line[40:44].strip(), line[44:50].strip()
        i += 1
        if arr != "N=" or typ not in ("I", "R"):
            continue
        size = int(line[50:].split()[0])
        per = 6 if typ == "I" else 5
        vals: List[str] = []
        for _ in range((size + per - 1) // per):
            vals += lines[i].split()
            i += 1
        vals = vals[:size]
        arrays[obj] = np.array([int(v) for v in vals]) if typ == "I" else np.array([_fortran_float(v) for v in vals])
    return arrays


def read_fchk_coefficients(path: str, shell_l: Sequence[int]) -> np.ndarray:
    """``Alpha MO coefficients`` as an n x n matrix (column k = orbital k) in
    the Cartesian component order of this package, for a basis whose shells
    have angular momenta ``shell_l`` in file order (d rows permuted as
    MolecularSystem_changeOrbitalOrder does from FCHK to LOWDIN)."""
    if any(l > 2 for l in shell_l):
        raise WfnFormatError("shells above d are outside this package")
    n = sum((l + 1) * (l + 2) // 2 for l in shell_l)
    arr = read_fchk_arrays(path)
    if "Alpha MO coefficients" not in arr:
        raise WfnFormatError("%s has no Alpha MO coefficients" % path)
    c = arr["Alpha MO coefficients"]
    if c.size != n * n:
        raise WfnFormatError("%s: %d coefficients for %d functions" % (path, c.size, n))
    C = c.reshape((n, n), order="F").copy()
    perm = []
    f = 0
    for l in shell_l:
        size = (l + 1) * (l + 2) // 2
        perm += [f + k for k in (FCHK_D_TO_LOWDIN if l == 2 else range(size))]
        f += size
    return C[perm, :]
