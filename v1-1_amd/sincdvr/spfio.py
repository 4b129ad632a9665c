"""Orbital files in the reference's format (Bin/spfs.bin and the like).

S/loadstuff.f90 writes a Fortran unformatted sequential file with two records:
  1. spf_header_write (:155-177): integer dims(3), num, cflag   (cflag 1 = complex)
  2. write_spf (:132-151): the orbitals as one array (dims(1)*dims(2)*dims(3), num),
     complex*16 (cflag 1) or real*8 (cflag 0), column-major;
each record framed by 4-byte length markers (the gfortran / flang default).
Reading (load_spfs0 / spf_read0, :247-430) takes min(num in file, requested)
orbitals and, per dimension with spfdimtype 1 (centred grids), chops or
zero-pads around the centre; with an odd/even size mismatch the odd grid
loses its point at the positive end, so an odd grid with a nucleus at 0 and
an even grid with the nucleus at -1/2 spacing map onto each other (:339-357).
"""
import struct

import numpy as np

_MARK = "<i"


def _record(f, payload):
    n = len(payload)
    f.write(struct.pack(_MARK, n))
    f.write(payload)
    f.write(struct.pack(_MARK, n))


def _read_record(f):
    head = f.read(4)
    if len(head) != 4:
        raise ValueError("truncated orbital file (record marker)")
    (n,) = struct.unpack(_MARK, head)
    if n < 0:
        raise ValueError("record split into sub-records: not an orbital file of this format")
    data = f.read(n)
    tail = f.read(4)
    if len(data) != n or len(tail) != 4 or struct.unpack(_MARK, tail)[0] != n:
        raise ValueError("corrupt orbital file record")
    return data


def write_spfs(path, spfs, dims):
    """spfs: (prod(dims), num) array, complex (cflag 1) or real (cflag 0)."""
    spfs = np.asarray(spfs)
    dims = tuple(int(d) for d in dims)
    if spfs.ndim == 1:
        spfs = spfs.reshape(-1, 1)
    if spfs.shape[0] != dims[0] * dims[1] * dims[2]:
        raise ValueError("orbital length does not match dims")
    cflag = 1 if np.iscomplexobj(spfs) else 0
    arr = np.asfortranarray(spfs, dtype=np.complex128 if cflag else np.float64)
    with open(path, "wb") as f:
        _record(f, struct.pack("<5i", dims[0], dims[1], dims[2], spfs.shape[1], cflag))
        _record(f, arr.tobytes(order="F"))


def read_header(path):
    with open(path, "rb") as f:
        h = _read_record(f)
    if len(h) != 20:
        raise ValueError("orbital file header is not dims(3), num, cflag")
    d0, d1, d2, num, cflag = struct.unpack("<5i", h)
    return (d0, d1, d2), num, cflag


def _ranges(rd, od, dimtype):
    """Input and output index ranges of one dimension (spf_read0 :330-372)."""
    if dimtype == 0:
        k = min(rd, od)
        return (0, k), (0, k)
    if dimtype == 2:
        if rd != od:
            raise ValueError("spfdimtype 2 requires equal dimensions (%d, %d)" % (rd, od))
        return (0, od), (0, od)
    if dimtype != 1:
        raise ValueError("spfdimtype not recognised")
    rr = oo = 0
    if (rd + od) % 2 == 1:
        rr, oo = -(rd % 2), -(od % 2)
    ibot = max(1, ((rd + rr) - (od + oo)) // 2 + 1)
    obot = max(1, ((od + oo) - (rd + rr)) // 2 + 1)
    k = min(od + oo, rd + rr)
    return (ibot - 1, ibot - 1 + k), (obot - 1, obot - 1 + k)


def read_spfs(path, outdims, nspf, dimtypes=(1, 1, 1)):
    """(prod(outdims), nspf) complex128 Fortran-ordered orbitals and the number
    loaded (min of the file's count and nspf; the rest stay zero)."""
    with open(path, "rb") as f:
        h = _read_record(f)
        d0, d1, d2, num, cflag = struct.unpack("<5i", h)
        body = _read_record(f)
    rd = (d0, d1, d2)
    npts = d0 * d1 * d2
    dt = np.complex128 if cflag else np.float64
    if len(body) != npts * num * np.dtype(dt).itemsize:
        raise ValueError("orbital record size does not match the header")
    src = np.frombuffer(body, dtype=dt).reshape((d0, d1, d2, num), order="F")
    nload = min(num, nspf)
    od = tuple(int(x) for x in outdims)
    out = np.zeros(od + (nspf,), dtype=np.complex128, order="F")
    rin, rout = zip(*[_ranges(rd[i], od[i], dimtypes[i]) for i in range(3)])
    out[rout[0][0]:rout[0][1], rout[1][0]:rout[1][1], rout[2][0]:rout[2][1], :nload] = \
        src[rin[0][0]:rin[0][1], rin[1][0]:rin[1][1], rin[2][0]:rin[2][1], :nload]
    return np.asfortranarray(out.reshape(od[0] * od[1] * od[2], nspf, order="F")), nload
