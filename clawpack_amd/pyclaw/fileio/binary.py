"""Clawpack binary frame files (upstream pyclaw/src/pyclaw/fileio/binary.py; the "binary64"
output format of Clawpack 5's Fortran codes).

A frame is three files in `path`:

* ``<prefix>.tNNNN``: the ascii header (time, num_eqn, nstates, num_aux, num_dim, num_ghost) plus a
  seventh line naming the data format, ``binary64``;
* ``<prefix>.qNNNN``: for each patch the ascii patch header (patch_number, AMR_level, the cell
  count, lower edge and cell width of each dimension, a blank line) and no data;
* ``<prefix>.bNNNN``: the data, raw little-endian float64: for each patch in order, the array of
  shape (num_eqn, n_x + 2 num_ghost, n_y + 2 num_ghost) in Fortran order (component fastest, then
  x, then y), ghost cells included.

Aux arrays, when written, go to ``<prefix>.aNNNN`` (headers) and ``<prefix>.baNNNN`` (data) in
the same layout.  The writer stores the state as PyClaw holds it, with no ghost cells, and says
so with num_ghost = 0; the reader strips whatever ghost layer num_ghost names, so it also reads
frames whose data carries ghost cells.  Values round-trip bit for bit.
"""
import os

import numpy as np

from . import ascii as _ascii

FORMAT_NAME = "binary64"


def _names(frame, file_prefix):
    n = str(frame).zfill(4)
    return {"t": "%s.t%s" % (file_prefix, n), "q": "%s.q%s" % (file_prefix, n),
            "b": "%s.b%s" % (file_prefix, n), "a": "%s.a%s" % (file_prefix, n),
            "ba": "%s.ba%s" % (file_prefix, n)}


def write(solution, frame, path, file_prefix="fort", write_aux=False, options=None, write_p=False):
    """Write `solution` as binary frame number `frame` into directory `path`."""
    if write_p:
        raise NotImplementedError("write_p (derived quantities) is not supported")
    os.makedirs(path, exist_ok=True)
    names = _names(frame, file_prefix)
    state = solution.state
    with open(os.path.join(path, names["t"]), "w") as f:
        f.write("%18.8e     time\n" % state.t)
        f.write("%5i                  num_eqn\n" % state.num_eqn)
        f.write("%5i                  nstates\n" % len(solution.states))
        f.write("%5i                  num_aux\n" % state.num_aux)
        f.write("%5i                  num_dim\n" % solution.domain.num_dim)
        f.write("%5i                  num_ghost\n" % 0)
        f.write("%s             file_format\n" % FORMAT_NAME)
    _write_pair(path, names["q"], names["b"], solution, lambda s: s.q)
    if write_aux and state.num_aux > 0:
        _write_pair(path, names["a"], names["ba"], solution, lambda s: s.aux)


def _write_pair(path, hname, dname, solution, getter):
    with open(os.path.join(path, hname), "w") as fh, open(os.path.join(path, dname), "wb") as fd:
        for state in solution.states:
            _ascii._write_header(fh, state.patch)
            a = np.asarray(getter(state), dtype="<f8")
            # Fortran order: component fastest, then x, then y
            fd.write(np.asfortranarray(a).tobytes(order="F"))


def read_t(frame, path="./_output", file_prefix="fort"):
    """(time, num_eqn, nstates, num_aux, num_dim, num_ghost, file_format) of a frame."""
    with open(os.path.join(path, _names(frame, file_prefix)["t"])) as f:
        vals = [line.split()[0] for line in f if line.strip()]
    head = _ascii.read_t(frame, path, file_prefix)
    fmt = vals[6] if len(vals) > 6 else "ascii"
    return head + (fmt,)


def read(solution_factory, frame, path="./_output", file_prefix="fort", read_aux=False, options=None):
    """Read a binary frame: returns (t, num_eqn, num_aux, num_dim, [(patch header dict, q)], aux)
    with q in PyClaw's layout (num_eqn, n_x, n_y); `solution_factory` is unused (kept for the
    upstream signature)."""
    t, num_eqn, nstates, num_aux, num_dim, num_ghost, fmt = read_t(frame, path, file_prefix)
    if fmt != FORMAT_NAME:
        raise IOError("frame %d in %s is %r, not %s" % (frame, path, fmt, FORMAT_NAME))
    names = _names(frame, file_prefix)
    patches = _read_pair(path, names["q"], names["b"], num_eqn, nstates, num_dim, num_ghost)
    aux = None
    if read_aux and num_aux > 0 and os.path.exists(os.path.join(path, names["ba"])):
        aux = _read_pair(path, names["a"], names["ba"], num_aux, nstates, num_dim, num_ghost)
    return t, num_eqn, num_aux, num_dim, patches, aux


def _read_pair(path, hname, dname, nv, nstates, num_dim, num_ghost):
    with open(os.path.join(path, hname)) as f:
        lines = f.read().split("\n")
    data = np.fromfile(os.path.join(path, dname), dtype="<f8")
    k = 0
    pos = 0
    out = []
    for _ in range(nstates):
        hdr, k = _ascii._read_header(lines, k, num_dim)
        n = [c + 2 * num_ghost for c in hdr["num_cells"]]
        count = nv * int(np.prod(n))
        if pos + count > data.size:
            raise IOError("%s holds %d values, the headers need at least %d" % (dname, data.size, pos + count))
        a = data[pos:pos + count].reshape([nv] + n, order="F")
        pos += count
        if num_ghost:
            a = a[(slice(None),) + tuple(slice(num_ghost, -num_ghost) for _ in range(num_dim))]
        out.append((hdr, np.asfortranarray(a)))
    if pos != data.size:
        raise IOError("%s holds %d values, the headers describe %d" % (dname, data.size, pos))
    return out
