"""Clawpack ascii frame files (upstream pyclaw/src/pyclaw/fileio/ascii.py).

A frame is two text files in `path`:

* ``<prefix>.tNNNN``: time, num_eqn, number of states (patches), num_aux, num_dim, num_ghost,
  one per line, the value first and its name after it;
* ``<prefix>.qNNNN``: for each patch a header (patch_number, AMR_level, the cell count of each
  dimension, the lower edge of each dimension, the cell width of each dimension, a blank line),
  then one line per cell with the num_eqn values of q, x fastest, a blank line after each row.

Values are written with %18.8e (about 9 significant digits), as the Clawpack tools expect;
reading parses the first token of each header line and the whitespace-separated values.
"""
import os

import numpy as np


def _frame_names(frame, file_prefix):
    return ("%s.t%s" % (file_prefix, str(frame).zfill(4)),
            "%s.q%s" % (file_prefix, str(frame).zfill(4)),
            "%s.a%s" % (file_prefix, str(frame).zfill(4)))


def write(solution, frame, path, file_prefix="fort", write_aux=False, options=None, write_p=False):
    """Write `solution` as frame number `frame` into directory `path`."""
    if write_p:
        raise NotImplementedError("write_p (derived quantities) is not supported")
    os.makedirs(path, exist_ok=True)
    tname, qname, aname = _frame_names(frame, file_prefix)
    state = solution.state
    with open(os.path.join(path, tname), "w") as f:
        f.write("%18.8e     time\n" % state.t)
        f.write("%5i                  num_eqn\n" % state.num_eqn)
        f.write("%5i                  nstates\n" % len(solution.states))
        f.write("%5i                  num_aux\n" % state.num_aux)
        f.write("%5i                  num_dim\n" % solution.domain.num_dim)
        f.write("%5i                  num_ghost\n" % 2)
    _write_array(os.path.join(path, qname), solution, lambda s: s.q)
    if write_aux and state.num_aux > 0:
        _write_array(os.path.join(path, aname), solution, lambda s: s.aux)


def _write_header(f, patch):
    """The patch header of a frame file (shared with the binary format)."""
    f.write("%5i                  patch_number\n" % patch.patch_index)
    f.write("%5i                  AMR_level\n" % patch.level)
    for d in patch.dimensions:
        f.write("%5i                  m%s\n" % (d.num_cells, d.name))
    for d in patch.dimensions:
        f.write("%18.8e     %slow\n" % (d.lower, d.name))
    for d in patch.dimensions:
        f.write("%18.8e     d%s\n" % (d.delta, d.name))
    f.write("\n")


def _read_header(lines, k, num_dim):
    """Parses the patch header starting at or after line k; returns (header dict, next line)."""
    while not lines[k].strip():
        k += 1
    hdr = {"patch_number": int(lines[k].split()[0]), "AMR_level": int(lines[k + 1].split()[0])}
    k += 2
    n = [int(lines[k + d].split()[0]) for d in range(num_dim)]
    k += num_dim
    lower = [float(lines[k + d].split()[0]) for d in range(num_dim)]
    k += num_dim
    delta = [float(lines[k + d].split()[0]) for d in range(num_dim)]
    k += num_dim
    hdr.update(num_cells=n, lower=lower, delta=delta)
    return hdr, k


def _write_array(fname, solution, getter):
    with open(fname, "w") as f:
        for state in solution.states:
            patch = state.patch
            _write_header(f, patch)
            a = np.asarray(getter(state))
            nv = a.shape[0]
            if patch.num_dim == 1:
                for i in range(a.shape[1]):
                    f.write("".join("%18.8e" % a[m, i] for m in range(nv)) + "\n")
            elif patch.num_dim == 2:
                # one line per cell, x fastest; a blank line closes each row (fixed y)
                rows = np.ascontiguousarray(np.transpose(a, (2, 1, 0)))  # (my, mx, nv)
                fmt = "%18.8e" * nv + "\n"
                for j in range(rows.shape[0]):
                    f.write("".join(fmt % tuple(cell) for cell in rows[j]))
                    f.write("\n")
            else:
                raise NotImplementedError("ascii frames of %d-dimensional patches" % patch.num_dim)


def read_t(frame, path="./_output", file_prefix="fort"):
    """(time, num_eqn, nstates, num_aux, num_dim, num_ghost) of a frame."""
    tname = _frame_names(frame, file_prefix)[0]
    with open(os.path.join(path, tname)) as f:
        vals = [line.split()[0] for line in f if line.strip()]
    t = float(vals[0])
    ints = [int(v) for v in vals[1:6]]
    while len(ints) < 5:
        ints.append(2)
    return (t,) + tuple(ints)


def read(solution_factory, frame, path="./_output", file_prefix="fort", read_aux=False, options=None):
    """Read a frame: returns (t, [(patch header dict, q array in pyclaw layout)], aux list or None).
    `solution_factory` is unused (kept for the upstream signature)."""
    t, num_eqn, nstates, num_aux, num_dim, num_ghost = read_t(frame, path, file_prefix)
    _, qname, aname = _frame_names(frame, file_prefix)
    patches = _read_array(os.path.join(path, qname), num_eqn, nstates, num_dim)
    aux = None
    if read_aux and num_aux > 0 and os.path.exists(os.path.join(path, aname)):
        aux = _read_array(os.path.join(path, aname), num_aux, nstates, num_dim)
    return t, num_eqn, num_aux, num_dim, patches, aux


def _read_array(fname, nv, nstates, num_dim):
    with open(fname) as f:
        lines = f.read().split("\n")
    k = 0
    out = []
    for _ in range(nstates):
        hdr, k = _read_header(lines, k, num_dim)
        n = hdr["num_cells"]
        ncell = int(np.prod(n))
        vals = []
        while len(vals) < ncell * nv:
            s = lines[k].split()
            k += 1
            if s:
                vals.extend(float(x) for x in s)
        a = np.array(vals).reshape(list(reversed(n)) + [nv])  # (my, mx, nv) for 2D
        a = np.asfortranarray(np.transpose(a, list(range(num_dim, -1, -1))))  # (nv, mx, my)
        out.append((hdr, a))
    return out
