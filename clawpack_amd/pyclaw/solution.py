"""Solution (upstream pyclaw/src/pyclaw/solution.py): a State on a Domain plus the time."""
import copy

import numpy as np

from .fileio import ascii, binary
from .geometry import Dimension, Domain, Patch
from .state import State


_FORMATS = {"ascii": ascii, "binary": binary}


def _format(name):
    if name not in _FORMATS:
        raise NotImplementedError("file_format %r (the 'ascii' and 'binary' frame formats are supported)" % name)
    return _FORMATS[name]


class Solution:
    """Solution(state, domain), Solution(state) (the state's patch as domain), or
    Solution(frame, path='./_output', file_format='ascii', file_prefix=None, read_aux=True) to
    read a frame written by Solution.write / Controller (file_format 'ascii' or 'binary')."""

    def __init__(self, *args, **kwargs):
        if len(args) >= 1 and isinstance(args[0], (int, np.integer)):
            self._read(int(args[0]), **kwargs)
            return
        if len(args) == 0 or not isinstance(args[0], State):
            raise TypeError("Solution(state, domain) or Solution(frame, path=...)")
        self.states = [args[0]]
        self.domain = args[1] if len(args) > 1 else Domain(args[0].patch)
        if not isinstance(self.domain, Domain):
            raise TypeError("Solution(state, domain): domain must be a Domain")

    def _read(self, frame, path="./_output", file_format="ascii", file_prefix=None, read_aux=True,
              options=None):
        t, num_eqn, num_aux, num_dim, patches, aux = _format(file_format).read(
            None, frame, path, file_prefix or "fort", read_aux, options)
        self.states = []
        for k, (hdr, q) in enumerate(patches):
            names = ["x", "y", "z"][:num_dim]
            dims = [Dimension(lo, lo + n * dl, n, name=nm)
                    for lo, dl, n, nm in zip(hdr["lower"], hdr["delta"], hdr["num_cells"], names)]
            patch = Patch(dims)
            patch.patch_index, patch.level = hdr["patch_number"], hdr["AMR_level"]
            st = State(patch, num_eqn, num_aux if aux is not None else 0)
            st.q[...] = q
            if aux is not None:
                st.aux[...] = aux[k][1]
            st.t = t
            self.states.append(st)
        self.domain = Domain(self.states[0].patch)

    def write(self, frame, path="./_output", file_format="ascii", file_prefix=None, write_aux=False,
              options=None, write_p=False):
        """Write frame number `frame` (upstream Solution.write) in each of the formats named by
        file_format ('ascii', 'binary', or a list of them)."""
        formats = [file_format] if isinstance(file_format, str) else list(file_format)
        writers = [_format(fmt) for fmt in formats]  # refuse an unknown format before writing any
        for w in writers:
            w.write(self, frame, path, file_prefix or "fort", write_aux, options, write_p)

    @property
    def state(self):
        return self.states[0]

    @property
    def patch(self):
        return self.domain.patch

    @property
    def t(self):
        return self.state.t

    @t.setter
    def t(self, value):
        for s in self.states:
            s.t = value

    @property
    def num_eqn(self):
        return self.state.num_eqn

    def __deepcopy__(self, memo):
        other = Solution(copy.copy(self.state), self.domain)
        return other
