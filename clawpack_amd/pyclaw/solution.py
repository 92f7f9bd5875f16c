"""Solution (upstream pyclaw/src/pyclaw/solution.py): a State on a Domain plus the time."""
import copy

from .geometry import Domain
from .state import State


class Solution:
    """Solution(state, domain) (or Solution(state) with the state's patch as domain)."""

    def __init__(self, *args):
        if len(args) == 0 or not isinstance(args[0], State):
            raise NotImplementedError("Solution(state, domain) is the supported form; reading a frame "
                                      "from disk is not part of this package")
        self.states = [args[0]]
        self.domain = args[1] if len(args) > 1 else Domain(args[0].patch)
        if not isinstance(self.domain, Domain):
            raise TypeError("Solution(state, domain): domain must be a Domain")

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
