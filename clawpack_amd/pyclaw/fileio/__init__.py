"""Frame files (upstream pyclaw/src/pyclaw/fileio/): the ascii and binary formats."""
from . import ascii  # noqa: F401
from . import binary  # noqa: F401
