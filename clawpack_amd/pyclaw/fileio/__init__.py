"""Frame files (upstream pyclaw/src/pyclaw/fileio/): the ascii format."""
from . import ascii  # noqa: F401
