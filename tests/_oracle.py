"""The oracle's ctypes binding lives beside the C sources (oracle/binding.py)."""
from oracle.binding import *  # noqa: F401,F403
from oracle.binding import MIN_GAMMA, MIN_LOGLIKE  # noqa: F401
