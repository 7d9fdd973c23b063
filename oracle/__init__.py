"""ORACLE - test infrastructure only (see oracle.h). Imported by tests, the
smoke check and the CPU-baseline leg of the benchmark; never by the product."""
from .binding import *  # noqa: F401,F403
