"""Python side of libqphixmilc_hip: ctypes binding of the C interface
(lib.py), synthetic-field generators (fields.py) and an object wrapper
(api.py).  All computation happens in the HIP library."""
from . import fields  # noqa: F401
from .lib import load, Lib, LIB_PATH  # noqa: F401
