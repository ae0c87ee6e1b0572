"""Host side of the MI355X RNS-CKKS hot path (ctypes over librnsckks.so's C ABI)."""
from .params import RnsCkksParams  # noqa: F401
