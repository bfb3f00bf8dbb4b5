"""Host-side Python plumbing of the MI355X camera path: the libvt.so binding and the seeded workloads.

The product's host API is the JavaScript TwinRenderer in js/renderer.mjs over the N-API addon; this package
serves tests and bench.py.
"""
from . import lib, scenes  # noqa: F401
