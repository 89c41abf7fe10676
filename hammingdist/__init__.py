"""``import hammingdist``: the reference's module name (python/hammingdist.cc:28).

Callers written against the reference's pybind11 module import this package
unchanged; every name is the MI355X engine's (hammingdist_amd.hammingdist,
the same pybind11 module with the reference's functions, argument names,
defaults, return types and error messages).  The engine's extras (C ABI
binding, batched mode, serving) stay in the hammingdist_amd package.
"""
import hammingdist_amd as _engine  # loads one HIP runtime, then the native module
from hammingdist_amd.hammingdist import *  # noqa: F401,F403

