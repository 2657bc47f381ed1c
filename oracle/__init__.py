"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes wrappers over
  liboracle.so        this project's C++ restatement of the reference CPU path
                      (snc_oracle.cpp, oracle_train.cpp), and
  _ref/libsnc_ref.so  the reference's own CPU sources compiled unchanged with
                      ref_shim.cpp (built only where the reference tree exists).
Only tests, the smoke check and the benchmark's CPU-baseline leg import this.
"""
from .oracle import *  # noqa: F401,F403
