"""cfd_amd -- the reference's smoke step (src/cfd of LCEBurton/python-game-engine)
on MI355X: the same operator / solver registry API with a ``"hip"`` backend
whose kernels are hand-written for gfx950 (python-game-engine_amd/csrc).

Importing ``cfd_amd.operators`` / ``cfd_amd.solvers`` registers the HIP
backend, as importing ``cfd.operators`` registers Numba in the reference
(src/cfd/operators/__init__.py:1-3, src/cfd/solvers/__init__.py:1-3).
"""
__version__ = "0.1.0"
