"""MI355X lattice-Boltzmann time step (D2Q9/D3Q19, BGK/TRT/MRT, Zou-He,
bounce-back, colour gradient) behind a plain C ABI (include/lbm.h).

`thesis_amd.lbm` is the host-side mirror of the reference solver's
interface; `thesis_amd.lattices` builds lattices in the latticegen2 layout.
"""
from . import lbm, lattices  # noqa: F401
