"""tlbm: an MI355X-native D3Q19 lattice Boltzmann step behind the reference's solver API.

``tlbm.lbm`` mirrors the reference module ``lbm.py`` (plugins, ``lbm_init``, ``lbm_step``,
``coarsen``, ``ManagedArray``, ``Solver``); the compute runs in ``libtlbm.so`` (hand-written
HIP for gfx950, C ABI in include/tlbm.h).
"""

from . import _capi  # noqa: F401
from . import adjoint, autotune, batch  # noqa: F401
from .lbm import (BGK, COLLISIONS, CX, CY, CZ, FIELDS, STORAGES, TRT, W, Box,  # noqa: F401
                  Cylinder, D3Q19, DeviceArray, DeviceField, Flow, Heat, HeatSink, HostBuffer,
                  LidY, ManagedArray, NoHeat, OutputArray, PeriodicY, Quiet, Report, SlipY,
                  SmagorinskyLES, SolidField, Solver, Sphere, array, coarsen, decode_populations,
                  default_tile_rows, empty, encode_populations, kernel_handle, lbm_init,
                  lbm_step, make_params, population_elems, set_launch, sync, tile_populations,
                  untile_populations, zeros)
