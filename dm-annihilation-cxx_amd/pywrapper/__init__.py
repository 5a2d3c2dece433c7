"""Python front end of the MI355X annihilation-map projector.

The reference's pywrapper (src/pywrapper/) only exposed a particle reader;
this package exposes the map path itself:
  capi      ctypes binding of libdmann.so (Projector, params, FITS I/O)
  synth     seeded synthetic NFW / Plummer particle sets
  autograd  torch.autograd.Function over the projector and its adjoint
  autotune  launch-shape / size-class autotuner
"""
from . import autograd, autotune, capi, synth  # noqa: F401
from .autograd import ProjectMap, project_map  # noqa: F401
from .capi import (FLUX_KINDS, PARTICLE_DTYPE, PREC_COMPENSATED, PREC_FP32, PREC_FP64, Batch, DmannError,  # noqa: F401
                   DmannParams, Projector, load_library, params_for_observer, params_from_ini,
                   read_fits, read_halo_file, read_particles, write_fits, write_ini)
