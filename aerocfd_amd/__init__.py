"""aerocfd_amd -- MI355X (gfx950) implementation of aerocfd's per-block
residual-and-update path.

The compute path is hand-written HIP in ``csrc/`` built into
``libaerocfd_amd.so`` (C ABI: ``include/aerocfd_amd.h``).  This package is the
host-side mirror of the reference interface: :class:`Solver` (resident
multi-block context) and stateless functions named after the reference
routines.  There is no CPU fallback: every call goes to the GPU library and
fails loudly when it is not built.
"""
from .solver import (FIELDS, STAGES, FlowConfig, Solver, crank2nd, find_dt, get_primitivevariables,
                     implicit1st, inviscid_flux_1, mcell_updates_per_s, sa_model, viscous_flux)

__all__ = ["FIELDS", "STAGES", "FlowConfig", "Solver", "crank2nd", "find_dt",
           "get_primitivevariables", "implicit1st", "inviscid_flux_1", "mcell_updates_per_s",
           "sa_model", "viscous_flux"]
