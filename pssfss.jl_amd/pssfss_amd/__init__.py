"""pssfss_amd — MI355X-native per-frequency MoM path of PSSFSS.jl.

Host-side mirror of the reference interface (Layer, RWGSheet, rectstrip,
setup_rwg, analyze) over the C-ABI library libpssfss_hip.so.
"""
from . import constants, rings, layers, sheets, rwg, modes, gsm, outputs, analyze, workloads  # noqa: F401
from . import adjoint, batch, meshsub  # noqa: F401
from .outputs import Result, outputs as output_functions, extract_result  # noqa: F401
from .layers import Layer, TE, TM  # noqa: F401
from .sheets import (RWGSheet, rectstrip, recttri, annular_ring, polyring, meander,  # noqa: F401
                     loadedcross, jerusalemcross, combine)
from .rwg import setup_rwg, RWGData  # noqa: F401
from .gsm import GSM, cascade, cascade_layer  # noqa: F401
