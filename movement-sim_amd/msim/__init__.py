"""msim — MI355X-native per-tick agent pipeline of movement-sim.

Host-side Python mirror of the C-ABI (libmsim.so) plus the seeded scenario
generators that feed it.  Import with movement-sim_amd/ on sys.path.
"""
from .types import ENTITY_DTYPE, PUSH_DTYPE, ROAD_DTYPE, push_consts  # noqa: F401
from ._lib import LIB_CONTRACT_PATH, LIB_PATH, Batch, MsimError, Simulation, load_library  # noqa: F401
