"""tdt: MI355X-native residual + Jameson RK4 engine for the 2-D / axisymmetric
turbulent Navier-Stokes solver, behind the C ABI in include/tdt.h."""
from .caseio import Flow, Mesh, load_case, read_flow, read_grid_bc, read_restart, write_case  # noqa: F401
from .lib import Batch, Solver, TdtError, load, primitives_from_state, version  # noqa: F401
