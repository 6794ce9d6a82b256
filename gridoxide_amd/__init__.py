"""gridoxide_amd: batched Newton-Raphson AC power flow on MI355X (HIP, gfx950).

The numeric path lives in ``libgridox_hip.so`` (C ABI: ``include/gridox_hip.h``); this package
holds the workload generator and the host mirror of the reference's solver interface.
"""
from .grid import PQ, PV, SLACK, Network, config_grid, generate_grid, scenario_loads, scenario_specs, without_branch  # noqa: F401
from .model import BatchResult, PowerFlowModel  # noqa: F401
from .pgm import PgmError, PgmModel, apply_batch_update, from_pgm, load_pgm_json  # noqa: F401
from .solver import (  # noqa: F401
    BatchError,
    BatchSolver,
    BusOutOfRange,
    BusOverride,
    BusType,
    Engine,
    IslandReport,
    IslandStatus,
    JacobianBackend,
    Mode,
    OutagesUnsupported,
    PersistentSolver,
    PowerFlowReport,
    Scenario,
    SolveStats,
    SolveStatus,
    Topology,
    YBus,
    build_ybus,
    uniform_load_scaling,
)
