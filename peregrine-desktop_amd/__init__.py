"""peregrine_desktop_amd -- MI355X-native value-iteration solve for Peregrine.

The package directory holds the HIP sources (csrc/), the built C-ABI
library (libperegrine_hip.so) and the host-side mirror of the reference's
solver interface. Load it by path (the directory name is not an identifier):

    spec = importlib.util.spec_from_file_location(
        "peregrine_desktop_amd", ROOT / "peregrine-desktop_amd" / "__init__.py",
        submodule_search_locations=[str(ROOT / "peregrine-desktop_amd")])
"""
from . import abi
from .model import (Grid6, Model, Params, control_set, cube_controls, ctl_soa, target_cells,
                    disc_target_cells)
from .occupancy import build_occupancy, rect
from .solver import (Context, HipSolver, default_context, gpu_budget, gpu_fits, solve_value,
                     sweep_gpu)
from . import configs

__all__ = [
    "abi", "Grid6", "Model", "Params", "control_set", "cube_controls", "ctl_soa", "target_cells",
    "disc_target_cells", "build_occupancy", "rect", "Context", "HipSolver", "default_context",
    "gpu_budget", "gpu_fits", "solve_value", "sweep_gpu", "configs",
]
