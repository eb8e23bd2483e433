"""filament.jl_amd -- MI355X-native per-z-step propagation hot path of
Filament.jl (RT, XYT and T geometries, and the spatial R and XY ones) behind
the C ABI of libfilament_hip.so.

Host side (this package): the reference's input conventions, FP64 table
construction, the run driver and plot.dat output.  Device side
(csrc/, built into libfilament_hip.so): the z-step and the field analyzer.
"""
from . import lib, physics, responses, workloads  # noqa: F401
from .solver import RModel, RTEnsemble, RTModel, StepGroup, TModel, XYModel, XYTModel  # noqa: F401
from . import corrector, driver, plotfield, stepping, tuning  # noqa: F401
from .stepping import CoarseStepper  # noqa: F401
from .driver import prepare, run  # noqa: F401
from .responses import (RK3, RK4, init_cubic, init_current_free, init_current_losses,  # noqa: F401
                        init_photoionization_avalanche, init_raman)

__all__ = ["RTModel", "RTEnsemble", "StepGroup", "TModel", "XYTModel", "RModel", "XYModel", "CoarseStepper", "lib", "physics", "responses", "workloads"]
