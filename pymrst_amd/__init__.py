"""pymrst_amd -- the fully-implicit TPFA Newton step of PyMRST, MI355X-native.

Host code in Python (PyTorch-ROCm for device memory and streams) over the pmx
HIP library (include/pmx.h), built for gfx950 by ``python -m
pymrst_amd.build``.  Nothing here falls back to the CPU: without the built
library or a GPU the device entry points raise.
"""
from .grid import (BARSA, CENTI, DARCY, DAY, MILLI, NORM_GRAVITY, POISE, Grid, Rock, Well, add_well,
                   cart_grid, compute_trans, make_rock, pore_volume, tensor_grid, tpfa_connections)

__all__ = ["BARSA", "CENTI", "DARCY", "DAY", "MILLI", "NORM_GRAVITY", "POISE", "Grid", "Rock", "Well",
           "add_well", "cart_grid", "compute_trans", "make_rock", "pore_volume", "tensor_grid",
           "tpfa_connections"]


def __getattr__(name):
    # device classes are imported lazily so that grid setup works without torch
    if name in ("TwoPhaseOilWaterModelAD", "OWState", "DeviceOWState", "CPRSolver", "SimpleADIFluid",
                "init_simple_adi_fluid", "simulate_schedule_ad", "simulateScheduleAD"):
        from . import ow
        return getattr(ow, name)
    if name in ("NonLinearSolver", "StepReport", "SimulationReport"):
        from . import nonlinear
        return getattr(nonlinear, name)
    if name in ("SimpleTimeStepSelectorAD", "FactorTimeStepSelectorAD", "IterationCountTimeStepSelectorAD",
                "StateChangeTimeStepSelectorAD", "rampup_timesteps", "rampupTimesteps"):
        from . import timesteps
        return getattr(timesteps, name)
    if name in ("ThreePhaseBlackOilModelAD", "BOState", "DeviceBOState", "DeckBlackOilFluid",
                "init_deck_blackoil_fluid"):
        from . import bo
        return getattr(bo, name)
    if name in ("ResultHandler", "simulate_with_output"):
        from . import results
        return getattr(results, name)
    raise AttributeError(name)
