"""Autotuner for the force kernel's launch shape: elements per chain patch and shots per wave.

The patch plan of a mesh (``SF2D_CHAIN_ELEMS``, elements per chain patch, 1..12) fixes how many element rows
one wave of the chain kernel holds and how many waves a colour launch has; the shots per wave of a batched
launch (``sf2d_mesh_set_shots_per_wave``) trade element-data reuse against waves in flight.  Which pair is
fastest depends on the mesh size, the number of shots and the device, so it is measured: every candidate runs
a short forward (``steps`` time steps from rest) on the workload's own mesh, model and shots, and the fastest
device time wins.  The colour schedule itself is the plan's greedy colouring (four colours on the structured
meshes), in launch order.

``tune(cfg, shots)`` returns the choice and the table; ``apply(choice)`` sets the environment that the library
reads when a handle is set up, so every simulation created afterwards in the process uses it.
"""
from __future__ import annotations

import os
from typing import Dict, Iterable, List, Optional, Sequence

from .config import SimulationConfig
from .driver import ShotBatch, Simulation

CHAIN_ELEMS = (4, 6, 8, 10, 12)
SHOTS_PER_WAVE = (1, 2, 4, 8, 16)


def _forward_ms(sim, steps: int, batch: bool) -> float:
    if batch:
        _, ms = sim.forward(timing=True)
        return ms
    sim.reset()
    _, t = sim.forward(nstep=steps)
    return t["loop_ms"]


def tune(cfg: SimulationConfig, shots: Optional[Sequence[list]] = None, steps: int = 20,
         chain_elems: Iterable[int] = CHAIN_ELEMS, shots_per_wave: Iterable[int] = SHOTS_PER_WAVE) -> Dict:
    """Time a ``steps``-step forward for every (elements per chain patch, shots per wave) candidate on this
    workload (a ShotBatch of ``shots`` when given, else a single simulation) and return
    ``{"chain_elems", "shots_per_wave", "ms_per_step", "table"}``; ``table`` lists every candidate."""
    small = cfg.replace(nstep=steps)
    batch = shots is not None
    groups: List[int] = sorted({min(g, len(shots)) for g in shots_per_wave}) if batch else [0]
    table = []
    saved = os.environ.get("SF2D_CHAIN_ELEMS")
    try:
        for n in chain_elems:
            os.environ["SF2D_CHAIN_ELEMS"] = str(n)
            sim = ShotBatch(small, shots) if batch else Simulation(small)
            try:
                if not sim.chain_kernel():
                    continue  # the mesh's patches are not chains: nothing to tune
                for g in groups:
                    if batch:
                        sim.set_shots_per_wave(g)
                    _forward_ms(sim, steps, batch)  # warm-up
                    ms = _forward_ms(sim, steps, batch)
                    table.append({"chain_elems": sim.chain_elems(), "shots_per_wave": g,
                                  "ms_per_step": ms / steps})
            finally:
                sim.close()
    finally:
        if saved is None:
            os.environ.pop("SF2D_CHAIN_ELEMS", None)
        else:
            os.environ["SF2D_CHAIN_ELEMS"] = saved
    if not table:
        return {"chain_elems": None, "shots_per_wave": None, "ms_per_step": None, "table": []}
    best = min(table, key=lambda r: r["ms_per_step"])
    return dict(best, table=table)


def apply(choice: Dict) -> None:
    """Make simulations created from now on use a tuned choice (environment read at set-up)."""
    if choice.get("chain_elems"):
        os.environ["SF2D_CHAIN_ELEMS"] = str(choice["chain_elems"])
    if choice.get("shots_per_wave"):
        os.environ["SF2D_SHOTS_PER_WAVE"] = str(choice["shots_per_wave"])
