"""Autotuner for the force kernel's launch shape: elements per chain patch and shots per wave.

The patch plan of a mesh (``sf2d_mesh_set_chain_elems``, elements per chain patch, 1..12) fixes how many element rows
one wave of the chain kernel holds and how many waves a colour launch has; the shots per wave of a batched
launch (``sf2d_mesh_set_shots_per_wave``) trade element-data reuse against waves in flight.  Which pair is
fastest depends on the mesh size, the number of shots and the device, so it is measured: every candidate runs
a short forward (``steps`` time steps from rest) on the workload's own mesh, model and shots, and the fastest
device time wins.  The colour schedule itself is the plan's greedy colouring (four colours on the structured
meshes), in launch order.

``tune(cfg, shots)`` returns the choice and the table; ``settings(choice)`` gives it as the ``chain_elems`` and
``shots_per_wave`` arguments of the handle it is for (``Simulation`` / ``ShotBatch``): per handle, nothing
process-wide.
"""
from __future__ import annotations

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
    for n in chain_elems:
        sim = ShotBatch(small, shots, chain_elems=n) if batch else Simulation(small, chain_elems=n)
        try:
            if not sim.chain_kernel():
                continue  # the mesh's patches are not chains: nothing to tune
            for g in groups:
                if batch:
                    sim.set_shots_per_wave(g)
                _forward_ms(sim, steps, batch)  # warm-up
                ms = _forward_ms(sim, steps, batch)
                table.append({"chain_elems": sim.chain_elems(), "shots_per_wave": g, "ms_per_step": ms / steps})
        finally:
            sim.close()
    if not table:
        return {"chain_elems": None, "shots_per_wave": None, "ms_per_step": None, "table": []}
    best = min(table, key=lambda r: r["ms_per_step"])
    return dict(best, table=table)


def settings(choice: Dict) -> Dict:
    """The tuned choice as constructor arguments of ``Simulation`` / ``ShotBatch`` (per handle: the library reads
    them when it builds that handle's plan and launch shape)."""
    return {"chain_elems": int(choice.get("chain_elems") or 0), "shots_per_wave": int(choice.get("shots_per_wave") or 0)}
