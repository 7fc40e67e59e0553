"""Snapshot export and import, and the CSV time series: the reference's on-disk formats.

* ``.npz`` with ``positions``, ``velocities``, ``densities``, ``pressures`` and ``time``, as
  written by the reference's Python visualiser (python/visualize.py:254-259).
* CSV rows ``time,particles,mass_error,energy,total_energy,avg_density,max_velocity`` as in
  examples/dam_break.cpp:97-101,161-166 (``energy`` is the energy error; ``avg_density``
  averages the capacity-length density vector, as the example does).

Both work with the ``sph`` module's ``Simulator`` and with the ctypes ``sphk.Engine``.
"""
from __future__ import annotations

import numpy as np


def _state(sim):
    if hasattr(sim, "get_positions"):  # sph.Simulator
        return (np.asarray(sim.get_positions(), np.float32), np.asarray(sim.get_velocities(), np.float32),
                np.asarray(sim.get_densities(), np.float32), np.asarray(sim.get_pressures(), np.float32),
                float(sim.get_current_time()))
    return (sim.positions(), sim.velocities(), sim.densities(), sim.pressures(), float(sim.time()))


def save_npz(sim, path: str) -> None:
    """Writes the current state in the reference's .npz layout."""
    pos, vel, rho, prs, t = _state(sim)
    np.savez(path, positions=pos, velocities=vel, densities=rho, pressures=prs, time=t)


def load_npz(path: str) -> dict:
    """Reads an .npz snapshot (no pickled objects are accepted)."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in ("positions", "velocities", "densities", "pressures", "time")}


def restore_npz(engine, path: str) -> dict:
    """Uploads the positions and velocities of a snapshot into a ``sphk.Engine`` (caller
    order); returns the snapshot.  Densities, pressures and the clock are not state of the
    step: the next step recomputes the first two, and the clock restarts at the upload."""
    snap = load_npz(path)
    engine.upload(snap["positions"], snap["velocities"])
    return snap


class CsvTimeSeries:
    """Appends one row per ``record`` call, in the example program's columns."""

    HEADER = "time,particles,mass_error,energy,total_energy,avg_density,max_velocity"

    def __init__(self, path: str):
        self.f = open(path, "w")
        self.f.write(self.HEADER + "\n")

    def record(self, sim) -> None:
        if hasattr(sim, "compute_conservation_errors"):  # sph.Simulator
            me, ee = sim.compute_conservation_errors()
            n = sim.get_particles().size()
            total = sim.get_total_energy()
        else:
            me, ee = sim.conservation_errors()
            n = sim.size()
            total = sim.total_energy()
        pos, vel, rho, _, t = _state(sim)
        vmax = float(np.sqrt((vel.astype(np.float64) ** 2).sum(1)).max()) if len(vel) else 0.0
        avg = float(np.float32(rho.sum(dtype=np.float32) / np.float32(max(len(rho), 1))))
        self.f.write(f"{t},{n},{me},{ee},{total},{avg},{vmax}\n")
        self.f.flush()

    def close(self) -> None:
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
