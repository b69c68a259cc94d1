"""pbdfluid - Python host interface of the MI355X-native PBF step.

The classes mirror the reference's step interface:
  Config      src/parser/Config.h:18-124   (getValue / setValue by dotted path, load, reload, write)
  Fluid       src/app/simulation/Fluid.h   (compute() = one step, Fluid.cu:14-97)
  Simulation  src/app/Simulation.h         (step() = compute() then device sync, Simulation.cu:16-19)
All work happens in libpbd.so (HIP kernels for gfx950 + C++ host) through the
C-ABI of include/pbd/pbd.h; there is no Python or CPU fallback.
"""
from __future__ import annotations

import os

import numpy as np

from . import capi
from .capi import (BUF, PBD_KERNEL_REFERENCE, PBD_KERNEL_TEXTBOOK, Context, PbdError, PbdParams, default_params,
                   generate_box, get_param, lattice_rest_density, load_params, set_param, write_params)

CONFIG_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "configs")


def config_path(name: str) -> str:
    """Path of a shipped config (reference.txt, dam_break_32k.txt, dam_break_1m.txt, dam_break_2m.txt)."""
    return os.path.join(CONFIG_DIR, name if name.endswith(".txt") else name + ".txt")


class Config:
    """Parameter set read from a config.txt (Config.h:21-111)."""

    def __init__(self, file: str | None = None, params: PbdParams | None = None):
        self.file = file
        if params is not None:
            self.params = params.copy()
        elif file is not None:
            self.params = load_params(file)
        else:
            self.params = default_params()

    def getValue(self, key: str, typ=str):
        v = get_param(self.params, key)
        if typ is str:
            return v
        if typ is bool:
            return float(v) != 0.0
        return typ(float(v)) if typ is int else typ(v)

    def setValue(self, value, key: str) -> None:
        if isinstance(value, bool):
            value = int(value)
        set_param(self.params, key, value)

    def load(self, file: str) -> None:
        self.file = file
        self.params = load_params(file)

    def reload(self) -> None:
        if self.file:
            self.params = load_params(self.file)

    def write(self, file: str | None = None) -> str:
        text = write_params(self.params)
        target = file or self.file
        if target:
            with open(target, "w") as fh:
                fh.write(text)
        return text


class Fluid:
    """The step object (Fluid : Cuda_Computable, Fluid.cu:4-97)."""

    def __init__(self, config: Config, capacity: int = 0):
        self.config = config
        self.ctx = Context(config.params, capacity)
        self._applied = config.params.copy()

    def _apply_config(self) -> None:
        # initializeFrame re-reads the configuration every frame (Initialize.h:20-71)
        if bytes(self._applied) != bytes(self.config.params):
            self.ctx.set_params(self.config.params)
            self._applied = self.config.params.copy()

    def compute(self) -> None:
        self._apply_config()
        self.ctx.compute()

    def step(self, frames: int = 1, graph: bool = False) -> None:
        self._apply_config()
        self.ctx.step(frames, graph)

    def add_particles(self, pos4, vel4=None, col4=None) -> None:
        self.ctx.add_particles(pos4, vel4, col4)

    def add_particle(self, pos, direction) -> None:
        self.ctx.add_particle(pos, direction)

    def spawn_box(self) -> None:
        self._apply_config()
        self.ctx.spawn_box()

    def clear_particles(self) -> None:
        self.ctx.clear()

    @property
    def number_of_particles(self) -> int:
        return self.ctx.n

    def download(self, which) -> np.ndarray:
        return self.ctx.download(which)

    def close(self) -> None:
        self.ctx.close()


class Simulation:
    """Simulation::step (Simulation.cu:16-19): one compute() then a device sync."""

    def __init__(self, config: Config, capacity: int = 0):
        self.fluid = Fluid(config, capacity)

    def step(self) -> None:
        self.fluid.compute()
        self.fluid.ctx.sync()


__all__ = [
    "BUF", "Config", "Context", "Fluid", "PbdError", "PbdParams", "Simulation", "PBD_KERNEL_REFERENCE",
    "PBD_KERNEL_TEXTBOOK", "capi", "config_path", "default_params", "generate_box", "get_param",
    "lattice_rest_density", "load_params", "set_param", "write_params",
]
