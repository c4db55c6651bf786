"""Solver registry keyed by the class attribute ``name`` (upstream ``solvers/registry.py:26-52``)."""

from __future__ import annotations

import logging
from typing import Any

from caustica_amd.registry import SOLVER_GROUP, PluginRegistry
from caustica_amd.solvers.base import SolverBase

log = logging.getLogger("caustica")


class SolverRegistry(PluginRegistry):
    def register(self, cls):
        name = getattr(cls, "name", None)
        if not name or not isinstance(name, str):
            raise ValueError(f"solver class {cls.__name__} must define a string 'name'")
        return self.add(name, cls)

    def _accept(self, ep_name: str, obj: Any) -> None:
        if isinstance(obj, type) and issubclass(obj, SolverBase):
            self.register(obj)
        else:
            log.warning("%s plugin '%s' is not a SolverBase subclass; ignored", self.group, ep_name)

    def collision_message(self, name: str, held: Any) -> str:
        return f"solver name '{name}' already registered by {held.__name__}"


solver_registry = SolverRegistry("solver", SOLVER_GROUP)
register = solver_registry.register
get = solver_registry.get
available = solver_registry.available
