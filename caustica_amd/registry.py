"""Name -> implementation registries with lazily scanned entry-point groups.

Same contract as upstream ``src/caustica/registry.py:112-267`` for the two axes this
package serves (solvers and compute backends): built-ins register through the same
door as plugins, the ``importlib.metadata`` entry-point group is scanned only on the
first lookup, registering a second implementation under a taken name is an error
(re-registering the same definition, e.g. after a module reload, is not), and a
failed lookup lists what is registered and the group to add one through.
The entry-point group names are the upstream ones, so an upstream plugin's
``pyproject.toml`` declaration reaches this package unchanged.
"""

from __future__ import annotations

import logging
from collections.abc import Callable
from importlib import metadata
from typing import Any, Generic, TypeVar

log = logging.getLogger("caustica")

T = TypeVar("T")

SOLVER_GROUP = "caustica.solvers"
BACKEND_GROUP = "caustica.backends"


class UnknownPluginError(KeyError):
    """Lookup miss; a KeyError whose str() is the plain message."""

    def __str__(self) -> str:
        return str(self.args[0]) if self.args else ""


def _identity(obj: Any) -> tuple[Any, Any] | None:
    mod = getattr(obj, "__module__", None)
    qual = getattr(obj, "__qualname__", None)
    if mod is None or qual is None or "<lambda>" in qual:
        return None
    return mod, qual


class PluginRegistry(Generic[T]):
    """name -> implementation behind a lazily scanned entry-point group."""

    def __init__(self, label: str, group: str) -> None:
        self.label = label
        self.group = group
        self._items: dict[str, T] = {}
        self._scanned = False

    # -- registration
    def add(self, name: str, obj: T) -> T:
        held = self._items.get(name)
        if held is not None and held is not obj:
            same = _identity(held) is not None and _identity(held) == _identity(obj)
            if not same:
                raise ValueError(self.collision_message(name, held))
        self._items[name] = obj
        return obj

    def collision_message(self, name: str, held: Any) -> str:
        return f"{self.label} name '{name}' already registered by {getattr(held, '__name__', held)!r}"

    def _accept(self, ep_name: str, obj: Any) -> None:
        self.add(ep_name, obj)

    # -- lookup
    def _scan(self) -> None:
        if self._scanned:
            return
        self._scanned = True
        try:
            eps = metadata.entry_points(group=self.group)
        except Exception as exc:  # noqa: BLE001 - a broken environment must not break lookups
            log.warning("entry-point scan of %s failed: %s", self.group, exc)
            return
        for ep in eps:
            if ep.name in self._items:
                continue
            try:
                self._accept(ep.name, ep.load())
            except Exception as exc:  # noqa: BLE001 - one broken plugin never breaks the rest
                log.warning("%s plugin '%s' failed to load: %s", self.group, ep.name, exc)

    def get(self, name: str) -> T:
        if name not in self._items:
            self._scan()
        try:
            return self._items[name]
        except KeyError:
            raise UnknownPluginError(
                f"unknown {self.label} {name!r}; available: {', '.join(self.available()) or '(none)'}. "
                f"Third-party {self.label}s are added through the '{self.group}' entry-point group."
            ) from None

    def available(self) -> list[str]:
        self._scan()
        return sorted(self._items)

    def __contains__(self, name: str) -> bool:
        self._scan()
        return name in self._items


class FactoryRegistry(PluginRegistry[Callable[..., Any]]):
    """Registry of zero-argument factories; ``@reg.register("name")`` decorator."""

    def register(self, name: str) -> Callable[[Callable[..., Any]], Callable[..., Any]]:
        def deco(fn: Callable[..., Any]) -> Callable[..., Any]:
            return self.add(name, fn)

        return deco
