"""Compute backends: a name -> factory registry and the ``hip`` backend.

Mirrors upstream ``src/caustica/core/backend.py`` (Backend value object :141-190,
registry :43-44, get_backend :236-270, factory contract checks :273-298). This
package registers exactly one backend, ``"hip"``: the MI355X engine behind
``libcaustica_hip.so``. ``"auto"`` resolves to it. There is no numpy/CPU backend
here, so a missing GPU or library is an error, never a silent fallback.
"""

from __future__ import annotations

import logging
from dataclasses import dataclass
from types import ModuleType
from typing import Any

import numpy as np

from caustica_amd.registry import BACKEND_GROUP, FactoryRegistry, UnknownPluginError

log = logging.getLogger("caustica")

backends: FactoryRegistry = FactoryRegistry("backend", BACKEND_GROUP)


class CausticaWarning(UserWarning):
    """Base category of the warnings this package emits."""


@dataclass(frozen=True)
class Backend:
    """A named backend. ``xp`` is the host array module used for setup arrays."""

    name: str
    xp: ModuleType
    device: int = 0

    @property
    def is_gpu(self) -> bool:
        return self.name == "hip"

    @property
    def native(self) -> bool:
        """True when solves run in libcaustica_hip (the fused HIP step loop)."""
        return self.name == "hip"

    def asarray(self, a: Any, dtype: Any = None) -> Any:
        return self.xp.asarray(a, dtype=dtype)

    def to_numpy(self, a: Any) -> np.ndarray:
        return np.asarray(a)

    def synchronize(self) -> None:
        """Solves synchronize inside their own session; nothing is pending here."""


@backends.register("hip")
def _hip_backend() -> Backend:
    """MI355X through libcaustica_hip; refuses with a fix-it message when unusable."""
    from caustica_amd import _native

    try:
        n = _native.device_count()
    except Exception as exc:  # library missing or HIP runtime error
        raise RuntimeError(
            f"Backend 'hip' is unusable: {exc}. Build the library with "
            f"`python -m caustica_amd.build` and run on a machine with an MI355X (gfx950)."
        ) from None
    if n < 1:
        raise RuntimeError("Backend 'hip' requested but no HIP device is visible.")
    return Backend("hip", np, 0)


def _unknown_backend(name: str) -> ValueError:
    return ValueError(
        f"Unknown backend name {name!r}. Available: {', '.join(backends.available())} "
        f"(plus 'auto', which selects 'hip'). Third-party backends are added through the "
        f"'{BACKEND_GROUP}' entry-point group."
    )


def check_backend_name(name: str) -> str:
    """Refuse an unregistered backend name early; returns it unchanged."""
    if name == "auto" or name in backends:
        return name
    raise _unknown_backend(name)


def get_backend(name: str = "auto") -> Backend:
    """Resolve a backend by name (``"auto"`` -> ``"hip"``)."""
    key = "hip" if name == "auto" else name
    try:
        factory = backends.get(key)
    except UnknownPluginError:
        raise _unknown_backend(name) from None
    return _checked(key, factory)


def _checked(name: str, factory: Any) -> Backend:
    backend = factory()
    if not isinstance(backend, Backend):
        raise TypeError(
            f"backend factory for '{name}' returned {type(backend).__name__}, not a Backend"
        )
    if backend.name != name:
        raise ValueError(
            f"backend factory for '{name}' returned a Backend named {backend.name!r}; the "
            f"registry key and Backend.name must match (runs are labelled by Backend.name)."
        )
    return backend
