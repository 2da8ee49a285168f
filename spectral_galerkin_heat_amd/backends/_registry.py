"""Name → factory registry of backends (sgh/backends/_registry.py:40-109)."""

from __future__ import annotations

from typing import Callable, Dict

from .base import MathBackend

_BACKEND_FACTORIES: Dict[str, Callable[[], MathBackend]] = {}


def register_backend(name: str):
    """Register a zero-argument backend factory; a duplicate name raises ``ValueError``."""

    def decorator(factory):
        if name in _BACKEND_FACTORIES:
            raise ValueError(
                f"Backend {name!r} is already registered "
                f"({_BACKEND_FACTORIES[name]!r}); cannot overwrite it.")
        _BACKEND_FACTORIES[name] = factory
        return factory

    return decorator


def get_backend(name: str = "mi355x") -> MathBackend:
    """Fresh backend by name; ``ValueError`` if unknown, ``ImportError`` if unusable here."""
    factory = _BACKEND_FACTORIES.get(name)
    if factory is None:
        available = ", ".join(sorted(_BACKEND_FACTORIES)) or "(none)"
        raise ValueError(f"Unknown backend: {name!r}. Registered backends: {available}.")
    return factory()


def registered_backends():
    return sorted(_BACKEND_FACTORIES)
