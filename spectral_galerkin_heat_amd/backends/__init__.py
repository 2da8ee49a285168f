"""Backend registry with the MI355X backend (sgh/backends/__init__.py)."""

from ._registry import get_backend, register_backend, registered_backends
from .base import MathBackend, to_host
from .mi355x_backend import MI355XBackend

__all__ = ["MI355XBackend", "MathBackend", "get_backend", "register_backend",
           "registered_backends", "to_host"]
