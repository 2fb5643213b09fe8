"""MI355X-native backend for GridapMHD's H1-HDiv assembly + block-preconditioned Newton-Krylov hot path."""
from . import meshers, fespaces, _lib  # noqa: F401
