from .grid import Grid  # noqa: F401
