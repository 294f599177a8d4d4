from .arrays import GArray  # noqa: F401
