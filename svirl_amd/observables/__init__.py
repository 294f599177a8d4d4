from .observables import Observables  # noqa: F401
from .vortex_detector import VortexDetector  # noqa: F401
