from .cg import CG  # noqa: F401
from .solvers import Solvers  # noqa: F401
from .td import TD  # noqa: F401
