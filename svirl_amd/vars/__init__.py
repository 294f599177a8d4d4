from .fixed_vortices import FixedVortices  # noqa: F401
from .vars import Vars  # noqa: F401
from .params import Params  # noqa: F401
