"""exactclosure.jl_amd -- MI355X-native DNS-aided LES hot path of ExactClosure.jl.

Modules mirror the reference's call surface:
  NavierStokes  (src/N.jl)        -> navierstokes
  Burgers       (src/Burgers.jl)  -> burgers
plus dnsaid (the RK4 DNS-aided LES driver), dnsaid_project (the reference's
five-variant forward-Euler comparison), cnn (the learned closure) and snapshot
(velocity fields on disk in the layout of the reference's saved array).  All compute runs in
libexactclosure.so (hand-written gfx950 HIP kernels behind a C ABI).
"""
from . import _abi  # noqa: F401
from . import navierstokes  # noqa: F401
from . import navierstokes as NavierStokes  # noqa: F401
from . import burgers  # noqa: F401
from . import burgers as Burgers  # noqa: F401
from . import dnsaid  # noqa: F401
from . import dnsaid_project  # noqa: F401
from . import cnn  # noqa: F401
from . import autotune  # noqa: F401
from . import snapshot  # noqa: F401

lib = _abi.lib
