"""brax_amd: an MI355X-native generalized-coordinates physics step.

The package mirrors the reference's generalized pipeline API
(brax/generalized/pipeline.py:32-94): ``brax_amd.generalized.pipeline.init``
and ``.step`` take a ``System`` and return a ``State``; the step itself runs in
hand-written HIP kernels behind the C-ABI declared in ``include/brax_gen.h``.
"""

from brax_amd import systems  # noqa: F401
from brax_amd.systems import System  # noqa: F401

__version__ = "0.1.0"
