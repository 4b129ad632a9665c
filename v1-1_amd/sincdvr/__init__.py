"""MI355X (gfx950) sinc-DVR orbital-operator library: Python host mirror of
the SINCDVR project interface over libsincdvr.so."""
from ._lib import EXPORTS, LIB_PATH, SdvError, load  # noqa: F401
from .ops import Batch, SincDVR, circulant_sizes  # noqa: F401
