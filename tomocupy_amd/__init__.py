"""tomocupy_amd: MI355X (gfx950) implementation of tomocupy's fourierrec path.

Python classes mirror the reference's interface for this path
(``FourierRec``, ``FBPFilter``, ``BackprojFunctions``, ``ProcFunctions``,
``GPURec``, ``GPURecSteps``, the phase filters of ``retrieve_phase``) on
torch GPU tensors; all compute runs in hand-written HIP kernels
of ``libtomohip.so`` through the C ABI in ``include/tomohip.h``.
"""
from .fourierrec import FourierRec
from .fbp_filter import FBPFilter
from .backproj_functions import BackprojFunctions
from .proc_functions import ProcFunctions

__all__ = ['FourierRec', 'FBPFilter', 'BackprojFunctions', 'ProcFunctions']
__version__ = '0.1.0'
