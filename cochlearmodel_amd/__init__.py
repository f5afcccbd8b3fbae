"""MI355X-native cochlear solver: the BM/OHC velocity step of
YonatanKoral/CochlearModel (Run_Fast_BM_Calculation=17), the IHC/lambda
filters and the JND reductions as gfx950 HIP kernels behind a C-ABI
(include/cochlea_amd.h).  Python here is a thin host over that C-ABI.
"""
from ._lib import CochleaError, lib  # noqa: F401

SECTIONS = 256
LAMBDA_COUNT = 3
