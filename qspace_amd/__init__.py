"""qspace_amd -- MI355X (gfx950) block-sparse symmetric tensor contraction.

The hot path of QSpace's contractQS and the two-site H*psi it drives,
implemented as HIP kernels behind a C-ABI (include/qsx.h, libqsx.so).
"""
from .qspace import QSpace, pack_itag, unpack_itag, parse_ctridx  # noqa: F401
from .api import contractQS, permuteQS, plusQS, orthoQS, ortho_args, itag_match, CtrPlan, DeviceBuffer, HPsi, su2_cg, su2_x3, su2_om_basis, su2_xom, make_opts  # noqa: F401
from ._capi import QsxError  # noqa: F401
