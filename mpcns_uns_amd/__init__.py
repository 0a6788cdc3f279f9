"""MI355X-native (gfx950) Hall-MHD step of MPCNS_Uns: HIP kernels behind a C-ABI (include/mhdg.h).

`Mesh` builds the signed cell complex, geometry and boundary data on the host; `Device` holds one mesh on
the GPU and runs the reference's mhd_common operations and the fused Mercury step there.
"""
from .mesh import Mesh, default_config, read_fields, write_fields, write_pvd, write_pvtu  # noqa: F401
from .device import Device  # noqa: F401
