"""gaia_amd: MI355X-native Vertex Block Descent hot path.

Host interface (mirrors the reference's VBDDynamics framework API) over the
C ABI of libgaia_vbd.so (include/gaia_vbd.h), whose device work runs in
hand-written HIP kernels for gfx950.
"""
from . import scenes  # noqa: F401
from ._lib import LIB_PATH, VBDError, header_functions, lib  # noqa: F401
from .vbd import (  # noqa: F401
    Batch,
    Context,
    Server,
    VBDPhysics,
    cloth_mesh_constants,
    color_graph,
    color_tet_mesh,
    color_tri_mesh,
    physics_params,
    state_read,
    state_write,
    tet_mesh_constants,
    write_ply,
)
