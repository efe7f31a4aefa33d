"""a5_streamer_amd: the a5_streamer hot path on MI355X (gfx950).

The product is the C-ABI library liba5g.so (declared in include/a5g.h). This
package is a thin ctypes binding over it whose Tree class mirrors the Afivo
calls the reference drivers make (a$D_init, a$D_set_base,
a$D_adjust_refinement, mg$D_fas_fmg, a$D_gc_tree, ...).
"""
from .lib import A5GError, Tree, lib_path, load, read_leaf_dump  # noqa: F401

__all__ = ["A5GError", "Tree", "load", "lib_path", "read_leaf_dump"]
