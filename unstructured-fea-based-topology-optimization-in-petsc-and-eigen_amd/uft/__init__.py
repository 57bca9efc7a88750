"""Python binding of libuft.so (the MI355X SIMP core) over its C-ABI.

Used by the tests and the benchmark; the production host side is the C++
drop-in in ../dropin.  Importing this module never falls back to a CPU
implementation: if libuft.so is missing or cannot be loaded it raises.
"""
from .capi import (Context, Params, UFTError, lib, lib_path, gen_box, box_bcs, VEC, SIZE,  # noqa: F401
                   FP64, MIXED, exported_symbols, header_symbols, bc_read_text, mesh_read_text,
                   vtu_write_host, node_tiles_host)
