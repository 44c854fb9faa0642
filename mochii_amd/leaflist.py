"""Leaf-list input files (the reference's par%amr_file) through the
library's reader, mc_leaf_list_read (generic_amr_read,
src/read_generic_amr.f90:34-237): text or FITS binary table, gzip or not.

read_leaf_list(path) -> dict(x, y, z, level, half, nH[, metallicity, xHI,
ndust]), coordinates recentred on the box centre as grid_create_amr does
(src/grid_mod_amr.f90:51-61); x, y, z, level, half are the arguments of
Model.grid_octree.  raster_order(...) is the file 'car' grid's permutation
from file row to raster slot (src/grid_mod_amr.f90:164-196)."""
import ctypes as C

import numpy as np

from .lib import MochiiError, check, library

_OPT = ("metallicity", "xHI", "ndust")


def read_leaf_list(path):
    L = library()
    n, half = C.c_int64(), C.c_double()
    has = (C.c_int32 * 3)()
    dp = C.POINTER(C.c_double)
    null = C.cast(None, dp)
    check(L.mc_leaf_list_read(str(path).encode(), 0, C.byref(n), C.byref(half), null, null, null,
                              C.cast(None, C.POINTER(C.c_int32)), null, null, null, null, has),
          "mc_leaf_list_read")
    nl = n.value
    out = {k: np.empty(nl) for k in ("x", "y", "z", "nH")}
    out["level"] = np.empty(nl, dtype=np.int32)
    opt = {k: np.empty(nl) for k, h in zip(_OPT, has) if h}

    def ptr(a):
        return a.ctypes.data_as(dp) if a is not None else null

    check(L.mc_leaf_list_read(str(path).encode(), nl, C.byref(n), C.byref(half), ptr(out["x"]),
                              ptr(out["y"]), ptr(out["z"]),
                              out["level"].ctypes.data_as(C.POINTER(C.c_int32)), ptr(out["nH"]),
                              ptr(opt.get("metallicity")), ptr(opt.get("xHI")),
                              ptr(opt.get("ndust")), has), "mc_leaf_list_read")
    out["half"] = half.value
    out.update(opt)
    return out


def raster_order(x, y, z, level, half):
    """Permutation p (leaf index -> file row) putting a single-level leaf
    list in raster order, ix fastest; the reference's checks raise here."""
    level = np.asarray(level)
    if level.size == 0 or np.any(level != level[0]):
        raise MochiiError("grid_type 'car' from a file needs a single-level leaf list")
    nxu = 2 ** int(level[0])
    if level.size != nxu ** 3:
        raise MochiiError(f"car grid from a file expects nleaf = (2^level)^3 = {nxu ** 3}, "
                          f"got {level.size}")
    dc = 2.0 * half / nxu

    def idx(c):
        return np.clip(((np.asarray(c) + half) / dc).astype(np.int64), 0, nxu - 1)

    slot = idx(x) + nxu * (idx(y) + nxu * idx(z))
    perm = np.full(level.size, -1, dtype=np.int64)
    if np.unique(slot).size != slot.size:
        raise MochiiError("car grid: two leaves map to one raster cell")
    perm[slot] = np.arange(level.size)
    return perm
