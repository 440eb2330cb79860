"""2D <-> 1D packing for the Python host mirror.

pack() follows convert_2d_1d_init (integration/ice_dyn_evp1d.F90:803-815),
convert_2d_1d_dyn (845-881, ghost uvel/vvel set to zero) and set_skipMe
(521-536); unpack() follows convert_1d_2d_dyn (884-947).  uvel_init/vvel_init
take the packed uvel/vvel as the reference's readers do (src/vars.F90:2120).
The production host does the same in Fortran (fortran/ice_dyn_evp1d.F90).
"""
from __future__ import annotations

import numpy as np

from . import FIELDS, NAVEL_FIELDS, NBR

DYN_2D = (
    "stressp_1", "stressp_2", "stressp_3", "stressp_4", "stressm_1", "stressm_2", "stressm_3",
    "stressm_4", "stress12_1", "stress12_2", "stress12_3", "stress12_4", "strength", "cdn_ocn",
    "aiu", "uocn", "vocn", "waterxU", "wateryU", "forcexU", "forceyU", "umassdti", "fmU",
    "strintxU", "strintyU", "Tbu", "Cb", "uvel", "vvel")


def pack(grid, ix: dict) -> dict:
    """1D state of a synth.Grid2D for the list ix (evp1d.build_index)."""
    na, navel = ix["na"], ix["navel"]
    i = ix["indxti"].astype(np.int64) - 1
    j = ix["indxtj"].astype(np.int64) - 1
    F = grid.fields
    f = {n: np.zeros(navel if n in NAVEL_FIELDS else na) for n in FIELDS}
    for n in DYN_2D:
        f[n][:na] = F[n][i, j]
    for n in ("uarear", "dxT", "dyT", "HTE", "HTN"):
        f[n][:] = F[n][i, j]
    f["HTEm1"][:] = F["HTE"][i - 1, j]
    f["HTNm1"][:] = F["HTN"][i, j - 1]
    f["uvel_init"][:] = f["uvel"][:na]
    f["vvel_init"][:] = f["vvel"][:na]
    return dict(
        na=na, navel=navel, fields=f,
        idx={n: np.ascontiguousarray(ix[n], dtype=np.int32) for n in NBR},
        skipT=(~grid.icetmask[i, j]).astype(np.int32),
        skipU=(~grid.iceumask[i, j]).astype(np.int32),
    )


def unpack(fields: dict, ix: dict, shape) -> dict:
    """2D fields (zero outside the list) from 1D fields."""
    i = ix["indxti"].astype(np.int64) - 1
    j = ix["indxtj"].astype(np.int64) - 1
    na = ix["na"]
    out = {}
    for n in DYN_2D:
        a = np.zeros(shape, order="F")
        a[i, j] = fields[n][:na]
        out[n] = a
    return out
