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


# ---------------------------------------------------------------------------
# v0 2D block arrays (one block): the reference's input_double_2d.bin fields
# ---------------------------------------------------------------------------
_V0_FROM = dict(strength="strength", uvel="uvel", vvel="vvel", dxt="dxT", dyt="dyT", uarear="uarear",
                cdn_ocn="cdn_ocn", aiX="aiu", uocn="uocn", vocn="vocn", waterx="waterxU", watery="wateryU",
                forcex="forcexU", forcey="forceyU", umassdti="umassdti", fm="fmU", strintx="strintxU",
                strinty="strintyU", Tbu="Tbu")


class Blocks2D:
    """The arrays an evp2d handle takes for one block covering a synth.Grid2D
    (attributes as Engine2D.load_state reads them)."""

    def __init__(self, grid, deltaminEVP: float = 1e-8):
        from . import FIELDS2D
        NX, NY = grid.NX, grid.NY
        self.nxb, self.nyb, self.nblocks = NX, NY, 1
        F = grid.fields
        hte, htn = F["HTE"], F["HTN"]
        htem1 = np.zeros_like(hte)
        htem1[1:, :] = hte[:-1, :]
        htnm1 = np.zeros_like(htn)
        htnm1[:, 1:] = htn[:, :-1]
        # the metric terms with the association of the 1D kernel (bench_v2b.F90:139-147),
        # so that both paths compute the same bits
        met = dict(cxp=1.5 * htn - 0.5 * htnm1, cyp=1.5 * hte - 0.5 * htem1,
                   cxm=-(1.5 * htnm1 - 0.5 * htn), cym=-(1.5 * htem1 - 0.5 * hte),
                   dxhy=0.5 * (hte - htem1), dyhx=0.5 * (htn - htnm1),
                   DminTarea=deltaminEVP * (F["dxT"] * F["dyT"]))
        for v in met.values():
            v[0, :] = 0.0
            v[:, 0] = 0.0
        self.f = {}
        for n in FIELDS2D[:42]:
            if n in met:
                a = met[n]
            elif n in _V0_FROM:
                a = F[_V0_FROM[n]]
            elif n in ("uvel_init", "vvel_init"):
                a = F[n[:4]]
            elif n in ("taubx", "tauby"):
                a = np.zeros((NX, NY))
            else:
                a = F[n]                           # the 12 stresses
            self.f[n] = np.asfortranarray(a, dtype=np.float64).reshape((NX, NY, 1), order="F")

        def cells(mask):
            jj, ii = np.nonzero(mask[1:NX - 1, 1:NY - 1].T)   # CICE order: j outer, i inner
            return (ii + 2).astype(np.int32), (jj + 2).astype(np.int32)
        ti, tj = cells(grid.icetmask)
        ui, uj = cells(grid.iceumask)
        m = NX * NY
        self.indxti = np.zeros((m, 1), np.int32)
        self.indxtj = np.zeros((m, 1), np.int32)
        self.indxui = np.zeros((m, 1), np.int32)
        self.indxuj = np.zeros((m, 1), np.int32)
        self.indxti[:ti.size, 0], self.indxtj[:tj.size, 0] = ti, tj
        self.indxui[:ui.size, 0], self.indxuj[:uj.size, 0] = ui, uj
        self.icellt = np.array([ti.size], np.int32)
        self.icellu = np.array([ui.size], np.int32)
        self.halo = None
