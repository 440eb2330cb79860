"""Python host mirror of the MI355X EVP1D engine (ctypes on libevp1d.so).

The product boundary is the C ABI in include/evp1d.h; the production host is
the Fortran layer in cice-evp1d_amd/fortran.  This module binds the same .so
for the verification harness and bench.py:

    from evp1d import Engine, set_evp_parameters, build_index
    p = set_evp_parameters(ndte=120)
    eng = Engine(na, navel, p)               # evp1d_create
    eng.set_index("ee", ee) ...              # 1-based, as findXinY produces
    eng.set_field("uvel", uvel) ...
    eng.run(120)                             # evp1d_run
    eng.get_field("stressp_1")

Names and meaning follow the reference (src/bench_v2b.F90, src/vars.F90,
integration/ice_dyn_evp1d.F90).  Every entry point raises Evp1dError with the
library's message on a nonzero status; nothing here falls back to a CPU path:
if the shared library is missing the import fails loudly.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent.parent
LIBDIR = PKG / "lib"

FIELDS = (
    "strength", "uvel", "vvel", "dxT", "dyT", "uarear", "cdn_ocn", "aiu", "uocn", "vocn",
    "waterxU", "wateryU", "forcexU", "forceyU", "umassdti", "fmU", "strintxU", "strintyU", "Tbu",
    "stressp_1", "stressp_2", "stressp_3", "stressp_4",
    "stressm_1", "stressm_2", "stressm_3", "stressm_4",
    "stress12_1", "stress12_2", "stress12_3", "stress12_4",
    "HTE", "HTN", "HTEm1", "HTNm1", "uvel_init", "vvel_init", "Cb",
    "str1", "str2", "str3", "str4", "str5", "str6", "str7", "str8",
)
FIELD_ID = {n: k for k, n in enumerate(FIELDS)}
NAVEL_FIELDS = frozenset({"uvel", "vvel", "str1", "str2", "str3", "str4", "str5", "str6", "str7", "str8"})
STRESS_FIELDS = FIELDS[19:31]
NBR = ("ee", "ne", "se", "nw", "sw", "sse")
MODES = {"launch": 0, "graph": 1, "persistent": 2, "auto": 3, "fused": 4}

# every symbol include/evp1d.h declares
EXPORTS = (
    "evp1d_set_evp_parameters", "evp1d_create", "evp1d_destroy", "evp1d_set_params",
    "evp1d_get_params", "evp1d_get_sizes", "evp1d_set_field", "evp1d_get_field",
    "evp1d_set_index", "evp1d_set_skip", "evp1d_set_halo", "evp1d_stress", "evp1d_stepu",
    "evp1d_halo", "evp1d_run", "evp1d_set_mode", "evp1d_get_mode", "evp1d_run_individual",
    "evp1d_synchronize", "evp1d_timer_start", "evp1d_timer_stop", "evp1d_index_count",
    "evp1d_index_build", "evp1d_last_error", "evp1d_build_info",
    "evp1d_create_batch", "evp1d_get_members", "evp1d_member_info", "evp1d_set_field_member",
    "evp1d_get_field_member", "evp1d_set_index_member", "evp1d_set_skip_member",
    "evp1d_autotune", "evp1d_set_tuning", "evp1d_get_tuning",
    "evp1d_adjoint_run", "evp1d_loss", "evp1d_loss_noalloc", "evp1d_tangent_run", "evp1d_set_adjoint_segment",
    "evp1d_get_adjoint_segment", "evp1d_get_gradient_storage", "evp1d_release_gradient_storage",
    "evp1d_set_tiling", "evp1d_get_tiling", "evp1d_tile_plan",
    "evp1d_set_fused_inplace", "evp1d_get_fused_inplace", "evp1d_autotune_fused",
    "evp1d_set_fused_depth", "evp1d_get_fused_depth", "evp1d_set_tile_order", "evp1d_get_tile_order",
    "evp1d_set_cache_hints", "evp1d_get_cache_hints",
    "evp1d_set_grid", "evp1d_put_2d", "evp1d_put_2d_async", "evp1d_get_2d", "evp1d_host_alloc", "evp1d_host_free",
    "evp2d_create", "evp2d_destroy", "evp2d_set_params", "evp2d_set_cells", "evp2d_set_halo",
    "evp2d_set_field", "evp2d_get_field", "evp2d_run", "evp2d_stress", "evp2d_stepu", "evp2d_halo",
    "evp2d_run_individual", "evp2d_synchronize", "evp2d_timer_start", "evp2d_timer_stop",
    "evp2d_get_cells", "evp2d_set_mode", "evp2d_get_mode", "evp2d_set_cache_hints", "evp2d_get_cache_hints",
)


class Evp1dError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[evp1d status {code}] {msg}")
        self.code = code


class Params(ctypes.Structure):
    """evp1d_params (include/evp1d.h); mirrors ice_dyn_shared, src/ice.F90:80-107."""

    _fields_ = [("ndte", ctypes.c_int32), ("reserved", ctypes.c_int32)] + [
        (n, ctypes.c_double) for n in (
            "capping", "e_factor", "epp2i", "deltaminEVP", "arlx", "arlx1i", "brlx", "denom1",
            "revp", "Ktens", "u0", "cosw", "sinw", "rhow")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)
_H = ctypes.c_void_p

_libs: dict = {}


def lib_path(fma: bool = False) -> Path:
    return LIBDIR / ("libevp1d_fma.so" if fma else "libevp1d.so")


def load(fma: bool = False, path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Load libevp1d(.so|_fma.so) from the package's lib/ (raises if absent).

    `path` binds another build of the same ABI (e.g. a tuning variant) in
    place of the default one for this process."""
    key = bool(fma)
    if path is None and key in _libs:
        return _libs[key]
    path = Path(path) if path is not None else lib_path(fma)
    if not path.exists():
        raise ImportError(f"{path} is not built: run __graft_entry__.build() or python -m evp1d.build")
    L = ctypes.CDLL(str(path))
    i32 = ctypes.c_int32
    sig = {
        "evp1d_set_evp_parameters": [i32, ctypes.c_double, ctypes.POINTER(Params)],
        "evp1d_create": [i32, i32, ctypes.POINTER(Params), ctypes.POINTER(_H)],
        "evp1d_destroy": [_H],
        "evp1d_set_params": [_H, ctypes.POINTER(Params)],
        "evp1d_get_params": [_H, ctypes.POINTER(Params)],
        "evp1d_get_sizes": [_H, _IP, _IP],
        "evp1d_set_field": [_H, i32, _DP, i32],
        "evp1d_get_field": [_H, i32, _DP, i32],
        "evp1d_set_index": [_H, i32, _IP, i32],
        "evp1d_set_skip": [_H, i32, _IP, i32],
        "evp1d_set_halo": [_H, i32, _IP, _IP],
        "evp1d_stress": [_H], "evp1d_stepu": [_H], "evp1d_halo": [_H],
        "evp1d_run": [_H, i32],
        "evp1d_set_mode": [_H, i32],
        "evp1d_get_mode": [_H, _IP],
        "evp1d_run_individual": [_H, i32, i32],
        "evp1d_synchronize": [_H],
        "evp1d_timer_start": [_H],
        "evp1d_timer_stop": [_H, ctypes.POINTER(ctypes.c_float)],
        "evp1d_index_count": [i32, i32, i32, _IP, _IP, _IP],
        "evp1d_index_build": [i32, i32, i32, _IP, i32, i32] + [_IP] * 9,
        "evp1d_last_error": [ctypes.c_char_p, i32],
        "evp1d_build_info": [_IP, _IP],
        "evp1d_create_batch": [i32, _IP, _IP, ctypes.POINTER(Params), ctypes.POINTER(_H)],
        "evp1d_get_members": [_H, _IP],
        "evp1d_member_info": [_H, i32, _IP, _IP, _IP, _IP],
        "evp1d_set_field_member": [_H, i32, i32, _DP, i32],
        "evp1d_get_field_member": [_H, i32, i32, _DP, i32],
        "evp1d_set_index_member": [_H, i32, i32, _IP, i32],
        "evp1d_set_skip_member": [_H, i32, i32, _IP, i32],
        "evp1d_autotune": [_H, i32, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
        "evp1d_set_tuning": [_H, i32, i32, i32, i32],
        "evp1d_get_tuning": [_H, _IP, _IP, _IP, _IP],
        "evp1d_adjoint_run": [_H, i32, i32, _DP, _DP, _DP, _DP, _DP, _DP, _DP],
        "evp1d_loss": [_H, _DP, _DP, _DP],
        "evp1d_loss_noalloc": [_H, _DP, _DP, _DP],
        "evp1d_set_adjoint_segment": [_H, i32],
        "evp1d_set_tile_order": [_H, i32],
        "evp1d_get_tile_order": [_H, _IP],
        "evp1d_get_adjoint_segment": [_H, _IP],
        "evp1d_get_gradient_storage": [_H, ctypes.POINTER(ctypes.c_int64)],
        "evp1d_release_gradient_storage": [_H],
        "evp1d_tangent_run": [_H, i32, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP],
        "evp1d_set_tiling": [_H, i32, i32, i32, i32],
        "evp1d_set_grid": [_H, i32, i32, _IP, i32],
        "evp1d_put_2d": [_H, i32, _IP, _DP],
        "evp1d_put_2d_async": [_H, i32, _IP, _DP],
        "evp1d_get_2d": [_H, i32, _IP, _DP],
        "evp1d_host_alloc": [ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)],
        "evp1d_host_free": [ctypes.c_void_p],
        "evp1d_get_tiling": [_H, _IP, _IP, _IP, _IP, _IP],
        "evp1d_set_fused_inplace": [_H, i32],
        "evp1d_get_fused_inplace": [_H, _IP],
        "evp1d_autotune_fused": [_H, i32, ctypes.POINTER(ctypes.c_float)],
        "evp1d_set_fused_depth": [_H, i32],
        "evp1d_get_fused_depth": [_H, _IP, _IP, _IP],
        "evp1d_set_cache_hints": [_H, i32],
        "evp1d_get_cache_hints": [_H, _IP, _IP],
        "evp2d_create": [i32, i32, i32, ctypes.POINTER(Params), ctypes.POINTER(_H)],
        "evp2d_destroy": [_H],
        "evp2d_set_params": [_H, ctypes.POINTER(Params)],
        "evp2d_set_cells": [_H, i32, i32, _IP, _IP, i32, _IP, _IP],
        "evp2d_set_halo": [_H, i32, _IP, _IP, _IP, _IP, i32, _IP, _IP],
        "evp2d_set_field": [_H, i32, _DP, ctypes.c_int64],
        "evp2d_get_field": [_H, i32, _DP, ctypes.c_int64],
        "evp2d_run": [_H, i32], "evp2d_stress": [_H], "evp2d_stepu": [_H], "evp2d_halo": [_H],
        "evp2d_run_individual": [_H, i32, i32],
        "evp2d_synchronize": [_H],
        "evp2d_timer_start": [_H],
        "evp2d_timer_stop": [_H, ctypes.POINTER(ctypes.c_float)],
        "evp2d_get_cells": [_H, _IP, _IP],
        "evp2d_set_mode": [_H, i32],
        "evp2d_get_mode": [_H, _IP],
        "evp2d_set_cache_hints": [_H, i32],
        "evp2d_get_cache_hints": [_H, _IP, _IP],
        "evp1d_tile_plan": [i32, i32] + [_IP] * 8 + [i32, i32, i32] + [_IP] * 5 + [i32, _IP, i32, _IP, _IP, i32,
                                                                                _IP, _IP],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _libs[key] = L
    return L


def _check(L, rc: int) -> None:
    if rc != 0:
        buf = ctypes.create_string_buffer(512)
        L.evp1d_last_error(buf, 512)
        raise Evp1dError(rc, buf.value.decode(errors="replace"))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def _ip(a: np.ndarray):
    return a.ctypes.data_as(_IP)


def set_evp_parameters(ndte: int, dt: float = 3600.0, fma: bool = False, **override) -> Params:
    """set_evp_parameters (src/ice.F90:109-130) plus optional overrides such as
    the four scalars readin_1d_v2 takes from the input file (vars.F90:2116)."""
    L = load(fma)
    p = Params()
    _check(L, L.evp1d_set_evp_parameters(int(ndte), float(dt), ctypes.byref(p)))
    for k, v in override.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def build_info(fma: bool = False) -> dict:
    L = load(fma)
    a, f = ctypes.c_int32(), ctypes.c_int32()
    _check(L, L.evp1d_build_info(ctypes.byref(a), ctypes.byref(f)))
    return {"abi": a.value, "fma": f.value}


def build_index(tmask: np.ndarray, nghost: int = 1, fma: bool = False) -> dict:
    """1D point list of a Fortran-ordered (nx, ny) T mask (rim included):
    evp1d_index_count / evp1d_index_build."""
    L = load(fma)
    nx, ny = tmask.shape
    m = np.asfortranarray(tmask != 0).astype(np.int32, order="F")
    na, navel = ctypes.c_int32(), ctypes.c_int32()
    _check(L, L.evp1d_index_count(nx, ny, nghost, _ip(m), ctypes.byref(na), ctypes.byref(navel)))
    na, navel = na.value, navel.value
    ti = np.zeros(max(na, 1), np.int32)
    tj = np.zeros(max(na, 1), np.int32)
    tij = np.zeros(max(navel, 1), np.int32)
    nb = {n: np.zeros(max(na, 1), np.int32) for n in NBR}
    _check(L, L.evp1d_index_build(nx, ny, nghost, _ip(m), na, navel, _ip(ti), _ip(tj), _ip(tij),
                                  *[_ip(nb[n]) for n in NBR]))
    out = dict(na=na, navel=navel, indxti=ti[:na], indxtj=tj[:na], indxtij=tij[:navel])
    out.update({n: nb[n][:na] for n in NBR})
    return out


def tile_plan(idx: dict, skipT, skipU, navel: int, tile_w: int = 32, tile_h: int = 16,
              max_entries: int = 2560, fma: bool = False) -> dict:
    """Host-only fused-subcycle tiling of a 1D list (evp1d_tile_plan).

    idx: the six 1-based neighbour lists.  Returns perm (slot -> 1-based
    point), counts, tiles (n x 14: a0,a1,b0,b1,c0,c1,h0,nh,k0,nk,sa,ia,sb,ib), list
    (halo slots), klist (constant slots), uloc (stepu slot x 4: own, nw, sse,
    sw LDS entries) and max_entries; ntiles 0 when the lists are no grid."""
    L = load(fma)
    skipT = np.ascontiguousarray(skipT, dtype=np.int32)
    skipU = np.ascontiguousarray(skipU, dtype=np.int32)
    na = len(skipT)
    lists = [np.ascontiguousarray(idx[n], dtype=np.int32) for n in NBR]
    perm = np.zeros(max(na, 1), np.int32)
    counts = np.zeros(4, np.int32)
    nt, mx, nl, nk = (ctypes.c_int32() for _ in range(4))
    args = [na, navel, *[_ip(a) for a in lists], _ip(skipT), _ip(skipU), tile_w, tile_h, max_entries, _ip(perm),
            _ip(counts), ctypes.byref(nt), ctypes.byref(mx)]
    _check(L, L.evp1d_tile_plan(*args, None, 0, None, 0, ctypes.byref(nl), None, 0, ctypes.byref(nk), None))
    tiles = np.zeros((max(nt.value, 1), 14), np.int32)
    hl = np.zeros(max(nl.value, 1), np.int32)
    kl = np.zeros(max(nk.value, 1), np.int32)
    uloc = np.zeros((max(int(counts[1] + counts[2]), 1), 4), np.int32)
    _check(L, L.evp1d_tile_plan(*args, _ip(tiles), nt.value, _ip(hl), nl.value, ctypes.byref(nl), _ip(kl),
                                nk.value, ctypes.byref(nk), _ip(uloc)))
    return dict(perm=perm[:na], counts=counts, ntiles=nt.value, max_entries=mx.value, tiles=tiles[:nt.value],
                list=hl[:nl.value], klist=kl[:nk.value], uloc=uloc[:int(counts[1] + counts[2])])


class Engine:
    """One evp1d handle (device-resident 1D state of one domain or batch)."""

    def __init__(self, na: int, navel: int, params: Params, fma: bool = False, members=None):
        """members: optional list of (na_k, navel_k) -> evp1d_create_batch."""
        self.L = load(fma)
        self.fma = fma
        h = _H()
        if members is None:
            self.na, self.navel = int(na), int(navel)
            _check(self.L, self.L.evp1d_create(self.na, self.navel, ctypes.byref(params), ctypes.byref(h)))
            self.members = [(self.na, self.navel)]
        else:
            mna = np.ascontiguousarray([m[0] for m in members], dtype=np.int32)
            mnv = np.ascontiguousarray([m[1] for m in members], dtype=np.int32)
            self.na, self.navel = int(mna.sum()), int(mnv.sum())
            _check(self.L, self.L.evp1d_create_batch(mna.size, _ip(mna), _ip(mnv), ctypes.byref(params),
                                                     ctypes.byref(h)))
            self.members = [(int(a), int(b)) for a, b in zip(mna, mnv)]
        self.h = h

    @classmethod
    def batch(cls, members, params: Params, fma: bool = False) -> "Engine":
        return cls(0, 0, params, fma=fma, members=members)

    def member_info(self, k: int) -> dict:
        a = [ctypes.c_int32() for _ in range(4)]
        _check(self.L, self.L.evp1d_member_info(self.h, k, *[ctypes.byref(x) for x in a]))
        return dict(zip(("na", "navel", "active_offset", "ghost_offset"), (x.value for x in a)))

    def set_field_member(self, k: int, name: str, a) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(self.L, self.L.evp1d_set_field_member(self.h, k, FIELD_ID[name], _dp(a), a.size))

    def get_field_member(self, k: int, name: str) -> np.ndarray:
        na_k, navel_k = self.members[k]
        out = np.empty(navel_k if name in NAVEL_FIELDS else na_k)
        _check(self.L, self.L.evp1d_get_field_member(self.h, k, FIELD_ID[name], _dp(out), out.size))
        return out

    def set_index_member(self, k: int, name: str, idx) -> None:
        a = np.ascontiguousarray(idx, dtype=np.int32)
        _check(self.L, self.L.evp1d_set_index_member(self.h, k, NBR.index(name), _ip(a), a.size))

    def set_skip_member(self, k: int, which: str, flags) -> None:
        a = np.ascontiguousarray(flags, dtype=np.int32)
        _check(self.L, self.L.evp1d_set_skip_member(self.h, k, {"skipT": 0, "skipU": 1}[which], _ip(a), a.size))

    def load_member(self, k: int, fields: dict, idx: dict, skipT, skipU) -> None:
        for n in NBR:
            self.set_index_member(k, n, idx[n])
        self.set_skip_member(k, "skipT", skipT)
        self.set_skip_member(k, "skipU", skipU)
        for n in FIELDS:
            if n in fields:
                self.set_field_member(k, n, fields[n])

    def autotune(self, reps: int = 20) -> dict:
        a, b = ctypes.c_float(), ctypes.c_float()
        _check(self.L, self.L.evp1d_autotune(self.h, reps, ctypes.byref(a), ctypes.byref(b)))
        out = self.get_tuning()
        out.update(stress_us=a.value, stepu_us=b.value)
        return out

    # gradients ----------------------------------------------------------
    def loss(self, u_target, v_target) -> float:
        ut = np.ascontiguousarray(u_target, dtype=np.float64)
        vt = np.ascontiguousarray(v_target, dtype=np.float64)
        out = ctypes.c_double()
        _check(self.L, self.L.evp1d_loss(self.h, _dp(ut), _dp(vt), ctypes.byref(out)))
        return out.value

    def loss_noalloc(self, u_target, v_target) -> float:
        """evp1d_loss_noalloc: the same L without allocations (fixed device
        scratch, na <= 4,194,304, slot layout already up to date)."""
        ut = np.ascontiguousarray(u_target, dtype=np.float64)
        vt = np.ascontiguousarray(v_target, dtype=np.float64)
        out = ctypes.c_double()
        _check(self.L, self.L.evp1d_loss_noalloc(self.h, _dp(ut), _dp(vt), ctypes.byref(out)))
        return out.value

    def adjoint_run(self, nsub: int, u_target, v_target, nsnap: int = 0) -> dict:
        """Loss after nsub subcycles and its gradient w.r.t. strength, cdn_ocn
        (per point), forcex, forcey (evp1d_adjoint_run).  nsnap: stored
        states (>= nsub: the whole trajectory; 0: whatever fits)."""
        ut = np.ascontiguousarray(u_target, dtype=np.float64)
        vt = np.ascontiguousarray(v_target, dtype=np.float64)
        g = {k: np.zeros(self.na) for k in ("strength", "cdn_ocn", "forcex", "forcey")}
        loss = ctypes.c_double()
        _check(self.L, self.L.evp1d_adjoint_run(self.h, int(nsub), int(nsnap), _dp(ut), _dp(vt),
                                                ctypes.byref(loss), _dp(g["strength"]), _dp(g["cdn_ocn"]),
                                                _dp(g["forcex"]), _dp(g["forcey"])))
        g["loss"] = loss.value
        return g

    def tangent_run(self, nsub: int, u_target, v_target, d_strength=None, d_cdn_ocn=None,
                    d_forcex=None, d_forcey=None) -> tuple:
        """(loss, directional derivative) along the given perturbation."""
        ut = np.ascontiguousarray(u_target, dtype=np.float64)
        vt = np.ascontiguousarray(v_target, dtype=np.float64)
        dirs = []
        for a in (d_strength, d_cdn_ocn, d_forcex, d_forcey):
            dirs.append(None if a is None else np.ascontiguousarray(a, dtype=np.float64))
        ptr = [(_dp(a) if a is not None else ctypes.cast(None, _DP)) for a in dirs]
        l, dl = ctypes.c_double(), ctypes.c_double()
        _check(self.L, self.L.evp1d_tangent_run(self.h, int(nsub), *ptr, _dp(ut), _dp(vt), ctypes.byref(l),
                                                ctypes.byref(dl)))
        return l.value, dl.value

    def set_tuning(self, stress_block=256, stress_xcd=0, stepu_block=256, stepu_xcd=0) -> None:
        _check(self.L, self.L.evp1d_set_tuning(self.h, stress_block, stress_xcd, stepu_block, stepu_xcd))

    def get_tuning(self) -> dict:
        a = [ctypes.c_int32() for _ in range(4)]
        _check(self.L, self.L.evp1d_get_tuning(self.h, *[ctypes.byref(x) for x in a]))
        return dict(zip(("stress_block", "stress_xcd", "stepu_block", "stepu_xcd"), (x.value for x in a)))

    def set_grid(self, nx: int, ny: int, indxtij) -> None:
        ij = np.ascontiguousarray(indxtij, dtype=np.int32)
        _check(self.L, self.L.evp1d_set_grid(self.h, nx, ny, _ip(ij), len(ij)))
        self._grid = (nx, ny)

    def put_2d(self, fields: dict) -> None:
        """fields: name -> (nx, ny) array (any order; sent in Fortran order)."""
        names = list(fields)
        ids = np.array([FIELD_ID[n] for n in names], dtype=np.int32)
        buf = np.stack([np.asfortranarray(fields[n], dtype=np.float64).ravel(order="F") for n in names])
        _check(self.L, self.L.evp1d_put_2d(self.h, len(names), _ip(ids), _dp(buf)))

    def get_2d(self, names) -> dict:
        nx, ny = self._grid
        names = list(names)
        ids = np.array([FIELD_ID[n] for n in names], dtype=np.int32)
        buf = np.zeros((len(names), nx * ny))
        _check(self.L, self.L.evp1d_get_2d(self.h, len(names), _ip(ids), _dp(buf)))
        return {n: buf[k].reshape((nx, ny), order="F") for k, n in enumerate(names)}

    def set_tiling(self, tile_w: int = 64, tile_h: int = 16, block: int = 512, xcd: int = 1) -> None:
        _check(self.L, self.L.evp1d_set_tiling(self.h, tile_w, tile_h, block, xcd))

    def get_tiling(self) -> dict:
        a = [ctypes.c_int32() for _ in range(5)]
        _check(self.L, self.L.evp1d_get_tiling(self.h, *[ctypes.byref(x) for x in a]))
        return dict(zip(("tile_w", "tile_h", "block", "xcd", "ntiles"), (x.value for x in a)))

    def set_fused_inplace(self, inplace: bool) -> None:
        _check(self.L, self.L.evp1d_set_fused_inplace(self.h, 1 if inplace else 0))

    def get_fused_inplace(self) -> bool:
        v = ctypes.c_int32()
        _check(self.L, self.L.evp1d_get_fused_inplace(self.h, ctypes.byref(v)))
        return bool(v.value)

    def set_fused_depth(self, depth: int) -> None:
        """Subcycles per fused launch: 1, or 2 (the pair form; evp1d_set_fused_depth)."""
        _check(self.L, self.L.evp1d_set_fused_depth(self.h, int(depth)))

    def set_tile_order(self, order: int) -> None:
        """Tile visiting order of the fused forms: 0 rows of tiles, 1 Hilbert
        curve over the tiles (evp1d_set_tile_order)."""
        _check(self.L, self.L.evp1d_set_tile_order(self.h, int(order)))

    def get_tile_order(self) -> int:
        o = ctypes.c_int32()
        _check(self.L, self.L.evp1d_get_tile_order(self.h, ctypes.byref(o)))
        return o.value

    def set_adjoint_segment(self, seg: int) -> None:
        """Subcycles per checkpointing unit of adjoint_run with fewer stored
        states than subcycles (evp1d_set_adjoint_segment; 0 = automatic)."""
        _check(self.L, self.L.evp1d_set_adjoint_segment(self.h, int(seg)))

    def get_adjoint_segment(self) -> int:
        s = ctypes.c_int32()
        _check(self.L, self.L.evp1d_get_adjoint_segment(self.h, ctypes.byref(s)))
        return s.value

    def gradient_storage(self) -> int:
        """Bytes of device memory that gradient calls keep with the handle."""
        b = ctypes.c_int64()
        _check(self.L, self.L.evp1d_get_gradient_storage(self.h, ctypes.byref(b)))
        return b.value

    def release_gradient_storage(self) -> None:
        """Give that memory back (evp1d_release_gradient_storage)."""
        _check(self.L, self.L.evp1d_release_gradient_storage(self.h))

    def get_fused_depth(self) -> dict:
        """{"depth": setting, "active": runs use the pair form, "block": its workgroup size}."""
        d, a, b = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(self.L, self.L.evp1d_get_fused_depth(self.h, ctypes.byref(d), ctypes.byref(a), ctypes.byref(b)))
        return {"depth": d.value, "active": bool(a.value), "block": b.value}

    def set_cache_hints(self, hints: int) -> None:
        """Pair form's cache hints: -1 chosen by size (default), 0 off, 1 on."""
        _check(self.L, self.L.evp1d_set_cache_hints(self.h, int(hints)))

    def get_cache_hints(self) -> dict:
        """{"hints": setting, "active": runs use them}."""
        s, a = ctypes.c_int32(), ctypes.c_int32()
        _check(self.L, self.L.evp1d_get_cache_hints(self.h, ctypes.byref(s), ctypes.byref(a)))
        return {"hints": s.value, "active": bool(a.value)}

    def autotune_fused(self, nsub: int = 100) -> dict:
        """Time the fused tile ladder x stress storage and the pair form's tiles
        x cache hints (evp1d_autotune_fused), keep the fastest; returns it with
        its time per subcycle."""
        us = ctypes.c_float()
        _check(self.L, self.L.evp1d_autotune_fused(self.h, nsub, ctypes.byref(us)))
        return dict(self.get_tiling(), inplace=self.get_fused_inplace(), depth=self.get_fused_depth()["depth"],
                    order=self.get_tile_order(), hints=self.get_cache_hints()["hints"], us=us.value)

    # lifetime ------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.evp1d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # data movement --------------------------------------------------------
    def length(self, name: str) -> int:
        return self.navel if name in NAVEL_FIELDS else self.na

    def set_field(self, name: str, a) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(self.L, self.L.evp1d_set_field(self.h, FIELD_ID[name], _dp(a), a.size))

    def get_field(self, name: str) -> np.ndarray:
        out = np.empty(self.length(name))
        _check(self.L, self.L.evp1d_get_field(self.h, FIELD_ID[name], _dp(out), out.size))
        return out

    def set_index(self, name: str, idx) -> None:
        a = np.ascontiguousarray(idx, dtype=np.int32)
        _check(self.L, self.L.evp1d_set_index(self.h, NBR.index(name), _ip(a), a.size))

    def set_skip(self, which: str, flags) -> None:
        a = np.ascontiguousarray(flags, dtype=np.int32)
        _check(self.L, self.L.evp1d_set_skip(self.h, {"skipT": 0, "skipU": 1}[which], _ip(a), a.size))

    def set_halo(self, dst, src) -> None:
        d = np.ascontiguousarray(dst, dtype=np.int32)
        s = np.ascontiguousarray(src, dtype=np.int32)
        if d.size != s.size:
            raise ValueError("dst and src differ in length")
        _check(self.L, self.L.evp1d_set_halo(self.h, d.size, _ip(d), _ip(s)))

    def set_params(self, p: Params) -> None:
        _check(self.L, self.L.evp1d_set_params(self.h, ctypes.byref(p)))

    def get_params(self) -> Params:
        p = Params()
        _check(self.L, self.L.evp1d_get_params(self.h, ctypes.byref(p)))
        return p

    def load_state(self, fields: dict, idx: dict, skipT, skipU) -> None:
        """Upload a whole 1D state (dicts keyed by the FIELDS / NBR names)."""
        for n in NBR:
            self.set_index(n, idx[n])
        self.set_skip("skipT", skipT)
        self.set_skip("skipU", skipU)
        for n in FIELDS:
            if n in fields:
                self.set_field(n, fields[n])

    # compute --------------------------------------------------------------
    def set_mode(self, mode: str) -> None:
        _check(self.L, self.L.evp1d_set_mode(self.h, MODES[mode]))

    def stress(self) -> None:
        _check(self.L, self.L.evp1d_stress(self.h))

    def stepu(self) -> None:
        _check(self.L, self.L.evp1d_stepu(self.h))

    def halo(self) -> None:
        _check(self.L, self.L.evp1d_halo(self.h))

    def run(self, nsub: int) -> None:
        _check(self.L, self.L.evp1d_run(self.h, int(nsub)))

    def run_individual(self, which: str, nsub: int) -> None:
        _check(self.L, self.L.evp1d_run_individual(self.h, {"stress": 0, "stepu": 1, "fused": 2}[which],
                                                   int(nsub)))

    def synchronize(self) -> None:
        _check(self.L, self.L.evp1d_synchronize(self.h))

    def timer_start(self) -> None:
        _check(self.L, self.L.evp1d_timer_start(self.h))

    def timer_stop(self) -> float:
        ms = ctypes.c_float()
        _check(self.L, self.L.evp1d_timer_stop(self.h, ctypes.byref(ms)))
        return ms.value


# ---------------------------------------------------------------------------
# the v0 2D block path (evp2d_*)
# ---------------------------------------------------------------------------
# include/evp1d.h enum evp2d_field: input_double_2d.bin's order (src/vars.F90:
# 299-305), then uvel_init, vvel_init, taubx, tauby and the str1..8 scratch
FIELDS2D = (
    "strength", "uvel", "vvel", "dxt", "dyt", "dxhy", "dyhx", "cxp", "cyp", "cxm", "cym",
    "DminTarea", "uarear", "cdn_ocn", "aiX", "uocn", "vocn", "waterx", "watery", "forcex", "forcey",
    "umassdti", "fm", "strintx", "strinty", "Tbu",
    "stressp_1", "stressp_2", "stressp_3", "stressp_4", "stressm_1", "stressm_2", "stressm_3", "stressm_4",
    "stress12_1", "stress12_2", "stress12_3", "stress12_4",
    "uvel_init", "vvel_init", "taubx", "tauby",
    "str1", "str2", "str3", "str4", "str5", "str6", "str7", "str8",
)
FIELD2D_ID = {n: k for k, n in enumerate(FIELDS2D)}


class Engine2D:
    """One evp2d handle: the reference's v0 arrays (nx_block, ny_block,
    max_blocks) on the device (src/evp_v0.F90, src/bench_v0.F90)."""

    def __init__(self, nx_block: int, ny_block: int, max_blocks: int, params: Params, fma: bool = False):
        self.L = load(fma)
        self.shape = (nx_block, ny_block, max_blocks)
        self.n = nx_block * ny_block * max_blocks
        self.h = _H()
        _check(self.L, self.L.evp2d_create(nx_block, ny_block, max_blocks, ctypes.byref(params),
                                           ctypes.byref(self.h)))

    def set_params(self, params: Params) -> None:
        _check(self.L, self.L.evp2d_set_params(self.h, ctypes.byref(params)))

    def set_cells(self, iblk: int, ti, tj, ui, uj) -> None:
        """1-based block iblk's T cells (ti, tj) and U cells (ui, uj)."""
        ti, tj, ui, uj = (np.ascontiguousarray(a, dtype=np.int32) for a in (ti, tj, ui, uj))
        _check(self.L, self.L.evp2d_set_cells(self.h, iblk, ti.size, _ip(ti), _ip(tj), ui.size, _ip(ui), _ip(uj)))

    def set_halo(self, ilo, jlo, ihi, jhi, src, dst) -> None:
        """src, dst: (3, ncopies) arrays of (i, j, block), 1-based."""
        lo = [np.ascontiguousarray(a, dtype=np.int32) for a in (ilo, jlo, ihi, jhi)]
        src = np.asfortranarray(src, dtype=np.int32)
        dst = np.asfortranarray(dst, dtype=np.int32)
        s = np.ascontiguousarray(src.T)   # (ncopies, 3): i, j, block per copy
        d = np.ascontiguousarray(dst.T)
        _check(self.L, self.L.evp2d_set_halo(self.h, lo[0].size, *[_ip(a) for a in lo], s.shape[0], _ip(s), _ip(d)))

    def set_field(self, name: str, a) -> None:
        a = np.asfortranarray(a, dtype=np.float64).reshape(-1, order="F")
        if a.size != self.n:
            raise ValueError(f"{name}: {a.size} values, need {self.n}")
        a = np.ascontiguousarray(a)
        _check(self.L, self.L.evp2d_set_field(self.h, FIELD2D_ID[name], _dp(a), a.size))

    def get_field(self, name: str) -> np.ndarray:
        out = np.zeros(self.n)
        _check(self.L, self.L.evp2d_get_field(self.h, FIELD2D_ID[name], _dp(out), out.size))
        return out.reshape(self.shape, order="F")

    def load_state(self, st) -> None:
        """Fields, cell lists and halo tables of an oracle.State2D-like object
        (attributes f, icellt, icellu, indxti.., halo)."""
        for n in FIELDS2D[:42]:
            self.set_field(n, st.f[n])
        for b in range(self.shape[2]):
            nt, nu = int(st.icellt[b]), int(st.icellu[b])
            self.set_cells(b + 1, st.indxti[:nt, b], st.indxtj[:nt, b], st.indxui[:nu, b], st.indxuj[:nu, b])
        if st.halo is not None:
            h = st.halo
            self.set_halo(h["ilo"], h["jlo"], h["ihi"], h["jhi"], h["src"], h["dst"])

    def run(self, nsub: int) -> None:
        _check(self.L, self.L.evp2d_run(self.h, nsub))

    def set_mode(self, mode: str) -> None:
        """"fused" (default): one tile kernel per subcycle; "plain": stress then stepu."""
        _check(self.L, self.L.evp2d_set_mode(self.h, {"plain": 0, "fused": 1}[mode]))

    def get_mode(self) -> str:
        m = ctypes.c_int32()
        _check(self.L, self.L.evp2d_get_mode(self.h, ctypes.byref(m)))
        return ("plain", "fused")[m.value]

    def set_cache_hints(self, hints: int) -> None:
        """Fused form's cache hints: -1 chosen by size (default), 0 off, 1 on."""
        _check(self.L, self.L.evp2d_set_cache_hints(self.h, int(hints)))

    def get_cache_hints(self) -> dict:
        """{"hints": setting, "active": fused runs use them}."""
        s, a = ctypes.c_int32(), ctypes.c_int32()
        _check(self.L, self.L.evp2d_get_cache_hints(self.h, ctypes.byref(s), ctypes.byref(a)))
        return {"hints": s.value, "active": bool(a.value)}

    def stress(self) -> None:
        _check(self.L, self.L.evp2d_stress(self.h))

    def stepu(self) -> None:
        _check(self.L, self.L.evp2d_stepu(self.h))

    def halo(self) -> None:
        _check(self.L, self.L.evp2d_halo(self.h))

    def run_individual(self, which: str, nsub: int) -> None:
        _check(self.L, self.L.evp2d_run_individual(self.h, {"stress": 0, "stepu": 1, "fused": 2}[which], nsub))

    def synchronize(self) -> None:
        _check(self.L, self.L.evp2d_synchronize(self.h))

    def timer_start(self) -> None:
        _check(self.L, self.L.evp2d_timer_start(self.h))

    def timer_stop(self) -> float:
        ms = ctypes.c_float()
        _check(self.L, self.L.evp2d_timer_stop(self.h, ctypes.byref(ms)))
        return ms.value

    def cells(self) -> tuple[int, int]:
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _check(self.L, self.L.evp2d_get_cells(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.evp2d_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
