"""Oracle for the EVP1D subcycle loop (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; it is the checker, never the thing measured or shipped.

Contents
  OracleLib          ctypes wrapper of lib/liboracle.so (C restatement of
                     src/bench_v2b.F90 stress/stepu, see evp1d_oracle.c)
  index_reference()  restatement of the reference's 1D-list construction:
                     calc_nActiveTU, calc_2d_indices_init, union, calc_navel,
                     setdiff, findXinY, convert_2d_1d_init
                     (integration/ice_dyn_evp1d.F90:494-519, 559-633, 743-816,
                     951-1087), written as the reference's two-pointer loops
  pack_dyn()/unpack_dyn()  convert_2d_1d_dyn / convert_1d_2d_dyn
                     (integration/ice_dyn_evp1d.F90:818-947) and set_skipMe
                     (521-536)
  write_inputs()/read_outputs()/write_outputs()
                     the reference binary formats of readin_1d_v2 and
                     writeout_1d (src/vars.F90:2101-2135, 2233-2304) and the
                     kernel_nml namelist (src/vars.F90:24-25)
Pinning: see evp1d_oracle.c; the Python index restatement is pinned against
the compiled reference's own input/output through the golden fixtures
(tests/golden) and checked against hand-built tiny grids in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBDIR = HERE / "lib"
REFDIR = HERE / "_ref"

# order of include/evp1d.h enum evp1d_field
FIELDS = (
    "strength", "uvel", "vvel", "dxT", "dyT", "uarear", "cdn_ocn", "aiu", "uocn", "vocn",
    "waterxU", "wateryU", "forcexU", "forceyU", "umassdti", "fmU", "strintxU", "strintyU", "Tbu",
    "stressp_1", "stressp_2", "stressp_3", "stressp_4",
    "stressm_1", "stressm_2", "stressm_3", "stressm_4",
    "stress12_1", "stress12_2", "stress12_3", "stress12_4",
    "HTE", "HTN", "HTEm1", "HTNm1", "uvel_init", "vvel_init", "Cb",
    "str1", "str2", "str3", "str4", "str5", "str6", "str7", "str8",
)
NAVEL_FIELDS = {"uvel", "vvel", "str1", "str2", "str3", "str4", "str5", "str6", "str7", "str8"}
STRESS_FIELDS = FIELDS[19:31]
NBR = ("ee", "ne", "se", "nw", "sw", "sse")


class Params(ctypes.Structure):
    """Same layout as evp1d_params (include/evp1d.h)."""

    _fields_ = [("ndte", ctypes.c_int32), ("reserved", ctypes.c_int32)] + [
        (n, ctypes.c_double) for n in (
            "capping", "e_factor", "epp2i", "deltaminEVP", "arlx", "arlx1i", "brlx", "denom1",
            "revp", "Ktens", "u0", "cosw", "sinw", "rhow")]


_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


class OracleLib:
    """The C restatement; `omp=True` loads the OpenMP build."""

    def __init__(self, omp: bool = False):
        path = LIBDIR / ("liboracle_omp.so" if omp else "liboracle.so")
        if not path.exists():
            subprocess.run(["make", "-C", str(HERE), "port"], check=True, capture_output=True)
        self.lib = ctypes.CDLL(str(path))
        self.lib.oracle_set_evp_parameters.argtypes = [ctypes.c_int32, ctypes.POINTER(Params)]
        self.lib.oracle_run.argtypes = [ctypes.POINTER(Params), ctypes.c_int32, ctypes.c_int32,
                                        ctypes.POINTER(_DP), ctypes.POINTER(_IP), _IP, _IP,
                                        ctypes.c_int32, _IP, _IP]

    def stress(self, p: Params, st: "State1D") -> None:
        """One oracle_stress call over 1..na (bench_v2b.F90 stress)."""
        f = st.f
        sig = (_DP * 12)(*[_dp(f[n]) for n in STRESS_FIELDS])
        strp = (_DP * 8)(*[_dp(f[f"str{k}"]) for k in range(1, 9)])
        self.lib.oracle_stress(ctypes.byref(p), 1, st.na, _ip(st.idx["ee"]), _ip(st.idx["ne"]),
                               _ip(st.idx["se"]), _dp(f["uvel"]), _dp(f["vvel"]), _dp(f["dxT"]),
                               _dp(f["dyT"]), _ip(st.skipT), _dp(f["strength"]), _dp(f["HTE"]),
                               _dp(f["HTN"]), _dp(f["HTEm1"]), _dp(f["HTNm1"]), sig, strp)

    def stepu(self, p: Params, st: "State1D") -> None:
        """One oracle_stepu call over 1..na (bench_v2b.F90 stepu)."""
        f = st.f
        strp = (_DP * 8)(*[_dp(f[f"str{k}"]) for k in range(1, 9)])
        self.lib.oracle_stepu(ctypes.byref(p), 1, st.na, _dp(f["cdn_ocn"]), _dp(f["aiu"]),
                              _dp(f["uocn"]), _dp(f["vocn"]), _dp(f["waterxU"]), _dp(f["wateryU"]),
                              _dp(f["forcexU"]), _dp(f["forceyU"]), _dp(f["umassdti"]), _dp(f["fmU"]),
                              _dp(f["uarear"]), _dp(f["strintxU"]), _dp(f["strintyU"]),
                              _dp(f["uvel_init"]), _dp(f["vvel_init"]), _dp(f["uvel"]), _dp(f["vvel"]),
                              strp, _ip(st.idx["nw"]), _ip(st.idx["sw"]), _ip(st.idx["sse"]),
                              _ip(st.skipU), _dp(f["Tbu"]), _dp(f["Cb"]))

    def params(self, ndte: int, **override) -> Params:
        p = Params()
        self.lib.oracle_set_evp_parameters(ndte, ctypes.byref(p))
        for k, v in override.items():
            setattr(p, k, v)
        return p

    def run(self, p: Params, st: "State1D", nsub: int, halo=None) -> None:
        """nsub subcycles in place on st (stress; stepu; halo)."""
        for name in FIELDS:
            a = st.f[name]
            assert a.dtype == np.float64 and a.flags.c_contiguous, name
        fptr = (_DP * len(FIELDS))(*[_dp(st.f[n]) for n in FIELDS])
        iptr = (_IP * 6)(*[_ip(st.idx[n]) for n in NBR])
        if halo is None:
            nh, hd, hs = 0, np.zeros(1, np.int32), np.zeros(1, np.int32)
        else:
            hd = np.ascontiguousarray(halo[0], dtype=np.int32)
            hs = np.ascontiguousarray(halo[1], dtype=np.int32)
            nh = hd.size
        self.lib.oracle_run(ctypes.byref(p), st.na, nsub, fptr, iptr, _ip(st.skipT), _ip(st.skipU),
                            nh, _ip(hd), _ip(hs))


class State1D:
    """1D arrays of one domain (copy semantics: .copy() gives a deep copy)."""

    def __init__(self, na: int, navel: int):
        self.na, self.navel = na, navel
        self.f = {n: np.zeros(navel if n in NAVEL_FIELDS else na) for n in FIELDS}
        self.idx = {n: np.ones(na, dtype=np.int32) for n in NBR}
        self.skipT = np.zeros(na, dtype=np.int32)
        self.skipU = np.zeros(na, dtype=np.int32)
        self.scalars = dict(capping=1.0, e_factor=0.25, epp2i=0.25, deltaminEVP=1e-8)

    def copy(self) -> "State1D":
        s = State1D(self.na, self.navel)
        s.f = {k: v.copy() for k, v in self.f.items()}
        s.idx = {k: v.copy() for k, v in self.idx.items()}
        s.skipT, s.skipU = self.skipT.copy(), self.skipU.copy()
        s.scalars = dict(self.scalars)
        return s


# ---------------------------------------------------------------------------
# 1D list construction, restated from integration/ice_dyn_evp1d.F90
# ---------------------------------------------------------------------------
def _union(x, y):
    """union (ice_dyn_evp1d.F90:591-633): merge of two sorted int vectors."""
    i = j = 0
    out = []
    while i < len(x) and j < len(y):
        if x[i] < y[j]:
            out.append(x[i]); i += 1
        elif x[i] > y[j]:
            out.append(y[j]); j += 1
        else:
            out.append(x[i]); i += 1; j += 1
    out.extend(x[i:])
    out.extend(y[j:])
    return out


def _setdiff(x, y):
    """setdiff (ice_dyn_evp1d.F90:951-998): symmetric difference of sorted vectors."""
    i = j = 0
    out = []
    while i < len(x) and j < len(y):
        if x[i] < y[j]:
            out.append(x[i]); i += 1
        elif x[i] > y[j]:
            out.append(y[j]); j += 1
        else:
            i += 1; j += 1
    out.extend(x[i:])
    out.extend(y[j:])
    return out


def _find_x_in_y(x, y, nx_len):
    """findXinY (ice_dyn_evp1d.F90:1000-1045), 1-based result."""
    indx = [0] * len(x)
    i, j1, j2 = 0, 0, nx_len
    ny_len = len(y)
    while i < len(x):
        y1 = y[j1] if j1 < ny_len else None
        y2 = y[j2] if j2 < ny_len else None
        if x[i] == y1:
            indx[i] = j1 + 1; i += 1; j1 += 1
        elif x[i] == y2:
            indx[i] = j2 + 1; i += 1; j2 += 1
        elif y1 is not None and x[i] > y1:
            j1 += 1
        elif y2 is not None and x[i] > y2:
            j2 += 1
        else:
            raise RuntimeError("findXinY: conditions not met")  # the reference stops here
    return indx


def index_reference(tmask: np.ndarray, nghost: int) -> dict:
    """1D list of a (nx, ny) Fortran-ordered T mask (nx, ny include the rim)."""
    nx, ny = tmask.shape
    na = int(np.count_nonzero(tmask))                               # calc_nActiveTU
    ti, tj = [], []
    for j in range(1 + nghost, ny + 1):                             # calc_2d_indices_init
        for i in range(1 + nghost, nx + 1):
            if tmask[i - 1, j - 1]:
                ti.append(i); tj.append(j)
    if len(ti) != na:
        raise ValueError("wet cells in the west/south rim")
    Iin = [i + (j - 1) * nx for i, j in zip(ti, tj)]
    Iee = [i - 1 + (j - 1) * nx for i, j in zip(ti, tj)]
    Ine = [i - 1 + (j - 2) * nx for i, j in zip(ti, tj)]
    Ise = [i + (j - 2) * nx for i, j in zip(ti, tj)]
    Inw = [i + 1 + (j - 1) * nx for i, j in zip(ti, tj)]
    Isw = [i + 1 + (j - 0) * nx for i, j in zip(ti, tj)]
    Isse = [i + (j - 0) * nx for i, j in zip(ti, tj)]
    u = _union(Iin, Iee)
    u = _union(u, Ine)
    u = _union(u, Ise)
    u = _union(u, Inw)
    u = _union(u, Isw)
    u = _union(u, Isse)
    navel = len(u)                                                  # calc_navel
    ghosts = _setdiff(u, Iin)
    indxtij = Iin + ghosts
    out = dict(na=na, navel=navel, indxti=np.array(ti, np.int32), indxtj=np.array(tj, np.int32),
               indxtij=np.array(indxtij, np.int64))
    for name, lst in zip(NBR, (Iee, Ine, Ise, Inw, Isw, Isse)):
        out[name] = np.array(_find_x_in_y(lst, indxtij, na), dtype=np.int32)
    return out


_DYN_2D = (
    "stressp_1", "stressp_2", "stressp_3", "stressp_4", "stressm_1", "stressm_2", "stressm_3",
    "stressm_4", "stress12_1", "stress12_2", "stress12_3", "stress12_4", "strength", "cdn_ocn",
    "aiu", "uocn", "vocn", "waterxU", "wateryU", "forcexU", "forceyU", "umassdti", "fmU",
    "strintxU", "strintyU", "Tbu", "Cb", "uvel", "vvel")


def pack_dyn(grid, ix: dict) -> State1D:
    """State1D from a synth.Grid2D: convert_2d_1d_init (803-815) for the static
    fields, convert_2d_1d_dyn (845-881) for the dynamic ones, set_skipMe
    (521-536) for the flags; uvel_init = uvel as readin_1d_v2 does."""
    na, navel = ix["na"], ix["navel"]
    st = State1D(na, navel)
    i = ix["indxti"].astype(np.int64) - 1
    j = ix["indxtj"].astype(np.int64) - 1
    F = grid.fields
    for n in _DYN_2D:
        st.f[n][:na] = F[n][i, j]
    st.f["uvel"][na:] = 0.0
    st.f["vvel"][na:] = 0.0
    st.f["uarear"][:] = F["uarear"][i, j]
    st.f["dxT"][:] = F["dxT"][i, j]
    st.f["dyT"][:] = F["dyT"][i, j]
    st.f["HTE"][:] = F["HTE"][i, j]
    st.f["HTN"][:] = F["HTN"][i, j]
    st.f["HTEm1"][:] = F["HTE"][i - 1, j]
    st.f["HTNm1"][:] = F["HTN"][i, j - 1]
    st.f["uvel_init"][:] = st.f["uvel"][:na]
    st.f["vvel_init"][:] = st.f["vvel"][:na]
    st.skipT[:] = (~grid.icetmask[i, j]).astype(np.int32)
    st.skipU[:] = (~grid.iceumask[i, j]).astype(np.int32)
    for n in NBR:
        st.idx[n][:] = ix[n]
    return st


def unpack_dyn(st: State1D, ix: dict, shape) -> dict:
    """convert_1d_2d_dyn (ice_dyn_evp1d.F90:884-947): 1D -> 2D, zero elsewhere."""
    i = ix["indxti"].astype(np.int64) - 1
    j = ix["indxtj"].astype(np.int64) - 1
    out = {}
    for n in _DYN_2D:
        a = np.zeros(shape, order="F")
        a[i, j] = st.f[n][:st.na]
        out[n] = a
    return out


# ---------------------------------------------------------------------------
# reference binary formats (src/vars.F90)
# ---------------------------------------------------------------------------
_IN_ORDER = ("strength", "uvel", "vvel", "dxT", "dyT", "uarear", "cdn_ocn", "aiu", "uocn", "vocn",
             "waterxU", "wateryU", "forcexU", "forceyU", "umassdti", "fmU", "strintxU", "strintyU",
             "Tbu") + STRESS_FIELDS


def input_chunks(st: State1D) -> dict:
    """The byte chunks of input_double_1d_v2.bin, input_integer_1d.bin and
    input_logical_1d.bin, in file order (readin_1d_v2, vars.F90:2101-2135).
    Logicals are 4-byte 0/1."""
    na, navel = st.na, st.navel

    def doubles():
        for n in _IN_ORDER:
            a = st.f[n]
            yield np.ascontiguousarray(a[:navel] if n in ("uvel", "vvel") else a[:na]).tobytes()
        s = st.scalars
        yield np.array([s["capping"], s["e_factor"], s["epp2i"], s["deltaminEVP"]]).tobytes()
        for n in ("HTE", "HTN", "HTEm1", "HTNm1"):
            yield np.ascontiguousarray(st.f[n][:na]).tobytes()

    def integers():
        for n in NBR:
            yield st.idx[n].astype("<i4").tobytes()

    def logicals():
        yield st.skipU.astype("<i4").tobytes()
        yield st.skipT.astype("<i4").tobytes()

    return {"input_double_1d_v2.bin": doubles, "input_integer_1d.bin": integers,
            "input_logical_1d.bin": logicals}


def input_md5s(st: State1D) -> dict:
    """md5 of each input file write_inputs would write, computed in memory."""
    out = {}
    for name, gen in input_chunks(st).items():
        h = hashlib.md5()
        for b in gen():
            h.update(b)
        out[name] = h.hexdigest()
    return out


def write_inputs(d: Path, st: State1D, ndte: int, nx_block: int = -1, ny_block: int = -1) -> None:
    """input_double_1d_v2.bin, input_integer_1d.bin, input_logical_1d.bin and
    kernel.nml as readin_1d_v2 / read_nml consume them (vars.F90:2101-2135,
    19-69)."""
    d = Path(d)
    d.mkdir(parents=True, exist_ok=True)
    na, navel = st.na, st.navel
    for name, gen in input_chunks(st).items():
        with open(d / name, "wb") as fh:
            for b in gen():
                fh.write(b)
    nb = int(na - np.count_nonzero(st.skipU))
    (d / "kernel.nml").write_text(
        "&kernel_nml\n"
        f" ndte = {ndte}\n testscale = 1\n binoutput = .true.\n lindividual = .false.\n"
        f" nx_block = {nx_block}\n ny_block = {ny_block}\n na = {na}\n nb = {nb}\n navel = {navel}\n/\n")


def read_inputs(d: Path, na: int, navel: int) -> State1D:
    """Inverse of write_inputs (same layout as readin_1d_v2)."""
    d = Path(d)
    st = State1D(na, navel)
    raw = np.fromfile(d / "input_double_1d_v2.bin", dtype="<f8")
    need = 2 * navel + (len(_IN_ORDER) - 2) * na + 4 + 4 * na
    if raw.size != need:
        raise ValueError(f"input_double_1d_v2.bin has {raw.size} doubles, expected {need}")
    p = 0
    for n in _IN_ORDER:
        m = navel if n in ("uvel", "vvel") else na
        st.f[n][:m] = raw[p:p + m]
        p += m
    st.scalars = dict(zip(("capping", "e_factor", "epp2i", "deltaminEVP"), raw[p:p + 4].tolist()))
    p += 4
    for n in ("HTE", "HTN", "HTEm1", "HTNm1"):
        st.f[n][:] = raw[p:p + na]
        p += na
    st.f["uvel_init"][:] = st.f["uvel"][:na]
    st.f["vvel_init"][:] = st.f["vvel"][:na]
    ints = np.fromfile(d / "input_integer_1d.bin", dtype="<i4").reshape(6, na)
    for k, n in enumerate(NBR):
        st.idx[n][:] = ints[k]
    logi = np.fromfile(d / "input_logical_1d.bin", dtype="<i4").reshape(2, na)
    st.skipU[:] = (logi[0] != 0)
    st.skipT[:] = (logi[1] != 0)
    return st


def output_bytes(st: State1D) -> tuple[bytes, bytes]:
    """Contents of output_stress_*_1d.bin and output_stepu_*_1d.bin
    (writeout_1d, vars.F90:2233-2304): the 12 stresses over 1..na, then
    strintx, strinty, uvel, vvel compacted over .not.skipU."""
    na = st.na
    stress = b"".join(np.ascontiguousarray(st.f[n][:na]).tobytes() for n in STRESS_FIELDS)
    keep = st.skipU == 0
    stepu = b"".join(np.ascontiguousarray(st.f[n][:na][keep]).tobytes()
                     for n in ("strintxU", "strintyU", "uvel", "vvel"))
    return stress, stepu


def md5(b: bytes) -> str:
    return hashlib.md5(b).hexdigest()


def run_reference(d: Path, binary: str = "evp_v2b_f0", threads: int = 1, timeout: float = 3600,
                  env_extra: dict | None = None) -> subprocess.CompletedProcess:
    """Run the compiled reference driver in directory d (inputs must exist).

    Writes output_stress_v2b_1d.bin / output_stepu_v2b_1d.bin there."""
    exe = REFDIR / binary
    if not exe.exists():
        raise FileNotFoundError(f"{exe} not built (make -C oracle ref)")
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    if env_extra:
        env.update(env_extra)
    return subprocess.run([str(exe)], cwd=str(d), env=env, capture_output=True, text=True,
                          timeout=timeout, check=True)


# ---------------------------------------------------------------------------
# v0 2D block path (src/bench_v0.F90, src/evp_v0.F90, src/vars.F90)
# ---------------------------------------------------------------------------
# Arrays of input_double_2d.bin in readin_2d's order (vars.F90:299-305), then
# the fields the driver derives or writes itself.
V0_FILE_FIELDS = (
    "strength", "uvel", "vvel", "dxt", "dyt", "dxhy", "dyhx", "cxp", "cyp", "cxm", "cym",
    "DminTarea", "uarear", "cdn_ocn", "aiX", "uocn", "vocn", "waterx", "watery", "forcex", "forcey",
    "umassdti", "fm", "strintx", "strinty", "Tbu") + STRESS_FIELDS
V0_FIELDS = V0_FILE_FIELDS + ("uvel_init", "vvel_init", "taubx", "tauby")


class State2D:
    """The v0 arrays: each field (nx_block, ny_block, nblocks) in Fortran
    order; cell lists (nx_block*ny_block, nblocks), 1-based, zero padded;
    optional halo description (readin_halo, vars.F90:80-99)."""

    def __init__(self, nxb: int, nyb: int, nblocks: int):
        self.nxb, self.nyb, self.nblocks = nxb, nyb, nblocks
        shape = (nxb, nyb, nblocks)
        self.f = {n: np.zeros(shape, order="F") for n in V0_FIELDS}
        m = (nxb * nyb, nblocks)
        self.indxti, self.indxtj, self.indxui, self.indxuj = (np.zeros(m, np.int32, order="F")
                                                              for _ in range(4))
        self.icellt = np.zeros(nblocks, np.int32)
        self.icellu = np.zeros(nblocks, np.int32)
        self.navel = 0
        self.scalars = dict(capping=1.0, e_factor=0.25, epp2i=0.25)
        self.halo = None   # dict(ilo, jlo, ihi, jhi, src (3, n), dst (3, n))

    def copy(self) -> "State2D":
        s = State2D(self.nxb, self.nyb, self.nblocks)
        s.f = {k: v.copy(order="F") for k, v in self.f.items()}
        for n in ("indxti", "indxtj", "indxui", "indxuj", "icellt", "icellu"):
            setattr(s, n, getattr(self, n).copy(order="F"))
        s.navel, s.scalars = self.navel, dict(self.scalars)
        s.halo = None if self.halo is None else {k: v.copy() for k, v in self.halo.items()}
        return s


def _metrics_v0(F: dict, deltaminEVP: float) -> dict:
    """dxhy, dyhx, cxp, cyp, cxm, cym, DminTarea on the 2D grid with the same
    expressions and association as the v2 kernel computes them per point
    from HTE, HTN and their i-1 / j-1 neighbours (bench_v2b.F90:139-147);
    zero on the first row and column, which have no such neighbours."""
    hte, htn = F["HTE"], F["HTN"]
    htem1 = np.zeros_like(hte)
    htem1[1:, :] = hte[:-1, :]
    htnm1 = np.zeros_like(htn)
    htnm1[:, 1:] = htn[:, :-1]
    out = dict(
        cxp=1.5 * htn - 0.5 * htnm1, cyp=1.5 * hte - 0.5 * htem1,
        cxm=-(1.5 * htnm1 - 0.5 * htn), cym=-(1.5 * htem1 - 0.5 * hte),
        dxhy=0.5 * (hte - htem1), dyhx=0.5 * (htn - htnm1),
        DminTarea=deltaminEVP * (F["dxT"] * F["dyT"]))
    for v in out.values():
        v[0, :] = 0.0
        v[:, 0] = 0.0
    return out


_V0_FROM_GRID = dict(strength="strength", uvel="uvel", vvel="vvel", dxt="dxT", dyt="dyT", uarear="uarear",
                     cdn_ocn="cdn_ocn", aiX="aiu", uocn="uocn", vocn="vocn", waterx="waterxU",
                     watery="wateryU", forcex="forcexU", forcey="forceyU", umassdti="umassdti", fm="fmU",
                     strintx="strintxU", strinty="strintyU", Tbu="Tbu")


def pack_v0(grid, nblocks: int = 1, deltaminEVP: float = 1e-8) -> State2D:
    """v0 arrays for a synth.Grid2D (nghost = 1).

    nblocks = 1: one block, the whole grid, lists over the interior in CICE
    order (j outer, i inner) of icetmask / iceumask cells.  nblocks > 1: the
    interior is split in x into nearly equal column bands; block b holds its
    band plus a one-cell halo, right-padded to the widest band.  Its T list
    spans its band and its east halo column (stepu at the band's east edge
    reads str there), its U list its band; the halo copy list fills each
    interior-facing halo column from the neighbour band after every subcycle
    and the outer ones from nothing (fill)."""
    if grid.nghost != 1:
        raise ValueError("pack_v0 needs nghost = 1")
    F = dict(grid.fields)
    F.update(_metrics_v0(grid.fields, deltaminEVP))
    NX, NY, nx = grid.NX, grid.NY, grid.nx
    edges = np.linspace(0, nx, nblocks + 1).round().astype(int)   # interior columns [e_b, e_b+1)
    widths = np.diff(edges)
    if widths.min() < 1:
        raise ValueError("more blocks than interior columns")
    nxb = int(widths.max()) + 2
    st = State2D(nxb, NY, nblocks)
    for b in range(nblocks):
        g0 = edges[b]                       # global 0-based column of block column 0 (the west halo)
        w = int(widths[b])
        cols = np.arange(g0, g0 + w + 2)    # west halo .. east halo, global 0-based
        for n in V0_FILE_FIELDS:
            src = F[_V0_FROM_GRID.get(n, n)] if n not in STRESS_FIELDS else F[n]
            st.f[n][:w + 2, :, b] = src[cols, :]
        ihi = w + 1                          # 1-based interior [2, ihi]
        t_hi = ihi + 1 if nblocks > 1 and b < nblocks - 1 else ihi

        def cells(mask, i_hi):
            # CICE order: j outer, i inner, over i in [2, i_hi], j in [2, NY-1]
            sub = mask[cols[1:i_hi], 1:NY - 1]           # block columns 2..i_hi, rows 2..NY-1
            jj, ii = np.nonzero(sub.T)                   # row-major in (j, i)
            return (ii + 2).astype(np.int32), (jj + 2).astype(np.int32)
        ti, tj = cells(grid.icetmask, t_hi)
        ui, uj = cells(grid.iceumask, ihi)
        st.icellt[b], st.icellu[b] = len(ti), len(ui)
        st.indxti[:len(ti), b], st.indxtj[:len(tj), b] = ti, tj
        st.indxui[:len(ui), b], st.indxuj[:len(uj), b] = ui, uj
    st.f["uvel_init"][:] = st.f["uvel"]
    st.f["vvel_init"][:] = st.f["vvel"]
    st.navel = int(np.count_nonzero(grid.tmask))
    if nblocks > 1:
        ilo = np.full(nblocks, 2, np.int32)
        jlo = np.full(nblocks, 2, np.int32)
        ihi = (widths + 1).astype(np.int32)
        jhi = np.full(nblocks, NY - 1, np.int32)
        src, dst = [], []
        for b in range(nblocks):
            for j in range(2, NY):
                # west halo column 1 <- east interior column of block b-1
                src.append((int(ihi[b - 1]), j, b) if b > 0 else (0, 0, 0))
                dst.append((1, j, b + 1))
                # east halo column ihi+1 <- west interior column of block b+1
                src.append((2, j, b + 2) if b < nblocks - 1 else (0, 0, 0))
                dst.append((int(ihi[b]) + 1, j, b + 1))
        st.halo = dict(ilo=ilo, jlo=jlo, ihi=ihi, jhi=jhi,
                       src=np.asfortranarray(np.array(src, np.int32).T),
                       dst=np.asfortranarray(np.array(dst, np.int32).T))
    return st


def v0_to_global(st: State2D, grid, name: str) -> np.ndarray:
    """The interior of every block of one field, put back on the (NX, NY) grid
    (rim from block data where the single-block layout has it)."""
    if st.nblocks == 1:
        return st.f[name][:, :, 0].copy(order="F")
    edges = np.linspace(0, grid.nx, st.nblocks + 1).round().astype(int)
    out = np.zeros((grid.NX, grid.NY), order="F")
    for b in range(st.nblocks):
        w = int(edges[b + 1] - edges[b])
        out[edges[b] + 1:edges[b] + 1 + w, :] = st.f[name][1:w + 1, :, b]
    return out


class OracleV0:
    """ctypes wrapper of the v0 restatement (evp2d_oracle.c) in liboracle.so."""

    def __init__(self):
        self.o = OracleLib()
        self.lib = self.o.lib

    def params(self, ndte: int, **override) -> Params:
        return self.o.params(ndte, **override)

    def run(self, p: Params, st: State2D, nsub: int) -> None:
        for n in V0_FIELDS:
            a = st.f[n]
            assert a.dtype == np.float64 and a.flags.f_contiguous, n
        fptr = (_DP * len(V0_FIELDS))(*[_dp(st.f[n]) for n in V0_FIELDS])
        scratch = np.zeros(8 * st.nxb * st.nyb)
        lists = [np.asfortranarray(getattr(st, n), dtype=np.int32) for n in ("indxti", "indxtj", "indxui", "indxuj")]
        if st.halo is None:
            nhb, lo = 0, [np.zeros(1, np.int32)] * 4
            nc, src, dst = 0, np.zeros(3, np.int32), np.zeros(3, np.int32)
        else:
            h = st.halo
            nhb = len(h["ilo"])
            lo = [np.ascontiguousarray(h[k], dtype=np.int32) for k in ("ilo", "jlo", "ihi", "jhi")]
            nc = h["src"].shape[1]
            src = np.asfortranarray(h["src"], dtype=np.int32)
            dst = np.asfortranarray(h["dst"], dtype=np.int32)
        self.lib.oracle_v0_run(ctypes.byref(p), st.nxb, st.nyb, st.nblocks, nsub, fptr,
                               _ip(np.ascontiguousarray(st.icellt)), _ip(np.ascontiguousarray(st.icellu)),
                               *[_ip(a) for a in lists], _dp(scratch), nhb, *[_ip(a) for a in lo], nc,
                               _ip(src), _ip(dst))


def params_v0(lib_or_params, st: State2D) -> Params:
    """set_evp_parameters then readin_2d's three file scalars (vars.F90:305)."""
    p = lib_or_params
    for k in ("capping", "e_factor", "epp2i"):
        setattr(p, k, st.scalars[k])
    return p


def write_inputs_v0(d: Path, st: State2D, ndte: int) -> None:
    """input_double_2d.bin, input_integer_2d.bin (readin_2d, vars.F90:295-316),
    with max_blocks > 1 the two halo files of readin_halo (vars.F90:80-99),
    and kernel.nml (vars.F90:19-69)."""
    d = Path(d)
    d.mkdir(parents=True, exist_ok=True)
    with open(d / "input_double_2d.bin", "wb") as fh:
        for n in V0_FILE_FIELDS:
            fh.write(np.asfortranarray(st.f[n]).tobytes(order="F"))
        s = st.scalars
        fh.write(np.array([s["capping"], s["e_factor"], s["epp2i"]]).tobytes())
    with open(d / "input_integer_2d.bin", "wb") as fh:
        for n in ("indxti", "indxtj", "indxui", "indxuj"):
            fh.write(np.asfortranarray(getattr(st, n), dtype="<i4").tobytes(order="F"))
        fh.write(st.icellt.astype("<i4").tobytes())
        fh.write(st.icellu.astype("<i4").tobytes())
        fh.write(np.array([st.navel], "<i4").tobytes())
    if st.nblocks > 1:
        h = st.halo
        with open(d / "base_output_integer_blockindices_ref_v0_2d.bin", "wb") as fh:
            fh.write(np.array([len(h["ilo"])], "<i4").tobytes())
            for k in ("ilo", "jlo", "ihi", "jhi"):
                fh.write(np.asarray(h[k], "<i4").tobytes())
        with open(d / "base_output_integer_halo_ref_v0_2d.bin", "wb") as fh:
            fh.write(np.array([h["src"].shape[1]], "<i4").tobytes())
            fh.write(np.asfortranarray(h["src"], dtype="<i4").tobytes(order="F"))
            fh.write(np.asfortranarray(h["dst"], dtype="<i4").tobytes(order="F"))
    (d / "kernel.nml").write_text(
        "&kernel_nml\n"
        f" ndte = {ndte}\n testscale = 1\n binoutput = .true.\n lindividual = .false.\n"
        f" nx_block = {st.nxb}\n ny_block = {st.nyb}\n max_blocks = {st.nblocks}\n nghost = 1\n"
        f" na = {int(st.icellt[0])}\n nb = {int(st.icellu[0])}\n navel = {st.navel}\n/\n")


def read_inputs_v0(d: Path, nxb: int, nyb: int, nblocks: int) -> State2D:
    """Inverse of write_inputs_v0 (halo files included when nblocks > 1)."""
    d = Path(d)
    st = State2D(nxb, nyb, nblocks)
    m = nxb * nyb * nblocks
    raw = np.fromfile(d / "input_double_2d.bin", dtype="<f8")
    if raw.size != len(V0_FILE_FIELDS) * m + 3:
        raise ValueError(f"input_double_2d.bin has {raw.size} doubles, expected {len(V0_FILE_FIELDS) * m + 3}")
    for k, n in enumerate(V0_FILE_FIELDS):
        st.f[n][:] = raw[k * m:(k + 1) * m].reshape((nxb, nyb, nblocks), order="F")
    st.scalars = dict(zip(("capping", "e_factor", "epp2i"), raw[-3:].tolist()))
    st.f["uvel_init"][:] = st.f["uvel"]
    st.f["vvel_init"][:] = st.f["vvel"]
    ints = np.fromfile(d / "input_integer_2d.bin", dtype="<i4")
    L = nxb * nyb * nblocks
    for k, n in enumerate(("indxti", "indxtj", "indxui", "indxuj")):
        setattr(st, n, np.asfortranarray(ints[k * L:(k + 1) * L].reshape((nxb * nyb, nblocks), order="F")))
    p = 4 * L
    st.icellt = ints[p:p + nblocks].copy()
    st.icellu = ints[p + nblocks:p + 2 * nblocks].copy()
    st.navel = int(ints[p + 2 * nblocks])
    if nblocks > 1:
        bi = np.fromfile(d / "base_output_integer_blockindices_ref_v0_2d.bin", dtype="<i4")
        nb = int(bi[0])
        lo = bi[1:].reshape(4, nb)
        hh = np.fromfile(d / "base_output_integer_halo_ref_v0_2d.bin", dtype="<i4")
        nc = int(hh[0])
        st.halo = dict(ilo=lo[0].copy(), jlo=lo[1].copy(), ihi=lo[2].copy(), jhi=lo[3].copy(),
                       src=np.asfortranarray(hh[1:1 + 3 * nc].reshape((3, nc), order="F")),
                       dst=np.asfortranarray(hh[1 + 3 * nc:1 + 6 * nc].reshape((3, nc), order="F")))
    return st


def output_bytes_v0(st: State2D) -> tuple[bytes, bytes]:
    """Contents of output_stress_v0_2d.bin and output_stepu_v0_2d.bin
    (writeout_2d, vars.F90:317-334): whole arrays, all blocks."""
    stress = b"".join(np.asfortranarray(st.f[n]).tobytes(order="F") for n in STRESS_FIELDS)
    stepu = b"".join(np.asfortranarray(st.f[n]).tobytes(order="F") for n in ("strintx", "strinty", "uvel", "vvel"))
    return stress, stepu
