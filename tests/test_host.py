"""CPU tests of the host side: C ABI exports, the 1D index builder, the
synthetic generator, the reference file formats and the 2D<->1D packing."""
import ctypes
import re
import tempfile
from pathlib import Path

import numpy as np
import pytest

import evp1d
from evp1d import layout, synth
from oracle import oracle as O

REPO = Path(__file__).resolve().parent.parent


def _declared_symbols():
    text = (REPO / "include" / "evp1d.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(evp[12]d_\w+)\s*\(", text)))


@pytest.mark.parametrize("fma", [False, True])
def test_library_exports_every_declared_symbol(fma):
    L = ctypes.CDLL(str(evp1d.lib_path(fma)))
    declared = _declared_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared) == set(evp1d.EXPORTS)


def test_gradient_storage_calls_reject_a_null_handle():
    """The storage query and release check their handle before any device
    call (EVP1D_EINVAL = 1)."""
    L = ctypes.CDLL(str(evp1d.lib_path(False)))
    L.evp1d_get_gradient_storage.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
    L.evp1d_release_gradient_storage.argtypes = [ctypes.c_void_p]
    b = ctypes.c_int64(-1)
    assert L.evp1d_get_gradient_storage(None, ctypes.byref(b)) == 1
    assert b.value == -1
    assert L.evp1d_release_gradient_storage(None) == 1


def test_stream_probe_library_rejects_unsupported_mixes():
    """The HBM streaming probe (bench.py --stream-probe) is built beside the
    library and refuses a read/write mix it has no kernel for before any
    HIP call (so this runs without a GPU)."""
    L = ctypes.CDLL(str(REPO / "cice-evp1d_amd" / "lib" / "libevp1d_probe.so"))
    f = L.evp1d_probe_stream
    f.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                  ctypes.POINTER(ctypes.c_double)]
    f.restype = ctypes.c_int
    us = ctypes.c_double()
    for nr, nw in ((0, 1), (7, 1), (1, 2), (3, 3)):
        assert f(nr, nw, 1, 1 << 20, 1, ctypes.byref(us)) == -1, (nr, nw)
    for per in (0, 2, 8):
        assert f(1, 1, per, 1 << 20, 1, ctypes.byref(us)) == -1, per
    assert f(1, 1, 4, (1 << 20) + 256, 1, ctypes.byref(us)) == -1   # not a multiple of 256 x 4
    assert f(1, 1, 1, 0, 1, ctypes.byref(us)) == -1
    assert f(1, 1, 1, 1 << 20, 0, ctypes.byref(us)) == -1
    g = L.evp1d_probe_stream_hints
    g.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                  ctypes.POINTER(ctypes.c_double)]
    g.restype = ctypes.c_int
    for hints in (-1, 2):
        assert g(1, 1, 1, hints, 1 << 20, 1, ctypes.byref(us)) == -1, hints
    assert g(7, 1, 1, 1, 1 << 20, 1, ctypes.byref(us)) == -1


def _c_prototypes():
    """name -> number of parameters of every function include/evp1d.h declares."""
    text = (REPO / "include" / "evp1d.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for name, args in re.findall(r"\b(evp[12]d_\w+)\s*\(([^;{]*?)\)\s*;", text):
        a = args.strip()
        out[name] = 0 if a in ("", "void") else a.count(",") + 1
    return out


def test_fortran_binding_matches_the_c_abi():
    """Every bind(C) interface of the Fortran C-ABI module names a symbol the
    header declares and the library exports, with as many arguments as the
    C prototype."""
    src = (REPO / "cice-evp1d_amd" / "fortran" / "evp1d_capi.F90").read_text()
    src = re.sub(r"&\s*\n\s*", " ", src)            # join continuation lines
    binds = re.findall(r"(?:function|subroutine)\s+\w+\s*\(([^)]*)\)\s*bind\s*\(\s*C\s*,\s*name\s*=\s*'(\w+)'",
                       src, flags=re.I)
    assert len(binds) >= 40
    protos = _c_prototypes()
    L = ctypes.CDLL(str(evp1d.lib_path(False)))
    for args, name in binds:
        assert name in protos, name
        assert hasattr(L, name), name
        n = 0 if not args.strip() else args.count(",") + 1
        assert n == protos[name], (name, n, protos[name])
    for name in ("evp1d_adjoint_run", "evp1d_tangent_run", "evp1d_set_adjoint_segment", "evp1d_run"):
        assert name in {b[1] for b in binds}, name


def test_build_info_flavours():
    assert evp1d.build_info(False) == {"abi": 1, "fma": 0}
    assert evp1d.build_info(True) == {"abi": 1, "fma": 1}


def test_c_abi_set_evp_parameters_matches_oracle():
    lib = O.OracleLib()
    for ndte in (1, 5, 120, 1000):
        a = evp1d.set_evp_parameters(ndte)
        b = lib.params(ndte)
        for name, _ in evp1d.Params._fields_:
            assert getattr(a, name) == getattr(b, name), (ndte, name)
    with pytest.raises(evp1d.Evp1dError):
        evp1d.set_evp_parameters(0)


def _mask(nx, ny, seed, frac):
    rng = np.random.default_rng(seed)
    m = np.zeros((nx, ny), dtype=bool, order="F")
    m[1:-1, 1:-1] = rng.random((nx - 2, ny - 2)) < frac
    return m


@pytest.mark.parametrize("nx,ny,seed,frac", [
    (8, 6, 1, 0.7), (30, 20, 2, 0.5), (50, 3, 3, 0.9), (3, 40, 4, 1.0), (25, 25, 5, 0.05)])
def test_index_builder_matches_reference_restatement(nx, ny, seed, frac):
    m = _mask(nx, ny, seed, frac)
    ref = O.index_reference(m, 1)
    got = evp1d.build_index(m, 1)
    assert (got["na"], got["navel"]) == (ref["na"], ref["navel"])
    for k in ("indxti", "indxtj", "indxtij", *O.NBR):
        assert np.array_equal(np.asarray(got[k], np.int64), np.asarray(ref[k], np.int64)), k


def test_index_builder_wet_east_and_north_rim_wraps_like_reference():
    """A wet east-rim cell (i = nx) takes neighbour linear indices that wrap
    into the next row, exactly as the reference's arithmetic does."""
    m = _mask(10, 9, 7, 0.8)
    m[-1, 3:6] = True          # east rim wet
    m[4:7, -1] = True          # north rim wet
    ref = O.index_reference(m, 1)
    got = evp1d.build_index(m, 1)
    for k in ("indxtij", *O.NBR):
        assert np.array_equal(np.asarray(got[k], np.int64), np.asarray(ref[k], np.int64)), k


def test_index_builder_rejects_wet_west_rim():
    m = _mask(10, 9, 8, 0.8)
    m[0, 4] = True
    with pytest.raises(evp1d.Evp1dError):
        evp1d.build_index(m, 1)
    with pytest.raises(ValueError):
        O.index_reference(m, 1)


def test_index_builder_empty_and_single():
    m = np.zeros((5, 5), dtype=bool, order="F")
    e = evp1d.build_index(m, 1)
    assert (e["na"], e["navel"]) == (0, 0)
    m[2, 2] = True
    s = evp1d.build_index(m, 1)
    assert (s["na"], s["navel"]) == (1, 7)
    assert all(s[k][0] > 1 for k in O.NBR)     # every neighbour is a ghost slot


def test_index_neighbours_point_to_the_right_cells():
    """Brute force: ee of (i,j) is the slot whose linear index is (i-1,j), etc."""
    m = _mask(12, 11, 9, 0.6)
    ix = evp1d.build_index(m, 1)
    nx = m.shape[0]
    lin = ix["indxtij"].astype(np.int64)
    off = {"ee": (-1, 0), "ne": (-1, -1), "se": (0, -1), "nw": (1, 0), "sw": (1, 1), "sse": (0, 1)}
    for k, (di, dj) in off.items():
        i = ix["indxti"].astype(np.int64) + di
        j = ix["indxtj"].astype(np.int64) + dj
        assert np.array_equal(lin[ix[k] - 1], i + (j - 1) * nx), k


def test_synth_is_deterministic_and_pinned():
    """The generator output is fixed: same bytes every run and on every host
    (only IEEE basic operations), pinned by the md5 of the 400x300 inputs."""
    import json
    g1 = synth.make_grid(40, 30, 123)
    g2 = synth.make_grid(40, 30, 123)
    for k in g1.fields:
        assert np.array_equal(g1.fields[k], g2.fields[k])
    rec = json.loads((REPO / "tests" / "golden" / "golden.json").read_text())["evp_c0"]
    w = synth.WORKLOADS["evp_c0"]
    g = synth.make_grid(w["nx"], w["ny"], w["seed"])
    st = O.pack_dyn(g, evp1d.build_index(g.tmask, 1))
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs(d, st, w["ndte"])
        for f, m in rec["inputs"].items():
            assert O.md5((Path(d) / f).read_bytes()) == m, f


def test_det_exp_accuracy():
    x = np.linspace(-30, 5, 20001)
    rel = np.abs(synth.det_exp(x) / np.exp(x) - 1.0)
    assert rel.max() < 4e-16


def test_normal_moments():
    z = synth.normal(1, 2, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01 and np.abs(z).max() <= 6.0


def test_reference_input_roundtrip():
    g = synth.make_grid(30, 20, 4)
    st = O.pack_dyn(g, O.index_reference(g.tmask, 1))
    st.scalars["capping"] = 0.5
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs(d, st, 7)
        back = O.read_inputs(d, st.na, st.navel)
        nml = (Path(d) / "kernel.nml").read_text()
    assert "ndte = 7" in nml and f"na = {st.na}" in nml and f"navel = {st.navel}" in nml
    for k in O.FIELDS:
        if k.startswith("str") and k[3:].isdigit():
            continue          # str1..8 are not part of the input file
        if k in ("Cb",):
            continue
        assert np.array_equal(back.f[k], st.f[k]), k
    assert back.scalars == st.scalars
    assert np.array_equal(back.skipT, st.skipT) and np.array_equal(back.skipU, st.skipU)


def test_output_format_compacts_stepu_over_active_u():
    st = O.State1D(5, 7)
    st.skipU[:] = [0, 1, 0, 1, 1]
    st.f["uvel"][:5] = [1, 2, 3, 4, 5]
    a, b = O.output_bytes(st)
    assert len(a) == 12 * 5 * 8
    arr = np.frombuffer(b)
    assert arr.size == 4 * 2 and np.array_equal(arr[4:6], [1.0, 3.0])


def test_product_pack_equals_reference_pack():
    g = synth.make_grid(45, 33, 11)
    ix = evp1d.build_index(g.tmask, 1)
    p = layout.pack(g, ix)
    r = O.pack_dyn(g, O.index_reference(g.tmask, 1))
    for k in O.FIELDS:
        assert np.array_equal(p["fields"][k], r.f[k]), k
    assert np.array_equal(p["skipT"], r.skipT) and np.array_equal(p["skipU"], r.skipU)
    back = layout.unpack(p["fields"], ix, g.tmask.shape)
    ocean = g.tmask
    for k in ("uvel", "stressp_2", "forcexU"):
        assert np.array_equal(back[k][ocean], g.fields[k][ocean]), k


_OOM_CHILD = r"""
import ctypes, resource, sys
# blocks above 128 KiB from mmap, returned to the system when freed, so the
# mapped size below is what the process really holds (M_MMAP_THRESHOLD = -3)
ctypes.CDLL("libc.so.6").mallopt(-3, 128 * 1024)
sys.path[:0] = sys.argv[1:3]
import numpy as np
import evp1d
from evp1d import synth
def lists(nx, ny):
    g = synth.make_grid(nx, ny, 5)
    ix = evp1d.build_index(g.tmask, g.nghost)
    return ix, np.zeros(int(ix["na"]), np.int32), int(ix["navel"])
small, big = lists(60, 50), lists(2000, 1600)
evp1d.tile_plan(small[0], small[1], small[1], small[2])   # library loaded, planner works
ix, skip, navel = big
# the address space from now on: what is mapped plus a little, so the
# planner's tables for the large lists (hundreds of MB) cannot be allocated
with open("/proc/self/statm") as fh:
    vm = int(fh.read().split()[0]) * resource.getpagesize()
resource.setrlimit(resource.RLIMIT_AS, (vm + (8 << 20), resource.RLIM_INFINITY))
try:
    evp1d.tile_plan(ix, skip, skip, navel)
    print("NO-ERROR")
except evp1d.Evp1dError as e:
    print("CODE", e.code, str(e)[:120])
"""


def test_planner_host_allocation_failure_is_a_status_code(tmp_path):
    """A host allocation failure inside the tile planner (on its worker
    threads or the caller) comes back as EVP1D_ENOMEM through the C ABI; it
    does not terminate the process.  The child process caps its own address
    space after planning small lists, then plans large ones."""
    import subprocess
    import sys
    from pathlib import Path
    repo = Path(__file__).resolve().parent.parent
    r = subprocess.run([sys.executable, "-c", _OOM_CHILD, str(repo / "cice-evp1d_amd"), str(repo)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout[-500:], r.stderr[-1500:])
    assert "CODE 4" in r.stdout, r.stdout + r.stderr[-800:]


def test_capping_fast_path_is_bit_identical(tmp_path):
    """visc_replpress skips the division of its (1 - capping) term where the
    term is a signed zero; tests/cpp/capping_term_check.cpp compares every
    output bit with the division on edge values (signed zeros, subnormals,
    huge, infinities, NaN) and random values, for capping 1, 0.5, 0 and 2."""
    import subprocess
    from pathlib import Path
    repo = Path(__file__).resolve().parent.parent
    exe = tmp_path / "capping_term_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{repo / 'cice-evp1d_amd' / 'csrc'}", str(repo / "tests" / "cpp" / "capping_term_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
