"""CPU tests of the v0 2D block oracle (oracle/evp2d_oracle.c).

Pins:
  * the compiled reference v0 (flang, f0 flags): both output files byte for
    byte on every committed 2D golden case (one to four blocks, halo copy
    lists, no ice, non-default file scalars) and the md5 sums at 400x300;
  * the 1D restatement (v2b) on the same grid: v0 and v2 must agree bit for
    bit at every active point (the reference's own v0 == v2 md5 claim,
    README.md:70-83), which checks the 1D list and index construction against
    a plain 2D stencil;
  * a split into blocks with a halo copy list must reproduce the single-block
    run on the interior.
Planted mistakes (a shifted halo source, a re-associated metric term, a T
list without the east halo column) must each be caught.
"""
import json
import tempfile
from pathlib import Path

import numpy as np
import pytest

import evp1d
from evp1d import synth
from oracle import oracle as O
from gpu_common import GOLD, golden_cases_v0, load_golden_v0


def _params(V, st, ndte):
    p = V.params(ndte)
    for k in ("capping", "e_factor", "epp2i"):
        setattr(p, k, st.scalars[k])
    return p


@pytest.mark.parametrize("case", golden_cases_v0())
def test_v0_oracle_matches_reference_golden(case):
    st, ndte, want_a, want_b = load_golden_v0(case)
    V = O.OracleV0()
    V.run(_params(V, st, ndte), st, ndte)
    a, b = O.output_bytes_v0(st)
    assert O.md5(a) == O.md5(want_a)
    assert O.md5(b) == O.md5(want_b)


@pytest.mark.parametrize("case", golden_cases_v0())
def test_v0_writer_reproduces_the_reference_input_files(case):
    st, ndte, _, _ = load_golden_v0(case)
    z = np.load(GOLD / f"{case}.npz")
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs_v0(Path(d), st, ndte)
        for k in z.files:
            if k.startswith("in_") and not k.endswith("nml"):
                name = k[3:][::-1].replace("_", ".", 1)[::-1]
                assert (Path(d) / name).read_bytes() == z[k].tobytes(), name


def test_v0_oracle_c0_md5():
    rec = json.loads((GOLD / "golden.json").read_text())["evp_c0_v0"]
    g = synth.make_grid(rec["nx"], rec["ny"], rec["seed"])
    st = O.pack_v0(g, 1)
    with tempfile.TemporaryDirectory() as d:
        O.write_inputs_v0(Path(d), st, rec["ndte"])
        for f, m in rec["inputs"].items():
            assert O.md5((Path(d) / f).read_bytes()) == m, f
    V = O.OracleV0()
    V.run(_params(V, st, rec["ndte"]), st, rec["ndte"])
    a, b = O.output_bytes_v0(st)
    assert O.md5(a) == rec["output_stress_md5"]
    assert O.md5(b) == rec["output_stepu_md5"]


def _v0_vs_v2(g, ndte, pack=O.pack_v0):
    st = pack(g, 1)
    V = O.OracleV0()
    V.run(_params(V, st, ndte), st, ndte)
    ix = evp1d.build_index(g.tmask, g.nghost)
    s1 = O.pack_dyn(g, ix)
    lib = O.OracleLib()
    lib.run(lib.params(ndte), s1, ndte)
    i = ix["indxti"].astype(np.int64) - 1
    j = ix["indxtj"].astype(np.int64) - 1
    keep = s1.skipU == 0
    bad = []
    for n in O.STRESS_FIELDS:
        if not np.array_equal(st.f[n][:, :, 0][i, j].view(np.uint64), s1.f[n][:s1.na].view(np.uint64)):
            bad.append(n)
    for n2, n1 in (("uvel", "uvel"), ("vvel", "vvel"), ("strintx", "strintxU"), ("strinty", "strintyU")):
        if not np.array_equal(st.f[n2][:, :, 0][i, j][keep].view(np.uint64), s1.f[n1][:s1.na][keep].view(np.uint64)):
            bad.append(n2)
    return bad


@pytest.mark.parametrize("nx,ny,seed,ndte,kw", [
    (40, 30, 31, 25, {}), (64, 20, 32, 10, dict(land_frac=0.5, sigma=3.0)),
    (24, 24, 33, 40, dict(land_frac=0.0, ice_prob=1.0)), (30, 12, 34, 5, dict(ice_prob=0.3))])
def test_v0_equals_v2_on_the_same_grid(nx, ny, seed, ndte, kw):
    g = synth.make_grid(nx, ny, seed, **kw)
    assert _v0_vs_v2(g, ndte) == []


def test_cross_check_catches_a_reassociated_metric():
    """cyp = HTE + 0.5*(HTE - HTEm1) equals 1.5*HTE - 0.5*HTEm1 mathematically
    but rounds differently at some cells: the bitwise v0/v2 comparison must
    see it."""
    g = synth.make_grid(40, 30, 35)

    def pack_bad(grid, nb):
        st = O.pack_v0(grid, nb)
        hte = grid.fields["HTE"]
        htem1 = np.zeros_like(hte)
        htem1[1:, :] = hte[:-1, :]
        cyp = hte + 0.5 * (hte - htem1)
        cyp[0, :] = 0.0
        cyp[:, 0] = 0.0
        st.f["cyp"][:, :, 0] = cyp
        return st
    assert _v0_vs_v2(g, 10, pack_bad) != []


def _interior_diff(g, a, b, names):
    out = []
    for n in names:
        A, B = O.v0_to_global(a, g, n), O.v0_to_global(b, g, n)
        if not np.array_equal(A[1:-1, 1:-1].view(np.uint64), B[1:-1, 1:-1].view(np.uint64)):
            out.append(n)
    return out


@pytest.mark.parametrize("nblocks", [2, 3, 5])
def test_blocks_with_halo_copies_equal_one_block(nblocks):
    g = synth.make_grid(45, 28, 36)
    V = O.OracleV0()
    one = O.pack_v0(g, 1)
    many = O.pack_v0(g, nblocks)
    V.run(_params(V, one, 15), one, 15)
    V.run(_params(V, many, 15), many, 15)
    assert _interior_diff(g, one, many, O.STRESS_FIELDS + ("uvel", "vvel", "strintx", "strinty")) == []


def test_block_check_catches_a_shifted_halo_source():
    g = synth.make_grid(45, 28, 36)
    V = O.OracleV0()
    one = O.pack_v0(g, 1)
    many = O.pack_v0(g, 3)
    src = many.halo["src"]
    k = np.flatnonzero(src[2] > 0)
    src[1, k] += 1        # every neighbour copy reads the cell a row further north
    V.run(_params(V, one, 15), one, 15)
    V.run(_params(V, many, 15), many, 15)
    assert _interior_diff(g, one, many, ("uvel", "vvel")) != []


def test_block_check_catches_a_t_list_without_the_east_halo_column():
    g = synth.make_grid(45, 28, 36)
    V = O.OracleV0()
    one = O.pack_v0(g, 1)
    many = O.pack_v0(g, 3)
    b = 0
    ihi = int(many.halo["ihi"][b])
    nt = int(many.icellt[b])
    keep = many.indxti[:nt, b] <= ihi
    ti, tj = many.indxti[:nt, b][keep], many.indxtj[:nt, b][keep]
    many.indxti[:, b] = 0
    many.indxtj[:, b] = 0
    many.indxti[:ti.size, b], many.indxtj[:tj.size, b] = ti, tj
    many.icellt[b] = ti.size
    V.run(_params(V, one, 15), one, 15)
    V.run(_params(V, many, 15), many, 15)
    assert _interior_diff(g, one, many, ("uvel", "vvel")) != []


def test_package_block_builder_equals_the_oracles():
    """evp1d.layout.Blocks2D (what bench.py feeds the 2D engine) builds the
    same arrays and lists as oracle.pack_v0."""
    from evp1d import layout
    g = synth.make_grid(37, 29, 61)
    a = layout.Blocks2D(g)
    b = O.pack_v0(g, 1)
    for n in O.V0_FIELDS:
        assert np.array_equal(a.f[n].view(np.uint64), b.f[n].view(np.uint64)), n
    for n in ("indxti", "indxtj", "indxui", "indxuj", "icellt", "icellu"):
        assert np.array_equal(getattr(a, n), getattr(b, n)), n
