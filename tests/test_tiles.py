"""CPU tests of the fused-subcycle tiling (evp1d_tile_plan, host only).

The fused kernel trusts three things about a plan, each re-derived here in
numpy from the neighbour lists alone: (1) the slot order is class-major
(stress only | stress and stepu | stepu only | neither), so the unfused
kernels' ranges stay valid; (2) every active T point is owned (stress stored)
by exactly one tile and every active U point is stepped by exactly one tile;
(3) each stepu read goes to an LDS entry that holds the right slot: its own,
nw, sse and sw neighbour (bench_v2b.F90:526-531), evaluated by the tile when
that neighbour is an active T point, read from memory otherwise; (4) the
owned points that other tiles evaluate as halo ("shared") lead their tile's
owned ranges and carry consecutive ids in slot order, since during a run
their stresses live in a table indexed by those ids.
"""
import os
from pathlib import Path

import numpy as np
import pytest

import evp1d
from evp1d import synth


def _lists(nx, ny, seed, frac, pT=0.7):
    rng = np.random.default_rng(seed)
    m = np.zeros((nx, ny), dtype=bool, order="F")
    m[1:-1, 1:-1] = rng.random((nx - 2, ny - 2)) < frac
    ix = evp1d.build_index(m)
    na = ix["na"]
    skipT = (rng.random(na) > pT).astype(np.int32)
    skipU = (rng.random(na) > pT).astype(np.int32)
    return ix, skipT, skipU


def _classes(skipT, skipU):
    t, u = skipT == 0, skipU == 0
    return np.where(t & ~u, 0, np.where(t & u, 1, np.where(u, 2, 3)))


def check_plan(plan, ix, skipT, skipU, tile_w, tile_h):
    """Independent check of a plan; raises AssertionError on the first fault."""
    na, navel = ix["na"], ix["navel"]
    perm = plan["perm"] - 1
    assert np.array_equal(np.sort(perm), np.arange(na))
    cls = _classes(skipT, skipU)
    cnt = np.bincount(cls, minlength=4)
    assert np.array_equal(plan["counts"], cnt)
    assert np.all(np.diff(cls[perm]) >= 0), "slot order is not class-major"
    inv = np.empty(na, np.int64)
    inv[perm] = np.arange(na)

    def slot(q1):          # 1-based list entry -> slot
        q = np.asarray(q1, np.int64) - 1
        return np.where(q < na, inv[np.minimum(q, na - 1)], q)

    t_end, u_begin, u_end = cnt[0] + cnt[1], cnt[0], cnt[0] + cnt[1] + cnt[2]
    owned = np.zeros(na, np.int32)
    stepped = np.zeros(na, np.int32)
    nw, sse, sw = slot(ix["nw"]), slot(ix["sse"]), slot(ix["sw"])
    shared = np.unique(plan["list"])                       # id -> slot
    sid = np.full(na, -1, np.int64)
    sid[shared] = np.arange(len(shared))
    for a0, a1, b0, b1, c0, c1, h0, nh, k0, nk, sa, ia, sb, ib in plan["tiles"]:
        for lo, hi, ns, i0 in ((a0, a1, sa, ia), (b0, b1, sb, ib)):
            want = np.where(np.arange(lo, hi) < lo + ns, i0 + np.arange(hi - lo), -1)
            assert np.array_equal(sid[lo:hi], want), "shared points do not lead their range"
        assert 0 <= a0 <= a1 <= cnt[0] and cnt[0] <= b0 <= b1 <= t_end and t_end <= c0 <= c1 <= u_end
        halo = plan["list"][h0:h0 + nh]
        kon = plan["klist"][k0:k0 + nk]
        table = np.concatenate([np.arange(a0, a1), np.arange(b0, b1), halo, kon])
        assert len(table) <= plan["max_entries"] <= (tile_w + 1) * (tile_h + 1)
        assert len(np.unique(table)) == len(table)
        owned[a0:a1] += 1
        owned[b0:b1] += 1
        assert np.all(halo < t_end) and not np.any((halo >= a0) & (halo < a1) | (halo >= b0) & (halo < b1))
        assert np.all((kon >= t_end) & (kon < navel))
        us = np.concatenate([np.arange(b0, b1), np.arange(c0, c1)])
        stepped[us] += 1
        if len(us) == 0:
            continue
        p = perm[us]
        loc = plan["uloc"][us - u_begin]
        assert np.all(loc < len(table))
        assert np.array_equal(table[loc[:, 0]], us)
        assert np.array_equal(table[loc[:, 1]], nw[p])
        assert np.array_equal(table[loc[:, 2]], sse[p])
        assert np.array_equal(table[loc[:, 3]], sw[p])
    assert np.all(owned[:t_end] == 1) and np.all(owned[t_end:] == 0)
    assert np.all(stepped[u_begin:u_end] == 1) and np.all(stepped[:u_begin] == 0) and np.all(stepped[u_end:] == 0)


@pytest.mark.parametrize("nx,ny,seed,frac", [(40, 30, 1, 0.8), (70, 20, 2, 0.6), (25, 60, 3, 1.0), (50, 50, 4, 0.3)])
@pytest.mark.parametrize("tile", [(32, 16), (8, 4), (5, 7), (1, 1), (64, 1)])
def test_plan_is_consistent(nx, ny, seed, frac, tile):
    ix, skipT, skipU = _lists(nx, ny, seed, frac)
    plan = evp1d.tile_plan(ix, skipT, skipU, ix["navel"], *tile)
    assert plan["ntiles"] > 0
    check_plan(plan, ix, skipT, skipU, *tile)


def test_tiles_are_grid_boxes_on_a_connected_domain():
    nx, ny, (tw, th) = 70, 45, (16, 8)
    m = np.zeros((nx, ny), dtype=bool, order="F")
    m[1:-1, 1:-1] = True
    ix = evp1d.build_index(m)
    z = np.zeros(ix["na"], np.int32)
    plan = evp1d.tile_plan(ix, z, z, ix["navel"], tw, th)
    check_plan(plan, ix, z, z, tw, th)
    perm = plan["perm"] - 1
    for a0, a1, b0, b1, *_ in plan["tiles"]:
        pts = perm[b0:b1]
        i, j = ix["indxti"][pts], ix["indxtj"][pts]
        assert i.max() - i.min() < tw and j.max() - j.min() < th
    # interior tiles are full boxes; halo per tile is one row and one column
    full = [t for t in plan["tiles"] if t[3] - t[2] == tw * th]
    assert full and all(t[7] == tw + th + 1 for t in full)


def test_plan_follows_a_permuted_list():
    ix, skipT, skipU = _lists(45, 35, 7, 0.75)
    na = ix["na"]
    rng = np.random.default_rng(8)
    sigma = rng.permutation(na)                 # new position k holds old point sigma[k]
    pos = np.empty(na, np.int64)
    pos[sigma] = np.arange(na)
    ix2 = dict(ix)
    for n in evp1d.NBR:
        v = ix[n][sigma].astype(np.int64) - 1
        ix2[n] = (np.where(v < na, pos[np.minimum(v, na - 1)], v) + 1).astype(np.int32)
    a = evp1d.tile_plan(ix, skipT, skipU, ix["navel"], 8, 8)
    b = evp1d.tile_plan(ix2, skipT[sigma], skipU[sigma], ix["navel"], 8, 8)
    check_plan(b, ix2, skipT[sigma], skipU[sigma], 8, 8)
    assert b["ntiles"] > 0 and abs(b["ntiles"] - a["ntiles"]) <= a["ntiles"] // 4


def test_lists_that_are_no_grid_give_no_tiles():
    ix, skipT, skipU = _lists(30, 30, 9, 0.9)
    bad = dict(ix)
    bad["nw"] = ix["nw"].copy()
    bad["nw"][[3, 10]] = bad["nw"][[10, 3]]      # two points now disagree about their east neighbour
    plan = evp1d.tile_plan(bad, skipT, skipU, ix["navel"], 8, 8)
    assert plan["ntiles"] == 0
    cls = _classes(skipT, skipU)
    assert np.all(np.diff(cls[plan["perm"] - 1]) >= 0)


def test_tile_too_large_for_lds_gives_no_tiles():
    ix, skipT, skipU = _lists(80, 80, 10, 1.0, pT=1.0)
    plan = evp1d.tile_plan(ix, skipT, skipU, ix["navel"], 60, 60, max_entries=2560)
    assert plan["ntiles"] == 0


def test_checker_catches_a_planted_mistake():
    """A wrong LDS entry (an nw read sent to the sse neighbour) must fail."""
    ix, skipT, skipU = _lists(40, 30, 11, 0.8)
    plan = evp1d.tile_plan(ix, skipT, skipU, ix["navel"], 8, 8)
    check_plan(plan, ix, skipT, skipU, 8, 8)
    bad = dict(plan)
    bad["uloc"] = plan["uloc"].copy()
    bad["uloc"][:, [1, 2]] = bad["uloc"][:, [2, 1]]
    with pytest.raises(AssertionError):
        check_plan(bad, ix, skipT, skipU, 8, 8)


def test_checker_catches_a_wrong_shared_prefix():
    """A tile whose shared prefix is one point short must fail."""
    ix, skipT, skipU = _lists(40, 30, 12, 0.9, pT=1.0)
    plan = evp1d.tile_plan(ix, skipT, skipU, ix["navel"], 8, 8)
    check_plan(plan, ix, skipT, skipU, 8, 8)
    bad = dict(plan)
    bad["tiles"] = plan["tiles"].copy()
    t = int(np.argmax(bad["tiles"][:, 12]))
    assert bad["tiles"][t, 12] > 0
    bad["tiles"][t, 12] -= 1
    with pytest.raises(AssertionError):
        check_plan(bad, ix, skipT, skipU, 8, 8)


def test_plan_on_the_bench_grid_has_small_halo():
    g = synth.make_grid(200, 160, 1001)
    ix = evp1d.build_index(g.tmask, g.nghost)
    from evp1d import layout
    s = layout.pack(g, ix)
    plan = evp1d.tile_plan(s["idx"], s["skipT"], s["skipU"], s["navel"], 32, 16)
    check_plan(plan, ix, s["skipT"], s["skipU"], 32, 16)
    nT = plan["counts"][0] + plan["counts"][1]
    assert plan["tiles"][:, 7].sum() < 0.2 * nT


def test_host_check_catches_corrupted_plans(tmp_path):
    """Build tests/cpp/plan_check_corruptions.cpp against the planner source
    and run it on a synthetic grid: every corruption of a valid plan (LDS
    entries, ownership ranges, shared ids, permutation, constant entries)
    must be reported by check_plan."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    repo = Path(__file__).resolve().parent.parent
    csrc = repo / "cice-evp1d_amd" / "csrc"
    exe = tmp_path / "corrupt"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", f"-I{csrc}", str(repo / "tests" / "cpp" / "plan_check_corruptions.cpp"),
                    str(csrc / "evp1d_tiles.cpp"), "-o", str(exe)], check=True, capture_output=True)
    g = synth.make_grid(300, 200, 71)
    ix = evp1d.build_index(g.tmask, g.nghost)
    from evp1d import layout
    s = layout.pack(g, ix)
    na, navel = s["na"], s["navel"]
    skip = (s["skipT"] != 0).astype(np.uint8) | ((s["skipU"] != 0).astype(np.uint8) << 1)
    with open(tmp_path / "in.bin", "wb") as fh:
        fh.write(np.array([na, navel], np.int32).tobytes())
        for n in evp1d.NBR:
            fh.write((s["idx"][n].astype(np.int32) - 1).tobytes())
        fh.write(skip.tobytes())
    r = subprocess.run([str(exe), str(tmp_path / "in.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "NOT CAUGHT" not in r.stdout and "good: ''" in r.stdout


def test_pair_plan_schedule_and_check(tmp_path):
    """Build tests/cpp/pair_plan_check.cpp against the planner source and run
    it on a synthetic grid: the two-subcycle tables exist for several tile
    shapes within their ring bounds, a CPU model of the pair kernel's
    schedule through them reproduces two plain subcycles exactly (integer
    hash kernels), and every corruption of a valid pair plan is reported by
    check_pairs."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    repo = Path(__file__).resolve().parent.parent
    csrc = repo / "cice-evp1d_amd" / "csrc"
    exe = tmp_path / "pairs"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", f"-I{csrc}", str(repo / "tests" / "cpp" / "pair_plan_check.cpp"),
                    str(csrc / "evp1d_tiles.cpp"), "-o", str(exe)], check=True, capture_output=True)
    g = synth.make_grid(300, 200, 72)
    ix = evp1d.build_index(g.tmask, g.nghost)
    from evp1d import layout
    s = layout.pack(g, ix)
    na, navel = s["na"], s["navel"]
    skip = (s["skipT"] != 0).astype(np.uint8) | ((s["skipU"] != 0).astype(np.uint8) << 1)
    with open(tmp_path / "in.bin", "wb") as fh:
        fh.write(np.array([na, navel], np.int32).tobytes())
        for n in evp1d.NBR:
            fh.write((s["idx"][n].astype(np.int32) - 1).tobytes())
        fh.write(skip.tobytes())
    r = subprocess.run([str(exe), str(tmp_path / "in.bin")], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "NOT CAUGHT" not in r.stdout and "good: ''" in r.stdout
    assert r.stdout.count(": 0 differing outputs") == 10   # five tiles, with and without a copy list


def test_hilbert_tile_order(tmp_path):
    """Build tests/cpp/tile_order_check.cpp against the planner source: the
    Hilbert tile order (tile_ids order 1) is a continuous curve of
    4-neighbours on power-of-two grids, keeps tile membership, and on a real
    grid its tiling and pair tables pass the host checks."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    repo = Path(__file__).resolve().parent.parent
    csrc = repo / "cice-evp1d_amd" / "csrc"
    exe = tmp_path / "order"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", f"-I{csrc}", str(repo / "tests" / "cpp" / "tile_order_check.cpp"),
                    str(csrc / "evp1d_tiles.cpp"), "-o", str(exe)], check=True, capture_output=True)
    g = synth.make_grid(300, 200, 73)
    ix = evp1d.build_index(g.tmask, g.nghost)
    from evp1d import layout
    s = layout.pack(g, ix)
    na, navel = s["na"], s["navel"]
    skip = (s["skipT"] != 0).astype(np.uint8) | ((s["skipU"] != 0).astype(np.uint8) << 1)
    with open(tmp_path / "in.bin", "wb") as fh:
        fh.write(np.array([na, navel], np.int32).tobytes())
        for n in evp1d.NBR:
            fh.write((s["idx"][n].astype(np.int32) - 1).tobytes())
        fh.write(skip.tobytes())
    r = subprocess.run([str(exe), str(tmp_path / "in.bin")], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all good" in r.stdout and "order 1:" in r.stdout


def test_host_planning_clean_under_sanitizers(tmp_path):
    """The planner checks (pair tables with the schedule model and the
    corruptions, tile plan corruptions, Hilbert order) and a fuzz of the 1D
    index builder, built host-only with AddressSanitizer and UBSan: every
    program must pass and no sanitizer may report."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    repo = Path(__file__).resolve().parent.parent
    csrc = repo / "cice-evp1d_amd" / "csrc"
    san = ["-O1", "-g", "-std=c++17", "-pthread", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=undefined", f"-I{csrc}", f"-I{repo / 'include'}"]
    g = synth.make_grid(300, 200, 72)
    ix = evp1d.build_index(g.tmask, g.nghost)
    from evp1d import layout
    s = layout.pack(g, ix)
    skip = (s["skipT"] != 0).astype(np.uint8) | ((s["skipU"] != 0).astype(np.uint8) << 1)
    with open(tmp_path / "in.bin", "wb") as fh:
        fh.write(np.array([s["na"], s["navel"]], np.int32).tobytes())
        for n in evp1d.NBR:
            fh.write((s["idx"][n].astype(np.int32) - 1).tobytes())
        fh.write(skip.tobytes())
    progs = {"pair_plan_check": csrc / "evp1d_tiles.cpp", "plan_check_corruptions": csrc / "evp1d_tiles.cpp",
             "tile_order_check": csrc / "evp1d_tiles.cpp", "index_fuzz": csrc / "evp1d_index.cpp"}
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    for name, lib in progs.items():
        exe = tmp_path / name
        subprocess.run(["g++", *san, str(repo / "tests" / "cpp" / f"{name}.cpp"), str(lib), "-o", str(exe)],
                       check=True, capture_output=True)
        r = subprocess.run([str(exe), str(tmp_path / "in.bin")], capture_output=True, text=True, timeout=300,
                           env=env)
        out = r.stdout + r.stderr
        assert r.returncode == 0, (name, out[-2000:])
        assert "AddressSanitizer" not in out and "runtime error" not in out, (name, out[-2000:])

