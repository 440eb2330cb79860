"""GPU tests of the device-side 2D <-> 1D copies (evp1d_set_grid,
evp1d_put_2d, evp1d_get_2d), against the host packing that follows
convert_2d_1d_dyn / convert_1d_2d_dyn (integration/ice_dyn_evp1d.F90:818-947;
evp1d/layout.py and the oracle's pack_dyn restate them)."""
import numpy as np
import pytest

import evp1d
from evp1d import layout, synth
from gpu_common import bits

pytestmark = pytest.mark.gpu
OUT = ("stressp_1", "stressp_2", "stressp_3", "stressp_4", "stressm_1", "stressm_2", "stressm_3",
       "stressm_4", "stress12_1", "stress12_2", "stress12_3", "stress12_4", "strintxU", "strintyU",
       "Cb", "uvel", "vvel")


def _grid(nx, ny, seed):
    g = synth.make_grid(nx, ny, seed)
    rng = np.random.default_rng(seed)
    for n in layout.DYN_2D[:12] + ("strintxU", "strintyU", "Cb"):   # nonzero everywhere, land included
        g.fields[n] = np.asfortranarray(rng.standard_normal(g.tmask.shape) * 10.0)
    return g


def _engine(s, ndte):
    p = evp1d.set_evp_parameters(ndte)
    eng = evp1d.Engine(s["na"], s["navel"], p)
    for n in evp1d.NBR:
        eng.set_index(n, s["idx"][n])
    eng.set_skip("skipT", s["skipT"])
    eng.set_skip("skipU", s["skipU"])
    return eng


@pytest.mark.parametrize("nx,ny,seed,mode", [(50, 40, 91, "fused"), (70, 33, 92, "graph"), (31, 64, 93, "fused")])
def test_put_run_get_2d_equals_host_packing(gpu, nx, ny, seed, mode):
    g = _grid(nx, ny, seed)
    ix = evp1d.build_index(g.tmask, g.nghost)
    s = layout.pack(g, ix)
    ndte = 12
    with _engine(s, ndte) as a, _engine(s, ndte) as b:
        a.set_mode(mode)
        b.set_mode(mode)
        for n in evp1d.FIELDS:
            a.set_field(n, s["fields"][n])
        for n in ("uarear", "dxT", "dyT", "HTE", "HTN", "HTEm1", "HTNm1"):   # static metrics as the host does
            b.set_field(n, s["fields"][n])
        b.set_grid(g.tmask.shape[0], g.tmask.shape[1], ix["indxtij"])
        dyn = {n: g.fields[n] for n in layout.DYN_2D}
        dyn["uvel_init"] = g.fields["uvel"]
        dyn["vvel_init"] = g.fields["vvel"]
        b.put_2d(dyn)
        for n in evp1d.FIELDS:            # every 1D position, ghosts included, bit for bit
            assert np.array_equal(bits(a.get_field(n)), bits(b.get_field(n))), n
        a.run(ndte)
        b.run(ndte)
        got = b.get_2d(OUT)
    i = ix["indxti"].astype(np.int64) - 1
    j = ix["indxtj"].astype(np.int64) - 1
    outside = np.ones(g.tmask.shape, bool)
    outside[i, j] = False
    with _engine(s, ndte) as c:           # host unpack of the same run
        c.set_mode(mode)
        for n in evp1d.FIELDS:
            c.set_field(n, s["fields"][n])
        c.run(ndte)
        for n in OUT:
            want = g.fields[n].copy()
            want[i, j] = c.get_field(n)[:ix["na"]]
            assert np.array_equal(bits(got[n]), bits(want)), n
            assert np.array_equal(bits(got[n][outside]), bits(g.fields[n][outside])), n


def test_pack2d_argument_checks(gpu):
    g = _grid(30, 20, 94)
    ix = evp1d.build_index(g.tmask, g.nghost)
    s = layout.pack(g, ix)
    nx, ny = g.tmask.shape
    with _engine(s, 3) as e:
        with pytest.raises(evp1d.Evp1dError):
            e.put_2d({"uvel": g.fields["uvel"]})          # no grid yet
        with pytest.raises(evp1d.Evp1dError):
            e.set_grid(nx, ny, ix["indxtij"][:-1])      # wrong length
        bad = ix["indxtij"].copy()
        bad[0] = nx * ny + 1
        with pytest.raises(evp1d.Evp1dError):
            e.set_grid(nx, ny, bad)                     # index outside the grid
        e.set_grid(nx, ny, ix["indxtij"])
        with pytest.raises(evp1d.Evp1dError):
            e.get_2d(["uvel"])                          # never put
    with evp1d.Engine.batch([(s["na"], s["navel"])] * 2, evp1d.set_evp_parameters(3)) as bt:
        with pytest.raises(evp1d.Evp1dError):
            bt.set_grid(nx, ny, ix["indxtij"])
