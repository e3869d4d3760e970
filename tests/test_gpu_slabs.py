"""Pairs of passes split over x-slabs (slab_planes > 0): the odd pass on slab k, then the even pass on slab
k - 1, as a wavefront.  Every cell computes what it computes in a whole-lattice pass, so the fields must
be the same bits as with slab_planes = 0, for any slab size, sweep direction, launch shape, graph
chunking, with slots of uneven lattices and a slot initialised late (the raw first thermal pass); one slot
is also checked against the oracle."""
import numpy as np
import pytest

from oracle import oracle as orc
from test_gpu_parity import assert_match, coupled_params

pytestmark = pytest.mark.gpu


def run_batch(params, slab_planes, steps_before=20, steps_after=40, **shape):
    from xadvcool_amd.engine import LatticeBatch
    cap = tuple(max(p["mask"].shape[i] for p in params) for i in range(3))
    b = LatticeBatch(slots=len(params), cap_shape=cap, fluid=True, thermal=True, slab_planes=slab_planes, **shape)
    try:
        for s, p in enumerate(params):
            b.load(s, p["mask"], heat=p["heat"], tau=p["tau"], tau_th_fluid=p["ttf"], tau_th_solid=p["tts"],
                   heat_power=p["P"], T_in=p["T_in"], u_in=p["u_in"])
        for s in range(len(params) - 1):
            b.init(s)
        b.set_active(len(params) - 1, False)
        b.step(steps_before)
        b.init(len(params) - 1)  # a refill at an even pass count
        b.step(steps_after)
        b.readout(with_stats=True)
        out = [tuple(t.cpu().numpy() for t in b.fields(s)) for s in range(len(params))]
        return out, b.stats()
    finally:
        b.close()


def designs():
    return [coupled_params(seed=20, shape=(22, 12, 10)), coupled_params(seed=21, shape=(17, 12, 9)),
            coupled_params(seed=22, shape=(25, 10, 10)), coupled_params(seed=23, shape=(19, 11, 8))]


@pytest.mark.parametrize("slab,shape", [(1, dict(graph_chunk=10)), (3, dict(graph_chunk=0)),
                                        (5, dict(graph_chunk=20, cells_per_thread=2, xcd_remap=True)),
                                        (2, dict(graph_chunk=10, fluid_storage="dense", block=128))])
def test_slab_pairs_give_the_bits_of_whole_passes(gpu, slab, shape):
    ps = designs()
    print(f"[slabs] slab_planes {slab} {shape}", flush=True)
    ref, st_ref = run_batch(ps, 0, **shape)
    got, st_got = run_batch(ps, slab, **shape)
    for s, (a, c) in enumerate(zip(got, ref)):
        for name, x, y in zip(("rho", "u", "T"), a, c):
            assert np.array_equal(x, y), f"slot {s} {name} differs with slab_planes={slab}"
    assert np.array_equal(st_got, st_ref)


def test_slab_pairs_match_the_oracle(gpu):
    p = coupled_params(seed=24, shape=(23, 12, 10))
    print("[slabs] oracle", flush=True)
    (got,), _ = run_batch([p], 2, steps_before=0, steps_after=60, graph_chunk=10)
    st = orc.State(p["mask"].shape, np.float32)
    st.T[...] = p["T_in"]
    st.init_thermal(mask=p["mask"])
    st.init_fluid()
    orc.run(st, p["mask"], 60, fluid=1, thermal=1, tau=p["tau"], u_in=p["u_in"], rho_out=1.0,
            tau_th_fluid=p["ttf"], tau_th_solid=p["tts"], heat_power=p["P"], T_in=p["T_in"], heat=p["heat"])
    assert_match("rho", got[0], st.rho)
    assert_match("u", got[1], st.u)
    assert_match("T", got[2], st.T)


def test_odd_step_counts_fall_back_to_whole_passes(gpu):
    """A run of an odd number of steps takes whole passes (the 13 after the refill; the 8 before go as slab
    pairs): same bits as whole passes throughout."""
    ps = designs()[:2]
    print("[slabs] odd step counts", flush=True)
    ref, _ = run_batch(ps, 0, steps_before=8, steps_after=13, graph_chunk=0)
    got, _ = run_batch(ps, 3, steps_before=8, steps_after=13, graph_chunk=0)
    for a, c in zip(got, ref):
        for x, y in zip(a, c):
            assert np.array_equal(x, y)
