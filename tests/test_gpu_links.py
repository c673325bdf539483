"""GPU parity of interpolated bounce-back links and the momentum-exchange force.

A channel with an immersed circle / sphere (the reference cylinder script's boundary
order) runs through lbm_set_links + lbm_step_db and through the oracle; populations
must be bit-identical after every step, the force of the last step must agree to
rtol 1e-12 (FP64 sums in different orders), and the macroscopic fields / totals
(which re-apply the links) must agree as in the plain parity tests.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import _util as U

pytestmark = pytest.mark.gpu

CASES = [("D2Q9", 0, 0), ("D2Q9", 0, 1), ("D2Q9", 1, 1), ("D2Q9", 2, 0),
         ("D3Q19", 0, 0), ("D3Q19", 1, 1), ("D3Q27", 0, 1), ("D3Q27", 2, 1)]


@pytest.mark.parametrize("lat,storage,ddf", CASES)
def test_links_match_oracle(gpu, lat, storage, ddf):
    print(f"case links {lat} storage={storage} ddf={ddf}", flush=True)
    desc, f, bc, flags, nodes, qdist = U.obstacle_case(lat, storage=storage, ddf=ddf, seed=storage * 7 + ddf)
    tau = 0.6
    run = U.DeviceRun(desc, f, bc, flags)
    run.set_links(nodes, qdist)
    a, b = f.copy(), f.copy()
    for step in range(4):
        run.step_db(tau)
        force = O.step_db_links(desc, b, a, bc, flags, tau, nodes, qdist)
        a, b = b, a
        got = run.f_host()
        assert np.array_equal(got.view(np.uint8), a.view(np.uint8)), f"populations differ after step {step + 1}"
    gforce = run.link_force()
    scale = np.max(np.abs(force)) + 1e-30
    assert np.all(np.abs(gforce - force) <= 1e-12 * scale), (gforce, force)
    rho, u = run.macroscopic()
    orho, ou = O.macroscopic(desc, a, bc, flags, links=(nodes, qdist))
    assert np.max(np.abs(rho - orho)) <= 1e-10 and np.max(np.abs(u - ou)) <= 1e-10
    tot = run.reduce()
    otot = O.reduce_moments(desc, a, bc, flags, links=(nodes, qdist))
    assert np.all(np.abs(tot - otot) <= 1e-10 * np.maximum(1.0, np.abs(otot)))


def test_links_kbc_les(gpu):
    desc, f, bc, flags, nodes, qdist = U.obstacle_case("D3Q19", storage=0, ddf=1, seed=11, collision=2, les=1)
    run = U.DeviceRun(desc, f, bc, flags)
    run.set_links(nodes, qdist)
    a, b = f.copy(), f.copy()
    for _ in range(3):
        run.step_db(0.52)
        force = O.step_db_links(desc, b, a, bc, flags, 0.52, nodes, qdist)
        a, b = b, a
    assert np.array_equal(run.f_host().view(np.uint8), a.view(np.uint8))
    assert np.allclose(run.link_force(), force, rtol=1e-12, atol=1e-15)


def test_links_clear_and_errors(gpu):
    from mojo_lbm_amd import _lib

    desc, f, bc, flags, nodes, qdist = U.obstacle_case("D2Q9")
    run = U.DeviceRun(desc, f, bc, flags)
    with pytest.raises(_lib.LbmError):
        run.link_force()  # no step with links yet
    with pytest.raises(_lib.LbmError):
        run.set_links(nodes[::-1].copy(), qdist)  # unsorted
    run.set_links(nodes, qdist)
    run.step_db(0.6)
    run.link_force()
    # clearing the links gives the plain step back
    _lib.check(run.L.lbm_set_links(run.h.ptr, None, None, 0))
    run2 = U.DeviceRun(desc, run.f_host(), bc, flags)
    run.step_db(0.6)
    run2.step_db(0.6)
    assert np.array_equal(run.f_host(), run2.f_host())
