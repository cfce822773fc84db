"""The committed config-1 fixtures (tests/golden/cfg1_seed0.npz, written by
tests/golden/make_fixtures.py): the generator, the oracle and the device must
keep reproducing them."""

import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_fixtures as mf  # noqa: E402
from srmg import configs, solver  # noqa: E402
from srmg.solver import F64  # noqa: E402

FIX = np.load(os.path.join(HERE, "golden", "cfg1_seed0.npz"), allow_pickle=False)


def _valid(a, p):
    G, M = p.b.G, p.M
    return a[(slice(G, G + M),) * p.D]


def test_generator_reproduces_the_fixture_tree_and_charge():
    p = configs.make_cfg1(0)
    for k, v in mf.flat_records(p.tree, p.conn).items():
        np.testing.assert_array_equal(v, FIX[k], err_msg=k)
    np.testing.assert_array_equal(_valid(p.b.u, p), FIX["b"])
    assert p.s == FIX["s"] and p.h == FIX["h"]


def test_oracle_reproduces_the_fixture_history():
    p = configs.make_cfg1(0)
    hist, u = mf.oracle_history(p)
    np.testing.assert_array_equal(hist, FIX["history"])
    np.testing.assert_array_equal(u, FIX["u"])
    # the reference's coarsest-level treatment on a 64 x 128 level 1 converges slowly
    assert 0.05 < hist[-1] < 0.2 and np.all(np.diff(hist[1:]) < 0)


@pytest.mark.gpu
def test_device_reproduces_the_fixture_history():
    p = configs.make_cfg1(0)
    geom, disc = solver.geometry_from_code(p.geom), solver.LaplacianDiscretization(2, 2)
    ctx = solver.Context(p.tree, p.conn, p.M)
    u, b, r, u1, z = (ctx.field(F64) for _ in range(5))
    host = p.b.like()
    _valid(host.u, p)[...] = FIX["b"]
    b.upload(host)
    solver.fill_ghost_all(u, p.bc)
    r0 = solver.resnorm(r, b, z, p.s, geometry=geom, bc=p.bc, lpldisc=disc)
    hist = [solver.resnorm(r, b, u, p.s, geometry=geom, bc=p.bc, lpldisc=disc) / r0]
    for _ in range(int(FIX["cycles"])):
        solver.vcycle(u, b, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2)
        hist.append(solver.resnorm(r, b, u, p.s, geometry=geom, bc=p.bc, lpldisc=disc) / r0)
    out = p.b.like()
    u.download(out)
    np.testing.assert_array_equal(_valid(out.u, p), FIX["u"])
    # the norms are sums in another fixed order than the oracle's
    np.testing.assert_allclose(hist, FIX["history"], rtol=1e-12, atol=0)
