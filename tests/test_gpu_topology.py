"""Topologies of the app's own presets (src/core/testConfigs.ts:102-224):
the 3-D 'stress' grid (200 vertices of a jittered 6x6x6 lattice, degree up
to 6) and 'random' graphs (random points, random edges, isolated vertices
joined to a neighbour). Their 32-vertex tiles touch up to ~140 edges, beyond
the assembly's LDS staging of 36, so the Gram matrix of these tiles comes
from the chunked path. Energy, gradient and A against the oracle at 1e-12,
A exactly symmetric, and one Sobolev step against the oracle's (the
reference's full saddle LDL^T): tau and projection iterations exact,
vertices 1e-9."""
import sys

import numpy as np
import pytest

from oracle import rc_oracle as rco
from rcurves import core, curves

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def max_tile_edges(V, e, tile=32):
    m = 0
    for t in range(0, V, tile):
        m = max(m, int(np.sum(((e >= t) & (e < t + tile)).any(axis=1))))
    return m


GRAPHS = {
    "stress": lambda: curves.stress_grid(seed=1),
    "random-30-50": lambda: curves.random_graph(30, 50, seed=1),
    "random-20-30": lambda: curves.random_graph(20, 30, seed=2),
    "random-500-1000": lambda: curves.random_graph(500, 1000, seed=1),
    "random-1000-3000": lambda: curves.random_graph(1000, 3000, seed=1),
}


@pytest.mark.parametrize("name", list(GRAPHS))
def test_preset_energy_gradient_gram(name):
    v, e = GRAPHS[name]()
    V = len(v)
    assert curves.connected_components(V, e) == 1
    mt = max_tile_edges(V, e)
    log(name, "V", V, "E", len(e), "max edges per 32-vertex tile", mt)
    if len(e) > 40:
        assert mt > 40   # the chunked assembly path is exercised
    en = core.calculate_energy(v, e, None, 3.0, 6.0, 1e-10)
    assert abs(en - rco.energy(v, e)) <= 1e-12 * abs(en)
    g = core.gradient_analytical(v, e, None, 3.0, 6.0, 1e-10)
    assert rel(g, rco.gradient(v, e)) <= 1e-12
    A = core.assemble_a(v, e, None, 3.0, 6.0, 1e-10)
    Ao = rco.assemble("A", v, e)
    log(name, "A rel err", rel(A, Ao))
    assert rel(A, Ao) <= 1e-12
    assert np.array_equal(A, A.T)
    assert np.abs(A.sum(1)).max() <= 1e-10 * np.abs(A).max()


@pytest.mark.parametrize("name", ["stress", "random-30-50", "random-500-1000"])
def test_preset_sobolev_step_vs_oracle(name):
    v, e = GRAPHS[name]()
    x0 = rco.barycenter_target(v, e)
    L0 = rco.total_length(v, e)
    r = core.sobolev_step_set(v, e, None, [core.barycenter_block(x0), core.total_length_block(L0)],
                              core.SobolevStepOptions(projection_mode="frozen", collect_field=True))
    out, gt, st = rco.sobolev_step(v, e, rco.ConstraintSet().barycenter(x0).total_length(L0),
                                   rco.default_opts())
    log(name, r["stats"], st["tau"], st["projection_iterations"], rel(r["vertices"], out),
        rel(r["descentField"], gt))
    assert r["accepted"] == st["accepted"]
    assert r["stats"]["tau"] == st["tau"]
    assert r["stats"]["projectionIterations"] == st["projection_iterations"]
    assert abs(r["energy"] - st["energy"]) <= 1e-10 * abs(st["energy"])
    assert rel(r["vertices"], out) <= 1e-9
    assert rel(r["descentField"], gt) <= 1e-9
