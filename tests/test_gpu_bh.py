"""Barnes-Hut tangent-point energy on the device (rc_energy_bh) against the
reference's stage-2 oracle vectors (tests/golden/bh_goldens.json, made by
tests/golden/make_bh_goldens.py) and the oracle's C restatement: energy to
1e-12 and the traversal counts (visits, approximated edge references, direct
pairs) exactly; theta = 0 reproduces the exact energy; the reference's gates
(oracle/check_bh_stage2.py): clustering fires and the energy error stays
within 5e-2 at (theta, theta_t) = (0.5, 0.75) and shrinks at (0.25, 0.5).
Block-cluster-tree products of the edge kernels and the Gram apply likewise
against the reference vectors and the oracle (block counts exact)."""
import json
import os
import sys
import time

import numpy as np
import pytest

from oracle import rc_oracle as rco
from rcurves import core, curves

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_bh_energy_vs_reference_vectors():
    G = json.load(open(os.path.join(HERE, "golden", "bh_goldens.json")))
    for c in G["cases"]:
        sh = G["shapes"][c["shape"]]
        log("bh", c["shape"], c["theta"], c["theta_t"], c["leaf_size"])
        en, st = core.bh_energy(sh["vertices"], sh["edges"], 3.0, 6.0, 1e-10, c["theta"],
                                c["leaf_size"], c["theta_t"])
        assert abs(en - c["energy"]) <= 1e-12 * abs(c["energy"]), (c, en)
        assert (st["visits"], st["approximated_edge_refs"], st["direct_pairs"]) == (
            c["visits"], c["approximated_edge_refs"], c["direct_pairs"]), (c, st)


@pytest.mark.parametrize("n", [2048, 8192])
def test_bh_energy_vs_oracle_at_size(n):
    v, e = curves.fourier_curve(n, seed=1)
    exact = core.calculate_energy(v, e, None, 3.0, 6.0, 1e-10)
    errs = {}
    for theta, theta_t, leaf in ((0.5, 0.75, 4), (0.25, 0.5, 4), (0.0, 0.0, 8)):
        log("bh", n, theta, theta_t, leaf)
        t0 = time.perf_counter()
        en, st = core.bh_energy(v, e, 3.0, 6.0, 1e-10, theta, leaf, theta_t)
        log(f"  {1e3 * (time.perf_counter() - t0):.1f} ms (tree on the host + traversal)", st)
        eo, so = rco.energy_bh(v, e, theta, theta_t, leaf)
        assert abs(en - eo) <= 1e-12 * abs(eo)
        assert (st["visits"], st["approximated_edge_refs"], st["direct_pairs"]) == so
        errs[theta] = abs(en - exact) / abs(exact)
    assert errs[0.0] <= 1e-12
    assert 0.0 < errs[0.25] < errs[0.5] <= 5e-2


def test_bh_rejects_bad_parameters():
    v, e = curves.trefoil_jittered(32, 0.01, seed=1)
    for theta, tt, leaf in ((-1.0, 0.5, 4), (float("nan"), 0.5, 4), (0.5, 0.75, 0)):
        with pytest.raises(core.L.RcError):
            core.bh_energy(v, e, 3.0, 6.0, 1e-10, theta, leaf, tt)


def frel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_bct_products_vs_reference_vectors():
    G = json.load(open(os.path.join(HERE, "golden", "bh_goldens.json")))
    for c in G["bct"]:
        sh = G["shapes"][c["shape"]]
        log("bct", c["shape"], c["kind"], c["theta"], c["leaf_size"])
        args = dict(theta=c["theta"], leaf_size=c["leaf_size"], theta_t=c["theta_t"])
        p1, st = core.bct_matvec_kernel(sh["vertices"], sh["edges"], 1e-10, c["psi1"], c["kind"],
                                        **args)
        p3, _ = core.bct_matvec_kernel(sh["vertices"], sh["edges"], 1e-10, c["psi3"], c["kind"],
                                       **args)
        bu = core.bct_apply_b_scalar(sh["vertices"], sh["edges"], 1e-10, c["u"], c["kind"], **args)
        assert frel(p1, c["phi1"]) <= 1e-12 and frel(p3, c["phi3"]) <= 1e-12, c["shape"]
        assert frel(bu, c["Bu"]) <= 1e-12
        s = c["stats"]
        assert (st["leaves"], st["admissible_blocks"], st["direct_blocks"], st["direct_pairs"]) == (
            s["leaves"], s["admissible_blocks"], s["direct_blocks"], s["direct_pairs"])


@pytest.mark.parametrize("kind", ["high", "low"])
def test_bct_products_vs_oracle_at_size(kind):
    v, e = curves.fourier_curve(2048, seed=2)
    m = len(e)
    psi = np.sin(0.3 * np.arange(m)) + 0.5
    u = np.cos(0.05 * np.arange(len(v)))
    exact, _ = rco.bct_matvec(v, e, psi, kind, 0.0, 0.0, 8)
    errs = {}
    for theta, tt, leaf in ((0.5, 0.75, 4), (0.25, 0.5, 8), (0.0, 0.0, 8)):
        log("bct size", kind, theta, leaf)
        t0 = time.perf_counter()
        ph, st = core.bct_matvec_kernel(v, e, 1e-10, psi, kind, theta, leaf, tt)
        log(f"  {1e3 * (time.perf_counter() - t0):.1f} ms (block lists on the host + products)",
            st)
        po, so = rco.bct_matvec(v, e, psi, kind, theta, tt, leaf)
        assert frel(ph, po) <= 1e-12
        assert (st["leaves"], st["admissible_blocks"], st["direct_blocks"],
                st["direct_pairs"]) == so
        err = frel(ph, exact)
        errs[theta] = err
        if theta == 0.0:
            assert err <= 1e-12 and st["admissible_blocks"] == 0
        else:
            # the reference gates only that blocks are admitted (check_bh_stage2.py)
            assert st["admissible_blocks"] > 0 and 0.0 < err < 1.0
        bu = core.bct_apply_b_scalar(v, e, 1e-10, u, kind, theta, leaf, tt)
        assert frel(bu, rco.bct_apply_b(v, e, u, kind, theta, tt, leaf)) <= 1e-12
    log("  relative errors against the exact product", errs)
    assert errs[0.25] < errs[0.5]


def test_bh_gradient_vs_reference_vectors():
    G = json.load(open(os.path.join(HERE, "golden", "bh_goldens.json")))
    for c in G["bh_gradients"]:
        sh = G["shapes"][c["shape"]]
        v, e = np.asarray(sh["vertices"]), np.asarray(sh["edges"], np.int32)
        log("bh gradient", c["shape"], c["theta"], c["leaf_size"])
        g = core.bh_gradient_fd(v, e, 3.0, 6.0, 1e-10, c["theta"], c["leaf_size"], c["theta_t"],
                                c["h"])
        E = rco.energy(v, e)
        m = len(e)
        # the device and the reference sum each energy in different orders:
        # 2 sqrt(P) quanta of ulp(E)/h bound the two differences' rounding
        tol = 2.0 * np.sqrt(m * (m - 3)) * float(np.spacing(E)) / c["h"]
        assert np.abs(g - np.asarray(c["dE"])).max() <= tol, c["shape"]
        if c["theta"] == 0.0:   # no cluster: the exact energy's forward difference
            ge = core.gradient_finite_diff(v, e, None, 3.0, 6.0, 1e-10, c["h"])
            assert np.abs(g - ge).max() <= tol
