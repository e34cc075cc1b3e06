"""hipGraph replay of the kernel stages (GM_OPT_GRAPH): every replay, with eager batches and
synchronisations of the program between them, gives the eager result bit for bit; a change of shape
captures anew; and the per-call time of a small batch is printed eager against replayed."""
import time

import numpy as np
import pytest

from gemini_unified_amd import GreenFKEngine

pytestmark = pytest.mark.gpu

SIGMA, EPS = 5.0 / 2048.0, 1e-6


def test_graph_replays_match_eager(c1_setup):
    nem, nodes, jre = c1_setup
    eng = GreenFKEngine(nem, nodes, variant="exact")
    fa, fb = [30 / 2048.0, 61 / 2048.0], [45 / 2048.0]
    ref_a = eng.gfk(fa, 0, 40, SIGMA, EPS, jre, 1)
    ref_b = eng.gfk(fb, 100, 24, SIGMA, EPS, jre, 1)
    ref_m = eng.minors(fa, 0, 40, SIGMA, EPS, jre, jre)
    eng.configure(graph=True)
    for rep in range(3):
        print(f"graph replay {rep}", flush=True)
        r = eng.gfk(fa, 0, 40, SIGMA, EPS, jre, 1)
        np.testing.assert_array_equal(r["gf"], ref_a["gf"])
        np.testing.assert_array_equal(r["gfre"], ref_a["gfre"])
        np.testing.assert_array_equal(r["stats"], ref_a["stats"])
        # eager work of another shape between the replays
        eng.configure(graph=False)
        rb = eng.gfk(fb, 100, 24, SIGMA, EPS, jre, 1)
        np.testing.assert_array_equal(rb["gf"], ref_b["gf"])
        eng.configure(graph=True)
    # same shapes, other frequencies: the replay reads the new uploads
    fc = [33 / 2048.0, 58 / 2048.0]
    eng.configure(graph=False)
    ref_c = eng.gfk(fc, 0, 40, SIGMA, EPS, jre, 1)
    eng.configure(graph=True)
    np.testing.assert_array_equal(eng.gfk(fc, 0, 40, SIGMA, EPS, jre, 1)["gf"], ref_c["gf"])
    # minors batches (no UJGF stages) in graph mode
    for rep in range(2):
        m = eng.minors(fa, 0, 40, SIGMA, EPS, jre, jre)
        for k in ("u_sph", "d_sph", "det_sph", "u_tor"):
            np.testing.assert_array_equal(m[k], ref_m[k])
    eng.close()


def test_graph_small_batch_latency(c1_setup):
    nem, nodes, jre = c1_setup
    eng = GreenFKEngine(nem, nodes, variant="fast")
    f = [50 / 2048.0]
    out = {}
    for g in (False, True):
        eng.configure(graph=g)
        eng.gfk(f, 200, 16, SIGMA, EPS, jre, 1, want_gf=False)
        t0 = time.perf_counter()
        for _ in range(20):
            eng.gfk(f, 200, 16, SIGMA, EPS, jre, 1, want_gf=False)
        out[g] = (time.perf_counter() - t0) / 20 * 1e3
    print(f"16-degree batch: {out[False]:.3f} ms per call eager, {out[True]:.3f} ms replayed", flush=True)
    eng.close()
