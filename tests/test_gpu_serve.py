"""Serving with continuous batching on the GPU: requests of different sizes arrive while others
run, join and leave the batch between substeps, and each result is bit-identical to the same
request run alone (with the Jacobi or the multigrid-preconditioned pressure PCG)."""
import numpy as np
import pytest

from heatflow_mpm import HeatflowMPM
from heatflow_mpm.scenes import random_scene
from heatflow_mpm.serve import GpuBackend, Request, Server, summarize

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cg", ["warp", "mg"])
def test_served_results_equal_the_request_run_alone(gpu, cg):
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(6):
        _, kw = random_scene(rng, [24, 32, 40][i % 3], name=f"q{i}")
        reqs.append(Request(i, kw, int(rng.integers(3, 9)), 0.02 * i))
    srv = Server(GpuBackend(cg=cg), keep_results=True, max_active=8)
    res = srv.run(reqs)
    joined = [e for e in srv.log if e[1] == "admit" and e[3] > 1]
    print("admissions into a running batch:", len(joined), summarize(res), flush=True)
    assert len(res) == len(reqs)
    for r, q in zip(res, reqs):
        alone = HeatflowMPM(**q.kwargs, cg=cg)
        alone.run(q.dt, q.substeps)
        assert np.array_equal(alone.grid_T(), r.fields["T"]), r.rid
        assert np.array_equal(alone.temperatures(), r.fields["temperatures"]), r.rid
        assert alone.stats()["p_iters_total"] == r.stats["p_iters_total"]
        alone.close()


def test_served_lanes_equal_the_request_run_alone(gpu):
    """Size-class lanes (a batch of small scenes and one of large ones, each on its own stream and host
    thread, running at the same time): every result is still bit-identical to the request run alone."""
    rng = np.random.default_rng(7)
    reqs = []
    for i in range(6):
        _, kw = random_scene(rng, [24, 32, 48][i % 3], name=f"l{i}")
        reqs.append(Request(i, kw, int(rng.integers(3, 9)), 0.01 * i))
    srv = Server(GpuBackend(cg="mg"), keep_results=True, lanes=[(32, 8), (48, 1)])
    res = srv.run(reqs)
    assert len(res) == len(reqs)
    for r, q in zip(res, reqs):
        alone = HeatflowMPM(**q.kwargs, cg="mg")
        alone.run(q.dt, q.substeps)
        assert np.array_equal(alone.grid_T(), r.fields["T"]), r.rid
        assert np.array_equal(alone._scene.get_field("U"), r.fields["U"]), r.rid
        alone.close()
