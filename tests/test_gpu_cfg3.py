"""The 512-update batch on nb = 2048 tiles with ranks up to 256 per tile
(stacked ranks 71-506, sigma_k = 10^(-12 k / r), acc 1e-8, maxrank 512):
every update of one plan against the oracle.

Bars: status 0 everywhere; each update's rank within one of the oracle's
and its product within 10*acc (relative Frobenius) of the oracle's product;
a sample within 10*acc of the dense update.  This exercises the fused
register QR at nb = 2048 (four rows per thread), the core SVD beyond the
register-resident shapes (K 193-506) and the compact-WY reassembly with the
16-column solve (K > 256) at scale.
"""
import numpy as np
import pytest

from conftest import rel_fro

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.timeout(900)
def test_batch_512_nb2048_vs_oracle(hc, orc):
    from hcore_amd.tiles import TrailingStep
    from test_gpu_coverage import _run_batch, _oracle_batch, _compare_batch, _dense
    step = TrailingStep.batch_512()
    ra, rb, rc = step.ranks()
    _say("case: 512 updates nb 2048, K %d-%d" % ((ra + rc).min(), (ra + rc).max()))
    status, ranks, outs, _ = _run_batch(hc, step, hc.FP64, 1e-8, np.float64)
    assert np.all(status == 0)
    oranks, oouts = _oracle_batch(orc, step, 1e-8, np.float64)
    _compare_batch(step, ranks, outs, oranks, oouts, 1e-8, "nb 2048")
    for u in range(0, len(step.pairs), 101):
        i, j = step.pairs[u]
        E = _dense(-1.0, *step.panel[i], *step.panel[j], 1.0, *step.C[u])
        assert rel_fro(outs[u][0] @ outs[u][1].T, E) <= 10 * 1e-8
