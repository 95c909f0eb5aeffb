"""HIP graphs of work blocks (kinc_sim_graph_*): several blocks recorded into
one graph and replayed three times, with the program's own eager work (other
blocks on the same handle) and synchronisations between the replays; the
outputs are poisoned before every replay and must equal the eager run's bit
for bit each time.  Both the one-stream Pearson path (the queue-free launch
sequence of config A) and the persistent form, whose bins fork onto
auxiliary streams and join back, are recorded."""
import numpy as np
import pytest

from kinc_amd import _native, gemgen
from parity_util import assert_block_equal

pytestmark = pytest.mark.gpu


def _outputs(torch, n, count, S, MC):
    dev = "cuda:0"
    return (torch.empty((n, count), dtype=torch.int8, device=dev),
            torch.empty((n, count, S), dtype=torch.int8, device=dev),
            torch.empty((n, count, MC), dtype=torch.float32, device=dev),
            torch.empty((n, count, MC), dtype=torch.int16, device=dev))


def _host(o, i):
    K, L, C, I = (a[i].cpu().numpy() for a in o)
    return K, L, C, I.view(np.uint16)


@pytest.mark.parametrize("G,S,flags", [
    (300, 120, dict(clusmethod="none", corrmethod="pearson")),
    (200, 200, dict(clusmethod="gmm", corrmethod="spearman", crit="BIC", preout=True,
                    postout=True)),
])
def test_graph_replays_equal_eager(G, S, flags):
    import torch
    torch.cuda.set_device(0)
    x = gemgen.modules(G, S, seed=41, nan_rate=0.2)
    p = _native.make_params(batch_capacity=4000, **flags)
    MC = p.max_clusters
    count = 4000
    starts = [0, 4000, 8000]
    eng = _native.Engine(x, p)
    try:
        eager = _outputs(torch, 3, count, S, MC)
        for i, st in enumerate(starts):
            eng.execute_device(st, count, *(t[i].data_ptr() for t in eager))
        eng.synchronize()
        want = [_host(eager, i) for i in range(3)]
        rec = _outputs(torch, 3, count, S, MC)
        eng.graph_begin()
        for i, st in enumerate(starts):
            eng.execute_device(st, count, *(t[i].data_ptr() for t in rec))
        g = eng.graph_end()
        other = _outputs(torch, 1, count, S, MC)
        try:
            for rep in range(3):
                for t in rec:
                    t.fill_(-3)
                torch.cuda.synchronize()
                g.launch(None)
                torch.cuda.synchronize()
                for i in range(3):
                    assert_block_equal(_host(rec, i), want[i], S, "replay %d block %d" % (rep, i))
                # eager work on the same handle between replays
                eng.execute_device(12000, count, *(t[0].data_ptr() for t in other))
                eng.synchronize()
        finally:
            g.close()
        with pytest.raises(_native.KincError):
            eng.graph_end()            # nothing is being recorded
    finally:
        eng.close()
