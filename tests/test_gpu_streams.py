"""Independent plans on several streams inside one captured graph: the
capturing stream forks into the other contexts' streams and joins them
again (bsz_stream_order), so the plans run concurrently.  Every replay is
checked against eager runs of the same plans, with the inputs rewritten
between replays, so a replay that skipped a branch or read stale inputs
would show."""

import numpy as np
import pytest

from bitstreamzoo_amd import api, dag_gen, tuning

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("serial", [False, True])
def test_forked_streams_replay_every_plan(serial):
    ctxs = [api.Context(0) for _ in range(3)]
    progs = [dag_gen.generate(300 + 100 * i, seed=30 + i) for i in range(6)]
    sizes = [1 << 14, (1 << 15) + 77, 1 << 16, 5000, 1 << 16, 12345]
    sets = [[dag_gen.text_case(n, seed=40 + 10 * k + i) for i, n in enumerate(sizes)] for k in range(2)]
    residents, plans, groups = [], [], [[] for _ in ctxs]
    try:
        for i, p in enumerate(progs):
            c = ctxs[i % len(ctxs)]
            r = tuning.Resident(c, [sets[0][i]])
            pl = api.Plan(c, api.Program(p.source), tuning.shape_case([sets[0][i]]), {"chunk_rows": 4})
            r.bind(pl)
            residents.append(r)
            plans.append(pl)
            groups[i % len(ctxs)].append(pl)

        def load(k):
            for i, r in enumerate(residents):
                for (_, name, _j), d in r.bufs.items():
                    d.write(sets[k][i].inputs[name])
            for c in ctxs:
                c.sync()

        def outputs():
            out = []
            for i, (p, pl) in enumerate(zip(progs, plans)):
                nw = (sets[0][i].bitlength + 63) // 64
                out.append([pl.download_output(0, o, nw) for o in p.outputs])
            return out

        eager = []
        for k in range(2):
            load(k)
            for pl in plans:
                pl.run()
            for c in ctxs:
                c.sync()
            eager.append(outputs())
        assert any(not np.array_equal(a, b) for a, b in zip(eager[0][0], eager[1][0]))
        ctx0 = ctxs[0]
        others = ctxs[1:]

        def batch():
            if serial:
                prev = ctx0
                for c, g in zip(ctxs, groups):
                    if c is not prev:
                        c.order_after(prev)
                    for pl in g:
                        pl.run()
                    prev = c
            else:
                for c in others:
                    c.order_after(ctx0)
                for g in groups:
                    for pl in g:
                        pl.run()
            for c in others:
                ctx0.order_after(c)

        graph = ctx0.capture(batch)
        for k in (1, 0, 1):
            load(k)
            graph.launch(1)
            ctx0.sync()
            got = outputs()
            for i in range(len(progs)):
                for a, b in zip(got[i], eager[k][i]):
                    assert np.array_equal(a, b), (k, i)
        del graph
        ms = tuning.time_runs_streams(ctx0, list(zip(ctxs, groups)), 3, 1, serial=serial)
        assert ms > 0
    finally:
        for pl in plans:
            pl.close()
        for r in residents:
            r.free()
        for c in ctxs:
            c.close()


def test_stream_order_needs_two_contexts():
    c = api.Context(0)
    try:
        with pytest.raises(api.BszError):
            c.order_after(c)
    finally:
        c.close()
