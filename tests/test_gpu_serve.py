"""The continuous-batching server (zg_server_*): every request's results are
the same whether it is answered alone or batched with others, and equal the
CPU oracle's."""
import threading
import time

import numpy as np
import pytest

import zindex_ref as zr
from test_gpu_parity import build, make_colors, make_reads, oracle_for, say

zindex_amd = pytest.importorskip("zindex_amd")

pytestmark = pytest.mark.gpu


def split_requests(seed, reads, n_requests):
    """Cut the read list into n_requests consecutive requests of random sizes
    (at least one empty request, one single-read request)."""
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.choice(np.arange(1, len(reads)), n_requests - 1, replace=False))
    bounds = [0] + cuts.tolist() + [len(reads)]
    reqs = [reads[a:b] for a, b in zip(bounds[:-1], bounds[1:])]
    reqs.insert(3, [])
    reqs.insert(5, reqs[5][:1])
    return reqs


def serve_all(idx, requests, **cfg):
    out = []
    with zindex_amd.Server(idx, **cfg) as srv:
        tickets = [srv.submit(r) for r in requests]
        for t in tickets:
            out.append(srv.wait(t))
        stats = srv.stats()
    return out, stats


@pytest.fixture(scope="module")
def serve_case():
    colors = make_colors(71, 12, 30000)
    reads = make_reads(72, colors, 600)
    idx = build(colors, fingerprint_bits=8, hashes=4, slot_scale=0.7)
    idx.load()
    return colors, reads, idx


def test_server_mixed_equals_alone_and_oracle(serve_case):
    colors, reads, idx = serve_case
    requests = split_requests(73, reads, 40)
    say(f"server: {len(requests)} requests, {len(reads)} reads")
    mixed, st = serve_all(idx, requests, slots=2, max_requests=8)
    say(f"server stats {st}")
    assert st["requests"] == len(requests)
    assert st["max_batch_requests"] > 1, "no request was batched with another"
    alone, st1 = serve_all(idx, requests, slots=1, max_requests=1)
    assert st1["max_batch_requests"] == 1 and st1["batches"] == len(requests)
    orc = oracle_for(idx)
    for i, (req, m, a) in enumerate(zip(requests, mixed, alone)):
        tm, cm, am, sm = m
        ta, ca, aa, sa = a
        assert sm["status"] == 0 and sm["n_reads"] == len(req)
        assert np.array_equal(tm, ta) and np.array_equal(cm, ca) and np.array_equal(am, aa), i
        # a request alone through the batch query
        t, c, agg = idx.query_reads(req, records=True)
        assert np.array_equal(tm, t) and np.array_equal(cm, c) and np.array_equal(am, agg), i
        assert sm["seeds"] == int(t.sum())
        if req:
            buf, off = zindex_amd.pack_reads(req)
            et, ec = orc.query_reads(buf, off, threads=zr.cpu_count())
            assert np.array_equal(tm, et) and np.array_equal(cm.astype(np.uint64), ec), i


def test_server_concurrent_submitters_and_poll(serve_case):
    """Requests submitted from several threads while batches run."""
    colors, reads, idx = serve_case
    requests = split_requests(74, reads, 24)
    expect = [idx.query_reads(r, records=True) for r in requests]
    results = [None] * len(requests)
    with zindex_amd.Server(idx, slots=3, max_requests=4) as srv:
        def client(ids):
            for i in ids:
                t = srv.submit(requests[i])
                results[i] = srv.wait(t)
        threads = [threading.Thread(target=client, args=(ids,))
                   for ids in np.array_split(np.arange(len(requests)), 4)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        # poll API: a request answered without wait()
        t = srv.submit(requests[0])
        deadline = time.time() + 60
        got = None
        while got is None and time.time() < deadline:
            got = srv.poll(t)
            time.sleep(0.001)
        assert got is not None and got[3]["status"] == 0 and got[3]["ticket"] == t
        assert np.array_equal(got[1], expect[0][1])
    for i, (got, exp) in enumerate(zip(results, expect)):
        assert np.array_equal(got[0], exp[0]) and np.array_equal(got[1], exp[1]), i
        assert np.array_equal(got[2], exp[2]), i


@pytest.mark.parametrize("mode", ["minimizers", "findere", "randstrobes", "long_kmers",
                                  "spaced_hpc"])
def test_server_seed_modes_mixed_equals_alone(mode):
    colors = make_colors(75, 6, 40000)
    reads = make_reads(76, colors, 240, lo=30, hi=3000)
    cfg = {"minimizers": dict(index_mode=mode, k=49, minimizer_size=31),
           "findere": dict(index_mode=mode, k=31, findere_z=8, homopolymer_compression=1),
           "randstrobes": dict(index_mode="minimizers", k=31, minimizer_size=19,
                               randstrobemers=1),
           "long_kmers": dict(index_mode="kmers", k=45),
           "spaced_hpc": dict(index_mode="kmers", k=31, homopolymer_compression=1,
                              seed_pattern="###_##_#_##_#_###_#_##_#_##_###")}[mode]
    idx = build(colors, fingerprint_bits=16, hashes=3, **cfg)
    idx.load()
    requests = split_requests(77, reads, 20)
    mixed, st = serve_all(idx, requests, slots=2, max_requests=5)
    assert st["max_batch_requests"] > 1
    orc = oracle_for(idx)
    for i, (req, m) in enumerate(zip(requests, mixed)):
        t, c, agg = idx.query_reads(req, records=True)
        assert np.array_equal(m[0], t) and np.array_equal(m[1], c) and np.array_equal(m[2], agg), i
        if req:
            buf, off = zindex_amd.pack_reads(req)
            et, ec = orc.query_reads(buf, off, threads=zr.cpu_count())
            assert np.array_equal(m[0], et) and np.array_equal(m[1].astype(np.uint64), ec), i


@pytest.mark.parametrize("colors_n,recover", [(600, True), (12, False)])
def test_server_sweep_batches_equal_alone_and_oracle(colors_n, recover, monkeypatch):
    # an index tuned to the sorted sweep: every batch runs as the sweep
    # (ZG_SERVE_SWEEP=1) from the slot's reserved chunk buffers, several
    # chunks per batch; each request's results equal the read-ordered query
    # alone and the oracle, mixed and alone, with recovery postings at 600
    # colours (the bit-plane lookup and its recovery pass)
    colors = make_colors(80 + colors_n, colors_n, 3000 if colors_n > 100 else 30000,
                         rates=(0.0, 0.01))
    reads = make_reads(81, colors, 240, lo=30, hi=3000)
    idx = build(colors, fingerprint_bits=8, hashes=4, slot_scale=0.6)
    idx.load(recover=recover)
    orc = oracle_for(idx, recover)
    requests = split_requests(82, reads, 16)
    expect = [idx.query_reads(r, records=True) for r in requests]  # read order
    idx.tuning = {"span": 1024, "tiles_per_wave": 1, "prefetch": 1, "waves_per_cu": 0,
                  "nontemporal": 0, "sweep": 1}
    monkeypatch.setenv("ZG_SERVE_SWEEP", "1")
    monkeypatch.setenv("ZG_SWEEP_CHUNK_POSITIONS", "20000")
    mixed, st = serve_all(idx, requests, slots=2, max_requests=5)
    alone, st1 = serve_all(idx, requests, slots=1, max_requests=1)
    say(f"sweep server: {st}")
    assert st["sweep_batches"] > 0 and st1["sweep_batches"] == len([r for r in requests if r])
    assert st["max_batch_requests"] > 1
    for i, (req, m, a, e) in enumerate(zip(requests, mixed, alone, expect)):
        for got in (m, a):
            assert got[3]["status"] == 0
            assert np.array_equal(got[0], e[0]) and np.array_equal(got[1], e[1]), i
            assert np.array_equal(got[2], e[2]), i
        if req:
            buf, off = zindex_amd.pack_reads(req)
            et, ec = orc.query_reads(buf, off, threads=zr.cpu_count())
            assert np.array_equal(m[0], et) and np.array_equal(m[1].astype(np.uint64), ec), i


def test_server_pinned_outputs_and_budget(serve_case):
    colors, reads, idx = serve_case
    C = idx.meta["n_colors"]
    req = reads[10:60]
    totals = zindex_amd.HostBuffer((len(req),), np.uint64)
    counts = zindex_amd.HostBuffer((len(req), C), np.uint32)
    with zindex_amd.Server(idx, slots=1, slot_reads=64, slot_bases=1 << 20) as srv:
        t = srv.submit(req, out=(totals.array, counts.array))
        got = srv.wait(t)
        exp = idx.query_reads(req, records=True)
        assert np.array_equal(totals.array, exp[0]) and np.array_equal(counts.array, exp[1])
        assert np.array_equal(got[2], exp[2])
        # a request larger than a slot is refused at submission
        with pytest.raises(zindex_amd.ZindexError) as e:
            srv.submit(reads[:100])
        assert e.value.code == -7
    totals.free()
    counts.free()
    # the index plus the slots must fit the HBM budget
    with pytest.raises(zindex_amd.ZindexError) as e:
        zindex_amd.Server(idx, hbm_budget=1 << 20)
    assert e.value.code == -7
