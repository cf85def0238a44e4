"""GPU step server: admission by HBM budget, batched stepping of several
clients' worlds, per-request latency statistics.  A client's worlds must
evolve exactly as the same worlds in a sim of their own."""

import threading

import numpy as np
import pytest

from helpers import action_stream

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    from arc import native, server
    return native, server


def test_admission_by_budget(mods):
    native, server = mods
    per = native.Sim(1).bytes_per_world()
    sv = server.StepServer(1000 * (per + 96))
    assert 990 <= sv.capacity <= 1000
    a = sv.open(600, seed0=0, difficulty=0.5)
    with pytest.raises(server.AdmissionError):
        sv.open(600)
    b = sv.open(sv.capacity - 600, seed0=7)
    assert sv.used == sv.capacity
    sv.close(a)
    assert sv.used == sv.capacity - 600
    with pytest.raises(KeyError):  # a closed client's id stays unknown
        sv.range(a)
    c = sv.open(500)
    assert c not in (a, b)  # ids are never reused, so a stale id cannot name c
    assert sv.range(c)[0] == 0 and sv.range(b)[0] == 600
    with pytest.raises(KeyError):
        sv.range(a)
    with pytest.raises(server.AdmissionError):  # beyond the device terrain's range
        sv.open(10, difficulty=7.0)
    sv.close_server()


def test_clients_match_standalone_sims(mods):
    native, server = mods
    sizes, seeds, steps = [37, 100, 5], [0, 1000, 77], 3
    per = native.Sim(1).bytes_per_world()
    sv = server.StepServer(400 * (per + 96))
    cids = [sv.open(n, seed0=s, difficulty=0.5) for n, s in zip(sizes, seeds)]
    refs = []
    for n, s in zip(sizes, seeds):
        r = native.Sim(n)
        ids = np.arange(n, dtype=np.int32)
        r.reset(ids, (ids + s).astype(np.uint32), np.full(n, 0.5))
        refs.append(r)
    acts = [action_stream(n, steps, seed=i + 1).astype(np.float32) for i, n in enumerate(sizes)]
    for k in range(steps):
        tickets = [sv.submit(c, acts[i][k]) for i, c in enumerate(cids)]
        for t in tickets:
            assert sv.wait(t) > 0.0
        for i, r in enumerate(refs):
            r.step(acts[i][k].astype(np.float64))
    for i, c in enumerate(cids):
        root, _, _, _ = refs[i].observe()
        assert np.array_equal(sv.observe(c), root), f"client {i}"
    st = sv.stats()
    assert st["requests"] == steps * len(sizes) and st["world_steps"] == steps * sum(sizes)
    assert st["batches"] >= steps and st["p99_ms"] >= st["p50_ms"] > 0.0
    sv.close_server()


def test_concurrent_clients(mods):
    """Client threads submitting and waiting concurrently; every request is
    stepped once and in order (the per-client step count is the time)."""
    native, server = mods
    per = native.Sim(1).bytes_per_world()
    sv = server.StepServer(2000 * (per + 96))
    sizes = [1, 64, 300, 700, 9, 128]
    cids = [sv.open(n, seed0=10 * i) for i, n in enumerate(sizes)]
    steps = 6
    errors = []

    def run(i):
        try:
            t = action_stream(sizes[i], 1, seed=i)[0].astype(np.float32)
            for _ in range(steps):
                sv.wait(sv.submit(cids[i], t))
        except Exception as e:  # reported below
            errors.append(e)

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(sizes))]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors
    st = sv.stats()
    assert st["requests"] == steps * len(sizes)
    assert st["world_steps"] == steps * sum(sizes)
    print(f"server: {st}")
    sv.close_server()


def test_pipelined_clients_share_batches(mods):
    """Clients that keep two requests in flight: the dispatcher starts a batch
    as soon as every open client has a request, so nearly every batch holds
    all of them (one is always queued behind the running one), and each
    client's worlds still evolve exactly as in a sim of their own."""
    native, server = mods
    sizes, steps = [512, 1024, 64, 2048], 12
    per = native.Sim(1).bytes_per_world()
    sv = server.StepServer(4000 * (per + 96))
    cids = [sv.open(n, seed0=100 * i, difficulty=0.5) for i, n in enumerate(sizes)]
    acts = [action_stream(n, steps, seed=20 + i).astype(np.float32) for i, n in enumerate(sizes)]
    sv.reset_stats()
    errors = []

    def run(i):
        try:
            q = [sv.submit(cids[i], acts[i][0])]
            for k in range(1, steps):
                q.append(sv.submit(cids[i], acts[i][k]))
                sv.wait(q.pop(0))
            sv.wait(q.pop(0))
        except Exception as e:  # reported below
            errors.append(e)

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(sizes))]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors
    st = sv.stats()
    assert st["requests"] == steps * len(sizes) and st["world_steps"] == steps * sum(sizes)
    print(f"pipelined clients: {st['batches']:.0f} batches, {st['requests_per_batch']:.2f} requests per batch, "
          f"p50 {st['p50_ms']:.3f} ms, p99 {st['p99_ms']:.3f} ms")
    # the first requests may start a batch of their own; afterwards every batch holds every client's next step
    assert st["requests_per_batch"] >= 0.75 * len(sizes)
    for i, n in enumerate(sizes):
        r = native.Sim(n)
        ids = np.arange(n, dtype=np.int32)
        r.reset(ids, (ids + 100 * i).astype(np.uint32), np.full(n, 0.5))
        for k in range(steps):
            r.step(acts[i][k].astype(np.float64))
        assert np.array_equal(sv.observe(cids[i]), r.observe()[0]), f"client {i}"
    sv.close_server()


def test_device_submission_matches_host(mods):
    """Targets handed over in device memory (a libarcsim buffer or a torch
    tensor) step the worlds exactly as the same targets handed over from host
    memory, with both kinds in the same batches and several device requests
    of one client in flight at once."""
    import torch
    native, server = mods
    n, steps = 50, 3
    per = native.Sim(1).bytes_per_world()
    sv = server.StepServer(200 * (per + 96))
    a = sv.open(n, seed0=3, difficulty=0.7)
    b = sv.open(n, seed0=3, difficulty=0.7)
    acts = action_stream(n, steps, seed=9).astype(np.float32)
    bufs = [native.DeviceBuffer.from_array(np.ascontiguousarray(acts[k])) for k in range(steps - 1)]
    last = torch.from_numpy(np.ascontiguousarray(acts[steps - 1])).to("cuda")
    torch.cuda.synchronize()
    with pytest.raises(ValueError):
        sv.submit_device(b, last[:-1])
    host = np.ascontiguousarray(acts[0])
    with pytest.raises(RuntimeError):  # a host address is refused before any kernel could read it
        sv.submit_device(b, host.ctypes.data)
    ta = [sv.submit(a, acts[k]) for k in range(steps)]
    tb = [sv.submit_device(b, x) for x in bufs] + [sv.submit_device(b, last)]
    for t in ta + tb:
        assert sv.wait(t) > 0.0
    ref = native.Sim(n)
    ids = np.arange(n, dtype=np.int32)
    ref.reset(ids, (ids + 3).astype(np.uint32), np.full(n, 0.7))
    for k in range(steps):
        ref.step(acts[k].astype(np.float64))
    root = ref.observe()[0]
    assert np.array_equal(sv.observe(a), root)
    assert np.array_equal(sv.observe(b), root)
    assert sv.stats()["requests"] == 2 * steps
    sv.close_server()


def test_open_close_while_stepping(mods):
    """Clients opened, stepped and closed by several threads while another
    client steps: every request completes, no range is handed out twice while
    in use, and the worlds of the client that stays open evolve exactly as the
    same worlds in a sim of their own."""
    native, server = mods
    per = native.Sim(1).bytes_per_world()
    sv = server.StepServer(600 * (per + 96))
    n, steps = 50, 4
    keep = sv.open(n, seed0=7, difficulty=0.5)
    acts = action_stream(n, steps, seed=3).astype(np.float32)
    errors, ranges = [], []
    lock = threading.Lock()

    def churn(i):
        try:
            m = 20 + i
            t = action_stream(m, 1, seed=i)[0].astype(np.float32)
            for r in range(5):
                c = sv.open(m, seed0=100 * i + r)
                with lock:
                    ranges.append(sv.range(c))
                for _ in range(2):
                    sv.wait(sv.submit(c, t))
                sv.close(c)
        except Exception as e:  # reported below
            errors.append(e)

    th = [threading.Thread(target=churn, args=(i,)) for i in range(6)]
    for x in th:
        x.start()
    for k in range(steps):
        sv.wait(sv.submit(keep, acts[k]))
    for x in th:
        x.join()
    assert not errors
    s0, c0 = sv.range(keep)
    assert all(s + c <= s0 or s >= s0 + c0 for s, c in ranges)  # never inside the open client's range
    assert sv.used == n
    ref = native.Sim(n)
    ids = np.arange(n, dtype=np.int32)
    ref.reset(ids, (ids + 7).astype(np.uint32), np.full(n, 0.5))
    for k in range(steps):
        ref.step(acts[k].astype(np.float64))
    assert np.array_equal(sv.observe(keep), ref.observe()[0])
    assert sv.stats()["requests"] == steps + 6 * 5 * 2
    sv.close_server()
