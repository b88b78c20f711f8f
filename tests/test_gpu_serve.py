"""The serving loop on the GPU: decode and encode requests submitted alone
and in a mix give the oracle's samples and bytes (each result is the same
alone and in the mix); admission by device-memory footprint holds requests
until earlier results are released; requests that can never run fail at
once with the reference's errors.
"""

import time

import numpy as np
import pytest

import oracle_lib as O
import prores_amd as P

pytestmark = pytest.mark.gpu


def to_p(frame):
    return P.Frame(frame.width, frame.height, frame.chroma, frame.bit_depth, frame.data.copy(),
                   None if frame.alpha is None else frame.alpha.copy(), frame.interlace,
                   P.Metadata(*frame.metadata))


def with_alpha(f, depth):
    mx = (1 << depth) - 1
    yy, xx = np.mgrid[0:f.height, 0:f.width]
    d = np.hypot((xx - f.width * 0.5) / (0.3 * f.width), (yy - f.height * 0.5) / (0.3 * f.height))
    f.alpha = (np.clip((1.0 - d) * 4.0, 0, 1) * mx).round().astype(np.uint16).ravel()
    return f


@pytest.fixture(scope="module")
def requests_():
    """(kind, input, expected) for a mixed set."""
    reqs = []
    dec_specs = [(720, 486, 2, 10, 3, 1, 1, 0), (640, 360, 2, 10, 1, 2, 0, 0), (512, 288, 3, 12, 5, 3, 0, 2),
                 (1280, 720, 2, 10, 0, 0, 0, 0), (333, 75, 3, 12, 4, 3, 2, 1)]
    for i, (w, h, ch, bd, prof, log2, il, atype) in enumerate(dec_specs):
        f = O.test_frame(w, h, ch, bd, 10 + 3 * i, 300 + i)
        f.interlace = il
        if atype:
            with_alpha(f, bd)
        pkt = O.encode(f, O.config(prof, log2_slice_mbs=log2, alpha=atype or 2))
        reqs.append(("decode", pkt, O.decode_result(pkt)))
    enc_specs = [(640, 360, 2, 10, 3, 3, 0, 0, False), (512, 288, 3, 12, 4, 2, 0, 2, True),
                 (720, 486, 2, 10, 2, 1, 1, 0, True)]
    for i, (w, h, ch, bd, prof, log2, il, atype, mats) in enumerate(enc_specs):
        f = O.test_frame(w, h, ch, bd, 8 + 5 * i, 400 + i)
        f.interlace = il
        if atype:
            with_alpha(f, bd)
        rng = np.random.default_rng(i)
        luma = [int(v) for v in rng.integers(4, 33, 64)] if mats else None
        chroma = [int(v) for v in rng.integers(4, 33, 64)] if mats else None
        want = O.encode(f, O.config(prof, log2_slice_mbs=log2, alpha=atype or 2, luma=luma, chroma=chroma))
        cfg = P.Config(prof, log2_slice_mbs=log2, alpha=atype or 2, luma_matrix=luma, chroma_matrix=chroma)
        reqs.append(("encode", (cfg, to_p(f)), want))
    return reqs


def submit(srv, req):
    kind, inp, _ = req
    return srv.submit_decode(inp) if kind == "decode" else srv.submit_encode(*inp)


def check(res, req):
    kind, inp, want = req
    if kind == "decode":
        if want[0] != "Ok":
            assert res.error is not None and (res.error.kind, res.error.msg) == (want[0], want[1])
            return None
        assert res.error is None, res.error
        got = res.samples()
        np.testing.assert_array_equal(got[:want[1].data.size], want[1].data)
        if want[1].alpha is not None:
            np.testing.assert_array_equal(res.alpha()[:want[1].alpha.size], want[1].alpha)
        return got.tobytes()
    assert res.error is None, res.error
    got = res.packet()
    assert got == want
    return got


def drain(srv, n, timeout=60.0):
    out = {}
    t0 = time.time()
    while len(out) < n:
        assert time.time() - t0 < timeout, f"only {len(out)} of {n} results"
        for r in srv.poll(64, 100):
            out[r.id] = r
    return out


def test_alone_and_in_the_mix_are_the_same(requests_):
    with P.Server() as srv:
        alone = []
        for i, req in enumerate(requests_):
            print("case: serve alone", i, req[0], flush=True)
            rid = submit(srv, req)
            res = drain(srv, 1)[rid]
            assert res.batch_frames == 1
            alone.append(check(res, req))
            srv.release(rid)
        print("case: serve mix", flush=True)
        ids = [submit(srv, req) for req in requests_ * 3]
        got = drain(srv, len(ids))
        for k, rid in enumerate(ids):
            req = requests_[k % len(requests_)]
            assert check(got[rid], req) == alone[k % len(requests_)]
            srv.release(rid)
        st = srv.stats()
        assert st["completed"] == st["submitted"] == len(ids) + len(requests_)
        assert st["held_bytes"] == 0 and st["pending"] == 0
        assert st["decode_batches"] >= 1 and st["encode_batches"] >= 1


def test_uploads_staged_while_a_batch_runs(requests_):
    """Batches of at most two frames and a burst of requests: the ones that
    arrive while a batch is on the GPU are uploaded beside it (on the
    stager's stream, the next batch waiting on their events), and every
    result equals the one the request gets alone."""
    with P.Server() as srv:
        alone = []
        for req in requests_:
            rid = submit(srv, req)
            alone.append(check(drain(srv, 1)[rid], req))
            srv.release(rid)
    with P.Server(max_batch_frames=2) as srv:
        print("case: serve burst, batches of two", flush=True)
        ids = [submit(srv, req) for req in requests_ * 4]
        got = drain(srv, len(ids))
        for k, rid in enumerate(ids):
            assert check(got[rid], requests_[k % len(requests_)]) == alone[k % len(requests_)]
            srv.release(rid)
        st = srv.stats()
        print("stats:", st, flush=True)
        assert st["max_batch_frames"] <= 2
        assert st["staged_ahead"] >= 1
        assert st["held_bytes"] == 0 and st["pending"] == 0


def test_admission_waits_for_release(requests_):
    """With room for one request, the second waits until the first result is
    released."""
    req = requests_[1]
    with P.Server(hbm_budget=64 << 20) as probe:
        rid = submit(probe, req)
        drain(probe, 1)
        need = probe.stats()["peak_held_bytes"]
        probe.release(rid)
    print("case: admission budget", need, flush=True)
    with P.Server(hbm_budget=need + (need // 8)) as srv:
        a = submit(srv, req)
        b = submit(srv, req)
        first = drain(srv, 1)
        assert list(first) == [a]
        time.sleep(0.2)
        assert srv.poll(8, 200) == []
        assert srv.stats()["pending"] == 1
        srv.release(a)
        second = drain(srv, 1)
        assert list(second) == [b]
        check(second[b], req)
        srv.release(b)
        assert srv.stats()["peak_held_bytes"] <= need + need // 8


def test_requests_that_cannot_run_fail_at_once(requests_):
    with P.Server(hbm_budget=2 << 20) as srv:
        big = submit(srv, requests_[3])  # 1280x720 needs more than 2 MiB
        bad = srv.submit_decode(b"\x00\x00\x00\x10icpf")
        refused_cfg = P.Config(P.Profile.Hq)
        refused = srv.submit_encode(refused_cfg, requests_[6][1][1])  # a 4:4:4 frame for a 4:2:2 profile
        got = drain(srv, 3, timeout=10)
        assert isinstance(got[big].error, P.ConfigError)
        want = O.decode_result(b"\x00\x00\x00\x10icpf")
        assert (got[bad].error.kind, got[bad].error.msg) == (want[0], want[1])
        try:
            P.Encoder(refused_cfg).encode(requests_[6][1][1])
            raise AssertionError("expected a Config error")
        except P.ConfigError as e:
            assert got[refused].error.msg == e.msg
        for rid in got:
            srv.release(rid)


def test_close_with_results_not_polled(requests_):
    """Closing the server while results wait to be polled, some polled and
    not released, and one refused at submit frees every request once."""
    for rep in range(3):
        print("case: close with pending results", rep, flush=True)
        srv = P.Server()
        ids = [srv.submit_decode(x) if kind == "decode" else srv.submit_encode(*x) for kind, x, want in requests_]
        ids.append(srv.submit_decode(b"\x00\x00\x00\x08icpf"))  # refused at once
        deadline = time.time() + 60
        while srv.stats()["completed"] < len(ids) and time.time() < deadline:
            time.sleep(0.01)
        assert srv.stats()["completed"] == len(ids)
        if rep == 1:
            got = srv.poll(2, 0)  # polled, never released
            assert len(got) == 2
        srv.close()
