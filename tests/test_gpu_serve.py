"""Serving loop on the GPU: continuous batching, resident spectra, admission by budget.

Each result must be bitwise the same alone and in the mix, and the same as a batched plan's."""
import threading
import time

import numpy as np
import pytest

from alignment_projects_amd import abi, synth, xcorr

pytestmark = pytest.mark.gpu

FIELDS = ("r", "dx", "dy", "peak", "sub_x", "sub_y", "status", "lag_x", "lag_y", "nx", "ny")


def _same(a, b):
    return all(getattr(a, k) == getattr(b, k) for k in FIELDS)


def _grid_tiles(ctx, n=5, w=384, h=64, seed=11):
    """A row of n tiles cut from one field with overlaps; neighbouring requests share tiles."""
    rng = np.random.default_rng(seed)
    F = synth.quantize(synth.bandlimited(rng, h + 16, n * (w // 2) + w + 16, 128, 40), np.uint8)
    tiles = []
    for i in range(n):
        x0 = 8 + i * (w // 2) + int(rng.integers(-3, 4))
        y0 = 8 + int(rng.integers(-3, 4))
        tiles.append(ctx.tile_put(np.ascontiguousarray(F[y0:y0 + h, x0:x0 + w])))
    return tiles


def test_alone_vs_mix_vs_plan(ctx):
    tiles = _grid_tiles(ctx)
    rgn = abi.rgn_big_enough(16, 384 * 64 // 8)
    pairs = []
    for i in range(len(tiles) - 1):
        pairs.append(xcorr.make_pair(tiles[i], tiles[i + 1], rgn=rgn))
        pairs.append(xcorr.make_pair(tiles[i + 1], tiles[i], rgn=rgn, mode=abi.XC_MODE_NCC, radius=300))
    pairs.append(xcorr.make_pair(tiles[0], tiles[2], mode=abi.XC_MODE_S, rgn=rgn, cnt=abi.cnt_enough_points(1000)))
    plan_res = ctx.plan(pairs).run()
    srv = xcorr.Server(ctx, hbm_budget=1 << 30, max_batch=4, streams=3, max_wait_us=100)
    tickets = [srv.submit(p) for p in pairs]
    mix = [srv.wait(t)[0] for t in tickets]
    alone = []
    for p in pairs:
        alone.append(srv.wait(srv.submit(p))[0])
    st = srv.stats()
    srv.close()
    for i in range(len(pairs)):
        assert _same(mix[i], alone[i]), (i, mix[i].as_dict(), alone[i].as_dict())
        assert _same(mix[i], plan_res[i]), (i, mix[i].as_dict(), plan_res[i].as_dict())
        assert mix[i].status == abi.XC_OK
    assert st["completed"] == 2 * len(pairs) and st["spectra_hits"] > 0
    for t in tiles:
        t.release()


def test_concurrent_submitters_and_small_budget(ctx):
    tiles = _grid_tiles(ctx, n=6, seed=12)
    rgn = abi.rgn_big_enough(16, 384 * 64 // 8)
    pairs = [xcorr.make_pair(tiles[i], tiles[(i + 1) % 6], rgn=rgn) for i in range(6)] * 4
    ref = ctx.plan(pairs[:6]).run()
    # a budget that holds only a few requests at a time: admission must wait for memory, not fail
    srv = xcorr.Server(ctx, hbm_budget=24 << 20, max_batch=8, streams=2, max_wait_us=50)
    out = [None] * len(pairs)

    def worker(k0):
        for k in range(k0, len(pairs), 3):
            out[k] = srv.wait(srv.submit(pairs[k]))[0]
            time.sleep(0.001)

    ths = [threading.Thread(target=worker, args=(k,)) for k in range(3)]
    for t in ths:
        t.start()
    for t in ths:
        t.join(timeout=60)
    st = srv.stats()
    srv.close()
    for k, r in enumerate(out):
        assert r is not None and _same(r, ref[k % 6]), k
    assert st["arena_peak"] <= 24 << 20
    for t in tiles:
        t.release()


def test_budget_rejects_oversized_request(ctx):
    a = ctx.tile_put(np.zeros((64, 384), np.uint8) + 7)
    srv = xcorr.Server(ctx, hbm_budget=1 << 20, max_batch=4, streams=1)
    with pytest.raises(xcorr.XcorrError) as e:
        srv.submit(xcorr.make_pair(a, a))
    assert e.value.status == abi.XC_E_BUDGET
    assert srv.stats()["rejected"] == 1
    srv.close()
    a.release()


def test_failed_batch_completes_with_error_and_recovers(ctx, monkeypatch):
    """A batch that fails after its requests left the queue (its pinned staging refused, forced from the
    host by the server's fault-injection setting) completes them with the error instead of leaving their
    waiters blocked, and gives back its memory: the spectra it was to produce are dropped, so the same pair
    submitted again is computed afresh and matches a plan bit for bit."""
    tiles = _grid_tiles(ctx, n=3, seed=13)
    rgn = abi.rgn_big_enough(16, 384 * 64 // 8)
    p = xcorr.make_pair(tiles[0], tiles[1], rgn=rgn)
    ref = ctx.plan([p]).run()[0]
    monkeypatch.setenv("XC_SERVE_FAIL_PINNED", "1")
    srv = xcorr.Server(ctx, hbm_budget=1 << 30, max_batch=4, streams=1, max_wait_us=0)
    monkeypatch.delenv("XC_SERVE_FAIL_PINNED")
    r, _ = srv.wait(srv.submit(p), raise_errors=False)
    assert r.status == abi.XC_E_NOMEM
    st = srv.stats()
    assert st["failed"] == 1 and st["arena_bytes"] == 0, st
    again, _ = srv.wait(srv.submit(p))
    assert _same(again, ref), (again.as_dict(), ref.as_dict())
    st = srv.stats()
    assert st["completed"] == 1 and st["spectra_misses"] == 4, st
    srv.close()
    for t in tiles:
        t.release()


def test_deadline_times_out_queued_requests(ctx):
    """Requests that wait in the queue past the server's deadline complete with XC_E_TIMEOUT (forced from
    the host: a 1 us deadline against a 200 ms wait for batch company); with a generous deadline the same
    requests run."""
    tiles = _grid_tiles(ctx, n=3, seed=14)
    rgn = abi.rgn_big_enough(16, 384 * 64 // 8)
    pairs = [xcorr.make_pair(tiles[0], tiles[1], rgn=rgn), xcorr.make_pair(tiles[1], tiles[2], rgn=rgn)]
    srv = xcorr.Server(ctx, hbm_budget=1 << 30, max_batch=64, streams=1, max_wait_us=200000, deadline_us=1)
    out = [srv.wait(t, raise_errors=False)[0] for t in [srv.submit(q) for q in pairs]]
    assert [r.status for r in out] == [abi.XC_E_TIMEOUT] * 2
    assert srv.stats()["timed_out"] == 2
    srv.close()
    ref = ctx.plan(pairs).run()
    srv = xcorr.Server(ctx, hbm_budget=1 << 30, max_batch=64, streams=1, max_wait_us=1000, deadline_us=10_000_000)
    out = [srv.wait(t)[0] for t in [srv.submit(q) for q in pairs]]
    assert all(_same(a, b) for a, b in zip(out, ref))
    assert srv.stats()["timed_out"] == 0
    srv.close()
    for t in tiles:
        t.release()


def test_cancel_queued_request(ctx):
    """A request cancelled while queued completes with XC_E_CANCELLED; the others of its would-be batch
    run as usual; a finished request can no longer be cancelled."""
    tiles = _grid_tiles(ctx, n=4, seed=15)
    rgn = abi.rgn_big_enough(16, 384 * 64 // 8)
    pairs = [xcorr.make_pair(tiles[i], tiles[i + 1], rgn=rgn) for i in range(3)]
    ref = ctx.plan(pairs).run()
    srv = xcorr.Server(ctx, hbm_budget=1 << 30, max_batch=64, streams=1, max_wait_us=2_000_000)
    tk = [srv.submit(q) for q in pairs]
    assert srv.cancel(tk[1]) == abi.XC_OK
    assert srv.cancel(tk[1]) == abi.XC_E_NOTREADY  # done (cancelled) already
    r1, _ = srv.wait(tk[1], raise_errors=False)
    assert r1.status == abi.XC_E_CANCELLED
    r0, _ = srv.wait(tk[0])
    while srv.poll(tk[2]) is None:
        time.sleep(0.001)
    assert srv.cancel(tk[2]) == abi.XC_E_NOTREADY
    r2, _ = srv.wait(tk[2])
    assert _same(r0, ref[0]) and _same(r2, ref[2])
    assert srv.cancel(tk[0]) == abi.XC_E_ARG  # collected by wait: unknown now
    st = srv.stats()
    assert st["cancelled"] == 1 and st["completed"] == 2, st
    srv.close()
    for t in tiles:
        t.release()


def test_masked_odd_width_tiles_serve_matches_plan(ctx):
    """Masked tiles and unmasked tiles of odd width (rows at odd addresses): a plan transforms their rows with
    the 8-bit row r2c's general form, the server with the general r2c; every result bitwise the same."""
    rng = np.random.default_rng(31)
    pairs, tiles = [], []
    for k, (w, h) in enumerate([(301, 97), (257, 130), (384, 64), (199, 211)]):
        F = synth.quantize(synth.bandlimited(rng, h + 24, w + 24, 128, 40), np.uint8)
        A = np.ascontiguousarray(F[4:4 + h, 6:6 + w])
        B = np.ascontiguousarray(F[9:9 + h, 2:2 + w])
        if k % 2 == 0:  # an elliptical hole in each mask
            yy, xx = np.mgrid[0:h, 0:w]
            ma = (((xx - w / 3) / (w / 6)) ** 2 + ((yy - h / 2) / (h / 5)) ** 2 > 1).astype(np.uint8)
            mb = (((xx - 2 * w / 3) / (w / 7)) ** 2 + ((yy - h / 3) / (h / 6)) ** 2 > 1).astype(np.uint8)
            ta, tb = ctx.tile_put(A, mask=ma), ctx.tile_put(B, mask=mb)
        else:
            ta, tb = ctx.tile_put(A), ctx.tile_put(B)
        tiles += [ta, tb]
        pairs.append(xcorr.make_pair(ta, tb, cnt=abi.cnt_enough_points(w * h // 4)))
        pairs.append(xcorr.make_pair(tb, ta, rgn=abi.rgn_min(w // 2, h // 2)))
    plan = ctx.plan(pairs)
    assert any(b["r2c_np"] > 0 for b in plan.buckets())  # the 8-bit row r2c
    ref = plan.run()
    srv = xcorr.Server(ctx, hbm_budget=1 << 30, max_batch=4, streams=2, max_wait_us=100)
    got = [srv.wait(srv.submit(p))[0] for p in pairs]
    srv.close()
    for i in range(len(pairs)):
        assert _same(got[i], ref[i]), (i, got[i].as_dict(), ref[i].as_dict())
        assert ref[i].status == abi.XC_OK
    plan.destroy()
    for t in tiles:
        t.release()
