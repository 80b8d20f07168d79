"""The serving scheduler on the CPU, over a host-memory device.

The scheduler of lyap_server_* (csrc/serve_core.cpp) is the code the GPU
server runs; here its device is the fake of tests/fake_device/fake_server.cpp:
memory is host memory, a launch writes every pixel's lambda as its index and
every voxel as its index, counts how often each was written, and finishes on
a virtual clock (automatic mode) or when the test says so (manual mode).
So admission under the HBM budget, deadline order, starvation, out-of-order
completion, cancellation, queue and memory limits, a launch that overruns
its slice and launch failures are checked deterministically without a GPU.
The reference has no request queue (one view per launch,
lyap_interactive.cu:711; one volume per process, lyap_calculate.cu:72).
"""
import ctypes as C
import functools
import os
import subprocess

import numpy as np
import pytest

import lyapunov3d_amd as lyap

HERE = os.path.dirname(os.path.abspath(__file__))
FAKE_DIR = os.path.join(HERE, "fake_device")


class FakeOptions(C.Structure):
    _fields_ = [("manual", C.c_int32), ("reserved", C.c_uint32), ("device_bytes", C.c_size_t),
                ("ms_per_launch", C.c_double), ("tick_ms", C.c_double)]


class LogEntry(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("slot", C.c_uint32), ("launch", C.c_uint32), ("kind", C.c_int32),
                ("begin", C.c_uint32), ("count", C.c_uint32), ("t", C.c_double)]


@pytest.fixture(scope="module")
def lib():
    subprocess.run(["make", "-s", "-C", FAKE_DIR, os.path.join(FAKE_DIR, "libfake_server.so")], check=True)
    L = C.CDLL(os.path.join(FAKE_DIR, "libfake_server.so"))
    vp = C.c_void_p
    L.fake_server_create.argtypes = [vp, vp, C.POINTER(vp)]
    L.fake_last_error.restype = C.c_char_p
    L.fake_last_error.argtypes = [vp]
    for name in ("fake_slot_count",):
        getattr(L, name).argtypes = [vp]
    L.fake_slot_kind.argtypes = [vp, C.c_uint]
    L.fake_running.argtypes = [vp, C.c_uint]
    L.fake_complete.argtypes = [vp, C.c_uint, C.c_int]
    L.fake_fail_next_launch.argtypes = [vp, C.c_int]
    L.fake_push_overrun.argtypes = [vp, C.c_double]
    L.fake_advance.argtypes = [vp, C.c_double]
    L.fake_rate.restype = C.c_double
    L.fake_rate.argtypes = [vp, C.c_int]
    for name in ("fake_pool_bytes", "fake_device_in_use"):
        getattr(L, name).restype = C.c_size_t
        getattr(L, name).argtypes = [vp]
    L.fake_live_bytes_all.restype = C.c_size_t
    L.fake_log.restype = C.c_uint
    L.fake_log.argtypes = [vp, vp, C.c_uint]
    L.fake_coverage.restype = C.c_size_t
    L.fake_coverage.argtypes = [vp, C.c_uint64, vp, C.c_size_t]
    L.lyap_server_submit.argtypes = [vp, vp, C.POINTER(C.c_uint64)]
    L.lyap_server_progress.argtypes = [vp, C.c_double]
    L.lyap_server_poll.argtypes = [vp, C.c_uint64, vp]
    L.lyap_server_release.argtypes = [vp, C.c_uint64]
    L.lyap_server_cancel.argtypes = [vp, C.c_uint64]
    L.lyap_server_clock_ms.restype = C.c_double
    L.lyap_server_clock_ms.argtypes = [vp]
    L.lyap_server_stats.argtypes = [vp, vp]
    L.lyap_server_finished.argtypes = [vp, vp, C.c_uint, C.POINTER(C.c_uint)]
    L.lyap_server_drain.argtypes = [vp, C.c_double]
    L.lyap_server_destroy.argtypes = [vp]
    return L


class Fake:
    """A scheduler over the fake device; calls return the library's status."""

    def __init__(self, L, manual=False, device_bytes=0, ms_per_launch=0.0, **cfg):
        self.L = L
        cfg.setdefault("slice_ms", 1e-6)       # one 64-ray piece or one z slice per launch
        cfg.setdefault("chunk_rays", 64)
        self.cfg = lyap.server_config(**cfg)
        opt = FakeOptions(int(manual), 0, device_bytes, ms_per_launch, 0.0)
        h = C.c_void_p()
        rc = L.fake_server_create(C.byref(self.cfg), C.byref(opt), C.byref(h))
        assert rc == lyap.OK, rc
        self.h = h
        self.keep = []                          # caller-owned outputs stay alive

    def submit(self, req):
        t = C.c_uint64()
        rc = self.L.lyap_server_submit(self.h, C.byref(req), C.byref(t))
        return rc, t.value

    def ok_submit(self, req):
        rc, t = self.submit(req)
        assert rc == lyap.OK, (rc, self.error())
        return t

    def progress(self, wait_ms=0.0):
        return self.L.lyap_server_progress(self.h, wait_ms)

    def poll(self, t):
        r = lyap.Result()
        assert self.L.lyap_server_poll(self.h, t, C.byref(r)) == lyap.OK
        return r

    def stats(self):
        s = lyap.ServerStats()
        assert self.L.lyap_server_stats(self.h, C.byref(s)) == lyap.OK
        return s

    def finished(self):
        buf = (C.c_uint64 * 256)()
        n = C.c_uint()
        assert self.L.lyap_server_finished(self.h, buf, 256, C.byref(n)) == lyap.OK
        return [buf[i] for i in range(n.value)]

    def release(self, t):
        return self.L.lyap_server_release(self.h, t)

    def cancel(self, t):
        return self.L.lyap_server_cancel(self.h, t)

    def drain(self, timeout_ms=0.0):
        return self.L.lyap_server_drain(self.h, timeout_ms)

    def now(self):
        return self.L.lyap_server_clock_ms(self.h)

    def complete(self, slot, status=lyap.OK):
        return self.L.fake_complete(self.h, slot, status)

    def running(self, slot):
        return bool(self.L.fake_running(self.h, slot))

    def nslots(self):
        return self.L.fake_slot_count(self.h)

    def log(self):
        n = self.L.fake_log(self.h, None, 0)
        buf = (LogEntry * max(n, 1))()
        self.L.fake_log(self.h, buf, n)
        return [buf[i] for i in range(n)]

    def coverage(self, t):
        n = self.L.fake_coverage(self.h, t, None, 0)
        out = np.zeros(n, np.uint8)
        if n:
            self.L.fake_coverage(self.h, t, out.ctypes.data, n)
        return out

    def error(self):
        return self.L.fake_last_error(self.h).decode()

    def close(self):
        if self.h:
            self.L.lyap_server_destroy(self.h)
            self.h = None


SEQ = "ABBAC"


@functools.lru_cache(maxsize=None)
def params():
    # built on first use, not at import: a GPU run that collects this module
    # must not load the library before torch has set up the device
    return lyap.make_params(settle=40, accum=200)


def view(w, h, deadline=0.0, lam=None, rgba=None):
    v = lyap.make_view(lyap.LyapCam(), params(), SEQ, w, h, rgba.ctypes.data if rgba is not None else 0,
                       lam=lam.ctypes.data if lam is not None else 0)
    return lyap.view_request(v, deadline_ms=deadline)


def volume(n, deadline=0.0, out=None):
    return lyap.volume_request(params(), SEQ, (2, 2, 2), (4, 4, 4), n,
                               out=out.ctypes.data if out is not None else 0, deadline_ms=deadline)


def host_array(addr, dtype, count):
    buf = (C.c_char * (np.dtype(dtype).itemsize * count)).from_address(addr)
    return np.frombuffer(buf, dtype=dtype, count=count).copy()


def first_issue_order(fk):
    seen = []
    for e in fk.log():
        if e.ticket not in seen:
            seen.append(e.ticket)
    return seen


def test_every_pixel_and_voxel_written_once(lib):
    """Views of ragged sizes and volumes, server-owned and caller-owned, cut
    into 64-ray pieces and single z slices over three mixed slots: every
    pixel and voxel is written exactly once, at its own place."""
    fk = Fake(lib, render_slots=2, volume_slots=1)
    own_lam = np.full(40 * 24, -1, np.float32)
    own_rgba = np.zeros(40 * 24 * 4, np.uint8)
    own_vol = np.full(6 * 5 * 7, -1, np.float32)
    sizes = [(17, 9), (64, 64), (100, 37), (8, 8), (1, 1), (9, 65)]
    views = {fk.ok_submit(view(w, h, deadline=3.0 * i)): (w, h) for i, (w, h) in enumerate(sizes)}
    t_own = fk.ok_submit(view(40, 24, deadline=1.0, lam=own_lam, rgba=own_rgba))
    vols = {fk.ok_submit(volume((5, 7, 11), deadline=2.0)): (5, 7, 11),
            fk.ok_submit(volume((16, 16, 16))): (16, 16, 16)}
    t_vown = fk.ok_submit(volume((6, 5, 7), out=own_vol))
    assert fk.drain() == lyap.OK
    st = fk.stats()
    n = len(views) + len(vols) + 2
    assert (st.submitted, st.completed, st.queued, st.active) == (n, n, 0, 0)
    pieces = fk.log()
    assert len(pieces) > n and st.launches > n
    assert {e.slot for e in pieces} == {0, 1, 2}
    for e in pieces:
        if e.kind == lyap.REQ_VIEW:
            assert e.begin % 64 == 0 and e.count % 64 == 0
    for t, (w, h) in views.items():
        r = fk.poll(t)
        assert r.status == lyap.OK and r.rays == w * h
        assert (fk.coverage(t) == 1).all()
        lam = host_array(r.lambda_, np.float32, w * h)
        assert np.array_equal(lam, np.arange(w * h, dtype=np.float32))
        rgba = host_array(r.rgba, np.uint8, 4 * w * h).reshape(-1, 4)
        assert (rgba[:, 0] == t & 0xff).all() and (rgba[:, 2] == 7).all()
    assert (fk.coverage(t_own) == 1).all()
    assert np.array_equal(own_lam, np.arange(40 * 24, dtype=np.float32))
    assert (own_rgba.reshape(-1, 4)[:, 3] == 255).all()
    for t, shape in vols.items():
        r = fk.poll(t)
        nv = int(np.prod(shape))
        assert r.status == lyap.OK and r.voxels == nv and r.evals == nv
        assert (fk.coverage(t) == 1).all()
        assert np.array_equal(host_array(r.lambda_, np.float32, nv), np.arange(nv, dtype=np.float32))
    assert np.array_equal(own_vol, np.arange(own_vol.size, dtype=np.float32))
    assert (fk.coverage(t_vown) == 1).all()
    # views and volumes shared launches
    by_launch = {}
    for e in pieces:
        by_launch.setdefault(e.launch, set()).add(e.kind)
    assert st.resident_bytes > 0
    for t in list(views) + list(vols) + [t_own, t_vown]:
        assert fk.release(t) == lyap.OK
    assert fk.stats().resident_bytes == 0
    assert lib.fake_pool_bytes(fk.h) > 0        # buffers go back to the pool, not freed
    fk.close()
    assert lib.fake_live_bytes_all() == 0


def test_separate_volume_slots(lib):
    """LYAP_SERVER_SEPARATE_VOLUMES: render slots carry only view pieces and
    volume slots only z slabs."""
    fk = Fake(lib, render_slots=2, volume_slots=2, flags=lyap.SERVER_SEPARATE_VOLUMES)
    kinds = [lib.fake_slot_kind(fk.h, i) for i in range(fk.nslots())]
    assert kinds == [1, 1, 2, 2]
    for i in range(4):
        fk.ok_submit(view(24, 24, deadline=1.0 + i))
        fk.ok_submit(volume(6, deadline=1.0 + i))
    assert fk.drain() == lyap.OK
    for e in fk.log():
        assert (e.slot < 2) == (e.kind == lyap.REQ_VIEW)
    fk.close()


def test_earliest_deadline_first(lib):
    """With both slots busy, queued requests start in deadline order; equal
    deadlines in submission order; requests without a deadline last."""
    fk = Fake(lib, manual=True, render_slots=1, volume_slots=1)
    b0, b1 = fk.ok_submit(view(8, 8, 0.5)), fk.ok_submit(view(8, 8, 0.5))
    fk.progress()
    assert fk.running(0) and fk.running(1)
    none = fk.ok_submit(view(8, 8))
    late = fk.ok_submit(view(8, 8, 40.0))
    early1 = fk.ok_submit(view(8, 8, 10.0))
    mid = fk.ok_submit(volume((4, 4, 1), 30.0))
    early2 = fk.ok_submit(view(8, 8, 10.0))
    for _ in range(5):
        fk.complete(0)
        fk.progress()
    while fk.drain(1.0) == lyap.ERR_PENDING:
        for s in range(2):
            if fk.running(s):
                fk.complete(s)
    assert first_issue_order(fk) == [b0, b1, early1, early2, mid, late, none]
    fk.close()


def test_deadline_order_does_not_starve_a_volume(lib):
    """A stream of short-deadline views keeps both slots busy; a volume with a
    longer deadline still gets its slices once the views' deadlines pass its
    own, and finishes while the stream goes on."""
    fk = Fake(lib, ms_per_launch=0.5, render_slots=1, volume_slots=1)
    vol = fk.ok_submit(volume((8, 8, 6), deadline=5.0))
    views = []
    for _ in range(200):
        views.append(fk.ok_submit(view(8, 8, deadline=1.0)))
        assert fk.progress(0.5) == lyap.OK
        if fk.poll(vol).status == lyap.OK:
            break
    r = fk.poll(vol)
    assert r.status == lyap.OK, "the volume starved"
    assert r.finish_ms < 20.0
    assert (fk.coverage(vol) == 1).all()
    # the stream kept going while the volume ran
    vol_t = [e.t for e in fk.log() if e.ticket == vol]
    assert any(e.kind == lyap.REQ_VIEW and vol_t[0] <= e.t <= vol_t[-1] for e in fk.log())
    fk.close()


def test_budget_holds_later_requests_behind_a_waiting_one(lib):
    """A volume that does not fit the HBM budget yet is not passed over by
    later small views that would fit (no starvation); a request with an
    earlier deadline that fits is still admitted; after a release the volume
    goes first."""
    view_bytes = 8 * 8 * (4 + 4)
    vol_bytes = 8 * 8 * 8 * 4
    fk = Fake(lib, manual=True, hbm_budget_bytes=vol_bytes + view_bytes // 2, render_slots=1, volume_slots=1)
    a = fk.ok_submit(view(8, 8, 1.0))
    fk.progress()
    v = fk.ok_submit(volume(8, 5.0))
    b = fk.ok_submit(view(8, 8, 10.0))
    c = fk.ok_submit(view(8, 8, 10.0))
    fk.progress()
    st = fk.stats()
    assert (st.queued, st.active) == (3, 1), "later views were admitted ahead of the waiting volume"
    fk.complete(0)
    fk.progress()
    assert fk.poll(a).status == lyap.OK
    assert fk.stats().queued == 3             # a's outputs are still resident
    assert fk.drain() == lyap.ERR_BUDGET      # nothing runs, nothing fits until a release
    assert fk.release(a) == lyap.OK
    fk.progress()
    issued = [e.ticket for e in fk.log()]
    assert v in issued and b not in issued
    order = []
    for _ in range(100):
        if fk.drain(1.0) == lyap.OK:
            break
        for s in range(2):
            if fk.running(s):
                fk.complete(s)
        for t in fk.finished():
            order.append(t)
            fk.release(t)
    order += fk.finished()
    assert order.index(v) < order.index(b) < order.index(c)
    assert fk.stats().peak_resident_bytes <= vol_bytes + view_bytes // 2
    fk.close()


def test_earlier_deadline_may_pass_a_waiting_request(lib):
    fk = Fake(lib, manual=True, hbm_budget_bytes=4096 + 512, render_slots=1, volume_slots=1)
    a = fk.ok_submit(view(8, 8, 1.0))
    fk.progress()
    v = fk.ok_submit(volume(8, 50.0))       # 2048 B: does not fit next to a
    d = fk.ok_submit(view(8, 8, 2.0))       # earlier deadline, fits
    fk.progress()
    issued = [e.ticket for e in fk.log()]
    assert d in issued and v not in issued and fk.stats().active == 2
    assert fk.poll(v).status == lyap.ERR_PENDING
    status = {}
    for _ in range(100):
        for s in range(2):
            if fk.running(s):
                fk.complete(s)
        fk.progress()
        for x in fk.finished():
            status[x] = fk.poll(x).status
            fk.release(x)
        if len(status) == 3:
            break
    assert status == {a: lyap.OK, d: lyap.OK, v: lyap.OK}
    fk.close()


def test_out_of_order_completion(lib):
    """Pieces of one view on two slots finish in either order; the view ends
    when its last piece does, and lyap_server_finished lists requests in the
    order they ended."""
    fk = Fake(lib, manual=True, render_slots=1, volume_slots=1)
    t = fk.ok_submit(view(16, 16, 1.0))          # four 64-ray pieces
    u = fk.ok_submit(view(8, 8, 2.0))
    fk.progress()
    assert fk.running(0) and fk.running(1)
    for _ in range(3):                           # slot 1 runs pieces 1, 2, 3 while slot 0 holds piece 0
        assert fk.complete(1) == lyap.OK
        fk.progress()
    assert fk.poll(t).status == lyap.ERR_PENDING
    assert [(e.slot, e.begin) for e in fk.log() if e.ticket == t] == [(0, 0), (1, 64), (1, 128), (1, 192)]
    fk.complete(1)                               # u's piece went to slot 1 after t's last
    fk.progress()
    assert fk.poll(u).status == lyap.OK and fk.poll(t).status == lyap.ERR_PENDING
    fk.complete(0)
    fk.progress()
    assert fk.poll(t).status == lyap.OK
    assert fk.finished() == [u, t]
    assert (fk.coverage(t) == 1).all()
    fk.close()


def test_cancel(lib):
    """Cancelling stops further pieces; the request ends CANCELLED once its
    pieces on the device are done. A queued request ends at once, and never
    takes memory."""
    fk = Fake(lib, manual=True, render_slots=1, volume_slots=1)
    t = fk.ok_submit(view(32, 8, 1.0))           # four pieces
    fk.progress()
    q = fk.ok_submit(volume(8, 9.0))
    assert fk.cancel(t) == lyap.OK
    assert fk.cancel(q) == lyap.OK
    assert fk.poll(q).status == lyap.ERR_CANCELLED and fk.poll(q).start_ms == 0
    assert fk.poll(t).status == lyap.ERR_PENDING
    fk.complete(0)
    fk.progress()
    assert fk.poll(t).status == lyap.ERR_PENDING
    fk.complete(1)
    fk.progress()
    assert fk.poll(t).status == lyap.ERR_CANCELLED
    assert len([e for e in fk.log() if e.ticket == t]) == 2
    assert q not in [e.ticket for e in fk.log()]
    st = fk.stats()
    assert (st.cancelled, st.completed, st.queued, st.active) == (2, 0, 0, 0)
    assert fk.cancel(t) == lyap.OK               # finished: no-op
    assert fk.cancel(12345) == lyap.ERR_NOT_FOUND
    assert fk.release(t) == lyap.OK and fk.stats().resident_bytes == 0
    assert fk.release(t) == lyap.ERR_NOT_FOUND
    fk.close()


def test_release_of_a_running_request_is_refused(lib):
    fk = Fake(lib, manual=True)
    t = fk.ok_submit(view(8, 8))
    fk.progress()
    assert fk.release(t) == lyap.ERR_PENDING
    fk.close()


def test_launch_failure_at_issue(lib):
    """A launch that cannot be issued ends the requests it carried with its
    error (after their other pieces finish); the server goes on."""
    fk = Fake(lib, manual=True, render_slots=1, volume_slots=1)
    lib.fake_fail_next_launch(fk.h, lyap.OK)
    lib.fake_fail_next_launch(fk.h, lyap.ERR_DEVICE)
    t = fk.ok_submit(view(16, 8, 1.0))           # piece 0 on slot 0, piece 1 fails on slot 1
    ok = fk.ok_submit(view(8, 8, 5.0))
    assert fk.progress() == lyap.ERR_DEVICE
    assert "injected launch failure" in fk.error()
    assert fk.poll(t).status == lyap.ERR_PENDING  # piece 0 still on the device
    fk.complete(0)
    fk.progress()
    assert fk.poll(t).status == lyap.ERR_DEVICE
    while fk.poll(ok).status == lyap.ERR_PENDING:
        for s in range(2):
            if fk.running(s):
                fk.complete(s)
        fk.progress()
    assert fk.poll(ok).status == lyap.OK
    st = fk.stats()
    assert (st.failed, st.completed) == (1, 1)
    fk.close()


def test_device_error_at_retire(lib):
    """A launch that fails on the device ends its requests with
    LYAP_ERR_DEVICE, also one whose other piece succeeded."""
    fk = Fake(lib, manual=True, render_slots=1, volume_slots=1)
    t = fk.ok_submit(view(16, 8, 1.0))
    fk.progress()
    fk.complete(0, lyap.OK)
    fk.complete(1, lyap.ERR_DEVICE)
    assert fk.progress() == lyap.ERR_DEVICE
    assert fk.poll(t).status == lyap.ERR_DEVICE
    assert fk.stats().failed == 1
    fk.close()


def test_device_out_of_memory(lib):
    """The device has less memory than the budget: pooled buffers are freed to
    make room; a request that cannot fit with nothing else resident ends with
    LYAP_ERR_DEVICE; the others complete."""
    view_bytes = 16 * 16 * 8
    fk = Fake(lib, device_bytes=3 * view_bytes, render_slots=1, volume_slots=1)
    ts = [fk.ok_submit(view(16, 16, 1.0 + i)) for i in range(3)]
    assert fk.drain() == lyap.OK
    for t in ts:
        fk.release(t)
    assert lib.fake_pool_bytes(fk.h) == 3 * view_bytes
    big = fk.ok_submit(view(24, 24, 1.0))        # needs 4608 B: only after the pool is freed
    assert fk.drain() == lyap.OK
    assert fk.poll(big).status == lyap.OK
    assert lib.fake_pool_bytes(fk.h) == 0
    fk.release(big)
    huge = fk.ok_submit(volume(13, 1.0))          # 8788 B: more than the device
    small = fk.ok_submit(view(8, 8, 2.0))
    assert fk.drain() == lyap.OK
    assert fk.poll(huge).status == lyap.ERR_DEVICE
    assert fk.poll(small).status == lyap.OK
    assert "out of device memory" in fk.error()
    fk.close()


def test_queue_limit(lib):
    fk = Fake(lib, manual=True, max_queued=2, render_slots=1, volume_slots=1)
    fk.ok_submit(view(8, 8))
    fk.ok_submit(view(8, 8))
    fk.progress()                                # both admitted and running
    fk.ok_submit(view(8, 8))
    fk.ok_submit(view(8, 8))
    rc, _ = fk.submit(view(8, 8))
    assert rc == lyap.ERR_QUEUE_FULL and "wait for admission" in fk.error()
    fk.complete(0)
    fk.progress()                                # one queued request admitted
    assert fk.submit(view(8, 8))[0] == lyap.OK
    fk.close()


def test_overrun_shrinks_the_next_slices(lib):
    """A launch that takes 20x its estimate raises the measured rate by
    0.75 + 0.25 * 20, so later slices take fewer rays; the other slot keeps
    issuing while the long launch runs."""
    fk = Fake(lib, render_slots=1, volume_slots=1, slice_ms=0.02, chunk_rays=64)
    ts = [fk.ok_submit(view(512, 512, 1.0)) for _ in range(2)]
    fk.progress()
    r0 = lib.fake_rate(fk.h, lyap.REQ_VIEW)
    # the first launches finish at their estimate: the rate stays
    lib.fake_advance(fk.h, 10.0)
    fk.progress()
    assert lib.fake_rate(fk.h, lyap.REQ_VIEW) == pytest.approx(r0)
    before = max(e.count for e in fk.log())
    lib.fake_push_overrun(fk.h, 20.0)
    lib.fake_advance(fk.h, 10.0)
    fk.progress()                                 # retires both, issues the overrunning launch first
    long_launch = fk.log()[-2]
    assert long_launch.t == pytest.approx(fk.now())
    for _ in range(100):
        if lib.fake_rate(fk.h, lyap.REQ_VIEW) != pytest.approx(r0):
            break
        assert fk.progress(1.0) == lyap.OK
    r1 = lib.fake_rate(fk.h, lyap.REQ_VIEW)
    assert r1 == pytest.approx(r0 * (0.75 + 0.25 * 20.0), rel=1e-9)
    end = long_launch.t + 20.0 * 0.02
    others = [e for e in fk.log() if e.slot != long_launch.slot and long_launch.t < e.t < end]
    assert len(others) >= 10, "the other slot waited for the long launch"
    fk.progress(1.0)
    after = [e.count for e in fk.log() if e.t >= end - 1e-9]
    assert after and max(after) < before
    assert fk.drain() == lyap.OK
    for t in ts:
        assert (fk.coverage(t) == 1).all()
    fk.close()


def test_submission_errors(lib):
    fk = Fake(lib, hbm_budget_bytes=1 << 12)
    assert fk.submit(view(0, 8))[0] == lyap.ERR_INVALID_ARGUMENT
    assert fk.error() == "empty view"
    bad = view(8, 8)
    bad.view.seq[0] = 7
    assert fk.submit(bad)[0] == lyap.ERR_BAD_SEQUENCE
    assert fk.submit(volume(64))[0] == lyap.ERR_BUDGET                 # 1 MiB > the whole budget
    assert fk.submit(volume((4096, 4096, 4096)))[0] == lyap.ERR_UNSUPPORTED
    r = view(8, 8)
    r.kind = 5
    assert fk.submit(r)[0] == lyap.ERR_INVALID_ARGUMENT
    assert fk.stats().submitted == 0
    fk.close()


def test_config_checks(lib):
    for kw in ({"render_slots": 5}, {"chunk_rays": 100}, {"flags": 8}):
        cfg = lyap.server_config(**kw)
        h = C.c_void_p()
        assert lib.fake_server_create(C.byref(cfg), None, C.byref(h)) == lyap.ERR_INVALID_ARGUMENT
