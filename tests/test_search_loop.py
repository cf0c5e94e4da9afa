"""search_with (the launch controller above the Engine) on a scripted engine and clock,
mirroring the reference's MockEngine tests (src/service/gpu.rs:3219-3364): iteration
growth, contiguous index ranges, hits, bench mode, cancel, time limit, progress and error
propagation. Runs the C++ header that the C ABI uses (meikey_amd/csrc/mk_search.h)."""
import ctypes

from tests import hostcheck as H

LAUNCH = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64))
CLOCK = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p)
REPORT = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double)
CANCEL = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)


class Script:
    """A fake engine: each launch of `iters` takes seconds_per_iter * iters of fake time."""

    def __init__(self, kpi=1000, seconds_per_iter=0.0, hit_at=None, fail_at=None, cancel_after=None):
        self.kpi, self.spi, self.hit_at, self.fail_at, self.cancel_after = kpi, seconds_per_iter, hit_at, fail_at, cancel_after
        self.t = 0.0
        self.calls = []
        self.reports = []

    def run(self, bench=False, target=0.2, time_limit=0.0, interval=5.0):
        def launch(_, base, iters, found, index):
            self.calls.append((base, iters))
            if self.fail_at is not None and len(self.calls) == self.fail_at:
                return 5
            self.t += self.spi * iters
            if self.hit_at is not None and base <= self.hit_at < base + self.kpi * iters:
                found[0] = 1
                index[0] = self.hit_at
            return 0

        def clock(_):
            return self.t

        def report(_, tries, elapsed):
            self.reports.append((tries, elapsed))

        def cancel(_):
            return int(self.cancel_after is not None and len(self.calls) >= self.cancel_after)

        out = (ctypes.c_uint64 * 5)()
        self._cb = (LAUNCH(launch), CLOCK(clock), REPORT(report), CANCEL(cancel))
        L = H.lib()
        L.hc_search_with.argtypes = [LAUNCH, CLOCK, REPORT, CANCEL, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        L.hc_search_with(*self._cb, None, self.kpi, int(bench), target, time_limit, interval, out)
        return {"status": out[0], "found": out[1], "index": out[2], "tries": out[3], "launches": out[4]}


def test_iterations_grow_by_4_up_to_2_16_and_ranges_are_contiguous():
    s = Script(kpi=1000, seconds_per_iter=0.0, cancel_after=12)
    r = s.run()
    iters = [c[1] for c in s.calls]
    # instant launches: x4 each time (scale clamp), capped at 2^16
    assert iters[:9] == [1, 4, 16, 64, 256, 1024, 4096, 16384, 65536]
    assert all(i == 65536 for i in iters[8:])
    base = 0
    for b, i in s.calls:
        assert b == base
        base += 1000 * i
    assert r["tries"] == base and r["found"] == 0


def test_iterations_follow_the_target():
    # each iteration takes 10 ms of fake time; target 200 ms -> about 20 iterations
    s = Script(kpi=10, seconds_per_iter=0.01, cancel_after=10)
    s.run(target=0.2)
    iters = [c[1] for c in s.calls]
    assert iters[:3] == [1, 4, 16]  # x4 per step at most
    assert all(18 <= i <= 22 for i in iters[4:])


def test_slow_launches_shrink_by_half_at_most():
    s = Script(kpi=10, seconds_per_iter=0.0, cancel_after=3)
    s.run()
    s2 = Script(kpi=10, seconds_per_iter=10.0, cancel_after=4)
    s2.run(target=0.2)
    assert [c[1] for c in s2.calls] == [1, 1, 1, 1]  # never below 1


def test_hit_stops_unless_bench():
    s = Script(kpi=1000, hit_at=5000)
    r = s.run()
    assert r["found"] == 1 and r["index"] == 5000 and r["tries"] == sum(1000 * i for _, i in s.calls)
    b = Script(kpi=1000, hit_at=5000, seconds_per_iter=0.001)
    r = b.run(bench=True, time_limit=1.0)
    assert r["found"] == 0 and len(b.calls) > 3 and b.t >= 1.0


def test_cancel_and_time_limit():
    s = Script(kpi=10, cancel_after=3)
    r = s.run()
    assert len(s.calls) == 3 and r["found"] == 0
    t = Script(kpi=10, seconds_per_iter=0.05)
    r = t.run(time_limit=2.0)
    assert t.t >= 2.0 and t.t - 0.05 * t.calls[-1][1] < 2.0


def test_progress_reports_every_interval():
    s = Script(kpi=10, seconds_per_iter=0.01, cancel_after=40)
    s.run(interval=1.0)
    assert len(s.reports) >= 3
    assert all(b[1] > a[1] for a, b in zip(s.reports, s.reports[1:]))


def test_engine_errors_propagate():
    s = Script(kpi=10, fail_at=3)
    r = s.run()
    assert r["status"] == 5 and len(s.calls) == 3


def test_search_with_on_a_host_engine(oracle):
    """The reference also runs search_with over a host-compiled engine (gpu.rs:2935-3215).
    Here each launch computes its keys with the kernel's own code on the CPU: the lane
    tile (comb W = 8) flags sign-unknown candidates, the full key with its x sign comes
    from the host-compiled comb, and the host confirmation accepts a hit. The controller
    must return the oracle's first hit of the index sequence."""
    import time
    w, kpi = 8, 64
    tab = H.table(w)
    key = bytes(range(7, 39))
    needle, fold, mode = b"k9", True, 0
    calls = []

    def launch(_, base, iters, found, index):
        calls.append((base, iters))
        for it in range(iters):
            first = base + it * kpi
            _, cand = H.lane(tab, w, key, 0, first, kpi, needle, fold, mode)
            for j in range(kpi):
                if not cand[j]:
                    continue
                i = first + j
                seed = H.seed(key, 0, i)
                pk = H.full_pk(tab, w, H.secret_scalar(seed))
                if H.confirm(key, 0, i, seed, pk, needle, fold, mode) == 1:
                    found[0], index[0] = 1, i
                    return 0
        return 0

    t0 = time.perf_counter()
    out = (ctypes.c_uint64 * 5)()
    cbs = (LAUNCH(launch), CLOCK(lambda _: time.perf_counter() - t0), REPORT(lambda *a: None), CANCEL(lambda _: 0))
    L = H.lib()
    L.hc_search_with.argtypes = [LAUNCH, CLOCK, REPORT, CANCEL, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
    L.hc_search_with(*cbs, None, kpi, 0, 0.2, 0.0, 5.0, out)
    status, found, index = out[0], out[1], out[2]
    assert status == 0 and found == 1
    assert [b for b, _ in calls] == sorted(b for b, _ in calls) and calls[0][0] == 0
    want = oracle.search(key, 0, 0, index + 1, needle, not fold, mode)
    assert want and want[0] == index
