"""sf_step_noalloc, the entry point that allocates nothing (needs an
MI355X): steps of one bubble on device-resident float arrays, its work
arrays carved from the library's fixed device array, against sf_step on a
one-bubble mixed-precision context from the same state. Two calls in a row,
the second on a smaller bubble and in place (outputs aliasing the inputs),
so the second reuses the array the first left filled."""
import numpy as np
import pytest

import soapbubble_amd as sb
from cases import make_case
from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("u_theta", "u_phi", "eta", "gamma")


def _reference(n, dt, st, ua, va, steps):
    with sb.SoapFilm(sb.bubble(n, dt), "mixed") as ctx:
        ctx.load_state(0, st)
        ctx.upload(0, "uair0", ua)
        ctx.upload(0, "vair0", va)
        ctx.step(steps)
        ctx.synchronize()
        return ctx.state(0), ctx.stats(0)


class _Hip:
    """Device buffers of the test through the HIP runtime the library itself
    links (libamdhip64.so.7, already loaded with it): a second runtime in the
    process (torch's own) would not see the GPU."""

    def __init__(self):
        import ctypes
        sb.load_library()
        self.ct = ctypes
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.owned = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        p = self.ct.c_void_p()
        assert self.lib.hipMalloc(self.ct.byref(p), a.nbytes) == 0
        self.owned.append(p)
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p.value, a.size

    def get(self, ptr, n):
        out = np.empty(n, dtype=np.float32)
        assert self.lib.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0   # device to host
        return out.astype(np.float64)

    def free(self):
        for p in self.owned:
            self.lib.hipFree(p)
        self.owned = []


def _noalloc(hip, n, dt, st, ua, va, steps, in_place):
    ins = [hip.put(st[k]) for k in FIELDS]
    air = [hip.put(ua), hip.put(va)]
    outs = ins if in_place else [hip.put(np.full(sz, np.nan)) for _, sz in ins]
    stats = sb.step_noalloc(sb.bubble(n, dt), sb.options("mixed"), 0.0, steps,
                            [p for p, _ in ins], [p for p, _ in air], [p for p, _ in outs])
    got = {k: hip.get(*outs[q]) for q, k in enumerate(FIELDS)}
    return got, stats


@pytest.mark.parametrize("calls", [((128, 0.004, 11, 6, False), (64, 0.005, 12, 4, True))])
def test_noalloc_matches_sf_step(calls):
    hip = _Hip()
    for n, dt, seed, steps, in_place in calls:
        print(f"noalloc {n}x{2 * n} dt {dt} seed {seed} steps {steps} in_place={in_place}",
              flush=True)
        st, ua, va = make_case(n, seed, u_scale=0.5)
        want, ws = _reference(n, dt, st, ua, va, steps)
        got, gs = _noalloc(hip, n, dt, st, ua, va, steps, in_place)
        assert gs["failed_solves"] == 0 and gs["max_residual"] <= 1e-10, gs
        assert ws["failed_solves"] == 0, ws
        assert gs["steps"] == steps
        for k in FIELDS:
            e = rel_l2(got[k], want[k])
            print(f"  {k}: rel L2 {e:.2e}", flush=True)
            assert e < 2e-6, (n, k, e)
        assert abs(gs["volume"] - ws["volume"]) <= 1e-6 * abs(ws["volume"])
    hip.free()
