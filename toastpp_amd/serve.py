"""Serving loop on the GPU library (tmi_server_*): continuous batching of
CalcFields requests.  Requests reference prepared solvers (matrices resident
on the device) and device buffers; admitted requests run concurrently on
their own streams, in arrival order, within an HBM budget."""
import ctypes as C

import numpy as np

from . import _lib as L


class Server:
    def __init__(self, slots=4, hbm_budget_gib=64.0):
        self._h = C.c_void_p()
        L.call("tmi_server_create", int(slots), float(hbm_budget_gib) * 2.0 ** 30, C.byref(self._h))
        self._pending = {}

    def submit(self, solver, nq, d_rhs, d_phi, ld):
        """solver: FwdSolver after Reset; d_rhs / d_phi: device pointers (int)
        to [n][ld] node-major arrays (d_phi holds the initial guess)."""
        rid = C.c_longlong()
        L.call("tmi_server_submit", self._h, solver._h, int(nq), C.c_void_p(d_rhs), C.c_void_p(d_phi), int(ld),
               C.byref(rid))
        self._pending[rid.value] = nq
        return rid.value

    def wait(self, rid, timeout=-1.0):
        """-> (stats [(iters, sweeps, relres, converged)], (queued, started,
        finished) seconds, status); raises TmiError on timeout or failure."""
        nq = self._pending[rid]
        stats = (L.ColStat * nq)()
        times = (C.c_double * 3)()
        status = C.c_int()
        L.call("tmi_server_wait", self._h, C.c_longlong(rid), float(timeout), C.cast(stats, C.c_void_p), times,
               C.byref(status))
        del self._pending[rid]
        if status.value not in (L.TMI_OK, L.TMI_ERR_NOT_CONVERGED):
            raise L.TmiError(status.value, L.lib().tmi_last_error().decode(errors="replace"))
        st = [(s.iters, s.sweeps, s.relres, bool(s.converged)) for s in stats]
        return st, tuple(times), status.value

    def counts(self):
        r, q, b = C.c_int(), C.c_int(), C.c_double()
        L.call("tmi_server_counts", self._h, C.byref(r), C.byref(q), C.byref(b))
        return r.value, q.value, b.value

    def clock(self):
        return L.lib().tmi_server_clock(self._h)

    def close(self):
        if self._h.value and L._lib is not None:
            L._lib.tmi_server_free(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        self.close()


def latency_stats(times):
    """p50 / p99 / mean of finished - queued (seconds)."""
    lat = np.array([t[2] - t[0] for t in times])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3)}
