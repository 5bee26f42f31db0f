"""ctypes binding of the serving entry points (nk_server_* in include/nunatak_hip.h) and the request-stream driver.

``Server`` wraps one nk_server: submit() queues an update_velocities! on copied inputs and returns at once;
wait() blocks for that request's (u, v) and its timing record. ``serve()`` replays a stream of requests at their
arrival times from one submitting thread while another collects results, so requests arrive while earlier ones
are still on the device, and reports per-request latency (finish − arrival) and throughput.
"""
import ctypes as ct
import sys
import threading
import time

import numpy as np

from . import abi
from .model import NunatakError, load


class Server:
    def __init__(self, slots=4, hbm_budget_bytes=0):
        self._L = load()
        cfg = abi.ServerConfig(slots=slots, pad=0, hbm_budget_bytes=hbm_budget_bytes)
        h = ct.c_void_p()
        st = self._L.nk_server_create(ct.byref(cfg), ct.byref(h))
        if st != abi.NK_OK:
            raise NunatakError(f"nk_server_create: {abi.STATUS_NAMES.get(st, st)}")
        self._h = h
        self._shapes = {}

    def close(self):
        if self._h and self._h.value:
            self._L.nk_server_destroy(self._h)
            self._h = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, grid, physics, options, fields):
        """Queue one solve; fields maps names to arrays shaped like the field ((ny, nx) or (nz, ny, nx))."""
        names = list(fields)
        arrays = []
        for name in names:
            sx, sy, sz = abi.packed_shape(grid, physics, name)
            a = np.empty((sz, sy, sx), dtype=np.float64)
            a[...] = fields[name]
            arrays.append(np.ascontiguousarray(a))
        ids = (ct.c_int32 * len(names))(*[abi.FIELD_ID[n] for n in names])
        ptrs = (ct.POINTER(ct.c_double) * len(names))(*[a.ctypes.data_as(ct.POINTER(ct.c_double)) for a in arrays])
        rid = ct.c_int64()
        st = self._L.nk_server_submit(self._h, ct.byref(grid), ct.byref(physics), ct.byref(options), len(names), ids, ptrs,
                                      ct.byref(rid))
        if st != abi.NK_OK:
            raise NunatakError(f"nk_server_submit: {abi.STATUS_NAMES.get(st, st)}: "
                               f"{self._L.nk_server_last_error(self._h).decode()}", st)
        self._shapes[rid.value] = (abi.packed_shape(grid, physics, "U"), abi.packed_shape(grid, physics, "V"))
        return rid.value

    def poll(self, rid):
        d = ct.c_int32()
        st = self._L.nk_server_poll(self._h, rid, ct.byref(d))
        if st != abi.NK_OK:
            raise NunatakError(f"nk_server_poll: {abi.STATUS_NAMES.get(st, st)}")
        return bool(d.value)

    def wait(self, rid):
        """(u, v, info) of a request; raises if its solve failed."""
        (ux, uy, _), (vx, vy, _) = self._shapes.pop(rid)
        u = np.empty((uy, ux))
        v = np.empty((vy, vx))
        info = abi.RequestInfo()
        st = self._L.nk_server_wait(self._h, rid, u.ctypes.data_as(ct.POINTER(ct.c_double)),
                                    v.ctypes.data_as(ct.POINTER(ct.c_double)), ct.byref(info))
        if st != abi.NK_OK:
            raise NunatakError(f"request {rid}: {abi.STATUS_NAMES.get(st, st)}: "
                               f"{self._L.nk_server_last_error(self._h).decode()}", st)
        return u, v, info.as_dict()


def serve(requests, slots=4, hbm_budget_bytes=0, keep_results=False, progress=None):
    """Replay requests [(arrival_s, grid, physics, options, fields), ...] (sorted by arrival) against a new server.

    Returns a dict with per-request records (arrival, submit, admit, start, finish in ms of the server clock,
    latency = finish − arrival), p50/p99 latency, solves/s over the span from the first arrival to the last finish,
    and the results when keep_results."""
    records = [None] * len(requests)
    results = [None] * len(requests)
    ids = [None] * len(requests)
    submitted = threading.Semaphore(0)
    errors = []
    with Server(slots=slots, hbm_budget_bytes=hbm_budget_bytes) as srv:
        t0 = time.perf_counter()
        offset_ms = [0.0]

        def submitter():
            try:
                for k, (arrival, g, p, o, f) in enumerate(requests):
                    delay = arrival - (time.perf_counter() - t0)
                    if delay > 0:
                        time.sleep(delay)
                    ids[k] = srv.submit(g, p, o, f)
                    submitted.release()
            except Exception as e:  # surfaced by the collector
                errors.append(e)
                for _ in range(len(requests)):
                    submitted.release()

        th = threading.Thread(target=submitter)
        th.start()
        for k in range(len(requests)):
            submitted.acquire()
            if errors:
                break
            u, v, info = srv.wait(ids[k])
            if k == 0:
                offset_ms[0] = info["submit_ms"] - 1e3 * requests[0][0]
            info["arrival_ms"] = offset_ms[0] + 1e3 * requests[k][0]
            info["latency_ms"] = info["finish_ms"] - info["arrival_ms"]
            records[k] = info
            if keep_results:
                results[k] = (u, v)
            if progress and (k + 1) % progress == 0:
                lat = [r["latency_ms"] for r in records[: k + 1]]
                print(f"[serve] {k + 1}/{len(requests)} done, p50 {np.percentile(lat, 50):.1f} ms", file=sys.stderr, flush=True)
        th.join()
    if errors:
        raise errors[0]
    lat = np.array([r["latency_ms"] for r in records])
    span_ms = max(r["finish_ms"] for r in records) - min(r["arrival_ms"] for r in records)
    out = {"records": records, "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "solves_per_s": len(records) / (span_ms / 1e3), "span_s": span_ms / 1e3}
    if keep_results:
        out["results"] = results
    return out
