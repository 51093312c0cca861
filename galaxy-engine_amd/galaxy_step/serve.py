"""ctypes mirror of include/ge_serve.h: the autotuner and the
continuous-batching request server (host C++ inside libge_physics.so)."""
import ctypes as C

import numpy as np

from .lib import (FIELDS_SPH, FIELDS_U, GE_OK, Context, GeError, GeParticles, PARTICLE_ORDER,
                  StepParams, load_library)

SERVE_EXPORTS = ["ge_autotune", "ge_server_default_config", "ge_server_create", "ge_server_destroy",
                 "ge_server_last_error", "ge_server_run", "ge_server_set_round_limit"]


class TuneResult(C.Structure):
    _fields_ = [("leaf_capacity", C.c_uint32), ("walk_block", C.c_uint32), ("ms_per_frame", C.c_float),
                ("reference_ms", C.c_float), ("table_ms", (C.c_float * 3) * 3)]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("hbm_budget_bytes", C.c_uint64), ("batch_capacity", C.c_uint32),
                ("max_systems", C.c_uint32), ("sph_slots", C.c_uint32), ("sph_capacity", C.c_uint32),
                ("round_frames", C.c_uint32), ("batching", C.c_int32), ("use_graph", C.c_int32),
                ("reserved", C.c_uint32 * 4)]


class Request(C.Structure):
    _fields_ = [("n", C.c_uint32), ("steps", C.c_int32), ("sph", C.c_int32), ("arrival_s", C.c_double),
                ("inp", GeParticles), ("out", GeParticles)]


class RequestResult(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
                ("status", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("particle_updates", C.c_uint64), ("completed", C.c_uint32),
                ("p50_s", C.c_double), ("p99_s", C.c_double), ("mean_s", C.c_double),
                ("solves_per_s", C.c_double), ("updates_per_s", C.c_double), ("max_resident", C.c_uint32),
                ("device_bytes", C.c_uint64), ("rounds", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_bound = False


def _lib():
    global _bound
    lib = load_library()
    if not _bound:
        P, vp = C.POINTER, C.c_void_p
        sig = {
            "ge_autotune": (C.c_int, [vp, P(StepParams), C.c_int32, P(TuneResult)]),
            "ge_server_default_config": (None, [P(ServerConfig)]),
            "ge_server_create": (C.c_int, [P(vp), P(ServerConfig), P(StepParams), P(StepParams)]),
            "ge_server_destroy": (C.c_int, [vp]),
            "ge_server_last_error": (C.c_char_p, [vp]),
            "ge_server_run": (C.c_int, [vp, C.c_uint32, P(Request), P(RequestResult), P(ServerStats)]),
            "ge_server_set_round_limit": (C.c_int, [vp, C.c_uint64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return lib


def autotune(ctx: Context, params, frames=3):
    """Runs ge_autotune on the particles ctx holds; returns (leaf, block, result)."""
    lib = _lib()
    r = TuneResult()
    ctx._check(lib.ge_autotune(ctx.h, C.byref(params), int(frames), C.byref(r)))
    return r.leaf_capacity, r.walk_block, r


def default_server_config(**kw):
    cfg = ServerConfig()
    _lib().ge_server_default_config(C.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, v)
    return cfg


def _fill(gp, arrays, state, n, fields):
    for f in fields:
        if f in state and state[f] is not None:
            dt = np.uint32 if f in FIELDS_U else np.float32
            a = np.ascontiguousarray(state[f], dtype=dt)
            assert len(a) == n, f
            arrays.append(a)
            setattr(gp, f, a.ctypes.data_as(C.POINTER(C.c_uint32 if f in FIELDS_U else C.c_float)))


class Server:
    """One ge_server.  run(requests) with requests = list of dicts
    {"state": {...}, "steps": k, "sph": 0/1, "arrival_s": t}; returns
    (outputs, results, stats)."""

    def __init__(self, cfg, gravity_params, sph_params=None):
        self.lib = _lib()
        h = C.c_void_p()
        rc = self.lib.ge_server_create(C.byref(h), C.byref(cfg), C.byref(gravity_params),
                                       C.byref(sph_params) if sph_params is not None else None)
        if rc != GE_OK:
            raise GeError(rc, "ge_server_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.ge_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_round_limit(self, max_rounds):
        """ge_server_set_round_limit: later runs stop after max_rounds batch
        rounds (0: no limit)."""
        rc = self.lib.ge_server_set_round_limit(self.h, int(max_rounds))
        if rc != GE_OK:
            raise GeError(rc, "ge_server_set_round_limit failed")

    def run(self, requests, out_fields=("x", "y", "vx", "vy", "ax", "ay", "id")):
        nreq = len(requests)
        reqs = (Request * max(nreq, 1))()
        res = (RequestResult * max(nreq, 1))()
        hold, outs = [], []
        for i, rq in enumerate(requests):
            st = rq["state"]
            n = len(st["x"])
            r = reqs[i]
            r.n, r.steps, r.sph, r.arrival_s = n, int(rq["steps"]), int(rq.get("sph", 0)), float(rq["arrival_s"])
            _fill(r.inp, hold, st, n, PARTICLE_ORDER)
            o = {f: np.zeros(n, np.uint32 if f in FIELDS_U else np.float32) for f in out_fields}
            _fill(r.out, hold, o, n, out_fields)
            outs.append(o)
        stats = ServerStats()
        rc = self.lib.ge_server_run(self.h, nreq, reqs, res, C.byref(stats))
        if rc != GE_OK:
            raise GeError(rc, self.lib.ge_server_last_error(self.h).decode())
        results = [{"arrival_s": res[i].arrival_s, "admit_s": res[i].admit_s, "done_s": res[i].done_s,
                    "status": res[i].status} for i in range(nreq)]
        del hold
        return outs, results, stats.as_dict()


__all__ = ["autotune", "default_server_config", "Server", "TuneResult", "ServerConfig", "ServerStats",
           "SERVE_EXPORTS", "FIELDS_SPH"]
