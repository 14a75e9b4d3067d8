"""Serving loop over the C ABI (jxrhip_server_*): admission by device
footprint, waves of persistent batches on four compute streams, continuous
batching of requests that arrive while others run, and p50/p99 latency.

The nearest caller in the reference is the batch decoder's grouped
submission (crates/jxr/src/batch/cuda.rs:20-156); admission and continuous
batching have no counterpart there."""
import ctypes as C
import threading
from dataclasses import dataclass
from typing import Dict, List, Optional

from . import abi
from .abi import check, load
from .decoder import HipDecodePlan, ResidentHipImage


@dataclass
class ServerResult:
    ticket: int
    tag: int
    status: int
    stage: int
    latency_ms: float
    queue_ms: float
    image: Optional[ResidentHipImage]  # when the request had no destination


def clock_ns() -> int:
    """The steady clock arrival times are measured on."""
    return load().jxrhip_server_clock_ns()


class HipServer:
    def __init__(self, session, hbm_budget_bytes: int = 0, max_inflight: int = 0, max_wave_images: int = 0,
                 max_waves: int = 0):
        cfg = abi.ServerConfig()
        cfg.hbm_budget_bytes = hbm_budget_bytes
        cfg.max_inflight = max_inflight
        cfg.max_wave_images = max_wave_images
        cfg.max_waves = max_waves
        h = C.c_void_p()
        check(load().jxrhip_server_create(session.handle, C.byref(cfg), C.byref(h)), "server_create")
        self._h = h
        self._s = session
        self._keep: Dict[int, HipDecodePlan] = {}  # plans of requests without a result yet
        # submit may run on several threads while another polls: a plan is
        # recorded under the lock its submission holds, so a result polled
        # at once still finds (and releases) it
        self._lock = threading.Lock()

    def submit(self, plan: HipDecodePlan, dst: Optional[int] = None, dst_bytes: int = 0, tag: int = 0,
               arrival_ns: int = 0) -> int:
        t = C.c_uint64()
        with self._lock:
            check(load().jxrhip_server_submit(self._h, plan.handle, C.c_void_p(dst) if dst else None,
                                              dst_bytes if dst else 0, tag, arrival_ns, C.byref(t)),
                  "server_submit")
            self._keep[t.value] = plan
        return t.value

    def poll(self, capacity: int = 256, timeout_us: int = 0) -> List[ServerResult]:
        buf = (abi.ServerResult * capacity)()
        n = load().jxrhip_server_poll(self._h, buf, capacity, timeout_us)
        out = []
        for r in buf[:n]:
            with self._lock:
                self._keep.pop(r.ticket, None)
            img = ResidentHipImage(self._s, C.c_void_p(r.image)) if r.image else None
            out.append(ServerResult(r.ticket, r.tag, r.status, r.stage, r.latency_ms, r.queue_ms, img))
        return out

    def drain(self):
        check(load().jxrhip_server_drain(self._h), "server_drain")

    def stats(self) -> dict:
        st = abi.ServerStats()
        check(load().jxrhip_server_get_stats(self._h, C.byref(st)), "server_stats")
        return {name: getattr(st, name) for name, _ in abi.ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            load().jxrhip_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass
