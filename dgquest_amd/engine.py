"""ctypes binding of the engine's runtime plumbing (include/dgq_engine.h):
events and device information used for timing and reporting."""
from __future__ import annotations

import ctypes as C
import os

from .quest import LIBDIR


class Engine:
    def __init__(self, libdir: str = LIBDIR):
        path = os.path.join(libdir, "libdgq_engine.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run dgquest_amd.build.build()")
        L = C.CDLL(path, mode=C.RTLD_GLOBAL)
        self.lib = L
        vp = C.c_void_p
        L.dq_last_error.restype = C.c_char_p
        L.dq_event_create.argtypes = [C.POINTER(vp)]
        L.dq_event_record.argtypes = [vp, vp]
        L.dq_event_elapsed_ms.argtypes = [vp, vp, C.POINTER(C.c_float)]
        L.dq_event_destroy.argtypes = [vp]
        L.dq_device_info.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_size_t),
                                     C.POINTER(C.c_size_t)]
        L.dq_device_sync.argtypes = []

        L.dq_compare_ws_bytes.restype = C.c_size_t
        L.dq_compare_ws_bytes.argtypes = [C.c_int]
        L.dq_compare_states.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, vp, vp, C.c_size_t, vp]
        L.dq_malloc.argtypes = [C.POINTER(vp), C.c_size_t]
        L.dq_free.argtypes = [vp]
        L.dq_memcpy_async.argtypes = [vp, vp, C.c_size_t, C.c_int, vp]
        L.dq_stream_sync.argtypes = [vp]

    def _ck(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.dq_last_error().decode())

    def event(self):
        e = C.c_void_p()
        self._ck(self.lib.dq_event_create(C.byref(e)))
        return e

    def record(self, ev, stream):
        self._ck(self.lib.dq_event_record(ev, C.c_void_p(stream)))

    def elapsed_ms(self, a, b) -> float:
        ms = C.c_float()
        self._ck(self.lib.dq_event_elapsed_ms(a, b, C.byref(ms)))
        return ms.value

    def compare_states(self, a, prec_a, b, prec_b, n):
        """<a|b>, <a|a>, <b|b> and max |a_i - b_i| of two device states of n
        qubits (dq_compare_states), e.g. from dgq_state_pointer."""
        L = self.lib
        wsb = L.dq_compare_ws_bytes(n)
        ws, out = C.c_void_p(), C.c_void_p()
        self._ck(L.dq_malloc(C.byref(ws), wsb))
        self._ck(L.dq_malloc(C.byref(out), 64))
        try:
            self._ck(L.dq_compare_states(a, prec_a, b, prec_b, n, out, ws, wsb, None))
            host = (C.c_double * 5)()
            self._ck(L.dq_memcpy_async(C.addressof(host), out, 40, 2, None))
            self._ck(L.dq_stream_sync(None))
        finally:
            L.dq_free(ws)
            L.dq_free(out)
        return {"inner": complex(host[0], host[1]), "norm_a": host[2], "norm_b": host[3], "max_abs_diff": host[4]}

    def sync(self):
        self._ck(self.lib.dq_device_sync())

    def info(self):
        name = C.create_string_buffer(64)
        cus = C.c_int()
        tot = C.c_size_t()
        free = C.c_size_t()
        self._ck(self.lib.dq_device_info(name, 64, C.byref(cus), C.byref(tot), C.byref(free)))
        return {"arch": name.value.decode(), "cus": cus.value, "total_mem": tot.value, "free_mem": free.value}
