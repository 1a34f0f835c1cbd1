"""Per-launch kernel timing through the library's HIP events (kfr_amd_timing_* in the C ABI).

    with kfr_amd.kernel_timing() as kt:
        plan.execute_batch(y, x, batch, False)
    for name, agg in kt.by_kernel().items():
        print(name, agg["launches"], agg["ms"], agg["bytes"])

Every transform-kernel launch made inside the block is bracketed by two events recorded on the stream
it is launched on; leaving the block waits for them. Do not open a block around a hipGraph capture.
"""
from __future__ import annotations

import ctypes as C
from collections import OrderedDict

from ._lib import check, lib


class kernel_timing:
    def __init__(self):
        self.records: list[tuple[str, int, float]] = []

    def __enter__(self) -> "kernel_timing":
        check(lib().kfr_amd_timing_begin() == 0)
        self.records = []
        return self

    def __exit__(self, *exc) -> None:
        L = lib()
        n = L.kfr_amd_timing_end()
        if exc[0] is not None:
            return
        check(n >= 0)
        name = C.create_string_buffer(256)
        nbytes = C.c_longlong()
        ms = C.c_double()
        for i in range(n):
            check(L.kfr_amd_timing_get(i, name, len(name), C.byref(nbytes), C.byref(ms)) == 0)
            self.records.append((name.value.decode(), int(nbytes.value), float(ms.value)))

    def by_kernel(self) -> "OrderedDict[str, dict]":
        """Per kernel, in order of total time: launches, total ms and total algorithmic bytes."""
        agg: dict[str, dict] = {}
        for nm, b, ms in self.records:
            a = agg.setdefault(nm, {"launches": 0, "ms": 0.0, "bytes": 0})
            a["launches"] += 1
            a["ms"] += ms
            a["bytes"] += b
        return OrderedDict(sorted(agg.items(), key=lambda kv: -kv[1]["ms"]))
