"""Optional progress reports of the oracle's long stages (test infrastructure only).

``hook`` is None (silent) or a callable taking one string; bench.py's full-size CPU
baseline sets it so that a run of several minutes reports each stage as it ends."""
import time

hook = None
_t0 = time.perf_counter()


def report(msg):
    if hook is not None:
        hook(f"[oracle {time.perf_counter() - _t0:8.1f} s] {msg}")
