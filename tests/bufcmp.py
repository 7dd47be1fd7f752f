"""Byte-for-byte comparison of large buffers (images, DOMs, dumps) with a
short failure message.

A plain ``assert a.tobytes() == b.tobytes()`` on megabyte buffers makes
pytest explain the failure with a full diff of the two byte strings, which
under ``-v`` runs for minutes without output; these helpers compare with
numpy and report only where the buffers differ.
"""
import numpy as np


def describe(a, b) -> str:
    """Empty when a and b hold the same bytes, else where they differ."""
    x = np.frombuffer(a if isinstance(a, (bytes, bytearray)) else np.ascontiguousarray(a).tobytes(), np.uint8)
    y = np.frombuffer(b if isinstance(b, (bytes, bytearray)) else np.ascontiguousarray(b).tobytes(), np.uint8)
    if x.size == y.size and np.array_equal(x, y):
        return ""
    n = min(x.size, y.size)
    bad = np.nonzero(x[:n] != y[:n])[0]
    where = f"first at byte {int(bad[0])}, last at {int(bad[-1])}" if bad.size else "common prefix equal"
    return f"{x.size} vs {y.size} bytes; {bad.size} differing bytes ({where})"


def assert_same(a, b, what: str) -> None:
    msg = describe(a, b)
    if msg:
        raise AssertionError(f"{what}: {msg}")
