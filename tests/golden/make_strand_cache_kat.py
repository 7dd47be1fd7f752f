"""Writes the .strands known-answer fixture: the byte sequence that the
reference's own loader test builds and parses (src/strand_cache.rs:254-281):
magic, version 1, scale 1.0, 3 vertices (0,0,0) (1,0,0) (1,1,0), one strand
{count 3, offset 0}, indices [0, 1, 2].  The expected parse is stored beside
it as JSON."""
import json
import struct
from pathlib import Path

HERE = Path(__file__).resolve().parent


def kat_bytes() -> bytes:
    b = bytearray(b"SSRSTRD\0")
    b += struct.pack("<I", 1)          # version
    b += struct.pack("<f", 1.0)        # scale
    b += struct.pack("<QQQ", 3, 1, 3)  # vertex, strand, index counts
    for v in [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (1.0, 1.0, 0.0)]:
        b += struct.pack("<fff", *v)
    b += struct.pack("<II", 3, 0)      # strand: count, offset
    b += struct.pack("<III", 0, 1, 2)  # indices
    return bytes(b)


if __name__ == "__main__":
    (HERE / "strand_cache_kat.strands").write_bytes(kat_bytes())
    (HERE / "strand_cache_kat.expected.json").write_text(json.dumps(
        {"vertices": [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0]], "strands": [[3, 0]],
         "indices": [0, 1, 2]}, indent=1) + "\n")
