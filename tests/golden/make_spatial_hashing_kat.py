"""Writes spatial_hashing_kat.json: the known answers of the reference's only
hot-path tests (src/physics/spatial_hashing.rs:602-632).

Those tests upload keys ``i % 256`` and vals ``i`` (spatial_hashing.rs:325-326),
sort them with ``pass_count`` 8-bit passes from either parity, run spatial
hashing and check the first ``min(n, 256)`` entries of start/close against the
closed forms at spatial_hashing.rs:551-585: cell c holds
``n / 256 + (n % 256 > c)`` keys, start is the exclusive prefix sum of those
counts and close is start + count - 1.  This script evaluates those closed
forms; it does not run the reference.
"""
import json
from pathlib import Path

CASES = [  # (number_of_keys, pass_count, read_from_even), spatial_hashing.rs:602-632
    (1, 1, True),
    (100, 4, True), (100, 4, False), (100, 3, True), (100, 3, False),
    (256, 4, True), (256, 4, False), (256, 3, True), (256, 3, False),
    (16_777_216, 4, True), (16_777_216, 4, False), (16_777_216, 3, True), (16_777_216, 3, False),
]


def expected(n):
    cells = min(n, 256)
    counts = [n // 256 + (1 if n % 256 > c else 0) for c in range(cells)]
    start, close, acc = [], [], 0
    for cnt in counts:
        start.append(acc)
        close.append(acc + cnt - 1)
        acc += cnt
    return start, close


def main():
    out = []
    for n, passes, even in CASES:
        s, c = expected(n)
        out.append({"number_of_keys": n, "pass_count": passes, "read_from_even": even,
                    "result_in_even": (passes + (0 if even else 1)) % 2 == 0,
                    "start": s, "close": c})
    path = Path(__file__).with_name("spatial_hashing_kat.json")
    path.write_text(json.dumps({"source": "src/physics/spatial_hashing.rs:551-632",
                                "cases": out}, separators=(",", ":")))
    print(path)


if __name__ == "__main__":
    main()
