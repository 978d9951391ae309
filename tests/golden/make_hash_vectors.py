"""Writes tests/golden/hash_vectors.json: raw 32-bit hash values of a fixed
key set, computed by the independent formulations in tests/hashref.py and a
bit-serial CRC written here from the reference's polynomials
(include/hash.hpp:143-207).  Run: python tests/golden/make_hash_vectors.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import hashref  # noqa: E402


def crc_bitwise(key, poly, width):
    mask = (1 << width) - 1
    crc = mask
    for i in range(4):
        crc ^= (key >> (8 * i)) & 0xFF
        for _ in range(8):
            crc = (crc >> 1) ^ poly if crc & 1 else crc >> 1
    return (crc ^ mask) & 0xFFFFFFFF


def main():
    rng = np.random.default_rng(20240601)
    keys = [0, 1, 2, 42, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFE, 0xFFFFFFFF, 0x87654321, 0xDEADBEEF]
    keys += rng.integers(0, 2**32, size=246, dtype=np.uint64).tolist()
    ka = np.array(keys, np.uint32)
    h1, h2, mu, ci = hashref.hash1(ka), hashref.hash2(ka), hashref.murmur(ka), hashref.city(ka)
    rows = []
    for i, k in enumerate(keys):
        rows.append({"key": int(k), "hash1": int(h1[i]), "hash2": int(h2[i]), "murmur": int(mu[i]),
                     "city": int(ci[i]), "crc64": crc_bitwise(int(k), 0x42F0E1EBA9EA3693, 64),
                     "crc32": crc_bitwise(int(k), 0x04C11D87, 32)})
    with open(os.path.join(HERE, "hash_vectors.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_hash_vectors.py", "vectors": rows}, f, indent=0)


if __name__ == "__main__":
    main()
