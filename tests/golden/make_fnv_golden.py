#!/usr/bin/env python3
"""Writes tests/golden/fnv_reference.npz: FNV-1a-64 values computed by the reference's own
engine::FnvHasher / engine::FnvBytes (src/ModArithmetic/Fnv.h, compiled into oracle/_ref) on
fixed inputs: the word list of a small matrix as the micro-benchmark checksum feeds it (rows,
cols, then the entries row-major, PerformanceSmokeValidation.cpp:117-125), and a few byte strings."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

if not O.ref_available():
    raise SystemExit("oracle/_ref/libhacsref.so is not built (make -C oracle ref)")
M = O.engine_matrix(7, 11, 2147483629, 0x05)
words = np.concatenate([np.array([7, 11], np.uint64), M.astype(np.uint64).ravel()])
strings = [b"", b"a", b"foobar", b"hacs"]
out = dict(matrix=M, words=words, words_fnv=np.array([O.ref_fnv_words(words)], np.uint64),
           strings=np.array([s.ljust(8, b"\0") for s in strings], dtype="S8"),
           string_lens=np.array([len(s) for s in strings], np.uint64),
           strings_fnv=np.array([O.ref_fnv_bytes(s) for s in strings], np.uint64))
np.savez_compressed(os.path.join(HERE, "fnv_reference.npz"), **out)
print("words fnv", hex(int(out["words_fnv"][0])), [hex(int(v)) for v in out["strings_fnv"]])
