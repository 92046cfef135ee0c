"""Writes tests/golden/kats.json: known-answer vectors for turbosort's sort path.

Every vector is data: inputs and expected outputs transcribed from the
reference crate's doctests and tests (the citation is stored with each
vector), plus published SplitMix64 outputs.  Expected outputs are the values
the reference asserts; where a reference test only asserts "sorted", the
expected list is the input sorted with Python's own integer/float ordering,
which is independent of the radix-key transforms under test.

Floats are stored as IEEE-754 bit patterns (hex strings) so NaN payloads and
-0.0 survive JSON.  Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import math
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))


def f32(x: float) -> str:
    return "0x%08x" % struct.unpack("<I", struct.pack("<f", x))[0]


def f64(x: float) -> str:
    return "0x%016x" % struct.unpack("<Q", struct.pack("<d", x))[0]


F32_NAN = "0x7fc00000"  # f32::NAN (sign clear)
F64_NAN = "0x7ff8000000000000"
F32_MIN = -3.4028234663852886e38
F32_MAX = 3.4028234663852886e38
F32_MIN_POS = 1.1754943508222875e-38
INF = math.inf

sort_kats = [
    dict(src="lib.rs:18-20", dtype="u32", input=[5, 3, 8, 1, 9, 2, 7, 4, 6], expected=[1, 2, 3, 4, 5, 6, 7, 8, 9]),
    dict(src="lib.rs:26-29", dtype="u32", with_buffer=True, input=[5, 3, 8, 1, 9], expected=[1, 3, 5, 8, 9]),
    dict(src="lib.rs:82-84", dtype="i32", input=[3, -1, 4, -1, 5, -9], expected=[-9, -1, -1, 3, 4, 5]),
    dict(src="lib.rs:92-96; README.md:36-41", dtype="f32",
         input=[f32(1.0), F32_NAN, f32(-0.0), f32(0.0), f32(-INF)],
         expected=[f32(-INF), f32(-0.0), f32(0.0), f32(1.0), F32_NAN]),
    dict(src="lib.rs:116-119; README.md:44-47", dtype="u64", with_buffer=True, input=[42, 7, 99, 1, 0],
         expected=[0, 1, 7, 42, 99]),
    dict(src="tests/correctness.rs:250-281", dtype="f32",
         input=[f32(v) for v in (1.0, -1.0, 0.0, -0.0, INF, -INF)] + [F32_NAN] +
               [f32(v) for v in (F32_MIN, F32_MAX, F32_MIN_POS, -F32_MIN_POS)],
         expected=[f32(v) for v in (-INF, F32_MIN, -1.0, -F32_MIN_POS, -0.0, 0.0, F32_MIN_POS, 1.0, F32_MAX, INF)] +
                  [F32_NAN]),
    dict(src="tests/correctness.rs:284-303", dtype="f64",
         input=[f64(v) for v in (1.0, -1.0, 0.0, -0.0, INF, -INF)] + [F64_NAN],
         expected=[f64(v) for v in (-INF, -1.0, -0.0, 0.0, 1.0, INF)] + [F64_NAN]),
    dict(src="tests/correctness.rs:308-311", dtype="i32", input=[2**31 - 1, -2**31, 0, -1, 1],
         expected=[-2**31, -1, 0, 1, 2**31 - 1]),
    dict(src="tests/correctness.rs:313-315", dtype="i64", input=[2**63 - 1, -2**63, 0, -1, 1],
         expected=[-2**63, -1, 0, 1, 2**63 - 1]),
    dict(src="tests/correctness.rs:317-319", dtype="i8", input=[127, -128, 0, -1, 1], expected=[-128, -1, 0, 1, 127]),
    dict(src="tests/correctness.rs:321-323", dtype="i16", input=[32767, -32768, 0, -1, 1],
         expected=[-32768, -1, 0, 1, 32767]),
]

# Saved proptest shrinks (tests/correctness.proptest-regressions:7-8): inputs
# only; the expected output is Python's sort of the same numbers.
_reg_f32 = [0.0, 0.0, 0.0, -1.6976289, -1.5736036e31]
_reg_f32_bits = [struct.unpack("<f", struct.pack("<f", v))[0] for v in _reg_f32]
sort_kats.append(dict(src="tests/correctness.proptest-regressions:7", dtype="f32",
                      input=[f32(v) for v in _reg_f32], expected=[f32(v) for v in sorted(_reg_f32_bits)]))
_reg_i32 = [0, 0, 0, 0, 0, 0, 0, 0, 423439103, 0, -459012928, -1, -1, -1, 0, -1329394665, -1, 0, 0, 0, 0,
            -1329394665, 0, 0, 0, 162020954, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 423439104, 126044353, 0, 0, 0, 0, 0, 0,
            -1, 0, 0, 0, 0, 0, 0, 0, 423439103, 423439103, 423439103, -1, 423439103, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
            0, 0, 0, 0, -2, -843883178, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, -1, -1, 0,
            0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
            0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, -1, 423439103, -1, 423439103, 423439103, -1, 423439103, -1, -1, -1,
            0, -1, -1, -1, -1, -1, 0, -2]
sort_kats.append(dict(src="tests/correctness.proptest-regressions:8", dtype="i32", input=_reg_i32,
                      expected=sorted(_reg_i32)))

# Stable by-key orders asserted by the reference.
order_kats = [
    dict(src="lib.rs:145-147; README.md:50-52 (people by age)", dtype="u32", keys=[31, 25, 31, 25],
         order=[1, 3, 0, 2]),
    dict(src="README.md:54-57 (then by name length)", dtype="u32", keys=[5, 4, 3, 5], order=[2, 1, 0, 3]),
    dict(src="by_key.rs:102-106 (words by length)", dtype="u32", keys=[4, 3, 5, 4], order=[1, 0, 3, 2]),
    dict(src="by_key.rs:127-134; README.md:60-66 (ids)", dtype="u32", keys=[30, 10, 20], order=[1, 2, 0],
         apply_to=[2, 0, 1], applied=[0, 1, 2]),
]

# Internal decisions of the reference (used by the CPU restatement only).
divert_kats = dict(
    log2_eighths=[dict(src="divert.rs:121", x=1 << 20, expect=160)],
    skippable=[
        dict(src="divert.rs:130", n_live=8, len=1 << 20, max=(1 << 20) // 256 + 64, expect=5),
        dict(src="divert.rs:138", n_live=4, len=1 << 16, max=(1 << 16) // 256, expect=0),
        dict(src="divert.rs:141", n_live=4, len=1024, max=4, expect=0),
        dict(src="divert.rs:143", n_live=8, len=1024, max=4, expect=6),
        dict(src="divert.rs:145", n_live=8, len=1 << 16, max=(1 << 16) - 1, expect=0),
        dict(src="divert.rs:146", n_live=0, len=1 << 16, max=0, expect=0),
        dict(src="divert.rs:147", n_live=3, len=1 << 16, max=0, expect=0),
    ],
)
histogram_kats = [
    dict(src="histogram.rs:140-143", gen="u64_byte2_top", expect="0x8000000000ff0000"),
    dict(src="histogram.rs:144", gen="u32_sevens", expect="0x0"),
    dict(src="histogram.rs:145-146", gen="u16_mul", expect=None),
    dict(src="histogram.rs:151-158", gen="u64_live_3_4", live=[3, 4]),
]

# Vigna's splitmix64.c with state 1234567: its first five outputs.
splitmix_kats = dict(seed=1234567, outputs=[6457827717110365317, 3203168211198807973, 9817491932198370423,
                                            4593380528125082431, 16408922859458223821])


def main() -> None:
    # The first 16 keys of the 1M-key u32 workload (SplitMix64 state 1, high
    # halves), so generators cannot drift.
    mask = (1 << 64) - 1

    def sm(seed, i):
        z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & mask
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        return z ^ (z >> 31)

    assert [sm(1234567, i) for i in range(5)] == splitmix_kats["outputs"]
    first16 = [sm(1, i) >> 32 for i in range(16)]
    out = dict(sort=sort_kats, order=order_kats, divert=divert_kats, histogram=histogram_kats,
               splitmix64=splitmix_kats, c1_first16=dict(seed=1, u32=first16))
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "kats.json"))


if __name__ == "__main__":
    main()
