"""Makes tests/golden/fp8_mfma_cases.npz: golden vectors of the FP8 matrix-core product
(v_mfma_scale_f32_16x16x128_f8f6f4 with e4m3fn operands) that pin the oracle's
restatement of the FP8 candidate pass (oracle/ssg_oracle.c or_fp8_mfma).

The operands are seeded random e4m3fn bytes of several distributions (every finite code,
normal values, positive values, half zeros, subnormal-heavy, the padded unit-vector rows
of the precision form with its -|x|^2/2 start values, random operand scales); the
outputs are what the instruction returns on a gfx950 GPU.
  python tests/golden/make_fp8_mfma.py build          (here: hipcc for gfx950)
  python tests/golden/make_fp8_mfma.py run OUT.npz    (on a machine with the GPU)
Load with numpy.load(..., allow_pickle=False): A (cases, 16, 128) uint8, B (cases, 128,
16) uint8, C and D (cases, 16, 16) float32, scale_a and scale_b (cases,) int32.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "_build", "fp8_mfma_cases")


def decode(b):
    """e4m3fn byte -> value (bias 7, subnormal quantum 2^-9, 0x7f/0xff NaN)."""
    b = np.asarray(b, np.uint8).astype(np.int64)
    s, e, m = b >> 7, (b >> 3) & 15, b & 7
    v = np.where(e == 0, np.ldexp(m / 8.0, -6), np.ldexp(1.0 + m / 8.0, e - 7))
    v = np.where((e == 15) & (m == 7), np.nan, v)
    return np.where(s == 1, -v, v)


_CODES = np.array([c for c in range(256) if c not in (0x7F, 0xFF)], np.uint8)
_POS = np.array([c for c in range(0x7F)], np.uint8)  # 0x00 .. 0x7e, increasing values


def encode(x):
    """Values with |x| <= 448 -> nearest e4m3fn byte, ties to the even code."""
    x = np.asarray(x, np.float64)
    vals = decode(_POS)
    a = np.abs(x)
    i = np.clip(np.searchsorted(vals, a), 1, len(vals) - 1)
    lo, hi = vals[i - 1], vals[i]
    pick_hi = (a - lo > hi - a) | ((a - lo == hi - a) & ((i & 1) == 0))
    code = np.where(pick_hi, i, i - 1).astype(np.uint8)
    return np.where(x < 0, code | 0x80, code).astype(np.uint8)


def cases():
    rng = np.random.default_rng(41)
    out = []

    def add(A, B, C, sa=127, sb=127):
        out.append((np.int32(sa), np.int32(sb), A.astype(np.uint8), B.astype(np.uint8), C.astype(np.float32)))

    def wideC():
        return (rng.standard_normal((16, 16)) * np.ldexp(1.0, rng.integers(-4, 21, (16, 16)))).astype(np.float32)

    # every finite code, uniformly
    add(rng.choice(_CODES, (16, 128)), rng.choice(_CODES, (128, 16)), wideC())
    add(rng.choice(_CODES, (16, 128)), rng.choice(_CODES, (128, 16)), np.zeros((16, 16), np.float32))
    # normal values
    add(encode(np.clip(rng.standard_normal((16, 128)) * 64, -448, 448)),
        encode(np.clip(rng.standard_normal((128, 16)) * 64, -448, 448)), wideC())
    # positive values
    add(encode(rng.uniform(0, 448, (16, 128))), encode(rng.uniform(0, 448, (128, 16))), wideC())
    # half zeros (both signs)
    za = rng.choice(_CODES, (16, 128))
    za[rng.random((16, 128)) < 0.5] = rng.choice(np.array([0x00, 0x80], np.uint8))
    zb = rng.choice(_CODES, (128, 16))
    zb[rng.random((128, 16)) < 0.5] = 0
    add(za, zb, wideC())
    # subnormal-heavy: exponent fields 0..2
    sub = np.array([c for c in _CODES if ((c >> 3) & 15) <= 2], np.uint8)
    add(rng.choice(sub, (16, 128)), rng.choice(np.concatenate([sub, _CODES]), (128, 16)),
        (rng.standard_normal((16, 16)) * 1e-4).astype(np.float32))
    # random operand scales
    add(encode(np.clip(rng.standard_normal((16, 128)) * 64, -448, 448)),
        encode(np.clip(rng.standard_normal((128, 16)) * 64, -448, 448)), wideC(),
        rng.integers(110, 141), rng.integers(110, 141))
    # the precision form: 96 values of unit-vector rows scaled to the e4m3fn range, zero
    # padded to 128, start value -|x|^2/2 in natural units, scales 2^-9 each
    for _ in range(2):
        x = rng.standard_normal((16, 96))
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        q = rng.standard_normal((16, 96))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        A = np.zeros((16, 128), np.uint8)
        B = np.zeros((128, 16), np.uint8)
        A[:, :96] = encode(np.clip(x * 512, -448, 448))
        B[:96, :] = encode(np.clip(q * 512, -448, 448)).T
        c = np.repeat((-0.5 * (x.astype(np.float32) ** 2).sum(1)).astype(np.float32)[:, None], 16, 1)
        add(A, B, c, 127 - 9, 127 - 9)
    return out


def build():
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", "-o", BIN, os.path.join(HERE, "fp8_mfma_cases.hip")],
                   check=True)


def run(out_path):
    cs = cases()
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.int32(len(cs)).tobytes())
            for sa, sb, A, B, C in cs:
                f.write(np.array([sa, sb], np.int32).tobytes() + A.tobytes() + B.tobytes() + C.tobytes())
        subprocess.run([BIN, fin, fout], check=True, timeout=120)
        D = np.fromfile(fout, np.float32).reshape(len(cs), 16, 16)
    np.savez_compressed(out_path, scale_a=np.array([c[0] for c in cs], np.int32),
                        scale_b=np.array([c[1] for c in cs], np.int32),
                        A=np.stack([c[2] for c in cs]), B=np.stack([c[3] for c in cs]),
                        C=np.stack([c[4] for c in cs]), D=D)
    print(f"wrote {out_path}: {len(cs)} cases")


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "build":
        build()
    elif len(sys.argv) == 3 and sys.argv[1] == "run":
        run(sys.argv[2])
    else:
        sys.exit(__doc__)
