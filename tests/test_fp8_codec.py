"""The engine's FP8 E4M3 codec (fuel_amd/csrc/fp8.h) against PyTorch's
float8_e4m3fn conversion (round to nearest even) and the scale rule, on the
host: the header is compiled into a small host program with hipcc."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "fuel_amd", "csrc")

PROG = r"""
#include "fp8.h"
#include <cstdio>
#include <vector>
int main(int argc, char** argv) {
  FILE* f = std::fopen(argv[1], "rb");
  long n = 0;
  std::fread(&n, 8, 1, f);
  std::vector<float> x(n);
  std::vector<int> e(n);
  std::fread(x.data(), 4, n, f);
  std::fread(e.data(), 4, n, f);
  std::fclose(f);
  std::vector<unsigned char> code(n);
  std::vector<int> se(n);
  std::vector<float> val(256);
  for (long i = 0; i < n; ++i) {
    code[i] = (unsigned char)fa::fp8_e4m3_encode(x[i], e[i]);
    se[i] = fa::fp8_scale_exp(x[i]);
  }
  for (int c = 0; c < 256; ++c) val[c] = fa::fp8_e4m3_value((unsigned)c);
  f = std::fopen(argv[2], "wb");
  std::fwrite(code.data(), 1, n, f);
  std::fwrite(se.data(), 4, n, f);
  std::fwrite(val.data(), 4, 256, f);
  std::fclose(f);
  return 0;
}
"""


@pytest.fixture(scope="module")
def codec(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("fp8")
    src = d / "codec.cpp"
    src.write_text(PROG)
    exe = d / "codec"
    subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O2", f"-I{CSRC}", str(src), "-o", str(exe)],
                   check=True, capture_output=True)

    def run(x, e):
        x = np.ascontiguousarray(x, np.float32)
        e = np.ascontiguousarray(e, np.int32)
        inp, out = d / "in.bin", d / "out.bin"
        with open(inp, "wb") as f:
            f.write(np.int64(x.size).tobytes() + x.tobytes() + e.tobytes())
        subprocess.run([str(exe), str(inp), str(out)], check=True)
        raw = np.fromfile(out, dtype=np.uint8)
        n = x.size
        code = raw[:n]
        se = raw[n:n + 4 * n].view(np.int32)
        val = raw[n + 4 * n:].view(np.float32)
        return code, se, val
    return run


def test_values_match_torch(codec):
    torch = pytest.importorskip("torch")
    _, _, val = codec(np.zeros(1, np.float32), np.zeros(1, np.int32))
    ref = torch.arange(256, dtype=torch.uint8).view(torch.float8_e4m3fn).float().numpy()
    finite = np.isfinite(ref)
    assert not finite[0x7F] and not finite[0xFF]
    assert np.array_equal(val[finite], ref[finite])


def test_encode_matches_torch_rne(codec):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(11)
    n = 400_000
    e = rng.integers(-20, 21, n).astype(np.int32)
    # scaled magnitudes across the whole range, subnormals, ties and the top code
    mag = np.exp2(rng.uniform(-14.0, np.log2(448.0), n))
    grid = rng.integers(0, 256, n).astype(np.uint8)
    gridv = torch.from_numpy(grid).view(torch.float8_e4m3fn).float().numpy()
    ok = np.isfinite(gridv) & (np.abs(gridv) <= 448)
    mag[: n // 4] = np.abs(gridv[: n // 4])  # exact codes
    mid = np.abs(gridv[n // 4: n // 2])
    mag[n // 4: n // 2] = mid * (1 + 2.0 ** -4)  # halfway above a normal code (a tie)
    mag[n // 2: n // 2 + 1000] = 2.0 ** -10  # tie to zero
    mag[n // 2 + 1000: n // 2 + 2000] = 3 * 2.0 ** -10  # tie between two subnormals
    sign = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    scaled = (sign * np.minimum(mag, 448.0)).astype(np.float32)
    keep = np.isfinite(scaled) & ok
    scaled, e = scaled[keep], e[keep]
    x = (scaled.astype(np.float64) * np.exp2(e.astype(np.float64))).astype(np.float32)
    scaled = (x.astype(np.float64) * np.exp2(-e.astype(np.float64))).astype(np.float32)  # exact: powers of two
    code, _, _ = codec(x, e)
    ref = torch.from_numpy(scaled).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    bad = code != ref
    # torch keeps the sign of a value that rounds to zero, as the codec does
    assert not bad.any(), (x[bad][:5], e[bad][:5], code[bad][:5], ref[bad][:5])


def test_scale_rule(codec):
    rng = np.random.default_rng(12)
    m = np.exp2(rng.uniform(-60, 60, 200_000)).astype(np.float32)
    m[:4] = [448.0, 449.0, 224.0, 1.75 * 2 ** 20]
    _, se, _ = codec(m, np.zeros(m.size, np.int32))
    scaled = m.astype(np.float64) * np.exp2(-se.astype(np.float64))
    assert (scaled <= 448).all() and (scaled > 224).all()
    _, se0, _ = codec(np.zeros(3, np.float32), np.zeros(3, np.int32))
    assert (se0 == 0).all()


def test_error_bound(codec):
    """|x - q(x) 2^e| <= 2^-4 |x| + 2^(e - 10) for every element of a scaled vector."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(13)
    v = (rng.standard_normal((2000, 96)) * np.exp2(rng.uniform(-30, 30, (2000, 1)))).astype(np.float32)
    v[:, :10] *= np.float32(1e-6)  # small elements next to large ones
    _, se, _ = codec(np.abs(v).max(1), np.zeros(2000, np.int32))
    e = np.repeat(se, 96).astype(np.int32)
    code, _, val = codec(v.ravel(), e)
    deq = val[code].astype(np.float64) * np.exp2(e.astype(np.float64))
    err = np.abs(deq - v.ravel().astype(np.float64))
    assert (err <= 2.0 ** -4 * np.abs(v.ravel()) + np.exp2(e - 10.0)).all()
