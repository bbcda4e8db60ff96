"""The certified shortlist's error bound rests on the accuracy of its MFMA instructions.

tests/gpu_probes/mfma_accuracy (built by tests/gpu_probes/build.py) runs chains of 1-24 instructions of
v_mfma_f32_16x16x32_bf16 and v_mfma_f32_16x16x32_fp8_fp8, and chains of 1, 3 and 6 of
v_mfma_f32_16x16x128_f8f6f4 (e4m3; the FP8 instruction of every width that is a multiple of
128), on seeded random operands and reports, per instruction and chain, the largest
|mfma - exact| / sum |products|. The bound takes 2 gamma_d |q^| |x^| for the accumulation
(BF16: within the sequential fp32 model) and, for FP8, 2^-10 |q^| |x^| more for the
instruction's own sum; this test holds both with a margin of at least 2."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_probes", "mfma_accuracy")


def test_mfma_accuracy_within_the_bound_terms():
    assert os.path.exists(PROBE), "the probe is not built (tests/gpu_probes/build.py builds it)"
    out = subprocess.run([PROBE], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    u = 2.0**-24
    for chain in (1, 4, 12, 24):
        d = 32 * chain
        gamma = d * u / (1 - d * u)
        assert r[f"bf16_{chain}"] <= gamma, r  # within the model the bound doubles (2 gamma_d)
        assert r[f"fp8_{chain}"] <= 2.0**-11, r  # half the FP8 term the bound adds
    for chain in (1, 3, 6):
        assert r[f"fp8x128_{chain}"] <= 2.0**-11, r
