"""Builds the MFMA accuracy probe (tests/gpu_probes/mfma_accuracy) that the GPU tests run; test
infrastructure, compiled with the engine's own HIP flags."""
import importlib.util
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "mfma_accuracy.hip")
PROBE = os.path.join(HERE, "mfma_accuracy")


def build():
    spec = importlib.util.spec_from_file_location("_vsg_build", os.path.join(ROOT, "vecsearch_amd", "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    b._run([os.path.join(b.ROCM, "bin", "hipcc"), *b.HIP_FLAGS, "-o", PROBE + ".tmp", SRC])
    os.replace(PROBE + ".tmp", PROBE)
    return PROBE


if __name__ == "__main__":
    print(build())
