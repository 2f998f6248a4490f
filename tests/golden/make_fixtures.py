"""Converts the reference's plaintext test fixtures (files/test/, float32
values printed as text) into tests/golden/fixtures.npz (float64, exactly the
doubles the reference's loadShareFromFile parses, util.cuh:271-289).

Run in a container that has the reference checkout:
    python tests/golden/make_fixtures.py /root/reference
Only data files are read; no reference code is imported or run.
"""
import os
import sys

import numpy as np

FILES = {
    # SecureML forward chain (Layer.cu:86-116, files/test/secureml)
    "secureml/input": "secureml_input",
    "secureml/weight0": "secureml_weight0",
    "secureml/bias0": "secureml_bias0",
    "secureml/weight2": "secureml_weight2",
    "secureml/bias2": "secureml_bias2",
    "secureml/weight4": "secureml_weight4",
    "secureml/bias4": "secureml_bias4",
    "secureml/outputlayer0": "secureml_out0",
    "secureml/outputlayer1": "secureml_out1",
    "secureml/outputlayer2": "secureml_out2",
    "secureml/outputlayer3": "secureml_out3",
    "secureml/outputlayer4": "secureml_out4",
    "secureml/outputlayer5": "secureml_out5",
    # single-layer fixtures (Layer.cu:52-84, 205-232, 298-339, 670-713)
    "test_fclayer_input": "fc_input",
    "test_fclayer_weight": "fc_weight",
    "test_fclayer_bias": "fc_bias",
    "test_fclayer_output": "fc_output",
    "test_relulayer_input": "relu_input",
    "test_relulayer_output": "relu_output",
    "test_maxpoollayer_input": "maxpool_input",
    "test_maxpoollayer_output": "maxpool_output",
    "test_convlayer_input": "conv_input",
    "test_convlayer_weight": "conv_weight",
    "test_convlayer_output": "conv_output",
}


def main(ref):
    base = os.path.join(ref, "files", "test")
    arrays = {}
    for rel, key in FILES.items():
        with open(os.path.join(base, rel)) as f:
            arrays[key] = np.array([float(t) for t in f.read().split()], np.float64)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.npz")
    np.savez_compressed(out, **arrays)
    print(out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
