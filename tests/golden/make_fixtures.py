"""Converts the reference's plaintext test fixtures (files/test/, float32
values printed as text) into tests/golden/fixtures.npz (float64, exactly the
doubles the reference's loadShareFromFile parses, util.cuh:271-289).

Run in a container that has the reference checkout:
    python tests/golden/make_fixtures.py /root/reference
Only data files (the fixtures and the model descriptions of files/models/,
copied as they are to tests/golden/models/) and the numeric brace lists of a
few GPU.cu / Layer.cu known-answer tests are read; no reference code is
imported, compiled or run.
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
    # backward fixtures (Layer.cu:118-180, 379-420, 772-840; files/test/secureml
    # gradin_layer*: dL/d(output of layer k) of one SecureML training step)
    "test_fclayer_gradin": "fc_gradin",
    "test_fclayer_gradout": "fc_gradout",
    "test_fclayer_newweight": "fc_newweight",
    "test_fclayer_newbias": "fc_newbias",
    "test_relulayer_gradin": "relu_gradin",
    "test_relulayer_gradout": "relu_gradout",
    "test_maxpoollayer_gradin": "maxpool_gradin",
    "test_maxpoollayer_gradout": "maxpool_gradout",
    "test_convlayer_gradin": "conv_gradin",
    "test_convlayer_gradout": "conv_gradout",
    "test_convlayer_gradweight": "conv_gradweight",
    "test_convlayer_newweight": "conv_newweight",
    "secureml/labels": "secureml_labels",
    "secureml/gradin_layer0": "secureml_gradin0",
    "secureml/gradin_layer1": "secureml_gradin1",
    "secureml/gradin_layer2": "secureml_gradin2",
    "secureml/gradin_layer3": "secureml_gradin3",
    "secureml/gradin_layer4": "secureml_gradin4",
    "secureml/gradin_layer5": "secureml_gradin5",
}

# Known-answer vectors written inline in the reference's src/test/GPU.cu: the
# brace lists of the named test are read as numbers (no code is run).
KATS = {
    ("GPU.cu", "GPUTest", "CutlassConvDgrad2"): {"delta": "dgrad2_delta", "filters": "dgrad2_filters",
                                                 "expected": "dgrad2_expected"},
    ("GPU.cu", "GPUTest", "CutlassConvWgrad2"): {"delta": "wgrad2_delta", "input": "wgrad2_input",
                                                 "expected": "wgrad2_expected"},
    # LayerNorm known answers (Layer.cu:887-1000): 2 rows of 48 features
    ("Layer.cu", "LayerTest", "LNForward"): {"input": "lnfwd_input", "expected": "lnfwd_expected"},
    ("Layer.cu", "LayerTest", "LNBackward"): {"input": "lnbwd_input", "delta": "lnbwd_delta",
                                              "expected": "lnbwd_expected", "expected_gamma": "lnbwd_gamma",
                                              "expected_beta": "lnbwd_beta"},
}


def kat_arrays(ref):
    import re

    out = {}
    for (fname, fixture, test), names in KATS.items():
        text = open(os.path.join(ref, "src", "test", fname)).read()
        start = text.index("TYPED_TEST(%s, %s)" % (fixture, test))
        end = text.find("TYPED_TEST(", start + 1)
        block = text[start:end if end > 0 else len(text)]
        for var, key in names.items():
            m = re.search(r"\b%s\s*=\s*\{([^}]*)\}" % var, block)
            out[key] = np.array([float(t) for t in m.group(1).replace(",", " ").split()], np.float64)
    return out


def main(ref):
    base = os.path.join(ref, "files", "test")
    arrays = {}
    for rel, key in FILES.items():
        with open(os.path.join(base, rel)) as f:
            arrays[key] = np.array([float(t) for t in f.read().split()], np.float64)
    arrays.update(kat_arrays(ref))
    here = os.path.dirname(os.path.abspath(__file__))
    out = os.path.join(here, "fixtures.npz")
    np.savez_compressed(out, **arrays)
    print(out, os.path.getsize(out), "bytes")
    # the model description files (data, files/models/*.json) as they are
    models = os.path.join(here, "models")
    os.makedirs(models, exist_ok=True)
    src = os.path.join(ref, "files", "models")
    for name in sorted(os.listdir(src)):
        if name.endswith(".json"):
            with open(os.path.join(src, name), "rb") as f, open(os.path.join(models, name), "wb") as g:
                g.write(f.read())
    print(models, len(os.listdir(models)), "model files")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
