"""loadModel's reading of the reference's model files (model.cu:24-115) through
the C ABI's parser (pir_model_read, host-only), checked against Python's own
JSON reader on every model description the reference ships
(tests/golden/models/, copied from files/models/), plus the error cases:
malformed JSON, missing fields and unknown layer types."""
import json
import os

import pytest

import piranha_amd as P
from piranha_amd import configs as C

MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")


def model_names():
    return sorted(n for n in os.listdir(MODELS) if n.endswith(".json"))


def python_doc(name):
    with open(os.path.join(MODELS, name)) as f:
        return json.load(f)


@pytest.mark.parametrize("name", model_names())
def test_reads_reference_model(name):
    path = os.path.join(MODELS, name)
    try:
        doc = python_doc(name)
    except json.JSONDecodeError:
        with pytest.raises(P.PirError, match="model file"):  # files/models/resnet18-imagenet.json is not JSON
            P.read_model(path)
        return
    ins, ncl, bs, spec = P.read_model(path)
    assert (ins, ncl, bs) == (doc["input_size"], doc["num_classes"], doc["batch_size"])
    want = [tuple(row) + (0,) * (8 - len(row)) for row in C.to_spec(C.from_model_json(doc))]
    assert spec == want


def write(tmp_path, text):
    p = tmp_path / "m.json"
    p.write_text(text)
    return str(p)


def test_formatting_and_escapes(tmp_path):
    text = ('\n{ "input_size" : 784 ,"num_classes":10, "batch_size": 2, "dataset": "M\\u004eIST \\"x\\"",\r\n'
            '\t"model": [ {"layer": "fc", "input_dim": 784, "output_dim": 10.0, "note": [true, false, null, -1.5e3]},'
            '{"layer":"relu","input_dim":10} ] }  ')
    ins, ncl, bs, spec = P.read_model(write(tmp_path, text))
    assert (ins, ncl, bs) == (784, 10, 2)
    assert spec == [(P.LAYER_FC, 784, 10, 0, 0, 0, 0, 0), (P.LAYER_RELU, 10, 0, 0, 0, 0, 0, 0)]


@pytest.mark.parametrize("text,msg", [
    ('{"input_size": 1, "num_classes": 1, "batch_size": 1, "dataset": "x", "model": [', "unexpected end"),
    ('{"input_size": 1, "num_classes": 1, "batch_size": 1, "dataset": "x", "model": []} x', "after the document"),
    ('{"input_size": 1, "num_classes": 1, "batch_size": 1, "dataset": "x"}', "no \"model\""),
    ('{"input_size": 1.5, "num_classes": 1, "batch_size": 1, "dataset": "x", "model": []}', "integer"),
    ('{"input_size": 1, "num_classes": 1, "batch_size": 1, "dataset": "x", "model": [{"layer": "fc"}]}',
     "input_dim"),
    ('{"input_size": 1, "num_classes": 1, "batch_size": 1, "dataset": "x", "model": [{"layer": "bn"}]}',
     "unknown layer type"),
    ('{"input_size": 1, "model": ' + "[" * 100 + "]" * 100 + "}", "nesting"),
    ('{"input_size": 1, "num_classes": 1, "batch_size": 1, "dataset": "x", "model": [{"layer": "cnn", '
     '"input_hw": [4], "in_channels": 1, "out_channels": 1, "filter_hw": [3, 3], "stride": 1, "padding": 0}]}',
     "not a pair"),
])
def test_malformed(tmp_path, text, msg):
    with pytest.raises(P.PirError, match=msg):
        P.read_model(write(tmp_path, text))


def test_missing_file(tmp_path):
    with pytest.raises(P.PirError, match="cannot open"):
        P.read_model(str(tmp_path / "absent.json"))
