"""Network layer specs for the measured configurations, in the layer-spec form
of pir_net_create (include/pir.h), plus a reader for the reference's model JSON
format (src/util/model.cu:24-115, files/models/*.json).

Layer specs: ("fc", in, out) | ("cnn", H, W, Cin, Cout, f, stride, pad) |
("relu", n_per_image) | ("maxpool", H, W, C, pool, stride) |
("avgpool", H, W, C, pool, stride) | ("ln", n_per_image) |
("res", H, W, in_planes, planes, num_blocks, stride, expansion).
"""
import json

from . import LAYER_AVGPOOL, LAYER_CNN, LAYER_FC, LAYER_LN, LAYER_MAXPOOL, LAYER_RELU, LAYER_RES

_TYPES = {"fc": LAYER_FC, "cnn": LAYER_CNN, "relu": LAYER_RELU, "maxpool": LAYER_MAXPOOL, "avgpool": LAYER_AVGPOOL,
          "ln": LAYER_LN, "res": LAYER_RES}


def to_spec(layers):
    """Symbolic layers -> integer rows for pir_net_create."""
    return [(_TYPES[l[0]],) + tuple(int(v) for v in l[1:]) for l in layers]


def input_size(layers):
    """Per-image input element count of the first layer."""
    l = layers[0]
    if l[0] in ("fc", "relu", "ln"):
        return l[1]
    return l[1] * l[2] * l[3]


def output_size(layers):
    """Per-image output element count of the last layer."""
    l = layers[-1]
    if l[0] == "fc":
        return l[2]
    if l[0] in ("relu", "ln"):
        return l[1]
    if l[0] == "res":
        H, W, _, planes, _, s, _ = l[1:]
        return ((H + 2 - 3) // s + 1) * ((W + 2 - 3) // s + 1) * planes
    if l[0] == "cnn":
        H, W, _, Cout, f, s, p = l[1:]
        return ((H + 2 * p - f) // s + 1) * ((W + 2 * p - f) // s + 1) * Cout
    H, W, C, pool, s = l[1:]
    return ((H - pool) // s + 1) * ((W - pool) // s + 1) * C


def from_model_json(doc, avgpool_as_maxpool=False):
    """Reads the reference's model JSON (model.cu:24-115).  Average-pool layers
    become max-pool layers of the same geometry when `avgpool_as_maxpool` (the
    max-pool variants of the VGG16 / AlexNet CIFAR models)."""
    if isinstance(doc, str):
        with open(doc) as f:
            doc = json.load(f)
    out = []
    for l in doc["model"]:
        t = l["layer"]
        if t == "fc":
            out.append(("fc", l["input_dim"], l["output_dim"]))
        elif t == "relu":
            out.append(("relu", l["input_dim"]))
        elif t == "ln":
            out.append(("ln", l["input_dim"]))
        elif t == "res":
            out.append(("res", l["input_hw"][0], l["input_hw"][1], l["in_planes"], l["out_planes"], l["num_blocks"],
                        l["stride"], l["expansion"]))
        elif t == "cnn":
            out.append(("cnn", l["input_hw"][0], l["input_hw"][1], l["in_channels"], l["out_channels"],
                        l["filter_hw"][0], l["stride"], l["padding"]))
        elif t == "maxpool" or (t == "averagepool" and avgpool_as_maxpool):
            out.append(("maxpool", l["input_hw"][0], l["input_hw"][1], l["in_channels"], l["pool_hw"][0],
                        l["stride"]))
        elif t == "averagepool":
            out.append(("avgpool", l["input_hw"][0], l["input_hw"][1], l["in_channels"], l["pool_hw"][0],
                        l["stride"]))
        else:
            raise ValueError(f"layer type {t!r} is not supported")
    return out


# files/models/secureml.json: FC 784-128-128-10 with ReLU after each FC
SECUREML = [("fc", 784, 128), ("relu", 128), ("fc", 128, 128), ("relu", 128), ("fc", 128, 10), ("relu", 10)]

# files/models/lenet.json
LENET = [
    ("cnn", 28, 28, 1, 20, 5, 1, 0), ("maxpool", 24, 24, 20, 2, 2), ("relu", 2880),
    ("cnn", 12, 12, 20, 50, 5, 1, 0), ("maxpool", 8, 8, 50, 2, 2), ("relu", 800),
    ("fc", 800, 500), ("relu", 500), ("fc", 500, 10), ("relu", 10),
]


def _vgg16():
    # files/models/vgg16-cifar10.json with its 2x2 stride-2 pools as max-pools
    L = []
    hw, cin = 32, 3
    blocks = [(64, 2), (128, 2), (256, 3), (512, 3), (512, 3)]
    for cout, nconv in blocks:
        for i in range(nconv):
            L.append(("cnn", hw, hw, cin, cout, 3, 1, 1))
            cin = cout
            if i < nconv - 1:
                L.append(("relu", hw * hw * cout))
        L.append(("maxpool", hw, hw, cout, 2, 2))
        hw //= 2
        L.append(("relu", hw * hw * cout))
    L += [("fc", 512, 256), ("relu", 256), ("fc", 256, 256), ("relu", 256), ("fc", 256, 10), ("relu", 10)]
    return L


VGG16_CIFAR = _vgg16()

# files/models/alexnet-cifar10.json with its pools as max-pools
ALEXNET_CIFAR = [
    ("cnn", 32, 32, 3, 96, 11, 4, 9), ("maxpool", 10, 10, 96, 3, 2), ("relu", 1536),
    ("cnn", 4, 4, 96, 256, 5, 1, 1), ("maxpool", 2, 2, 256, 2, 1), ("relu", 256),
    ("cnn", 1, 1, 256, 384, 3, 1, 1), ("relu", 384),
    ("cnn", 1, 1, 384, 384, 3, 1, 1), ("relu", 384),
    ("cnn", 1, 1, 384, 256, 3, 1, 1), ("relu", 256),
    ("fc", 256, 256), ("relu", 256), ("fc", 256, 256), ("relu", 256), ("fc", 256, 10), ("relu", 10),
]

MODELS = {"secureml": SECUREML, "lenet": LENET, "vgg16_cifar": VGG16_CIFAR, "alexnet_cifar": ALEXNET_CIFAR}


def uneven_query_batches(queries=64, seed=20):
    """Per-query batch sizes of the uneven-batch workload: `queries` concurrent
    AlexNet-CIFAR inferences, each with a batch drawn uniformly from 1..32 by
    numpy's default_rng(seed) (64 queries, 986 images for seed 20)."""
    import numpy as np

    return [int(b) for b in np.random.default_rng(seed).integers(1, 33, queries)]


def normal_weights(layers, sigma, seed, bias=0.01):
    """{layer: (weights, biases)} as doubles for every FC / CNN layer, weights
    drawn from N(0, sigma) by numpy's default_rng(seed) in layer order, FC
    biases `bias` (the reference's default, FCLayer.cu:46), CNN layers without
    bias; for pir_net_set_weights (layouts [in][out] and [Cout][f][f][Cin])."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = {}
    for i, l in enumerate(layers):
        if l[0] == "fc":
            out[i] = (rng.normal(0.0, sigma, l[1] * l[2]), np.full(l[2], bias))
        elif l[0] == "cnn":
            out[i] = (rng.normal(0.0, sigma, l[3] * l[4] * l[5] * l[5]), None)
    return out
