#!/usr/bin/env python3
"""Benchmark of the 3-party RSS secure-inference path on one MI355X.

Headline (`value`): 3PC VGG16-CIFAR-shape inference, batch 64, Z_2^64 with 26
fractional bits, all three parties on one GPU, int8-limb MFMA ring products and
fused protocol kernels replayed as one hipGraph per forward.  One step = one
forward over the batch, inputs resident in HBM.

Also reported on the same line:
  * plain_images_per_s: the same forward in the reference's form (plain 64-bit
    VALU GEMMs, one launch per DeviceData / RSS op, uint8-per-bit dReLU);
  * ring_gemm: the 3-term RSS local product A0B0+A0B1+A1B0 at M=N=K in
    {512..8192}, Gop/s = 6MNK/t, plain VALU vs int8-limb MFMA;
  * other_configs: SecureML MLP, LeNet and the uneven-batch AlexNet workload,
    and the headline forward on uniformly random sharings (below);
  * roofline of the dominant kernel (the limb implicit-GEMM convolution),
    timed per launch with HIP events; its HBM traffic per launch comes from
    the committed rocprofv3 counter summary under profiles/;
  * hbm_roofline: achieved HBM bandwidth of the fused ReLU kernel (standalone)
    and of the maxpool kernels (each with the ReLU fused after it);
  * fp64_error: the reconstructed VGG16 logits against an FP64 forward of the
    same network (weights decoded from the shares);
  * measured_peaks: HBM copy / read / write streams, the int8 MFMA rate and
    the 64-bit multiply-add rate, measured beside the datasheet figures;
  * cpu_baseline: the CPU oracle on a bounded sample of every configuration
    (VGG16 is the headline entry), host cores and CPU model.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

INT8_PEAK_TOPS = 5000.0  # MI355X dense int8 MFMA, 5.0 POP/s
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "limb_gemm_conv_vgg16_b64.md")


def kernel_name(P, v):
    """The kernel of a profile_gemm variant: limb_gemm_kernel<form, tile>."""
    return f"limb_gemm_kernel<{v % P.LIMB_WIDE}, {1 if v >= P.LIMB_WIDE else 2}>"


def read_traffic(kname):
    """HBM bytes per launch of kernel `kname` (e.g. "limb_gemm_kernel<3, 1>") in
    this workload, from the committed rocprofv3 request-size counter summary
    (a line `traffic_bytes_per_launch[<kname>] = N`), or None when it is absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                key = f"traffic_bytes_per_launch[{kname}] ="
                if line.strip().startswith(key):
                    return float(line.split("=", 1)[1].split()[0])
    except OSError:
        pass
    return None


def decode(shares, fp):
    """Party A's reconstruction (shareA + shareB + party C's shareA) as doubles."""
    v = (shares[0, 0] + shares[0, 1] + shares[2, 0]).view(np.int64)
    return v / float(1 << fp)


def fp64_forward(layers, batch, params, x):
    """The same network in plain FP64 (NHWC cross-correlation with zero padding,
    FC x.W + b, ReLU, max-pooling): the plaintext reference the reconstructed
    fixed-point outputs are reported against.  params[i] = (W, b) as doubles."""
    a = np.asarray(x, np.float64).reshape(batch, -1)
    for i, l in enumerate(layers):
        if l[0] == "cnn":
            H, W, Cin, Cout, f, s, p = l[1:]
            t = np.pad(a.reshape(batch, H, W, Cin), ((0, 0), (p, p), (p, p), (0, 0)))
            OH, OW = (H + 2 * p - f) // s + 1, (W + 2 * p - f) // s + 1
            win = np.lib.stride_tricks.sliding_window_view(t, (f, f), axis=(1, 2))[:, ::s, ::s][:, :OH, :OW]
            cols = win.transpose(0, 1, 2, 4, 5, 3).reshape(batch * OH * OW, f * f * Cin)
            a = (cols @ params[i][0].reshape(Cout, f * f * Cin).T).reshape(batch, -1)
        elif l[0] == "fc":
            a = a.reshape(batch, l[1]) @ params[i][0].reshape(l[1], l[2])
            if params[i][1] is not None:
                a = a + params[i][1]
        elif l[0] == "relu":
            a = np.maximum(a, 0.0)
        else:
            H, W, C, pool, s = l[1:]
            OH, OW = (H - pool) // s + 1, (W - pool) // s + 1
            t = a.reshape(batch, H, W, C)
            win = np.lib.stride_tricks.sliding_window_view(t, (pool, pool), axis=(1, 2))[:, ::s, ::s][:, :OH, :OW]
            a = win.max(axis=(-2, -1)).reshape(batch, -1)
    return a


def fp64_error(net, layers, batch, fp, x):
    """Max |decoded 3PC output - FP64 forward| over the final layer, with the
    weights decoded from the network's own shares."""
    params = {}
    for i, l in enumerate(layers):
        if l[0] in ("fc", "cnn"):
            w, b = net.weights(i)
            params[i] = (decode(w, fp), None if b is None else decode(b, fp))
    ref = fp64_forward(layers, batch, params, x).ravel()
    got = decode(net.activation(len(layers) - 1), fp)
    err = float(np.max(np.abs(got - ref)))
    return {"max_abs_err": err, "max_abs_ref": float(np.max(np.abs(ref))),
            "rel_err": err / max(float(np.max(np.abs(ref))), 1e-30), "outputs": int(ref.size)}


def hbm_roofline(P, layers, batch, layer_ms):
    """Achieved HBM bandwidth of the fused elementwise / comparison kernels.

    relu_kernel: the standalone fused ReLU (pir_rss_relu) on the first VGG16
    activation's size, 94 B per element: 5 of the 6 input share arrays (party
    B's second share never reaches the output), 6 output arrays and 6
    derivative bytes.  In the network the ReLU layers run inside the GEMM
    epilogue or the maxpool kernel before them.  maxpool: the maxpool layers
    (with the fused ReLU that follows each of them) from per-layer HIP-event
    times; a k x k window reads 6k^2 - (k^2 - 1) words (the selects never read
    the later operands' party-B second share) and writes 6 output arrays and
    6k^2 derivative bytes, plus 54 B for the fused ReLU's output."""
    out = {}
    L = P.lib()
    n = batch * 32 * 32 * 64
    x = P.to_dev(np.random.default_rng(3).integers(0, 2**64, (3, 2, n), dtype=np.uint64))
    y, d = P.Buffer((3, 2, n), np.uint64), P.Buffer((3, 2, n), np.uint8)
    for _ in range(3):
        P.check(L.pir_rss_relu(x.ptr, n, y.ptr, d.ptr, 1, None))
    P.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        P.check(L.pir_rss_relu(x.ptr, n, y.ptr, d.ptr, 1, None))
    P.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    gbs = 94.0 * n / (ms * 1e-3) / 1e9
    out["relu_kernel"] = {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                          "elements": n, "ms": ms}
    del x, y, d
    byts = mp_ms = 0.0
    for i, l in enumerate(layers):
        if l[0] != "maxpool":
            continue
        H, W, C, pool, s = l[1:]
        windows = batch * ((H - pool) // s + 1) * ((W - pool) // s + 1) * C
        fused = i + 1 < len(layers) and layers[i + 1][0] == "relu"
        byts += windows * (8.0 * (5 * pool * pool + 1) + 48.0 + 6.0 * pool * pool + (54.0 if fused else 0.0))
        mp_ms += float(layer_ms[i])
    if mp_ms > 0:
        gbs = byts / (mp_ms * 1e-3) / 1e9
        out["maxpool"] = {"achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                          "bytes_per_forward": byts, "ms_per_forward": mp_ms}
    return out


CALIB_BYTES = 1 << 30


def counter_calibration(P):
    """Known-size kernels that scale the HBM byte counters: the read-only
    stream (peak_read_kernel, reads CALIB_BYTES, writes nothing) and the
    write-only fill (peak_write_kernel, writes CALIB_BYTES, reads nothing),
    each launched twice (warm-up + one timed launch); the fill also runs once
    to prepare the read kernel's source."""
    for which, name in ((P.PEAK_HBM_READ, "peak_read_kernel"), (P.PEAK_HBM_WRITE, "peak_write_kernel")):
        P.measure_peak(which, CALIB_BYTES, reps=1)
        log(f"counter calibration: {name} over {CALIB_BYTES} B")


def measured_peaks(P):
    """Measured peaks beside the datasheet figures (include/pir.h
    pir_measure_peak): HBM copy / read / write streams over 1 GiB buffers, the
    int8 MFMA instruction rate on low- and on full-entropy operands (the latter
    is the power-limited rate that digit planes of random shares see), the
    64-bit multiply-add rate of the VALUs."""
    out = {}
    for key, which, unit in (("hbm_copy", P.PEAK_HBM_COPY, "GB/s"), ("hbm_read", P.PEAK_HBM_READ, "GB/s"),
                             ("hbm_write", P.PEAK_HBM_WRITE, "GB/s"), ("mfma_i8", P.PEAK_MFMA_I8, "TOP/s"),
                             ("mfma_i8_full_entropy", P.PEAK_MFMA_I8_RANDOM, "TOP/s"),
                             ("valu_u64_muladd", P.PEAK_VALU_U64, "G/s")):
        best = 0.0
        for _ in range(3):
            r, _ms = P.measure_peak(which, 1 << 30, reps=10)
            best = max(best, r)
        out[key] = {"value": best, "unit": unit}
    out["hbm_copy"]["datasheet"] = HBM_PEAK_GBS
    out["mfma_i8"]["datasheet"] = INT8_PEAK_TOPS
    out["mfma_i8_full_entropy"]["datasheet"] = INT8_PEAK_TOPS
    log("measured peaks: " + ", ".join(f"{k} {v['value']:.0f} {v['unit']}" for k, v in out.items()))
    return out


def time_graph(P, net, steps, warmup):
    for _ in range(warmup):
        net.replay()
    net.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        net.replay()
    net.synchronize()
    return (time.perf_counter() - t0) / steps


def time_eager(net, steps, warmup):
    for _ in range(warmup):
        net.forward_shares(None)
    net.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        net.forward_shares(None)
    net.synchronize()
    return (time.perf_counter() - t0) / steps


def make_net(P, C, layers, batch, fp, mode, seed=1, capture=True, autotune=True, sigma=None):
    """The network with the layers' default (He) weights, or N(0, sigma)
    weights and 0.01 FC biases when `sigma` is given."""
    net = P.Network(C.to_spec(layers), batch, fp, seed=seed, mode=mode)
    if sigma is not None:
        for i, (w, b) in C.normal_weights(layers, sigma, seed).items():
            net.set_weights(i, w, b)
    rng = np.random.default_rng(seed)
    net.forward(rng.uniform(0.0, 1.0, batch * C.input_size(layers)))  # setPublic of U[0,1) inputs
    net.synchronize()
    if autotune and mode == P.MODE_FUSED:
        net.autotune()  # one eager forward timing candidate GEMM plans; outputs unchanged
    if capture:
        net.capture()
    return net


def bench_random_shares(P, C, layers, batch, fp, steps, warmup, autotune=True, masked=False):
    """The headline forward with every weight and input share re-dealt on the
    GPU as a uniformly random sharing of the same values (the Philox dealer,
    pir_net_rerandomize): the same GEMMs and kernels on full-entropy digit
    planes, where the reference's placeholder precompute leaves two of a
    value's three components at zero.  masked: also the masked precompute
    mode (random truncation pairs, reshare masks, dReLU / selectShare
    randomness drawn in the kernels)."""
    net = make_net(P, C, layers, batch, fp, P.MODE_FUSED, capture=False, autotune=False)
    net.rerandomize(7)
    if masked:
        net.set_precompute(P.PRECOMPUTE_MASKED, 0x5EED, 0)
    net.forward_shares(None)
    net.synchronize()
    if autotune:
        net.autotune()
    net.capture()
    t = time_graph(P, net, steps, warmup)
    prof = net.profile_gemm()
    again = net.profile_gemm()
    prof = {v: (min(prof[v][0], again[v][0]),) + prof[v][1:] for v in prof}
    dom = max(prof, key=lambda v: prof[v][0])
    k_ms, k_ops, k_n = prof[dom]
    del net
    return {"images_per_s": batch / t, "ms_per_forward": t * 1e3,
            "dominant_kernel": {"kernel": kernel_name(P, dom), "launches": k_n,
                                "avg_launch_ms": k_ms / k_n, "achieved_tops": k_ops / (k_ms * 1e-3) / 1e12},
            "sharing": "uniformly random consistent RSS sharings of the same weights and inputs, dealt on the GPU",
            "precompute": "masked (Philox4x32-10)" if masked else "placeholder (the reference's constants)"}


def bench_uneven(P, C, steps, warmup, autotune=True, profile=False, per_query=False):
    """64 concurrent AlexNet-CIFAR queries with per-query batches drawn from
    1..32 (986 images), FP20: the queries concatenated along the batch axis
    into one forward, replayed from its hipGraph.  per_query: also time the
    ungrouped alternative, one forward (and graph) per query back to back."""
    qb = C.uneven_query_batches()
    total = int(sum(qb))
    layers = C.ALEXNET_CIFAR
    n = make_net(P, C, layers, total, 20, P.MODE_FUSED, autotune=autotune)
    tt = time_graph(P, n, steps, warmup)
    out = {"images_per_s": total / tt, "ms_per_forward": tt * 1e3, "queries": len(qb), "images": total,
           "grouping": "queries concatenated along the batch axis into one graph"}
    log(f"alexnet uneven: {total} images, {total / tt:.1f} images/s ({tt * 1e3:.3f} ms)")
    if profile:
        layer_ms = n.profile_layers()
        prof = n.profile_gemm()
        log("per-layer ms: " + ", ".join(f"{l[0]}{i}:{m:.3f}" for i, (l, m) in enumerate(zip(layers, layer_ms))))
        out["layer_ms"] = [float(m) for m in layer_ms]
        out["limb_gemm"] = {kernel_name(P, v): {
            "ms": ms, "launches": k, "TOP/s": ops / ms / 1e9, "frac_of_peak": ops / ms / 1e9 / INT8_PEAK_TOPS}
            for v, (ms, ops, k) in sorted(prof.items())}
        log("limb GEMM: " + ", ".join(f"{k} {v['launches']}x {v['ms']:.3f} ms {v['TOP/s']:.0f} TOP/s"
                                      for k, v in out["limb_gemm"].items()))
    del n
    if per_query:
        # the ungrouped alternative: one forward (its own graph) per query, run back to back
        t_of = {}
        for b in sorted(set(qb)):
            nb = make_net(P, C, layers, b, 20, P.MODE_FUSED, autotune=autotune)
            t_of[b] = time_graph(P, nb, 5, 3)
            del nb
        seq = sum(t_of[b] for b in qb)
        out["one_forward_per_query"] = {"images_per_s": total / seq, "ms_total": seq * 1e3,
                                        "ms_by_batch": {str(b): t * 1e3 for b, t in t_of.items()}}
        log(f"one forward per query, back to back: {seq * 1e3:.3f} ms, {total / seq:.1f} images/s "
            f"(concatenated: {tt / seq:.3f} of that time)")
    return out


def bench_training(P, C, layers, batch, fp, steps, warmup, autotune=True):
    """One training step (NeuralNetwork::forward + backward: reveal-softmax
    delta, every layer's backward and SGD update), fused limb-MFMA form,
    captured as one hipGraph; synthetic U[0,1) inputs and one-hot labels."""
    net = P.Network(C.to_spec(layers), batch, fp, seed=1, mode=P.MODE_FUSED)
    rng = np.random.default_rng(1)
    net.forward(rng.uniform(0.0, 1.0, batch * C.input_size(layers)))
    n_classes = C.output_size(layers)
    labels = np.zeros((batch, n_classes))
    labels[np.arange(batch), rng.integers(0, n_classes, batch)] = 1.0
    net.set_train_labels(labels.ravel())
    if autotune:
        net.autotune_train()
    net.capture_train()
    t = time_graph_fn(net.replay_train, net, steps, warmup)
    prof = net.profile_train_gemm()
    kern = {kernel_name(P, v): {"ms": ms, "launches": n, "TOP/s": ops / ms / 1e9}
            for v, (ms, ops, n) in sorted(prof.items())}
    fwd_ms, sm_ms, bwd_ms = net.profile_train_layers()
    log("training step per layer (ms) forward: " + ", ".join(f"{l[0]}{i}:{m:.3f}" for i, (l, m) in
                                                              enumerate(zip(layers, fwd_ms))))
    log(f"  softmax delta {sm_ms:.3f}; backward: " + ", ".join(f"{l[0]}{i}:{m:.3f}" for i, (l, m) in
                                                              enumerate(zip(layers, bwd_ms))))
    return {"workload": f"VGG16-CIFAR training step, batch {batch}: forward + reveal-softmax backward + SGD "
                        f"update (log_lr 5), hipGraph", "images_per_s": batch / t, "ms_per_step": t * 1e3,
            "steps": steps, "limb_gemm_per_step": kern,
            "eager_ms": {"forward": float(fwd_ms.sum()), "softmax_delta": sm_ms, "backward": float(bwd_ms.sum())}}


def time_graph_fn(replay, net, steps, warmup):
    for _ in range(warmup):
        replay()
    net.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        replay()
    net.synchronize()
    return (time.perf_counter() - t0) / steps


def bench_ring_gemm(P, sizes, reps):
    L = P.lib()
    out = []
    rng = np.random.default_rng(0)
    for n in sizes:
        A0, A1, B0, B1 = (rng.integers(0, 2**64, n * n, dtype=np.uint64) for _ in range(4))
        d = [P.to_dev(x) for x in (A0, A1, B0, B1)]
        Cp, Cl = P.Buffer((n * n,), np.uint64), P.Buffer((n * n,), np.uint64)
        # plain: localMatMul's three gemms + two accumulations
        P.check(L.pir_ring_gemm3(P.ALGO_PLAIN, n, n, n, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, Cp.ptr, None))
        P.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            P.check(L.pir_ring_gemm3(P.ALGO_PLAIN, n, n, n, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, Cp.ptr, None))
        P.synchronize()
        tp = (time.perf_counter() - t0) / reps
        # limb: B packed once (as for weights), then the K-concatenated product
        packed = P.Buffer((L.pir_limb_packed_bytes(n, n, 2),), np.uint8)
        b0, k0 = P._ptr_array([d[2].ptr, d[2].ptr])
        b1, k1 = P._ptr_array([d[3].ptr, None])
        a, k2 = P._ptr_array([d[0].ptr, d[1].ptr])
        P.synchronize()
        t0 = time.perf_counter()
        P.check(L.pir_limb_pack(b0, b1, 2, n, n, 1, packed.ptr, None))
        P.synchronize()
        tpack = time.perf_counter() - t0
        P.check(L.pir_limb_gemm(a, 2, n, n, 1, packed.ptr, n, Cl.ptr, 1, 0, None))
        P.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            P.check(L.pir_limb_gemm(a, 2, n, n, 1, packed.ptr, n, Cl.ptr, 1, 0, None))
        P.synchronize()
        tl = (time.perf_counter() - t0) / reps
        same = bool(np.array_equal(Cp.numpy(), Cl.numpy()))
        ops = 6.0 * n ** 3
        rec = {"n": n, "plain_gops": ops / tp / 1e9, "limb_gops": ops / tl / 1e9, "limb_pack_ms": tpack * 1e3,
               "limb_int8_tops": 144.0 * n ** 3 / tl / 1e12, "bit_exact_plain_vs_limb": same}
        rec["limb_mfma_frac"] = rec["limb_int8_tops"] / INT8_PEAK_TOPS
        log(f"ring gemm n={n}: plain {rec['plain_gops']:.1f} Gop/s, limb {rec['limb_gops']:.1f} Gop/s "
            f"({rec['limb_int8_tops']:.0f} int8 TOP/s), equal={same}")
        out.append(rec)
        del d, Cp, Cl, packed
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(C, vgg_batch):
    """The oracle (the reference's arithmetic restated in C++, g++ -O3, OpenMP
    over the host cores; plain uint64 GEMM loops, uint8-per-bit dReLU) on a
    bounded sample of every configuration: SecureML b128, LeNet b32 and VGG16
    b`vgg_batch` forwards, the uneven AlexNet workload's first 3 queries, and
    the 3-term ring product at 512..2048 in full and at 4096 on 256 of its
    rows.  The headline entry is VGG16."""
    import oracle as O

    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    table = {}

    def net_rate(name, layers, batch, fp, sigma, seed=1):
        if sigma is None:  # He init N(0, sqrt(2 / fan_in)), the layers' default
            rng = np.random.default_rng(seed)
            w = {}
            for i, l in enumerate(layers):
                if l[0] == "fc":
                    w[i] = (rng.normal(0, np.sqrt(2 / l[1]), l[1] * l[2]), np.full(l[2], 0.01))
                elif l[0] == "cnn":
                    w[i] = (rng.normal(0, np.sqrt(2 / (l[5] * l[5] * l[3])), l[3] * l[4] * l[5] * l[5]), None)
        else:
            w = C.normal_weights(layers, sigma, seed)
        params = {i: (O.set_public(a, fp), None if b is None else O.set_public(b, fp)) for i, (a, b) in w.items()}
        x = O.set_public(np.random.default_rng(seed).uniform(0, 1, batch * C.input_size(layers)), fp)
        t0 = time.perf_counter()
        O.net_forward(layers, batch, fp, params, x)
        dt = time.perf_counter() - t0
        table[name] = {"value": batch / dt, "unit": "images/s", "sample": f"batch {batch}, {dt:.2f} s"}
        log(f"CPU {name}: {batch / dt:.3f} images/s ({dt:.1f} s)")

    net_rate("secureml_b128_fp13", C.SECUREML, 128, 13, 0.1)
    net_rate("lenet_fp13", C.LENET, 32, 13, 0.1)
    qb = C.uneven_query_batches()[:3]
    net_rate("alexnet_uneven_fp20", C.ALEXNET_CIFAR, int(sum(qb)), 20, None)
    table["alexnet_uneven_fp20"]["sample"] += " (the first 3 queries concatenated)"
    net_rate("vgg16_fp26", C.VGG16_CIFAR, vgg_batch, 26, None)

    rng = np.random.default_rng(0)
    for n in (512, 1024, 2048, 4096):
        rows = n if n <= 2048 else 256
        A0, A1 = (rng.integers(0, 2**64, rows * n, dtype=np.uint64) for _ in range(2))
        B0, B1 = (rng.integers(0, 2**64, n * n, dtype=np.uint64) for _ in range(2))
        t0 = time.perf_counter()
        for A, B in ((A0, B0), (A0, B1), (A1, B0)):
            O.gemm(rows, n, n, A, 1, B, 1, 1)
        dt = time.perf_counter() - t0
        gops = 6.0 * rows * n * n / dt / 1e9
        table[f"ring_gemm_{n}"] = {"value": gops, "unit": "Gop/s (6MNK/t)",
                                   "sample": f"{rows} of {n} rows, {dt:.2f} s" if rows < n else f"{dt:.2f} s"}
        log(f"CPU ring gemm n={n}: {gops:.1f} Gop/s ({dt:.1f} s)")

    head = table["vgg16_fp26"]
    return {"value": head["value"], "unit": "images/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"VGG16-CIFAR 3PC forward, batch {vgg_batch}, FP26, CPU oracle (g++ -O3 OpenMP, "
                      f"plain uint64 GEMM loops, uint8-per-bit dReLU), {head['sample'].split(', ')[1]}",
            "table": table}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    # ~0.7 s of replays first: the GPU clocks ramp up over the first few hundred ms
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--gemm-sizes", default="512,1024,2048,4096,8192")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-batch", type=int, default=8)
    ap.add_argument("--calib", action="store_true",
                    help="also run the read-only and write-only kernels of known size (counter calibration)")
    ap.add_argument("--quick", action="store_true", help="headline only (for profiling runs)")
    ap.add_argument("--no-autotune", action="store_true", help="heuristic GEMM plans only")
    ap.add_argument("--tune-file", default=None,
                    help="GEMM plan table: loaded before the run when the file exists, saved after the headline "
                         "network is tuned (profiling runs that must use one plan per shape)")
    ap.add_argument("--train-only", action="store_true", help="VGG16 training step only (for profiling runs)")
    ap.add_argument("--random-sharings-only", action="store_true",
                    help="the VGG16 forward on uniformly random sharings only (for profiling runs)")
    ap.add_argument("--uneven-only", action="store_true",
                    help="the uneven-batch AlexNet workload only, with its per-layer and GEMM profile")
    ap.add_argument("--per-query", action="store_true",
                    help="with --uneven-only: also time one forward per query, back to back")
    ap.add_argument("--masked-only", action="store_true",
                    help="the VGG16 forward in the masked precompute mode only (for profiling runs)")
    ap.add_argument("--gemm-only", action="store_true",
                    help="the ring-GEMM sweep (--gemm-sizes) only (for profiling runs)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (all three parties share it); --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    import piranha_amd as P
    from piranha_amd import configs as C

    if args.calib:
        counter_calibration(P)
    if args.tune_file and os.path.exists(args.tune_file):
        P.autotune_load(args.tune_file)
        log(f"GEMM plans loaded from {args.tune_file}: {P.autotune_entries()} shapes")

    def save_plans():
        if args.tune_file and not os.path.exists(args.tune_file):
            P.autotune_save(args.tune_file)

    if args.train_only:
        tr = bench_training(P, C, C.VGG16_CIFAR, args.batch, 26, args.steps, args.warmup,
                            autotune=not args.no_autotune)
        log(f"VGG16 training: {tr['ms_per_step']:.2f} ms/step, {tr['images_per_s']:.1f} images/s")
        print(json.dumps({"training": tr}), flush=True)
        save_plans()
        return

    if args.uneven_only:
        un = bench_uneven(P, C, args.steps, args.warmup, autotune=not args.no_autotune, profile=True,
                          per_query=args.per_query)
        print(json.dumps({"alexnet_uneven_64q_fp20": un}), flush=True)
        save_plans()
        return

    if args.gemm_only:
        sizes = [int(v) for v in args.gemm_sizes.split(",") if v]
        print(json.dumps({"ring_gemm": bench_ring_gemm(P, sizes, reps=2)}), flush=True)
        return

    if args.random_sharings_only or args.masked_only:
        rs = bench_random_shares(P, C, C.VGG16_CIFAR, args.batch, 26, args.steps, args.warmup,
                                 autotune=not args.no_autotune, masked=args.masked_only)
        log(f"VGG16 on random sharings: {rs['ms_per_forward']:.3f} ms/forward, {rs['images_per_s']:.1f} images/s")
        print(json.dumps({"random_sharings": rs}), flush=True)
        save_plans()
        return

    layers, batch, fp = C.VGG16_CIFAR, args.batch, 26
    log(f"VGG16-CIFAR b{batch}: building (fused limb-MFMA, hipGraph)")
    net = make_net(P, C, layers, batch, fp, P.MODE_FUSED, autotune=not args.no_autotune)
    log(f"autotuned GEMM shapes: {P.autotune_entries()}")
    save_plans()
    t = time_graph(P, net, args.steps, args.warmup)
    images_s = batch / t
    log(f"VGG16 fused: {t * 1e3:.3f} ms/forward, {images_s:.1f} images/s")

    # dominant kernel: the limb implicit-GEMM convolution (Cin % 32 == 0 form),
    # every launch bracketed by HIP events on the network's stream
    prof = net.profile_gemm()
    for _ in range(2):
        again = net.profile_gemm()
        prof = {v: (min(prof[v][0], again[v][0]),) + prof[v][1:] for v in prof}
    layer_ms = net.profile_layers()
    dom = max(prof, key=lambda v: prof[v][0])
    k_ms, k_ops, k_n = prof[dom]
    avg_ms = k_ms / k_n
    achieved = k_ops / k_n / (avg_ms * 1e-3) / 1e12
    log("limb GEMM kernels: " + ", ".join(f"{kernel_name(P, v)} {prof[v][2]} launches {prof[v][0]:.3f} ms "
                                          f"{prof[v][1] / prof[v][0] / 1e9:.0f} TOP/s" for v in sorted(prof)))
    log("per-layer ms: " + ", ".join(f"{layers[i][0]}{i}:{m:.3f}" for i, m in enumerate(layer_ms)))
    kname = kernel_name(P, dom)
    roofline = {"bound": "mfma", "kernel": f"{kname} (3-party implicit-GEMM conv, int8 limbs)",
                "achieved": achieved, "peak": INT8_PEAK_TOPS, "unit": "TOP/s", "frac": achieved / INT8_PEAK_TOPS,
                "traffic": read_traffic(kname), "avg_launch_ms": avg_ms, "launches_per_forward": k_n,
                "ops_per_launch": k_ops / k_n,
                "share_of_forward": k_ms / float(layer_ms.sum())}
    hbm = hbm_roofline(P, layers, batch, layer_ms)
    peaks = measured_peaks(P)
    roofline["measured_peak"] = peaks["mfma_i8"]["value"]
    roofline["frac_of_measured_peak"] = achieved / peaks["mfma_i8"]["value"]
    for k, v in hbm.items():
        v["measured_copy_peak"] = peaks["hbm_copy"]["value"]
        v["frac_of_measured_copy"] = v["achieved"] / peaks["hbm_copy"]["value"]
        log(f"{k}: {v['achieved']:.0f} GB/s ({100 * v['frac']:.1f}% of HBM peak, "
            f"{100 * v['frac_of_measured_copy']:.1f}% of the measured copy)")
    xin = np.random.default_rng(1).uniform(0.0, 1.0, batch * C.input_size(layers))  # make_net's input
    acc = fp64_error(net, layers, batch, fp, xin)
    log(f"reconstructed logits vs FP64 forward: max abs err {acc['max_abs_err']:.3g} "
        f"(max |logit| {acc['max_abs_ref']:.3g})")
    del net

    result = {
        "metric": "3PC secure inference images/s (1 GPU) + ring-GEMM Gop/s, limb-MFMA vs plain",
        "value": images_s, "unit": "images/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"3PC VGG16-CIFAR inference, batch {batch}, 32x32x3 inputs U[0,1), "
                               "13 conv3x3 + 5 maxpool2x2 + FC 512-256-256-10, Z_2^64 FP26, "
                               "He-init weights, all 3 parties on one GPU, hipGraph forward",
                   "global_batch": batch, "parallelism": "none (1 GPU, 3 simulated parties)"},
        "roofline": roofline,
        "hbm_roofline": hbm,
        "measured_peaks": peaks,
        "fp64_error": acc,
    }

    if not args.quick:
        log("VGG16 plain one-launch-per-op form")
        pnet = make_net(P, C, layers, batch, fp, P.MODE_PLAIN, capture=False)
        tp = time_eager(pnet, args.plain_steps, 1)
        del pnet
        result["plain_images_per_s"] = batch / tp
        result["speedup_vs_plain"] = images_s / (batch / tp)
        log(f"VGG16 plain: {tp * 1e3:.1f} ms/forward, {batch / tp:.1f} images/s, speedup {result['speedup_vs_plain']:.1f}x")

        log("VGG16 training step")
        tr = bench_training(P, C, layers, batch, fp, max(10, args.steps // 4), 20, autotune=not args.no_autotune)
        result["training"] = tr
        log(f"VGG16 training: {tr['ms_per_step']:.2f} ms/step, {tr['images_per_s']:.1f} images/s; " +
            ", ".join(f"{k} {v['launches']}x {v['ms']:.3f} ms {v['TOP/s']:.0f} TOP/s"
                      for k, v in tr["limb_gemm_per_step"].items()))

        sizes = [int(s) for s in args.gemm_sizes.split(",") if s]
        result["ring_gemm"] = bench_ring_gemm(P, sizes, reps=2)
        for rec in result["ring_gemm"]:
            # the plain form's 64-bit multiply-adds (3 n^3) against the measured VALU rate
            rec["plain_valu_frac"] = rec["plain_gops"] / 2.0 / peaks["valu_u64_muladd"]["value"]
            rec["limb_frac_of_measured_mfma"] = rec["limb_int8_tops"] / peaks["mfma_i8"]["value"]

        other = {}
        for name, L, b, f in [("secureml_b128_fp13", C.SECUREML, 128, 13), ("lenet_b128_fp13", C.LENET, 128, 13)]:
            n = make_net(P, C, L, b, f, P.MODE_FUSED, sigma=0.1)  # N(0, 0.1) weights
            tt = time_graph(P, n, max(5, args.steps // 2), 2)
            other[name] = {"images_per_s": b / tt, "ms_per_forward": tt * 1e3}
            log(f"{name}: {b / tt:.1f} images/s")
            del n
        other["alexnet_uneven_64q_fp20"] = bench_uneven(P, C, max(5, args.steps // 2), 20,
                                                        autotune=not args.no_autotune)
        rs = bench_random_shares(P, C, layers, batch, fp, args.steps, 20, autotune=not args.no_autotune)
        dk = rs["dominant_kernel"]
        dk["frac_of_datasheet"] = dk["achieved_tops"] / INT8_PEAK_TOPS
        dk["frac_of_measured_peak"] = dk["achieved_tops"] / peaks["mfma_i8"]["value"]
        dk["frac_of_measured_full_entropy_peak"] = dk["achieved_tops"] / peaks["mfma_i8_full_entropy"]["value"]
        other["vgg16_b64_fp26_random_sharings"] = rs
        log(f"VGG16 on random sharings: {rs['images_per_s']:.1f} images/s")
        rm = bench_random_shares(P, C, layers, batch, fp, args.steps, 20, autotune=not args.no_autotune, masked=True)
        other["vgg16_b64_fp26_masked_precompute"] = rm
        log(f"VGG16 in the masked precompute mode: {rm['images_per_s']:.1f} images/s")
        result["other_configs"] = other

        log("CPU baseline (oracle)")
        result["cpu_baseline"] = cpu_baseline(C, args.cpu_batch)
        log(f"CPU: {result['cpu_baseline']['value']:.3f} images/s on {result['cpu_baseline']['cores']} threads")

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
