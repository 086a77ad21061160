"""The batched workload at its full size: 1,024 clouds with sizes log-uniform in
[2^10, 2^20], alternating uniform and 16-cluster Gaussian (sigma 4) points in
[0,1024)^3, 30-bit codes.  Every cloud's arrays, built in one segmented launch
per stage (eagerly and by HIP-graph replay, also after a new plan upload),
must equal the oracle's single-cloud build of the same points: the per-cloud
contract of the reference pipeline, src/main.rs:60-170.

With more than 64 clouds the segmented kernels' wave-parallel cloud search
(64 candidates per round) takes a second round, which smaller batches never
reach."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from gpu_util import dev

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]

KEYS = ("sorted", "unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parent", "edge_count",
        "offsets")
SEED = 0xC0FFEE04
N_CLOUDS = 1024


def config4_clouds(tclib):
    rng = np.random.default_rng(SEED)
    sizes = np.round(np.exp(rng.uniform(np.log(2 ** 10), np.log(2 ** 20), N_CLOUDS))).astype(np.uint32)
    clouds = []
    for i, s in enumerate(sizes):
        if i % 2 == 0:
            clouds.append(tclib.gen_uniform(int(s), SEED + i))
        else:
            clouds.append(tclib.gen_gaussian_mixture(int(s), SEED + i, 16, 4.0))
    return sizes, clouds


@pytest.fixture(scope="module")
def workload(oracle, tclib):
    sizes, clouds = config4_clouds(tclib)
    with ThreadPoolExecutor(16) as ex:
        want = list(ex.map(lambda c: oracle.build_tree(c, 0.0, 1024.0, 30), clouds))
    print(f"config-4 batch: {len(clouds)} clouds, {int(sizes.sum())} points", flush=True)
    return sizes, clouds, want


def check_all(got, want, tag):
    assert len(got) == len(want), tag
    bad = []
    for i, (g, w) in enumerate(zip(got, want)):
        for k in KEYS:
            if not np.array_equal(np.asarray(g[k]), np.asarray(w[k])):
                bad.append((i, k))
        if not np.array_equal(g["nodes"].view(np.uint8).reshape(-1), w["nodes"].view(np.uint8).reshape(-1)):
            bad.append((i, "nodes"))
    assert not bad, f"{tag}: {len(bad)} mismatching arrays, first {bad[:8]}"


def test_config4_batch_eager_matches_oracle(gpu_ctx, workload):
    import torch

    import tc

    sizes, clouds, want = workload
    total = int(sizes.sum())
    b = tc.Batch(gpu_ctx, N_CLOUDS, total, 30)
    try:
        pts = dev(np.concatenate(clouds))
        b.set(sizes)
        b.build(pts)
        torch.cuda.synchronize()
        check_all(b.download(), want, "eager")
    finally:
        b.close()


def test_config4_batch_graph_replays_match_oracle(gpu_ctx, workload):
    """Capture once, replay; then upload a new plan (the clouds in reverse
    order, from a second buffer of the same size) and replay again, with eager
    work and synchronisations between the replays."""
    import torch

    import tc

    sizes, clouds, want = workload
    total = int(sizes.sum())
    stream = torch.cuda.Stream()
    b = tc.Batch(gpu_ctx, N_CLOUDS, total, 30)
    try:
        buf = dev(np.concatenate(clouds))
        with torch.cuda.stream(stream):
            b.set(sizes, stream=stream)
            b.capture(buf, stream=stream)
            b.replay(stream=stream)
        stream.synchronize()
        check_all(b.download(stream), want, "replay 1")
        _ = torch.ones(4096, device="cuda").sum().item()  # unrelated eager work between replays
        buf.copy_(dev(np.concatenate(clouds[::-1])))
        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            b.set(sizes[::-1].copy(), stream=stream)
            b.replay(stream=stream)
        stream.synchronize()
        check_all(b.download(stream), want[::-1], "replay 2 (reversed plan)")
        with torch.cuda.stream(stream):
            b.replay(stream=stream)
        stream.synchronize()
        check_all(b.download(stream), want[::-1], "replay 3 (same plan)")
    finally:
        b.close()
