"""Table images on the GPU: round trips of every layout, and images of the
reference's own tables -- the CPU_version's packed words and the RSQF blocks
in the golden fixtures (written by the compiled reference code) -- loaded and
queried, or reproduced bit for bit by saving."""
import numpy as np
import pytest

import gpuquotientfilters_amd as G
import oracle as O
from conftest import golden
from gpu_helpers import dev, require_gpu, rsqf_lookup, say, sqf_lookup
from image_util import make_image

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


@pytest.mark.parametrize("q,r", [(16, 5), (12, 13)])
def test_sqf_round_trips(q, r):
    say("sqf image round trip", q, r)
    keys = O.mt19937(5, O.key_count(0.8, q))
    f = G.SQF(q, r)
    f.bulk_build(dev(keys), NoDuplicates=False)
    probe = np.concatenate([keys[:500], O.mt19937(6, 500)])
    want = sqf_lookup(f, probe)
    for layout in (G.IMAGE_SQF_GPU, G.IMAGE_SQF_CPU):
        img = G.image_save(f, layout)
        assert G.image_info(img) == (layout, q, r, G.HASH_AP)
        g = G.SQF(q, r)
        G.image_load(g, img)
        assert (g.table_bytes() == f.table_bytes()).all(), layout
        assert (sqf_lookup(g, probe) == want).all(), layout


@pytest.mark.parametrize("name", ["sqf_fnv_q12_r5_a0.6.npz", "sqf_fnv_q14_r5_a0.9.npz", "sqf_fnv_q12_r13_a0.8.npz"])
def test_sqf_cpu_layout_is_the_reference_table(name):
    """FNV build saved in the CPU layout == the reference CPU_version's words
    byte for byte; the reference's words loaded == the GPU build."""
    g = golden(name)
    q, r = int(g["q"]), int(g["r"])
    say("sqf image vs reference cpu words", name)
    G.set_sqf_hash(G.HASH_FNV)
    try:
        f = G.SQF(q, r)
        f.bulk_build(dev(g["keys"]), NoDuplicates=True)
        img = G.image_save(f, G.IMAGE_SQF_CPU)
        assert G.image_info(img) == (G.IMAGE_SQF_CPU, q, r, G.HASH_FNV)
        words = np.frombuffer(img[64:], dtype="<u4")
        assert len(words) == len(g["words"]) and (words == g["words"]).all()
        h = G.SQF(q, r)
        G.image_load(h, make_image(G.IMAGE_SQF_CPU, q, r, g["words"].astype("<u4").tobytes(), hash_kind=1))
        assert (h.table_bytes() == f.table_bytes()).all()
        assert (sqf_lookup(h, g["lookup_keys"]) == g["lookup_ret"]).all()
    finally:
        G.set_sqf_hash(G.HASH_AP)


@pytest.mark.parametrize("name", ["rsqf_r5_q10_a0.6.npz", "rsqf_r5_q12_a0.9.npz"])
def test_rsqf_images_of_reference_tables(name):
    g = golden(name)
    q, r = int(g["q"]), int(g["r"])
    say("rsqf image", name)
    for tab, res in (("table_random", "lookup_random"), ("table_ascending", "lookup_ascending")):
        f = G.RSQF(q, r)
        G.image_load(f, make_image(G.IMAGE_RSQF, q, r, g[tab].tobytes(), units=G.RSQF.calcNumBlocksGPU(q)))
        assert (f.dump() == g[tab]).all()
        assert (rsqf_lookup(f, g["lookup_keys"]) == g[res]).all()
        img = G.image_save(f)
        assert img[64:] == g[tab].tobytes()
        h = G.RSQF(q, r)
        G.image_load(h, img)
        assert (h.dump() == g[tab]).all()
