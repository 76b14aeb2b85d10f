"""Serving layer: the LRU scene store's policy on the CPU (a stand-in builder, no device), and on
the GPU the request loop's images equal each request rendered alone, under a budget small enough
to force evictions and rebuilds."""
import math

import numpy as np
import pytest

from rendertoy_amd import serving


class Fake:
    def __init__(self, n):
        self.nbytes = n
        self.pins = 0


def make_store(budget):
    built = []

    def build(spec):
        built.append(spec)
        return Fake(spec)
    return serving.SceneStore({k: 100 * (k + 1) for k in range(5)}, budget, build=build,
                              estimate=lambda spec: spec), built


def test_lru_evicts_least_recently_used_unpinned():
    st, built = make_store(budget=600)
    for k in (0, 1, 2):                        # 100 + 200 + 300 = 600 bytes resident
        st.release(st.acquire(k))
    assert st.bytes == 600 and st.misses == 3
    st.release(st.acquire(0))                  # hit: 0 becomes most recently used
    assert st.hits == 1
    st.release(st.acquire(3))                  # needs 400: evicts 1 (200) and 2 (300), keeps 0
    assert list(st.resident) == [0, 3] and st.evictions == 2 and st.bytes == 500
    assert built == [100, 200, 300, 400]


def test_pinned_scenes_are_not_evicted():
    st, _ = make_store(budget=700)
    a = st.acquire(0)                          # pinned (in a queued batch)
    st.release(st.acquire(1))
    st.release(st.acquire(2))                  # 600 resident
    st.acquire(3)                              # evicts 1 (oldest unpinned) and 2, not 0
    assert 0 in st.resident and 1 not in st.resident and 2 not in st.resident
    with pytest.raises(MemoryError):
        st.acquire(4)                          # 500 more cannot fit: 0 and 3 are pinned
    st.release(a)


def test_poisson_arrivals_and_request_mix():
    t = serving.poisson_arrivals(1000.0, 20000, 7)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1e-3) < 3e-5 and abs(gaps.std() - 1e-3) < 5e-5   # exponential
    reqs = serving.make_requests(3000, 64, 500.0, 11)
    eyes = np.array([r.eye for r in reqs])
    np.testing.assert_allclose(np.linalg.norm(eyes, axis=1), 3.0, rtol=1e-12)
    assert abs(eyes.mean(0)).max() < 0.15                   # uniform on the sphere
    assert {r.size for r in reqs} == {256, 512, 1024}
    assert min(r.scene for r in reqs) == 0 and max(r.scene for r in reqs) == 63


@pytest.mark.gpu
def test_served_images_equal_alone_renders_under_eviction(ren):
    import torch
    from rendertoy_amd import scenes
    specs = {}
    for k, f in enumerate((6, 20, 33)):
        d = scenes.icosphere(f, 500 + k)
        specs[k] = serving.SceneSpec([d], scenes.value_noise_rgba8(32, 600 + k) if k != 1 else None)
    sizes = {k: serving.scene_bytes_estimate(s.tris(), s.texture)[0] for k, s in specs.items()}
    budget = sizes[2] + sizes[1] + 4096       # never all three at once
    store = serving.SceneStore(specs, budget)
    reqs = serving.make_requests(24, 3, 1000.0, 21, sizes=(16, 40, 64))
    for r in reqs:
        r.keep_image = True
    srv = serving.RenderServer(store, max_batch=5)
    srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    assert store.evictions > 0 and store.hits > 0
    assert all(r.latency > 0 for r in reqs), [r.latency for r in reqs]
    assert sum(srv.batches) == len(reqs)
    for r in reqs:
        res = store.acquire(r.scene)
        img = res.raycaster.render(serving.camera_for(r.eye), r.size, r.size, textures=res.binding)
        store.release(res)
        assert torch.equal(img.view(torch.int32), r.image.view(torch.int32)), r
