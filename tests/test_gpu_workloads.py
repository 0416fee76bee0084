"""The batched and serving workloads the measurements quote, on the GPU at
full size, checked the way the reference's own results are defined: every
image equals the same view rendered alone (the plain one-view call), bit
for bit.

* batched: 4 resident scenes of 0.5/1.5/3/6 M Gaussians, the 24-request mix
  (640x480 .. 2064x2208, half stereo pairs, half foveated with three levels),
  autotuned, then replayed from its captured graph;
* serving: 8 headsets requesting stereo 2064x2208 frames (three levels) at
  90 Hz over 8 resident scenes of 1..6 M Gaussians (24 M), 0.3 s of arrivals,
  continuous batching on two workers (batches of up to two headsets).

Images are compared on the device (a served stereo frame is 109 MB).
"""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _same_bits(a: torch.Tensor, b: torch.Tensor) -> bool:
    return a.shape == b.shape and torch.equal(a.view(torch.int32), b.view(torch.int32))


@pytest.fixture
def fresh_handle_options():
    from foveatedhtgscudabackend_amd import _lib
    h = _lib.handle_for(0)
    bits, run = h.get_option(_lib.OPT_SORT_DIGIT_BITS), h.get_option(_lib.OPT_BLEND_XCD_RUN)
    yield _lib, h
    h.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH)
    h.set_option(_lib.OPT_SORT_DIGIT_BITS, bits)
    h.set_option(_lib.OPT_BLEND_XCD_RUN, run)
    h.set_option(_lib.OPT_BLEND_WORK_TILE, 1)


def test_batched_workload_equals_views_rendered_alone(fresh_handle_options):
    import bench
    from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer
    from foveatedhtgscudabackend_amd.batch import autotune, render_batch
    lib, h = fresh_handle_options
    scenes = bench.batched_scenes(0, 20.0)
    views = bench.batched_views(0, 8, 3)
    assert len(views) > 24 and len({(v.settings.width, v.settings.height) for v in views}) > 10
    tuned = autotune(scenes, views, frames=1)
    print("autotune:", tuned)
    h.set_option(lib.OPT_FRAME_MODE, lib.FRAME_GRAPH)
    for rep in range(3):  # exact (new shape), captured, replayed
        out = render_batch(scenes, views)
    torch.cuda.synchronize()
    h.set_option(lib.OPT_FRAME_MODE, lib.FRAME_EXACT)
    # the views alone with 8x8 work tiles, whatever work tiles the batch was tuned to
    h.set_option(lib.OPT_BLEND_WORK_TILE, 1)
    rast = {}
    for k, bv in enumerate(views):
        key = (bv.foveation_radius_px, bv.blend_width_px, bv.quarter_radius_px)
        r = rast.setdefault(key, FoveatedHTGSRasterizer(*key))
        alone = r.render(*scenes[bv.scene], bv.settings, False, False, False)
        torch.cuda.synchronize()
        assert _same_bits(out[k], alone), f"view {k} ({bv.settings.width}x{bv.settings.height}, {key})"
    del scenes
    torch.cuda.empty_cache()


def test_serving_workload_equals_frames_rendered_alone(fresh_handle_options):
    import bench
    from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer
    from foveatedhtgscudabackend_amd import serving as V
    lib, h = fresh_handle_options
    scenes = bench.serving_scenes(0, 20.0)
    radii = bench.foveation_radii(3, 2064, 100.0)
    headsets = [V.HeadsetSim(k, k, seed=700 + k) for k in range(8)]
    reqs = V.make_requests(headsets, 0.3, 2064, 2208, 100.0, foveation=radii)
    srv = V.Server(scenes, workers=2, max_batch_headsets=2, foveation=radii)
    for hs in headsets:
        assert srv.admit(hs.hid, hs.scene, 2064, 2208), f"headset {hs.hid} not admitted"
    srv.warm_up(reqs)
    stats = srv.run(reqs)
    print(stats)
    assert stats["requests"] == len(reqs) >= 8 * 26
    assert stats["served"] + stats["dropped"] == len(reqs) and stats["served"] > len(reqs) // 2
    h.set_option(lib.OPT_FRAME_MODE, lib.FRAME_EXACT)
    rast = FoveatedHTGSRasterizer(*radii)
    checked = 0
    for r in srv.done:
        if r.dropped:
            continue
        a, b = rast.render_stereo(*scenes[srv.admitted[r.headset]], *r.settings, False, False, False)
        torch.cuda.synchronize()
        for got, want in ((r.images[0], a), (r.images[1], b)):
            assert _same_bits(got, want), f"headset {r.headset} frame {r.frame}"
        r.images = None
        checked += 1
    assert checked == stats["served"]
    del srv, scenes
    torch.cuda.empty_cache()
