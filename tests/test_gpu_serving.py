"""Serving on the GPU: every served stereo frame equals the same request
rendered alone through the plain stereo call, whatever batch it was served in;
statistics account for every request."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("foveation", [(375, 125, 0), (80, 27, 140)], ids=["two_levels", "three_levels"])
def test_served_frames_equal_frames_rendered_alone(foveation):
    from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer
    from foveatedhtgscudabackend_amd import scene as S
    from foveatedhtgscudabackend_amd import serving as V
    scenes = [S.scene_to_torch(S.make_scene(n, seed=80 + k, cube=10.0), "cuda") for k, n in enumerate((20000, 35000))]
    headsets = [V.HeadsetSim(h, h % 2, seed=900 + h) for h in range(3)]
    reqs = V.make_requests(headsets, 0.25, 320, 256, 100.0, foveation=foveation)
    srv = V.Server(scenes, workers=2, max_batch_headsets=3, foveation=foveation)
    for hs in headsets:
        assert srv.admit(hs.hid, hs.scene, 320, 256)
    stats = srv.run(reqs)
    print(stats)
    assert stats["requests"] == len(reqs)
    assert stats["served"] + stats["dropped"] == len(reqs) and stats["served"] > 0
    rast = FoveatedHTGSRasterizer(*foveation)
    checked = 0
    for r in srv.done:
        if r.dropped:
            continue
        left, right = r.settings
        a, b = rast.render_stereo(*scenes[srv.admitted[r.headset]], left, right, False, False, False)
        torch.cuda.synchronize()
        for got, want in ((r.images[0], a), (r.images[1], b)):
            np.testing.assert_array_equal(got.cpu().numpy().view(np.uint32), want.cpu().numpy().view(np.uint32),
                                          err_msg=f"headset {r.headset} frame {r.frame}")
        checked += 1
    assert checked == stats["served"]


def test_admission_refuses_what_does_not_fit():
    from foveatedhtgscudabackend_amd import scene as S
    from foveatedhtgscudabackend_amd import serving as V
    scenes = [S.scene_to_torch(S.make_scene(1000, seed=1, cube=10.0), "cuda")]
    free, _ = torch.cuda.mem_get_info()
    srv = V.Server(scenes, workers=1, reserve_bytes=free + (1 << 30))
    assert not srv.admit(0, 0, 320, 256)
    assert srv.rejected == [0]
