"""fhtgs_sort_pairs_noalloc (the sort with the library's fixed per-device
scratch) through the C ABI on the GPU: on two calls in a row, the first at
the largest size it accepts and the second smaller, it sorts exactly as
fhtgs_sort_pairs (the handle's scratch) does and as a stable sort on the
low end_bit bits does, and reports no inconsistency; sizes past the limit are
rejected before any device work."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _pairs(n, key_bytes, end_bit, seed):
    rng = np.random.default_rng(seed)
    kd = np.uint16 if key_bytes == 2 else np.uint32
    keys = rng.integers(0, 1 << min(end_bit, 8 * key_bytes), size=n, dtype=np.uint64).astype(kd)
    keys[rng.random(n) < 0.1] = np.iinfo(kd).max  # sentinel keys, as the render's sort sees them
    vals = rng.permutation(n).astype(np.uint32)
    return keys, vals, kd


def _to_dev(a):
    view = {np.uint16: np.int16, np.uint32: np.int32}[a.dtype.type]
    return torch.from_numpy(a.view(view).copy()).cuda()


def test_noalloc_sort_equals_handle_sort_on_two_calls():
    from foveatedhtgscudabackend_amd import _lib
    h = _lib.handle_for(0)
    err = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    for call, (n, key_bytes, end_bit) in enumerate(((_lib.SORT_NOALLOC_MAX_PAIRS, 4, 17), (5001, 2, 12))):
        keys, vals, kd = _pairs(n, key_bytes, end_bit, 100 + call)
        ka, va = _to_dev(keys), _to_dev(vals)
        kb, vb = _to_dev(keys), _to_dev(vals)
        _lib.sort_pairs_noalloc(ka.data_ptr(), va.data_ptr(), n, key_bytes, end_bit, err.data_ptr(), stream)
        h.sort_pairs(kb.data_ptr(), vb.data_ptr(), n, key_bytes, end_bit, stream)
        torch.cuda.synchronize()
        assert int(err.item()) == 0, f"call {call}: the sort reported an inconsistency"
        assert torch.equal(ka, kb) and torch.equal(va, vb), f"call {call}: differs from fhtgs_sort_pairs"
        order = np.argsort(keys.astype(np.uint64) & ((1 << end_bit) - 1), kind="stable")
        np.testing.assert_array_equal(ka.cpu().numpy().view(kd), keys[order])
        np.testing.assert_array_equal(va.cpu().numpy().view(np.uint32), vals[order])
        err.fill_(7)


def test_noalloc_sort_rejects_sizes_past_its_scratch():
    from foveatedhtgscudabackend_amd import _lib
    with pytest.raises(ValueError, match="at most"):
        _lib.sort_pairs_noalloc(0, 0, _lib.SORT_NOALLOC_MAX_PAIRS + 1, 4, 17, None, 0)
