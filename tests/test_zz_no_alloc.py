"""exa_sampler_locate_noalloc, the allocation-free point location, called
through the C ABI against exa_sampler_locate (stream-ordered scratch) on the
same points: two calls in a row, the second smaller and on another stream,
so a queue count left by the first call would show; results bit-identical.
Sizes beyond EXA_LOCATE_NOALLOC_MAX are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from owlexastitcher_amd import ExaStitchSampler, lib
from owlexastitcher_amd._capi import EXA_ERR_INVALID
from tests import scenes

pytestmark = pytest.mark.gpu


def _p(t):
    return C.c_void_p(t.data_ptr())


def test_locate_noalloc_matches_locate(device):
    import torch
    sc = scenes.noise_small()
    sc.model.setNumGridCells((16, 16, 16))
    s = ExaStitchSampler()
    s.build(sc.model)
    L = lib()
    b = sc.cell_bounds
    side = torch.cuda.Stream()
    for n, seed, stream in [(1 << 20, 1, None), ((1 << 14) + 3, 2, side)]:
        print(f"case: locate_noalloc vs locate, noise_small scene, {n} points, seed {seed}, "
              f"{'side' if stream else 'default'} stream", flush=True)
        pos = np.random.default_rng(seed).uniform(b[:3] - 2, b[3:] + 2, (n, 3)).astype(np.float32)
        d_pos = torch.from_numpy(pos).to(device)
        ref = torch.full((n, 3), -7, dtype=torch.int32, device=device)
        got = torch.full((n, 3), 7, dtype=torch.int32, device=device)
        torch.cuda.synchronize()
        h = C.c_void_p(stream.cuda_stream) if stream is not None else None
        assert L.exa_sampler_locate(s.h, _p(d_pos), n, _p(ref), h) == 0
        assert L.exa_sampler_locate_noalloc(s.h, _p(d_pos), n, _p(got), h) == 0
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
        out = got.cpu().numpy()
        assert np.count_nonzero((out[:, 0] >= 0) & (out[:, 1] == -1)) > n // 100  # element stage ran
        assert np.count_nonzero(out[:, 0] < 0) > 0 and np.count_nonzero(out[:, 1] >= 0) > n // 10
    rc = L.exa_sampler_locate_noalloc(s.h, _p(d_pos), (1 << 24) + 1, _p(got), None)
    assert rc == EXA_ERR_INVALID
    assert b"EXA_LOCATE_NOALLOC_MAX" in L.exa_get_last_error()
    assert L.exa_sampler_locate_noalloc(s.h, _p(d_pos), 0, _p(got), None) == 0
