"""gf_transform_batch_noalloc against gf_transform_batch through the C ABI: same inputs, same
reservation, identical output descriptors and slot contents, on two calls in a row (the second
smaller), with the kernel's scratch in the library's static device array."""
import ctypes as ct

import numpy as np
import pytest

import graphflow_amd as g
from gpu_util import SEED, torch_cuda

pytestmark = pytest.mark.gpu


def _fill(pool, handles, fills, rng):
    for h, fill in zip(handles, fills):
        base = pool.base(h)
        pool.write(0, base, rng.random(fill, dtype=np.float32))
        pool.write(1, base, rng.integers(0, 2**32, fill, dtype=np.uint64).astype(np.uint32))
        pool.write(2, base, np.arange(fill, dtype=np.uint32))


def _cols(pool, outs, C):
    res = []
    for col in range(3):
        for o in outs:
            a = np.zeros(o[2], np.uint32)
            pool.read(col, o[0] * C, a)
            res.append(a)
    return np.concatenate(res) if res else np.zeros(0, np.uint32)


@pytest.mark.parametrize("B,bin_path,T", [(7, g.BIN_PATH_AUTO, 1024), (4096, g.BIN_PATH_AUTO, 1024),
                                          (300, g.BIN_PATH_HASH, 1024)])
def test_noalloc_equals_fused_stage_twice(B, bin_path, T):
    torch = torch_cuda()
    C = 1000
    node = g.IdMod(B, 0).desc()
    tun = g.Tuning(seed=SEED, batch_chunks=4, tile_items=T, bin_path=bin_path)
    pa = g.Pool(g.PAYLOAD_PARTICLE, 8192, C)
    pb = g.Pool(g.PAYLOAD_PARTICLE, 8192, C)
    stage = g.Stage(pa, node, 1, SEED, 4, tile_items=T, bin_path=bin_path)
    for call, fills in enumerate(([1000, 37, 613, 999], [5, 400])):
        rng_a, rng_b = np.random.default_rng(call), np.random.default_rng(call)
        ha = [pa.acquire() for _ in fills]
        hb = [pb.acquire() for _ in fills]
        _fill(pa, ha, fills, rng_a)
        _fill(pb, hb, fills, rng_b)
        # existing entry point: the stage reserves, launches once, returns the unused reservation
        stage.transform_batch([(h.index, h.generation, f, 0) for h, f in zip(ha, fills)])
        more, R = stage.advance()
        outs_a = stage.finish()
        # no-alloc entry point: the caller's device buffers and reservation
        bound = g.output_chunk_bound(sum(fills), C, B)
        res_h = [pb.acquire() for _ in range(bound)]
        chunks = torch.tensor([[h.index, h.generation, f, 0] for h, f in zip(hb, fills)], dtype=torch.int64)
        chunks = chunks.to(torch.int32).cuda()
        reserved = torch.tensor([[h.index, h.generation] for h in res_h], dtype=torch.int64).to(torch.int32).cuda()
        outputs = torch.full((bound, 4), -1, dtype=torch.int32, device="cuda")
        result = torch.full((2,), -1, dtype=torch.int32, device="cuda")
        tc = tun.c()
        rc = g.lib().gf_transform_batch_noalloc(pb.handle, ct.byref(node), 1, ct.byref(tc), chunks.data_ptr(), len(fills),
                                              reserved.data_ptr(), bound, outputs.data_ptr(), result.data_ptr(), None)
        assert rc == g.OK
        torch.cuda.synchronize()
        r = result.cpu().numpy().view(np.uint32)
        assert r[1] == 0 and r[0] == R == len(outs_a)
        outs_b = [tuple(int(x) for x in row) for row in outputs[:R].cpu().numpy().view(np.uint32)]
        # the same chunks in the same order: both pools reserved alike (the released tail of the first
        # reservation has a newer generation than the stage's cancelled one, so generations are
        # compared on the first call)
        assert [(o[0], o[2], o[3]) for o in outs_b] == [(o[0], o[2], o[3]) for o in outs_a]
        if call == 0:
            assert outs_b == outs_a
        assert (_cols(pa, outs_a, C) == _cols(pb, outs_b, C)).all()
        for h in reversed(res_h[R:]):
            pb.release(h)
        for o in outs_b:
            pb.release(g.ChunkHandle(o[0], o[1]))
        for o in outs_a:
            pa.release(g.ChunkHandle(o[0], o[1]))
        for h in ha:
            pa.release(h)
        for h in hb:
            pb.release(h)


def test_noalloc_rejects_batches_beyond_its_scratch():
    torch_cuda()
    pool = g.Pool(g.PAYLOAD_PARTICLE, 64, 8192)
    node = g.IdMod(65536, 0).desc()
    tc = g.Tuning(seed=SEED, batch_chunks=64, tile_items=256).c()
    # 64 full chunks of 32 tiles into 65536 bins: 2048 * 65536 count words, far above the scratch
    rc = g.lib().gf_transform_batch_noalloc(pool.handle, ct.byref(node), 1, ct.byref(tc), 1, 64, 1, 1, 1, 1, None)
    assert rc == g.BAD_ARGUMENT
