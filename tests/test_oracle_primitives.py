"""The CPU oracle pinned against the reference's own unit-test literals and published KATs."""
import hashlib
import itertools

import numpy as np
import pytest

import oracle as O


def test_output_chunk_bound_literals():
    # tests/unit/test_reservation.cpp:49-61
    cases = [((0, 128, 2), 0), ((0, 1, 1), 0), ((1, 1, 2), 1), ((17, 127, 2), 2), ((128, 128, 2), 2),
             ((4097, 128, 2), 34), ((4097, 1, 2), 4097), ((64, 128, 64), 64), ((1, 128, 64), 1)]
    for args, want in cases:
        assert O.output_chunk_bound(*args) == (O.OK, want), args


def test_output_chunk_bound_rejections():
    # tests/unit/test_reservation.cpp:64-71
    assert O.output_chunk_bound(4, 0, 2)[0] == O.BAD_ARGUMENT
    assert O.output_chunk_bound(4, 2, 0)[0] == O.BAD_ARGUMENT
    assert O.output_chunk_bound(2**64 - 1, 1, 2)[0] == O.OVERFLOW


def _distributions(n, bins):
    if bins == 1:
        yield (n,)
        return
    for i in range(n + 1):
        for rest in _distributions(n - i, bins - 1):
            yield (i,) + rest


def test_bound_covers_every_distribution():
    # exhaustive small-shape property of tests/unit/test_reservation.cpp:74-101
    for n in range(0, 11):
        for cap in range(1, 6):
            for bins in range(1, 5):
                _, bound = O.output_chunk_bound(n, cap, bins)
                exact_max = max(sum(-(-v // cap) for v in d) for d in _distributions(n, bins))
                assert exact_max <= bound, (n, cap, bins)
                # the bound is attained by the closed form's extremes
                assert bound == min(n, -(-n // cap) + min(bins, n) - 1) if n else bound == 0


def test_plan_dense_bins_literals():
    # tests/unit/test_dense_bin_table.cpp:18-59
    rc, base, fills, r = O.plan_dense_bins([3, 0, 2], 2)
    assert (rc, base, fills, r) == (O.OK, [0, 2, 2], [2, 1, 2], 3)
    rc, base, fills, r = O.plan_dense_bins([0, 0], 8)
    assert (rc, base, fills, r) == (O.OK, [0, 0], [], 0)
    rc, base, fills, r = O.plan_dense_bins([5], 5)
    assert (base, fills, r) == ([0], [5], 1)
    rc, base, fills, r = O.plan_dense_bins([4, 4], 4)
    assert (r, fills) == (2, [4, 4])
    rc, base, fills, r = O.plan_dense_bins([3, 2, 2], 2)
    assert r == 4 and O.output_chunk_bound(7, 2, 3)[1] == 6
    assert O.plan_dense_bins([], 1)[0] == O.BAD_ARGUMENT
    assert O.plan_dense_bins([1], 0)[0] == O.BAD_ARGUMENT


def test_build_dense_table_literal():
    import ctypes as C
    base = np.array([0, 2, 2], np.uint32)
    idx = np.array([10, 11, 12], np.uint32)
    out = np.zeros(6, np.uint32)
    u = C.POINTER(C.c_uint32)
    O.lib().or_build_dense_table(base.ctypes.data_as(u), 3, idx.ctypes.data_as(u), 3, out.ctypes.data_as(u))
    assert out.tolist() == [0, 2, 2, 10, 11, 12]


def test_tile_flattening():
    # C=1000, T=128: seven full tiles and a 104-item tail (tests/integration/test_dense_bins.cpp:195-218)
    rc, tiles = O.flatten_tiles([7], [1000], 128)
    assert rc == O.OK and len(tiles) == 8
    assert tiles[-1] == [7, 896, 104]
    assert all(t[2] == 128 for t in tiles[:-1])
    # partial chunks share one batch, zero fill contributes no tile (Execution.hpp:113-124)
    rc, tiles = O.flatten_tiles([3, 4, 5], [37, 0, 613], 128)
    assert [t[0] for t in tiles] == [3] + [5] * 5
    assert sum(t[2] for t in tiles) == 650
    assert tiles[1:] == [[5, o, min(128, 613 - o)] for o in range(0, 613, 128)]


def test_num_frames_literals():
    # tests/unit/test_launch_policy.cpp:27-57 (maxThreads 1024, maxBlocks 0x7FFFFFFF, 108 SMs)
    f = lambda w, t, cap=8, mb=0x7FFFFFFF, sms=108: O.num_frames_for(w, t, cap, 1024, mb, sms)[1]
    assert [f(0, 128), f(1, 128), f(17, 128), f(127, 128), f(128, 128), f(129, 128), f(4097, 128)] == [0, 1, 1, 1, 1, 2, 33]
    assert f(1 << 20, 32) == 8 * 108
    assert f(1 << 20, 128, mb=64) == 64
    assert f(1 << 20, 128, cap=1, sms=1) == 1
    assert f((0x7FFFFFFF + 1) * 128, 128, cap=2**32 - 1) == 0x7FFFFFFF
    assert O.num_frames_for(1, 0, 8, 1024, 64, 1)[0] == O.BAD_ARGUMENT
    assert O.num_frames_for(1, 2048, 8, 1024, 64, 1)[0] == O.BAD_ARGUMENT


def test_hash_table_geometry():
    # tests/integration/test_hash_bins.cpp:296-305: entries = nextPow2(slots * T)
    L = O.lib()
    assert L.or_next_pow2(2 * 64) == 128
    assert L.or_next_pow2(2 * 128) == 256
    assert L.or_next_pow2(4 * 128) == 512
    assert L.or_next_pow2(2 * 100) == 256
    # slot function is a bijection below `entries` (BinTable.hpp:84-93)
    slots = {L.or_hash_bin_slot(b, 255) for b in range(256)}
    assert len(slots) == 256


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10 (Salmon et al., SC'11)
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


M64 = (1 << 64) - 1


def test_mix64_splitmix64_published_outputs():
    # SplitMix64 (Steele, Lea, Flood, OOPSLA 2014; Vigna's reference splitmix64.c): every output is
    # the finalizer of state += 0x9E3779B97F4A7C15. The first five outputs from state 1234567 are
    # the generator's published test vector; a wrong shift or multiplier fails all of them.
    L = O.lib()
    published = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                 4593380528125082431, 16408922859458223821]
    state = 1234567
    for want in published:
        state = (state + 0x9E3779B97F4A7C15) & M64
        assert L.or_mix64(state) == want
    assert L.or_mix64(0) == 0


def _unxorshift(y, s):
    # inverse of x ^ (x >> s): fixed point of x = y ^ (x >> s), reached after 64 / s rounds
    x = y
    for _ in range(64 // s + 1):
        x = y ^ (x >> s)
    return x


def _unmix64(v):
    # the finalizer's inverse from its algebra: undo each xor-shift, multiply by the modular inverse
    v = _unxorshift(v, 31)
    v = (v * pow(0x94D049BB133111EB, -1, 1 << 64)) & M64
    v = _unxorshift(v, 27)
    v = (v * pow(0xBF58476D1CE4E5B9, -1, 1 << 64)) & M64
    return _unxorshift(v, 30)


def test_stream_key_bit_layout_through_the_inverse_mixer():
    # rng/ParticleRng.hpp:35-38,56-94: id in bits 0-31, event in 32-47, purpose in 48-63, the packed
    # position xored with mix64(seed) before the final mix. Undoing the final mix with the
    # independently derived inverse exposes the packed position, so each field's bit range is
    # checked on its own (a swapped field, a wrong shift or a missing seed mix fails).
    L = O.lib()
    assert all(_unmix64(L.or_mix64(v)) == v for v in (0, 1, 12345, 2**63, M64, 0xDEADBEEFCAFEF00D))
    for seed in (0, 42, M64):
        s = L.or_mix64(seed)
        pos = lambda i, e, p: _unmix64(L.or_stream_key(seed, i, e, p)) ^ s
        assert pos(0, 0, 0) == 0
        assert pos(7, 0, 0) == 7 and pos(0xFFFFFFFF, 0, 0) == 0xFFFFFFFF
        assert pos(0, 1, 0) == 1 << 32 and pos(0, 0xFFFF, 0) == 0xFFFF << 32
        assert pos(0, 0, 1) == 1 << 48 and pos(0, 0, 0xFFFF) == 0xFFFF << 48
        assert pos(3, 5, 9) == 3 | 5 << 32 | 9 << 48
        # fields are masked to their bit budget, never carried into the next field
        assert pos(2**32 + 3, 0, 0) == 3 and pos(0, 0x10005, 0) == 5 << 32
    # uniqueness property of tests/unit/test_philox.cpp:48-61 (64 ids x 64 events x 32 purposes)
    keys = {L.or_stream_key(42, i, e, p) for i in range(64) for e in range(64) for p in range(32)}
    assert len(keys) == 64 * 64 * 32


def _insert(keys, counts, bin_):
    import ctypes as C
    u = C.POINTER(C.c_uint32)
    return O.lib().or_hash_table_insert(keys.ctypes.data_as(u), counts.ctypes.data_as(u), len(keys), bin_)


def _table(entries):
    return np.full(entries, 0xFFFFFFFF, np.uint32), np.zeros(entries, np.uint32)


def test_hash_table_exhaustion():
    # tests/integration/test_hash_bins.cpp:504-546: HashBinTable<4> has nextPow2(2*4) = 8 entries,
    # so the ninth distinct bin of a tile finds the table full; bins already present still count
    L = O.lib()
    keys, counts = _table(L.or_next_pow2(2 * 4))
    assert len(keys) == 8
    assert [_insert(keys, counts, b) for b in range(8)] == [1] * 8
    assert all(_insert(keys, counts, b) == 0 for b in range(8, 16))
    assert _insert(keys, counts, 3) == 1
    assert sorted(keys.tolist()) == list(range(8))
    assert counts[keys.tolist().index(3)] == 2 and counts.sum() == 9


def test_hash_table_probe_sequence():
    # BinTable.hpp:90-121: home slot (bin * 0x9E3779B1) & mask, linear probing on collision
    L = O.lib()
    keys, counts = _table(16)
    home = lambda b: (b * 0x9E3779B1) & 15
    assert _insert(keys, counts, 5) == 1 and keys[home(5)] == 5
    # 5 + 16 has the same home slot (16 * odd = 0 mod 16): it lands one slot further
    assert home(21) == home(5)
    assert _insert(keys, counts, 21) == 1 and keys[(home(5) + 1) & 15] == 21
    assert _insert(keys, counts, 37) == 1 and keys[(home(5) + 2) & 15] == 37
    # the probe wraps past the last slot
    keys, counts = _table(8)
    b = next(x for x in range(100) if (x * 0x9E3779B1) & 7 == 7)
    assert _insert(keys, counts, b) == 1 and _insert(keys, counts, b + 8) == 1
    assert keys[7] == b and keys[0] == b + 8


def test_hash_table_counts_equal_a_brute_force_histogram():
    rng = np.random.default_rng(7)
    for entries, distinct in ((256, 128), (512, 300), (64, 64)):
        keys, counts = _table(entries)
        bins = rng.choice(1 << 20, distinct, replace=False)[rng.integers(0, distinct, 4 * entries)]
        assert all(_insert(keys, counts, int(b)) == 1 for b in bins)
        occ = keys != 0xFFFFFFFF
        got = dict(zip(keys[occ].tolist(), counts[occ].tolist()))
        v, c = np.unique(bins, return_counts=True)
        assert got == dict(zip(v.tolist(), c.tolist()))


def test_fmix32_known_values():
    # murmur3 fmix32 (smhasher): fmix32(0)=0, bijection on a sample, avalanche of 1
    L = O.lib()
    assert L.or_fmix32(0) == 0
    assert L.or_fmix32(1) == 0x514E28B7
    sample = {L.or_fmix32(v) for v in range(1 << 14)}
    assert len(sample) == 1 << 14


def test_uniform_conversions():
    L = O.lib()
    assert L.or_uniform_f32(0) == 2.0 ** -24
    assert L.or_uniform_f32(0xFFFFFFFF) == 1.0
    assert L.or_uniform_f64(0, 0) == 0.0
    assert L.or_uniform_f64(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53


def test_zipf_table_mass():
    # Zipf(s=1.1) over 2^20 values: top-1 mass 12.37 %, top-10 33.15 % (closed form sum_k k^-s)
    t = O.zipf_table(1 << 20, 1.1)
    assert abs(t[0] / t[-1] - 0.1237) < 5e-4
    assert abs(t[9] / t[-1] - 0.3315) < 5e-4
    L = O.lib()
    import ctypes as C
    p = t.ctypes.data_as(C.POINTER(C.c_double))
    assert L.or_zipf_sample(p, 1 << 20, 0, 0) == 0
    assert L.or_zipf_sample(p, 1 << 20, 0xFFFFFFFF, 0xFFFFFFFF) == (1 << 20) - 1


def test_partition_items_is_a_stable_counting_sort():
    rng = np.random.default_rng(1)
    n = 5000
    ids = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    node = dict(role=1, op="CLS_ID_MOD", bins=7, u=[0])
    rc, bins, tot, rank = O.partition_items(node, n, ids=ids)
    assert rc == O.OK
    assert (bins == ids % 7).all()
    order = np.argsort(bins, kind="stable")
    flat = np.concatenate([[0], np.cumsum(tot)[:-1]])[bins] + rank
    assert (np.argsort(flat) == order).all()
    assert tot.tolist() == np.bincount(bins, minlength=7).tolist()


def test_partition_items_rejects_out_of_range_bins():
    lut = [0, 1, 5]
    node = dict(role=1, op="CLS_LUT", bins=3, u=[3, 1, 0, 0])
    rc, *_ = O.partition_items(node, 3, ids=[0, 1, 2], gens=[0, 0, 0], lut=lut)
    assert rc == O.INVALID


def _kernel_axis_rows(n, bins, dist):
    ids = np.arange(n, dtype=np.uint64)
    if dist == "balanced":
        b = ids % bins
    elif dist == "one-hot":
        b = np.zeros_like(ids)
    else:
        b = np.where(ids % 8 == 0, ids % bins, 0)
    v, c = np.unique(b, return_counts=True)
    return b"".join(b"%d,%d\n" % (int(x), int(y)) for x, y in zip(v, c))


@pytest.mark.parametrize("n,bins,dist,sha", [
    (262144, 4096, "balanced", "06f1bd73ff252b1a8f2f51aa879ad8195ac6816d0627d8edb5071fe4dc441e55"),
    (262144, 65536, "balanced", "968f4a008df5ad231cfb5578843033a9d4f2a28233f8568f0c7c6ed51fd638c0"),
    (1000000, 64, "balanced", "f3a977d720ab00c985bb39e0fffab715f57696fe6a762412e14e3f3810ce3505"),
    (1000000, 64, "one-hot", "e59ef1c49b3c32c8714b7d10b1aabc17be313c9124ebee3c53e1337afe7678b4"),
    (1000000, 64, "skewed", "ab5932dacd593dc0ff4734c4903dd396da7d07f00c375aeefc0f83e6224200d2"),
    (1000000, 1024, "balanced", "8b946fbb08966136cf9f17d9547ceb262187e11781345e184da1c0798d2de9a6"),
])
def test_histogram_serialisation_known_answers(n, bins, dist, sha):
    # tools/BENCH_TABLE.md:215-222: the reference's 'w,c\n' ascending rows and their SHA-256
    assert hashlib.sha256(_kernel_axis_rows(n, bins, dist)).hexdigest() == sha


@pytest.mark.parametrize("mode", [O.MODE_RUNNER, O.MODE_STEP])
@pytest.mark.parametrize("n,bins,dist", [(20000, 64, 0), (20000, 64, 1), (20000, 64, 2), (8192, 1024, 0)])
def test_kernel_axis_pipeline_through_the_scheduler(mode, n, bins, dist):
    # Generate -> Partition(id % B, gen := bin) -> Sink(histogram of gen): the scheduler must deliver
    # every id exactly once with its closed-form bin (tests/integration/test_kernel_axes.cpp:119-145)
    g = [dict(role=0, op="SRC_PARTICLE_UNIFORM", bins=1, routes=[1]),
         dict(role=1, op="CLS_ID_MOD", upd="GEN_TO_BIN", bins=bins, routes=[2] * bins, u=[dist]),
         dict(role=2, op="SNK_HIST", bins=0, u=[2, 0, bins])]
    r = O.run(g, seed=42, items=n, capacity=1024, live=4096, batch=8, mode=mode, hist_size=bins)
    assert r["status"] == O.OK, r["message"]
    name = {0: "balanced", 1: "one-hot", 2: "skewed"}[dist]
    rows = O.hist_rows(r["hist"], 0)
    text = b"".join(b"%d,%d\n" % (int(v), int(c)) for v, c in rows)
    assert text == _kernel_axis_rows(n, bins, name)
    assert r["sunk"] == n


def test_scripted_routing_fixture():
    # tests/integration/test_scripted_routing.cpp:16-183 restated with LUT classifiers:
    # Filter -> A/B by (id, visit); A: +1000 on feedback, B: +1 on feedback; weights {0,1000,1,1001}
    # ids 0..3, visits 0..2 (gen counts visits). LUT rows: id major, visit minor.
    filt = [0, 0, 0,   0, 1, 1,   1, 1, 1,   0, 1, 0]
    a = [0, 0, 0,   1, 0, 0,   0, 0, 0,   1, 0, 0]
    b = [0, 0, 0,   0, 0, 0,   1, 0, 0,   0, 1, 0]
    lut = filt + a + b
    g = [dict(role=0, op="SRC_PARTICLE_UNIFORM", bins=1, routes=[1]),
         dict(role=1, op="CLS_LUT", bins=2, routes=[2, 3], u=[4, 3, 0, 0]),
         dict(role=1, op="CLS_LUT", upd="ADD_ON_BIN", bins=2, routes=[4, 1], u=[4, 3, 1, 12], f=[0.0, 1000.0]),
         dict(role=1, op="CLS_LUT", upd="ADD_ON_BIN", bins=2, routes=[4, 1], u=[4, 3, 1, 24], f=[0.0, 1.0]),
         dict(role=2, op="SNK_RECORD_WEIGHT", bins=0)]
    # the source draws a U(0,1] weight; the fixture starts at weight 0, so subtract the start weight
    for cap in (1, 3):
        for mode in (O.MODE_RUNNER, O.MODE_STEP):
            r = O.run(g, seed=42, items=4, capacity=cap, live=4, batch=2, mode=mode, record=True, lut=lut)
            assert r["status"] == O.OK, r["message"]
            w0 = np.array([O.lib().or_uniform_f32(O.stream_words(42, i, 0, 0)[0]) for i in range(4)], np.float32)
            final = r["record"].view(np.float32)
            assert (final - w0).round().tolist() == [0.0, 1000.0, 1.0, 1001.0]


@pytest.mark.parametrize("mode", [O.MODE_RUNNER, O.MODE_STEP])
def test_feedback_network_matches_schedule_free_model(mode):
    g = [dict(role=0, op="SRC_PARTICLE_UNIFORM", bins=1, routes=[1]),
         dict(role=1, op="CLS_WEIGHT_ROUTE", bins=3, routes=[4, 2, 3], u=[0], f=[1e-3]),
         dict(role=1, op="XF_SCALE_WEIGHT", bins=1, routes=[1], f=[0.5]),
         dict(role=1, op="XF_SCALE_WEIGHT", bins=1, routes=[1], f=[0.75]),
         dict(role=2, op="SNK_RECORD_WEIGHT", bins=0)]
    n = 30000
    r = O.run(g, seed=42, items=n, capacity=1024, live=2048, batch=8 if mode else 1, mode=mode, record=True)
    assert r["status"] == O.OK, r["message"]
    w, gens = O.feedback_scalar(42, n)
    assert (r["record"] == w).all()
    # S = P + K*R: Filter outputChunkBound(C, C, 3) = 3, A and B one each (Flow.hpp:580-626)
    assert r["pool_chunks"] == 2048 + 5
    assert r["sunk"] == n
    # mean passes: A-routed particles need ~9 halvings, B ~21 (feedback-network functors)
    assert 13.0 < gens.mean() < 17.0


def test_credit_pressure_shrinks_batch():
    # tests/integration/test_batching.cpp:550-575: 8 one-item chunks into a B=8 join with a tight pool
    g = [dict(role=0, op="SRC_PARTICLE_UNIFORM", bins=1, routes=[1]),
         dict(role=1, op="CLS_ID_MOD", bins=8, routes=[2] * 8, u=[0]),
         dict(role=1, op="CLS_ID_MOD", upd="GEN_TO_BIN", bins=8, routes=[3] * 8, u=[0]),
         dict(role=2, op="SNK_HIST", bins=0, u=[2, 0, 8])]
    r = O.run(g, seed=1, items=8, capacity=8, live=8, batch=8, num_chunks=12, mode=O.MODE_RUNNER,
              trace_cap=10000, hist_size=8)
    assert r["status"] == O.OK, r["message"]
    tr = r["trace"]
    # decode: batches of node 2 and their input chunk counts
    i, joins = 0, []
    while i < len(tr):
        node = tr[i] & 0xFFFF
        n_in = tr[i + 2]
        nb = tr[i + 3 + 3 * n_in]
        R = tr[i + 4 + 3 * n_in + nb]
        if node == 2:
            joins.append(n_in)
        i += 5 + 3 * n_in + nb + 4 * R
    assert max(joins) == 4 and len(joins) == 2
    assert r["hist"].tolist() == [1] * 8


def test_stall_is_reported():
    g = [dict(role=0, op="SRC_PARTICLE_UNIFORM", bins=1, routes=[1]),
         dict(role=1, op="CLS_ID_MOD", bins=8, routes=[2] * 8, u=[0]),
         dict(role=2, op="SNK_HIST", bins=0, u=[2, 0, 8])]
    r = O.run(g, seed=1, items=64, capacity=8, live=64, batch=1, num_chunks=3, mode=O.MODE_RUNNER, hist_size=8)
    assert r["status"] == O.STALL


def test_feedback_golden_fixture_headline():
    # tests/golden/feedback_N1e8.json holds the N=1e8 histograms of the four routing distributions
    # from the schedule-free model; recompute the headline (parity routing) case and its SHA
    import json
    import os

    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "feedback_N1e8.json")))
    case = fx["cases"]["parity"]
    w, _ = O.feedback_scalar(fx["seed"], fx["items"], 1e-3, 0, 2, (0.5, 0.75), threads=os.cpu_count() or 1)
    v, c = np.unique(w, return_counts=True)
    assert len(v) == case["rows"] and int(c.sum()) == case["sum"] == fx["items"]
    text = "".join(f"{a},{b}\n" for a, b in zip(v.tolist(), c.tolist())).encode()
    assert hashlib.sha256(text).hexdigest() == case["sha256"]
