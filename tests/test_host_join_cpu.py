"""HostHashMap (core/host_structs.hpp:9-78) on the host: inserts of the rows that pass the
build predicates, payloads or rids returned by probes, misses, capacity refusal."""
import numpy as np
import pytest


def test_host_hashmap_matches_device_semantics(golap):
    rng = np.random.default_rng(4)
    keys = rng.permutation(50_000).astype(np.int32) * 3 - 70_000
    attr = rng.integers(0, 25, keys.size).astype(np.int32)
    pay = rng.integers(-5, 1000, keys.size).astype(np.int32)
    m = golap.HostHashMap(keys)
    m.build([(attr, 3, 5)], payload=pay, threads=4)
    keep = (attr >= 3) & (attr <= 5)
    assert m.filled == int(keep.sum())
    probe = np.concatenate([keys, np.array([1, 2, 4, 2**31 - 1, -2**31 + 1], np.int32)])
    out, hits = m.probe(probe)
    want = np.where(keep, pay.astype(np.uint32).astype(np.uint64), np.uint64(2**64 - 1))
    assert np.array_equal(out[:keys.size], want) and hits == int(keep.sum())
    assert np.all(out[keys.size:] == np.uint64(2**64 - 1))
    m2 = golap.HostHashMap(keys)  # rids without payload
    m2.build()
    out2, _ = m2.probe(keys[::7])
    assert np.array_equal(out2, np.arange(0, keys.size, 7, dtype=np.uint64))
    m3 = golap.HostHashMap(keys, ht_size=10)  # more rows than ht_size qualify
    with pytest.raises(golap.GolapError):
        m3.build()
    with pytest.raises(ValueError):  # a predicate column shorter than the keys
        golap.HostHashMap(keys).build([(attr[:10], 0, 5)])
