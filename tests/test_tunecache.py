"""
The autotuner's JSON cache (fastbarnes/tunecache.py) on the CPU: keys ignore the order of a
class's fields and change with every input of the timing; the file round-trips, and a file of
another format or one that does not parse is ignored.  The GPU side (a second plan taking its
choices from the file) is tests/test_gpu_batch.py::test_autotune_cache_round_trip.
"""
import json

from fastbarnes import tunecache


def _key(**kw):
    a = dict(arch='gfx950:sramecc+:xnack-', compute_units=256, precision='fp32c', axis=0, length=2400,
             blocked=True, members=[((2400, 1200), 87, 4, 0), ((2400, 1200), 87, 4, 0), ((2400, 1200), 41, 3, 0)])
    a.update(kw)
    return tunecache.class_key(a['arch'], a['compute_units'], a['precision'], a['axis'], a['length'], a['blocked'],
                               a['members'])


def test_key_ignores_member_order_and_sees_each_input():
    k = _key()
    assert _key(members=[((2400, 1200), 41, 3, 0), ((2400, 1200), 87, 4, 0), ((2400, 1200), 87, 4, 0)]) == k
    others = [_key(arch='gfx942'), _key(compute_units=304), _key(precision='fp64'), _key(axis=1), _key(length=1200),
              _key(blocked=False), _key(members=[((2400, 1200), 87, 4, 0), ((2400, 1200), 41, 3, 0)]),
              _key(members=[((2400, 1200), 87, 4, 0)] * 2 + [((2400, 1200), 41, 4, 0)]),
              _key(members=[((2400, 1200), 87, 4, 0)] * 2 + [((2400, 1200), 41, 3, 1)]),
              _key(members=[((2400, 1200), 87, 4, 0)] * 2 + [((2400, 600), 41, 3, 0)])]
    assert len(set(others)) == len(others) and k not in others
    assert json.loads(k)['build'] == tunecache.build_id()


def test_file_round_trip_and_bad_files(tmp_path):
    path = tmp_path / 'sub' / 'tune.json'
    c = tunecache.TuneCache(str(path))
    assert c.get(_key()) is None
    c.put(_key(), 'prefix', 19, 512)
    c.put(_key(axis=1), 'slide', 39, 256)
    c.save()
    c2 = tunecache.TuneCache(str(path))
    assert c2.get(_key()) == dict(kind='prefix', seg=19, threads=512)
    assert c2.get(_key(axis=1)) == dict(kind='slide', seg=39, threads=256)
    path.write_text('{not json')
    assert tunecache.TuneCache(str(path)).entries == {}
    path.write_text(json.dumps(dict(version=tunecache.VERSION + 1, entries={_key(): dict(kind='slide', seg=39,
                                                                                          threads=256)})))
    assert tunecache.TuneCache(str(path)).entries == {}
    path.write_text(json.dumps(dict(version=tunecache.VERSION, entries={_key(): dict(kind='bogus', seg=39,
                                                                                      threads=256)})))
    assert tunecache.TuneCache(str(path)).get(_key()) is None
