"""
Autotuned pass choices kept in a JSON file, so that a plan of a batch seen before takes its
tuned passes without timing them again.

One entry per line-length class of a plan's pass (the fields whose lines along one axis have
the same length), keyed by everything the timing depended on: the device's architecture and
compute-unit count, a hash of the library build, the precision, the axis, the line length,
whether the blocked intermediate is allowed, and the multiset of the class's fields as
(grid size, half kernel size along the axis, num_iter, method).  Point counts are not part of
the key: the same grids with other stations (the next time step) reuse the entry.

    cache = TuneCache('~/.cache/fastbarnes/tune.json')
    batch.autotune(d_pts, d_val, d_out, stream, cache=cache)   # 'tuned' the first time, then 'cache'
"""
import hashlib
import json
import os

VERSION = 1

_build_id = None


def build_id():
    """Hash of the loaded library file (choices of another build may not exist in this one)."""
    global _build_id
    if _build_id is None:
        from fastbarnes import _lib
        _lib.lib()
        h = hashlib.sha1()
        with open(_lib.LIB_PATH, 'rb') as f:
            for block in iter(lambda: f.read(1 << 20), b''):
                h.update(block)
        _build_id = h.hexdigest()[:16]
    return _build_id


def class_key(arch, compute_units, precision, axis, length, blocked, members):
    """The cache key of one line-length class; members: (size, T, num_iter, method) per field."""
    counts = {}
    for m in members:
        counts[m] = counts.get(m, 0) + 1
    body = dict(arch=arch, cus=int(compute_units), build=build_id(), precision=precision, axis=int(axis),
                len=int(length), blocked=bool(blocked),
                fields=sorted([list(m[0]), int(m[1]), int(m[2]), int(m[3]), c] for m, c in counts.items()))
    return json.dumps(body, sort_keys=True, separators=(',', ':'))


class TuneCache:
    """Entries {key: {'kind': 'slide'|'prefix', 'seg': int, 'threads': int}} in a JSON file.
    A file of another format version, or one that does not parse, is ignored (and replaced on
    save)."""

    def __init__(self, path):
        self.path = os.path.expanduser(path)
        self.entries = {}
        try:
            with open(self.path) as f:
                d = json.load(f)
            if isinstance(d, dict) and d.get('version') == VERSION and isinstance(d.get('entries'), dict):
                self.entries = d['entries']
        except (OSError, ValueError):
            pass

    def get(self, key):
        e = self.entries.get(key)
        if not isinstance(e, dict) or e.get('kind') not in ('slide', 'prefix'):
            return None
        try:
            return dict(kind=e['kind'], seg=int(e['seg']), threads=int(e['threads']))
        except (KeyError, TypeError, ValueError):
            return None

    def put(self, key, kind, seg, threads):
        self.entries[key] = dict(kind=kind, seg=int(seg), threads=int(threads))

    def save(self):
        """Writes the file atomically (a temporary file renamed over it)."""
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = '%s.tmp%d' % (self.path, os.getpid())
        with open(tmp, 'w') as f:
            json.dump(dict(version=VERSION, entries=self.entries), f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)
