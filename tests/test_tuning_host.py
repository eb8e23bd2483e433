"""The tile table of the ensemble autotuner on the host: keys, apply, and the
JSON file round trip (no device needed; the timing itself is a GPU test in
tests/test_gpu_batch.py)."""
import types

import _helpers as H


class _Stub:
    """Stands in for a model: grid, batch, form, precision and set_tiling."""

    def __init__(self, Nr, Nt, B=1, form=3, precision=4):
        self.grid = types.SimpleNamespace(Nr=Nr, Nt=Nt)
        self.B, self.form, self.precision = B, form, precision
        self.tiles = []

    def set_tiling(self, t):
        self.tiles.append(t)


def test_keys_apply_and_file_round_trip(tmp_path):
    pkg = H.load_pkg()
    T = pkg.tuning
    a, b = _Stub(1024, 4096, B=10), _Stub(256, 1024, B=6, form=0)
    assert T.tile_key(a) != T.tile_key(b)
    assert T.tile_key(a) == T.tile_key(_Stub(1024, 4096, B=10))
    table = {T.tile_key(a): {"tiling": pkg.lib.TILING_128, "ms": {"128": 0.07, "64": 0.08}}}
    path = tmp_path / "tiles.json"
    T.save(table, path)
    back = T.load(path)
    assert back == table
    assert T.apply([a, b], back) == 1
    assert a.tiles == [pkg.lib.TILING_128] and b.tiles == []
