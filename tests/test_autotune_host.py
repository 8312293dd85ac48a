"""The autotuner's selection logic (the GPU timing itself is replaced by a
table): the fastest brick shape at the first chunk, then the fastest chunk
for that shape; unfit chunks are rejected."""
import pytest

from hrdayapy_amd import autotune


class _Eng:
    def __init__(self, brick):
        self.brick = brick


def test_tune_picks_fastest_shape_then_chunk(monkeypatch):
    times = {((4, 4, 4), 50): 90e-6, ((8, 4, 4), 50): 80e-6, ((8, 8, 4), 50): 85e-6,
             ((8, 8, 8), 50): 95e-6, ((8, 4, 4), 20): 82e-6, ((8, 4, 4), 10): 79e-6}
    monkeypatch.setattr(autotune, "time_steps", lambda eng, steps, chunk: times[(eng.brick, chunk)])
    monkeypatch.setattr(autotune.torch.cuda, "empty_cache", lambda: None)
    built = []

    def build(b):
        built.append(b)
        return _Eng(b)

    brick, chunk, table = autotune.tune(build, chunks=(50, 20, 10), steps=100)
    assert brick == (8, 4, 4) and chunk == 10
    assert built == list(autotune.BRICK_SHAPES)
    assert table["shape_us"] == {"4x4x4": 90.0, "8x4x4": 80.0, "8x8x4": 85.0, "8x8x8": 95.0}
    assert table["chunk_us"] == {50: 80.0, 20: 82.0, 10: 79.0}


def test_tune_rejects_odd_or_non_dividing_chunks():
    for bad in ((50, 30), (50, 25)):
        with pytest.raises(ValueError):
            autotune.tune(lambda b: _Eng(b), chunks=bad, steps=100)
