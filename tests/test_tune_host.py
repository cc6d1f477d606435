"""Host logic of the launch-shape tuner: interleaved timing rounds keep each candidate's best time (CPU)."""
from pgj import tune


def test_interleaved_rounds_keep_each_candidates_best(monkeypatch):
    order = []
    current = {}
    # candidate 'b' is faster, but the device is slow during the first round's timing of 'b' only
    script = {('a', 0): 2.0, ('b', 0): 9.0, ('a', 1): 2.1, ('b', 1): 1.5, ('a', 2): 2.2, ('b', 2): 1.6}
    rounds = {'a': 0, 'b': 0}

    def fake_time(launch, reps):
        cand = current['c']
        t = script[(cand, rounds[cand])]
        rounds[cand] += 1
        return t

    monkeypatch.setattr(tune, 'time_launch', fake_time)

    def select(c):
        current['c'] = c
        order.append(c)

    times = tune._fastest(['a', 'b'], select, None, 6, 3)
    assert order == ['a', 'b', 'a', 'b', 'a', 'b']
    assert times == {'a': 2.0, 'b': 1.5}
    assert min(times, key=times.get) == 'b'
