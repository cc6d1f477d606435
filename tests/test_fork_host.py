"""Host logic of the stream fork: order of records and waits, persistent events, nested-fork guard (CPU).

torch's stream and event objects are replaced by recorders, so no device is touched.
"""
import contextlib

import pytest

from pgj import _chunks


class _Rec:
    log = []

    def __init__(self, name):
        self.name = name


class FakeStream(_Rec):
    count = 0

    def __init__(self, name=None):
        FakeStream.count += 1
        super().__init__(name or f's{FakeStream.count}')

    def wait_event(self, ev):
        _Rec.log.append(('wait', self.name, ev.name))


class FakeEvent(_Rec):
    count = 0

    def __init__(self):
        FakeEvent.count += 1
        super().__init__(f'e{FakeEvent.count}')

    def record(self, st):
        _Rec.log.append(('record', self.name, st.name))


@pytest.fixture
def fake_cuda(monkeypatch):
    state = {'current': FakeStream('main'), 'capturing': False}
    cuda = _chunks.torch.cuda
    monkeypatch.setattr(cuda, 'Stream', FakeStream)
    monkeypatch.setattr(cuda, 'Event', FakeEvent)
    monkeypatch.setattr(cuda, 'current_stream', lambda: state['current'])
    monkeypatch.setattr(cuda, 'is_current_stream_capturing', lambda: state['capturing'])

    @contextlib.contextmanager
    def stream(st):
        prev = state['current']
        state['current'] = st
        try:
            yield
        finally:
            state['current'] = prev
    monkeypatch.setattr(cuda, 'stream', stream)
    _Rec.log.clear()
    _chunks.Fork._depth = 0
    return state


def test_split_forks_and_joins_with_the_same_events(fake_cuda):
    split = _chunks.StreamSplit(6, 6, 2)
    seen = []
    events = (split._fork.fork, tuple(split._fork.join))
    for _ in range(2):
        _Rec.log.clear()
        split.run(lambda l0, nl: seen.append((l0, nl, fake_cuda['current'].name)), 6)
        f, (j,) = events
        side = split._fork.side[0].name
        assert _Rec.log == [('record', f.name, 'main'), ('wait', side, f.name), ('record', j.name, side),
                            ('wait', 'main', j.name)]
    assert seen == [(0, 3, 'main'), (3, 3, side)] * 2
    assert (split._fork.fork, tuple(split._fork.join)) == events


def test_nested_fork_under_capture_is_refused(fake_cuda):
    outer = _chunks.Fork(1)
    inner = _chunks.Fork(1)
    noop = lambda: None
    fake_cuda['capturing'] = False
    outer.run([lambda: inner.run([noop, noop]), noop])     # eager nesting is allowed
    fake_cuda['capturing'] = True
    with pytest.raises(RuntimeError, match='nested'):
        outer.run([lambda: inner.run([noop, noop]), noop])
    assert _chunks.Fork._depth == 0
    outer.run([noop, noop])                                 # a single fork under capture is fine
    with pytest.raises(ValueError):
        outer.run([noop])
