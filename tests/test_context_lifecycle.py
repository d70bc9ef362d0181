"""Lifetime of the per-thread default contexts (host logic, no device): a
thread's context is its own, tensors and plans made on it keep it alive, and
it is destroyed once the thread has ended and the last of them is gone --
also when they die together in one garbage-collection pass.  The library is
replaced by a recorder of context and tensor handles."""
import ctypes as C
import gc
import threading

import pytest

from sparsebackends_amd import storage


class FakeLib:
    def __init__(self):
        self.next = 1000
        self.created, self.destroyed, self.freed = [], [], []
        self.log = []  # ("free" | "destroy", handle) in call order
        self.lock = threading.Lock()

    def sb_ctx_create(self, device, out):
        with self.lock:
            self.next += 1
            out._obj.value = self.next
            self.created.append(self.next)
        return 0

    def sb_ctx_destroy(self, h):
        self.destroyed.append(h.value)
        self.log.append(("destroy", h.value))
        return 0

    def sb_tensor_free(self, h):
        self.freed.append(h.value)
        self.log.append(("free", h.value))
        return 0

    def sb_ctx_graphs(self, h, mode):
        return 0


@pytest.fixture
def fake(monkeypatch):
    lib = FakeLib()
    monkeypatch.setattr(storage, "load", lambda: lib)
    yield lib
    storage._drain()  # destroy what the test left pending while the recorder is installed


def in_thread(fn):
    """runs fn in a new thread (so the test's contexts never reach the main
    thread's storage) and returns its result"""
    out = {}

    def run():
        out["v"] = fn()

    t = threading.Thread(target=run)
    t.start()
    t.join()
    return out["v"]


def test_thread_context_is_destroyed_after_the_thread(fake):
    h = in_thread(lambda: storage.context(0).value)
    gc.collect()
    in_thread(lambda: storage.context(0))  # any context() call destroys what is pending
    assert h in fake.destroyed


def test_a_new_thread_never_continues_an_ended_threads_context(fake):
    seen = [in_thread(lambda: storage.context(0).value) for _ in range(4)]
    assert len(set(seen)) == 4


def test_one_context_per_thread_and_device(fake):
    def two():
        return storage.context(0).value, storage.context(0).value, storage.context(1).value

    a, b, c = in_thread(two)
    assert a == b and a != c


def test_tensors_keep_their_context_alive(fake):
    def make():
        t = storage._Tensor(C.c_void_p(77), 0)
        return t, storage.context(0).value

    t, h = in_thread(make)
    in_thread(lambda: storage.context(0))
    assert h not in fake.destroyed  # the thread ended, its tensor lives
    del t
    in_thread(lambda: storage.context(0))
    assert 77 in fake.freed and h in fake.destroyed
    # the tensor was freed before its context was destroyed
    assert fake.log.index(("free", 77)) < fake.log.index(("destroy", h))


def test_context_dying_in_a_collection_with_its_tensors(fake):
    def make():
        a = storage._Tensor(C.c_void_p(91), 0)
        b = storage._Tensor(C.c_void_p(92), 0)
        a.peer, b.peer = b, a  # a reference cycle: only the collector frees them
        return storage.context(0).value

    h = in_thread(make)
    in_thread(lambda: storage.context(0))
    assert h not in fake.destroyed
    gc.collect()
    assert {91, 92} <= set(fake.freed)
    assert h not in fake.destroyed  # deferred past the collection pass
    in_thread(lambda: storage.context(0))
    assert h in fake.destroyed
    assert max(fake.log.index(("free", 91)), fake.log.index(("free", 92))) < fake.log.index(("destroy", h))


def test_binding_keeps_a_shared_context_alive(fake):
    made = {}
    ready, done = threading.Event(), threading.Event()

    def owner():
        made["h"] = storage.context(0)
        ready.set()

    def user():
        ready.wait()
        storage.bind_context(made["h"], 0)
        assert storage.context(0).value == made["h"].value
        done.set()
        storage.bind_context(None, 0)
        return storage.context(0).value

    t = threading.Thread(target=owner)
    t.start()
    t.join()
    mine = in_thread(user)
    assert done.is_set() and mine != made["h"].value
    h = made.pop("h").value
    gc.collect()
    in_thread(lambda: storage.context(0))
    assert h in fake.destroyed
