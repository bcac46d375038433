"""The inputs of a serving slot reach the library intact, and a bad water list is an error.

Concurrent serving of mixed sizes once ended in an illegal memory access.  The cause:
engine.Slot cleared its freshly allocated arrays, the water list `listttm2` among them, on
torch's default stream and uploaded the configuration on its own stream.  The default
(null) stream waits for every blocking stream of the device, which includes the library's
streams of all other running slots, while the slot's own stream does not wait for it; so a
slot created while other slots were busy could have its upload overwritten by the late
clear.  A listttm2 of zeros then gave every water the atom offset 0 - 1 = -1, and the
MB-pol kernels read and wrote one to nine elements before the position, force, energy and
heat-flux arrays.

Two fixes, each tested here on the CPU:
  * Slot buffers are allocated, cleared and loaded on the slot's own stream, which the
    host then synchronizes (ordering checked with a recording stand-in for torch);
  * both MB-pol entry points check listttm2 with one predicate (water_base) before any atom
    array is indexed with it, on the host for host arrays and on the device for device
    arrays; the host form is mbgpu_check_listttm2, exercised here.
"""
import contextlib
import importlib

import numpy as np
import pytest

INT_MAX = 2 ** 31 - 1


# ------------------------------------------------------------------ list check ----
def test_field_layout_lists_pass(pkg, waterbox):
    for nw in (1, 3, 64):
        lst = np.arange(1, 3 * nw + 1, dtype=np.int32)
        assert pkg.check_listttm2(lst, 3 * nw) == 0
        assert pkg.check_listttm2(lst, 4 * nw) == 0           # M sites after the waters
    box = waterbox.make_cluster(9, seed=3)
    assert pkg.check_listttm2(box.listttm2, box.natms) == 0
    # waters need not be in order or adjacent, only each water's three atoms
    lst = np.array([7, 8, 9, 1, 2, 3], dtype=np.int32)
    assert pkg.check_listttm2(lst, 9) == 0


def test_cleared_list_is_rejected(pkg):
    # the state a late clear left behind: every entry zero
    assert pkg.check_listttm2(np.zeros(3 * 64, np.int32), 256) == 1


@pytest.mark.parametrize("entries, natms, bad", [
    ([1, 2, 3, 4, 5, 6], 5, 2),             # last atom beyond natms
    ([1, 2, 3, 4, 5, 7], 9, 2),             # H2 not O + 2
    ([1, 3, 2], 9, 1),                      # H1 and H2 swapped
    ([0, 1, 2], 9, 1),                      # 0-based ids
    ([-1, 0, 1], 9, 1),
    ([INT_MAX, -INT_MAX - 1, -INT_MAX], 9, 1),   # a + 1 would overflow: must not wrap
    ([INT_MAX - 2, INT_MAX - 1, INT_MAX], INT_MAX, 0),   # largest valid ids
    ([4, 5, 6, 4, 5, 6], 6, 0),             # the check is per water, not a permutation test
])
def test_bad_entries_are_named(pkg, entries, natms, bad):
    assert pkg.check_listttm2(np.array(entries, np.int32), natms) == bad


def test_bad_sizes(pkg):
    assert pkg.check_listttm2(np.array([1, 2], np.int32), 9) == -1
    assert pkg.check_listttm2(np.zeros(0, np.int32), 9) == 0


# ------------------------------------------------------------ slot stream order ----
class _Stream:
    def __init__(self, name, log):
        self.name, self.log = name, log

    def synchronize(self):
        self.log.append(("sync", self.name, None))


class _Tensor:
    def __init__(self, fake, shape, tag):
        self.fake, self.tag = fake, tag
        self.shape = shape

    def __getitem__(self, i):
        return _Tensor(self.fake, None, f"{self.tag}[{i}]")

    def view(self, *shape):
        return _Tensor(self.fake, shape, f"{self.tag}[view]")

    def numpy(self):
        return self.host

    def copy_(self, src, non_blocking=False):
        self.fake.log.append(("copy", self.fake.current, self.tag.split("[")[0]))
        return self


class _FakeTorch:
    """Records, per slot array, on which stream it is cleared and written."""
    float64, int32 = "f64", "i32"

    def __init__(self):
        self.log = []
        self.current = "default"
        self.n = 0
        fake = self

        class _Cuda:
            @staticmethod
            @contextlib.contextmanager
            def stream(s):
                prev = fake.current
                fake.current = s.name
                try:
                    yield
                finally:
                    fake.current = prev

        self.cuda = _Cuda()

    def zeros(self, *shape, dtype=None, device=None):
        tag = f"t{self.n}"
        self.n += 1
        self.log.append(("zero", self.current, tag))
        return _Tensor(self, shape, tag)

    def empty(self, n, dtype=None, pin_memory=False):
        # host staging (not an array the library reads)
        t = _Tensor(self, (n,), "host")
        t.host = np.zeros(n)
        return t

    def as_tensor(self, a, dtype=None):
        return a


def _alloc_and_load(slot_cls, box):
    fake = _FakeTorch()
    s = slot_cls.__new__(slot_cls)
    s.torch, s.dev, s.box = fake, "cuda:0", box
    s.stream = _Stream("slot", fake.log)
    s._alloc(box)
    s.load(box)
    return fake.log


def _unordered_writes(log):
    """Arrays whose clear and upload are not on one stream followed by a host sync."""
    cleared = {tag: st for op, st, tag in log if op == "zero"}
    copies = [(st, tag) for op, st, tag in log if op == "copy"]
    bad = sorted({tag for st, tag in copies if cleared.get(tag) != st})
    if not copies or log[-1][:2] != ("sync", "slot") or any(st != "slot" for st, _ in copies):
        bad.append("<upload not on the slot stream, or not synchronized>")
    return bad


def test_slot_clears_and_uploads_on_its_own_stream(waterbox):
    engine = importlib.import_module("dlpoly2-mbpol-heat_amd.engine")
    box = waterbox.make_waterbox(8, 9.32, 2, seed=1)
    log = _alloc_and_load(engine.Slot, box)
    assert _unordered_writes(log) == []
    # every array the library reads is among those cleared and loaded: the configuration
    # buffer (x, y, z, v, ek are views of it) and the list
    assert sum(op == "zero" for op, _, _ in log) == 2
    assert {tag for op, _, tag in log if op == "copy"} == {tag for op, _, tag in log if op == "zero"}


def test_order_check_catches_a_clear_on_the_default_stream(waterbox):
    """The check above flags the arrangement that caused the fault (clear on torch's
    default stream, upload on the slot's)."""
    engine = importlib.import_module("dlpoly2-mbpol-heat_amd.engine")

    class OldSlot(engine.Slot):
        def _alloc(self, box):
            t = self.torch
            n, nr = box.natms, box.nreal
            self.x, self.y, self.z = (t.zeros(n) for _ in range(3))
            self.lst = t.zeros(nr)
            self.v = t.zeros(3, nr)
            self.ek = t.zeros(nr)

        def load(self, box):
            # the upload of that arrangement: every array on the slot's stream, then a sync
            t = self.torch
            with t.cuda.stream(self.stream):
                for a in (self.x, self.y, self.z, self.lst, self.v, self.ek):
                    a.copy_(None)
            self.stream.synchronize()

    box = waterbox.make_waterbox(8, 9.32, 2, seed=1)
    bad = _unordered_writes(_alloc_and_load(OldSlot, box))
    assert len(bad) == 6
