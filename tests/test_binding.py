"""Host-side checks of the Python binding that need no GPU: every buffer a
method hands to the library is validated first, so a CPU tensor, a tensor of
the wrong element width, a non-contiguous view or a buffer shorter than n
raises ValueError and never reaches a kernel as a bad address."""
import pytest
import torch


class FakeDev:
    """Stands in for a GPU tensor: the attributes _buf reads, nothing else."""

    def __init__(self, numel, esize=4, contiguous=True, addr=0x7F0000000000):
        self._n, self._e, self._c, self._a = numel, esize, contiguous, addr
        self.is_cuda = True

    def numel(self):
        return self._n

    def element_size(self):
        return self._e

    def is_contiguous(self):
        return self._c

    def data_ptr(self):
        return self._a


class RecordingLib:
    """Records library calls instead of making them (every call returns OK)."""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def f(*args):
            self.calls.append((name, args))
            return 0
        return f


@pytest.fixture
def fake_ctx(edpp):
    c = object.__new__(edpp.Context)
    c._L = RecordingLib()
    c.h = None
    yield c
    c.h = None


def test_valid_buffers_reach_the_library(fake_ctx):
    n = 100
    fake_ctx.gather(FakeDev(n), FakeDev(n), FakeDev(n))
    fake_ctx.scan(FakeDev(n, 8), FakeDev(n, 8), wide=True, total=FakeDev(1, 8))
    fake_ctx.split(FakeDev(n), FakeDev(n), 16, FakeDev(n), FakeDev(n), FakeDev(16))
    fake_ctx.join_count(FakeDev(n), FakeDev(7), FakeDev(1, 8))
    names = [c[0] for c in fake_ctx._L.calls]
    assert names == ["edpp_gather_i32", "edpp_scan_exclusive_i64", "edpp_split_i32", "edpp_hash_join_count"]
    # the address of each buffer, and n from the index array, are passed on
    args = fake_ctx._L.calls[0][1]
    assert args[1:5] == (0x7F0000000000, 0x7F0000000000, 0x7F0000000000, n)


@pytest.mark.parametrize("case", ["cpu", "width", "short", "strided"])
def test_gather_rejects_bad_buffers(fake_ctx, case):
    n = 64
    bad = {"cpu": torch.zeros(n, dtype=torch.int32),
           "width": FakeDev(n, esize=8),
           "short": FakeDev(n - 1),
           "strided": FakeDev(n, contiguous=False)}[case]
    for pos in range(3):
        bufs = [FakeDev(n), FakeDev(n), FakeDev(n)]
        bufs[pos] = bad
        with pytest.raises(ValueError):
            fake_ctx.gather(*bufs, n=n)
        with pytest.raises(ValueError):
            fake_ctx.scatter(*bufs, n=n)
    assert fake_ctx._L.calls == []


def test_wide_flag_must_match_the_element_width(fake_ctx):
    n = 32
    with pytest.raises(ValueError):  # int32 buffers passed as 64-bit values
        fake_ctx.gather(FakeDev(n, 4), FakeDev(n, 4), FakeDev(n), wide=True)
    with pytest.raises(ValueError):  # int64 buffers passed as 32-bit values
        fake_ctx.scan(FakeDev(n, 8), FakeDev(n, 8))
    with pytest.raises(ValueError):  # a 64-bit index array
        fake_ctx.gather(FakeDev(n, 8), FakeDev(n, 8), FakeDev(n, 8), wide=True)
    with pytest.raises(ValueError):
        fake_ctx.split(FakeDev(n, 4), FakeDev(n, 4), 4, wide=True)
    assert fake_ctx._L.calls == []


def test_every_method_checks_lengths(fake_ctx, edpp):
    n = 50
    short = FakeDev(n - 1)
    cases = [
        lambda: fake_ctx.stream_op(0, FakeDev(n), None, short, n=n),
        lambda: fake_ctx.sort(FakeDev(n), short, n=n),
        lambda: fake_ctx.sort(FakeDev(n), FakeDev(n), FakeDev(n), short, n=n),
        lambda: fake_ctx.access_scale(0, FakeDev(n), short, 1, n=n),
        lambda: fake_ctx.access_chunks(FakeDev(n), 10, FakeDev(4), FakeDev(1, 8)),  # 5 chunks, 4 order entries
        lambda: fake_ctx.access_chunks(FakeDev(n), 10, FakeDev(5), FakeDev(1, 4)),  # 32-bit count
        lambda: fake_ctx.join_count(FakeDev(n), FakeDev(n), FakeDev(1, 4)),
        lambda: fake_ctx.scan(FakeDev(n), short, n=n),
        lambda: fake_ctx.scan(FakeDev(n), FakeDev(n), total=FakeDev(0)),
        lambda: fake_ctx.scan_local(FakeDev(511), FakeDev(512), 1, 512),
        lambda: fake_ctx.split(FakeDev(n), FakeDev(n), 256, start=FakeDev(255)),
        lambda: fake_ctx.split(FakeDev(n), FakeDev(n), 4, FakeDev(n), short),
        lambda: fake_ctx.gen_i32(short, 1, n=n),
        lambda: fake_ctx.gen_i64(FakeDev(n, 4), 1),
        lambda: fake_ctx.gen_perm(short, 1, n=n),
        lambda: fake_ctx.gen_split_keys(FakeDev(n), short, 16, 0, 1),
        lambda: fake_ctx.call_features(edpp.Call(), FakeDev(edpp.NUM_FEATURES - 1)),
        lambda: edpp.make_job(edpp.JOB_SCAN, n, FakeDev(n), short),
        lambda: edpp.make_job(edpp.JOB_GATHER, n, FakeDev(n, 8), FakeDev(n, 8), 8, idx=short),
        lambda: edpp.make_job(edpp.JOB_SPLIT, n, FakeDev(n), FakeDev(n), 4, buckets=16, aux=FakeDev(8)),
        lambda: edpp.make_call(edpp.OP_GATHER, n, FakeDev(n), FakeDev(n), 4, idx=short),
        lambda: edpp.make_call(edpp.OP_SCAN, n, FakeDev(n, 4), FakeDev(n, 4), 8),
        lambda: fake_ctx.gather(FakeDev(n), FakeDev(n), FakeDev(n), n=-1),
    ]
    for i, f in enumerate(cases):
        with pytest.raises(ValueError):
            f()
        assert fake_ctx._L.calls == [], i


def test_optional_buffers_and_raw_addresses_pass_through(fake_ctx, edpp):
    n = 10
    fake_ctx.split(FakeDev(n), FakeDev(n), 2)  # no values, no starts
    fake_ctx.gather(0x1000, 0x2000, 0x3000, n=n)  # raw addresses: the caller vouches for them
    assert [c[0] for c in fake_ctx._L.calls] == ["edpp_split_i32", "edpp_gather_i32"]
    j = edpp.make_job(edpp.JOB_SCAN, n, FakeDev(n), FakeDev(n))
    assert j.n == n and j.aux is None
