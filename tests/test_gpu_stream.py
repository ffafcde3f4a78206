"""The hand-written streaming kernels behind the bandwidth ceilings
(rsort_stream_bandwidth): each form moves exactly the bytes it claims to,
so the ceilings measure real reads and writes."""
from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FILL = (0x01234567, 0x89ABCDEF, 0xFEDCBA98, 0x76543210)  # the write kernel's 16-byte pattern


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.init()


def words(n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randint(-(1 << 31), 1 << 31, (n,), dtype=torch.int64, device="cuda", generator=g).to(torch.int32)


@pytest.mark.parametrize("form", [0, 4, 8, 12])
@pytest.mark.parametrize("nbytes", [16, 4096 + 48, (1 << 22) + 16 * 37])
def test_copy_and_write(rs, form, nbytes):
    n = nbytes // 4
    src = words(n, 5)
    dst = torch.zeros(n, dtype=torch.int32, device="cuda")
    rs.stream_bandwidth(2 | form, dst.data_ptr(), src.data_ptr(), nbytes, 4)
    torch.cuda.synchronize()
    assert torch.equal(dst, src)
    rs.stream_bandwidth(1 | form, dst.data_ptr(), None, nbytes, 4)
    torch.cuda.synchronize()
    fill = torch.tensor(FILL, dtype=torch.int64).to(torch.int32).repeat(n // 4).cuda()
    assert torch.equal(dst, fill)
    sink = torch.zeros(1024, dtype=torch.int32, device="cuda")
    rs.stream_bandwidth(0 | form, sink.data_ptr(), src.data_ptr(), nbytes, 4)
    torch.cuda.synchronize()
    assert int(sink.abs().sum()) == 0  # random data never hits the read kernel's store pattern


@pytest.mark.parametrize("form", [0, 4])
@pytest.mark.parametrize("tiles", [1, 3, 64])
def test_scatter_is_the_pass_permutation(rs, form, tiles):
    """Tile t's run d (64 keys) lands at position t * 64 of bucket d's region."""
    n = tiles * 16384
    src = words(n, 9)
    dst = torch.zeros(n, dtype=torch.int32, device="cuda")
    rs.stream_bandwidth(3 | form, dst.data_ptr(), src.data_ptr(), n * 4, 2)
    torch.cuda.synchronize()
    expect = src.view(tiles, 256, 64).transpose(0, 1).reshape(-1)
    assert torch.equal(dst, expect)
