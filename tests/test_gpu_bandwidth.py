"""GPU: the bandwidth probes that bench.py prices its per-phase roofline
against.  The HBM copy must reproduce its input exactly and reject what it
cannot copy with 128-bit accesses; the LDS probe's folded per-lane sums
must equal their closed form, so that none of its reads was dropped."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_bw_copy_exact(pkg):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    src = torch.randn(4 * 100003, device="cuda")
    dst = torch.zeros_like(src)
    pkg.bw_copy(src, dst)
    torch.cuda.synchronize()
    assert torch.equal(src, dst)
    odd = torch.zeros(4 * 10 + 2, device="cuda")
    with pytest.raises(pkg.MpcError):
        pkg.bw_copy(odd, odd.clone())  # not a whole number of 128-bit words
    base = torch.zeros(4 * 11, device="cuda")
    with pytest.raises(pkg.MpcError):
        pkg.bw_copy(base[1:41], base[1:41].clone())  # 4-byte aligned only


def lds_expected(n_wg, reps):
    """Per lane: word w = (lane + 256 k + r) mod 1024 holds (4w, 4w+1, 4w+2,
    4w+3); a sums x + y = 8w + 1, b sums z + w = 8w + 5, mod 2^32."""
    lane = np.arange(256, dtype=np.uint64)
    a = np.zeros(256, np.uint64)
    b = np.zeros(256, np.uint64)
    for r in range(reps):
        for k in range(4):
            w = (lane + 256 * k + r) % 1024
            a += 8 * w + 1
            b += 8 * w + 5
    one = ((a % 2**32) ^ (b % 2**32)).astype(np.uint32)
    return np.tile(one, n_wg)


@pytest.mark.parametrize("n_wg,reps", [(1, 1), (3, 37), (300, 1030)])
def test_bw_lds_closed_form(pkg, n_wg, reps):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    out = torch.zeros(256 * n_wg, dtype=torch.int32, device="cuda")
    pkg.bw_lds(out, n_wg, reps)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, lds_expected(n_wg, reps))
