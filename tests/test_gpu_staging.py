"""Host <-> device copies through the pinned staging buffers (zscan.download,
zscan._upload_rows): chunking over several staging buffers, sizes that are
not multiples of 8 bytes, other dtypes, empty tensors, and float32 host rows
widened on the way in."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_download_matches_device(gpu, monkeypatch):
    import torch
    from redrock_amd import zscan
    monkeypatch.setattr(zscan, 'STAGE_BYTES', 1 << 20)     # many chunks on small tensors
    monkeypatch.setattr(zscan, '_STAGING', [])
    g = torch.Generator(device='cuda').manual_seed(5)
    ts = [torch.randn(3 * (1 << 17) + 5, dtype=torch.float64, device='cuda', generator=g),
          torch.randint(0, 255, (1_000_003,), dtype=torch.uint8, device='cuda', generator=g),
          torch.randint(-9, 9, (777, 333), dtype=torch.int32, device='cuda', generator=g),
          torch.empty(0, dtype=torch.float64, device='cuda'),
          torch.randn(17, 4, dtype=torch.float64, device='cuda', generator=g)]
    print('case: download of 5 tensors through 1 MB staging buffers', flush=True)
    got = zscan.download(ts)
    for t, h in zip(ts, got):
        assert h.shape == tuple(t.shape) and h.dtype == t.cpu().numpy().dtype
        assert np.array_equal(h, t.cpu().numpy())
    # arrays are ordinary memory, independent of the staging buffers
    again = zscan.download(ts[:1])
    assert np.array_equal(again[0], got[0]) and again[0].ctypes.data != got[0].ctypes.data


def test_upload_rows_matches_host(gpu, monkeypatch):
    import torch
    from redrock_amd import zscan
    monkeypatch.setattr(zscan, 'STAGE_BYTES', 1 << 16)
    monkeypatch.setattr(zscan, '_STAGING', [])
    rng = np.random.default_rng(7)
    rows = [rng.normal(size=(11, 97)).astype(np.float32) for _ in range(301)]
    dst = torch.empty((301, 11, 97), dtype=torch.float64, device='cuda')
    print('case: 301 float32 rows through 64 KB staging buffers', flush=True)
    zscan._upload_rows(dst, rows)
    want = np.stack(rows).astype(np.float64)
    assert np.array_equal(dst.cpu().numpy(), want)
    # float64 rows: copied by the library's threads (rr_gather_rows), several chunks
    rows64 = [rng.normal(size=(11, 97)) for _ in range(301)]
    dst64 = torch.empty((301, 11, 97), dtype=torch.float64, device='cuda')
    zscan._upload_rows(dst64, rows64)
    assert np.array_equal(dst64.cpu().numpy(), np.stack(rows64))
    # a non-contiguous float64 row among them: every row through numpy
    rows64[7] = np.asfortranarray(rows64[7])
    zscan._upload_rows(dst64, rows64)
    assert np.array_equal(dst64.cpu().numpy(), np.stack(rows64))
    # a 2-D host array of rows, and rows larger than a staging buffer
    big = rng.normal(size=(3, 9000))
    dst2 = torch.empty((3, 9000), dtype=torch.float64, device='cuda')
    zscan._upload_rows(dst2, big)
    assert np.array_equal(dst2.cpu().numpy(), big)
