"""The allocation-free entry point ebsd_find_bands_noalloc, called through the C ABI, against
ebsd_find_bands on the same plan: bit-identical outputs on two calls in a row (the second
with fewer patterns), and no change in the device's free memory across a call."""
import numpy as np
import pytest

from helpers import make_pair

pytestmark = pytest.mark.gpu

# (shape, dtype, extra plan keywords, background mode, first N, second N)
CASES = {
    "120-u16-fast": ((120, 120), "uint16", {}, "fast", 2300, 700),
    "240-u16-exact": ((240, 240), "uint16", dict(nTheta=360, nRho=180, nBands=12), "exact", 300, 41),
    "240-u8-fast": ((240, 240), "uint8", dict(nTheta=360, nRho=180, nBands=12), "fast", 260, 9),
    "60-u8-none": ((60, 60), "uint8", {}, None, 1500, 37),
    "80-f32-none": ((80, 80), "float32", {}, None, 300, 5),
}


@pytest.mark.parametrize("case", list(CASES))
def test_noalloc_equals_find_bands(case):
    import torch
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    shape, dtype, kw, mode, n1, n2 = CASES[case]
    print(f"no-alloc {case} N={n1},{n2}", flush=True)
    bd, _ = make_pair(shape, background_mode=mode or "fast", **kw)
    pats = kikuchi_patterns(n1, shape, dtype, seed=97, family="FCC", device="cuda")
    if mode:
        bd.collect_background(patsIn=pats[:64].cpu().numpy().astype(np.float64))
    plan = bd.device_plan()
    x, flags = bd._prepare(plan, pats)
    torch.cuda.synchronize()
    for n in (n1, n2):
        want = plan.find_bands(x[:n], flags=flags)
        got = plan.alloc_out(n)
        for v in got.values():
            v.fill_(-7)      # every output must be written by the call
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        plan.find_bands_noalloc(x[:n], out=got, flags=flags)
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0, "the no-allocation call changed free device memory"
        for k in want:
            a, b = got[k].cpu().numpy(), want[k].cpu().numpy()
            assert np.array_equal(a, b), (case, n, k)


def test_noalloc_rejects_bad_arguments():
    import ctypes
    import torch
    from pyebsdindex_amd._hip import HipLibraryError
    bd, _ = make_pair((60, 60))
    plan = bd.device_plan()
    x = torch.zeros((4, 60, 60), dtype=torch.uint8, device="cuda")
    b = plan.bands_struct(plan.alloc_out(4))
    s = torch.cuda.current_stream().cuda_stream
    call = plan.lib.ebsd_find_bands_noalloc
    assert call(plan.handle, x.data_ptr(), 0, -1, ctypes.byref(b), s) != 0      # negative N
    assert call(plan.handle, x.data_ptr(), 7, 4, ctypes.byref(b), s) != 0       # bad dtype
    assert call(plan.handle, None, 0, 4, ctypes.byref(b), s) != 0               # null patterns
    assert call(None, x.data_ptr(), 0, 4, ctypes.byref(b), s) != 0              # null plan
    assert call(plan.handle, x.data_ptr(), 0, 0, ctypes.byref(b), s) == 0       # nothing to do
    with pytest.raises(HipLibraryError, match="bad dtype"):
        plan.find_bands_noalloc(x, flags=0x40)
