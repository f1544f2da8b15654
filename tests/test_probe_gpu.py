"""Peak probes (libf2probe.so) on the device: each MFMA probe and the HBM read
probe run, report positive rates, and stay under the published peaks of the
MI355X (HBM 8 TB/s; dense MFMA bf16 2.5 PFLOP/s, fp32 157.3 TFLOP/s, fp64
78.6 TFLOP/s) with a 5% margin for the boost clock; the flop count of a launch
is the instruction count times its flops."""

import ctypes

import pytest
import torch

from forte2_amd import _native as nat

pytestmark = pytest.mark.gpu

PUBLISHED = {"bf16": 2500.0, "fp32": 157.3, "fp64": 78.6}


def _sink():
    lib = nat.probe_lib()
    n = ctypes.c_int64()
    nat.probe_check(lib.f2probe_sink_len(ctypes.byref(n)), "sink_len")
    return lib, torch.empty(n.value, dtype=torch.float32, device="cuda"), n.value


@pytest.mark.parametrize("kind", ["bf16", "fp32", "fp64"])
def test_mfma_probe(dev, kind):
    lib, sink, n = _sink()
    fl, ms = ctypes.c_double(), ctypes.c_double()
    iters = 2000
    st = torch.cuda.current_stream().cuda_stream
    nat.probe_check(lib.f2probe_mfma(nat.F2PROBE_KIND[kind], iters, ctypes.c_void_p(sink.data_ptr()), n,
                                     ctypes.c_void_p(st), ctypes.byref(fl), ctypes.byref(ms)), "mfma")
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    per = {"bf16": 2 * 16 * 16 * 32, "fp32": 2 * 16 * 16 * 4, "fp64": 2 * 16 * 16 * 4}[kind]
    assert fl.value == per * 4 * iters * cus * 8 * 4
    tf = fl.value / (ms.value * 1e-3) / 1e12
    print(f"   {kind}: {tf:.1f} TFLOP/s ({ms.value:.3f} ms)", flush=True)
    assert 0 < tf <= 1.05 * PUBLISHED[kind] * cus / 256


def test_read_probe(dev):
    lib, sink, n = _sink()
    buf = torch.ones((1 << 30) // 4, dtype=torch.float32, device="cuda")
    ms = ctypes.c_double()
    st = torch.cuda.current_stream().cuda_stream
    nat.probe_check(lib.f2probe_read(ctypes.c_void_p(buf.data_ptr()), buf.numel() * 4, 3,
                                     ctypes.c_void_p(sink.data_ptr()), n, ctypes.c_void_p(st),
                                     ctypes.byref(ms)), "read")
    gbs = buf.numel() * 4 / (ms.value * 1e-3) / 1e9
    print(f"   read: {gbs:.0f} GB/s", flush=True)
    assert 0 < gbs <= 8400
    # a misaligned buffer is rejected before any launch
    with pytest.raises(ValueError):
        nat.probe_check(lib.f2probe_read(ctypes.c_void_p(buf.data_ptr() + 4), 32, 1,
                                         ctypes.c_void_p(sink.data_ptr()), n, ctypes.c_void_p(st),
                                         ctypes.byref(ms)), "misaligned")
