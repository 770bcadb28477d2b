"""Host side of letkf_amd.core: a job descriptor keeps its device buffers alive
(it holds raw pointers that the asynchronous kernels read later)."""
import gc
import weakref

from letkf_amd import core


class _Buf:
    def __init__(self, addr):
        self.addr = addr

    def data_ptr(self):
        return self.addr


def test_make_job_keeps_buffers_alive():
    refs = []

    def mk(a):
        b = _Buf(a)
        refs.append(weakref.ref(b))
        return b

    job = core.make_job(k=4, npt=1, hdxf=mk(0x1000), rdiag=mk(0x2000), dep=mk(0x3000),
                        loc_off=mk(0x4000), loc_idx=mk(0x5000), loc_rho=mk(0x6000),
                        transm=mk(0x7000))
    gc.collect()
    assert all(r() is not None for r in refs)
    assert job.hdxf == 0x1000 and job.transm == 0x7000
    del job
    gc.collect()
    assert all(r() is None for r in refs)
