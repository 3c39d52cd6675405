"""Devices (hoomd/device.py).

Only ``GPU`` exists: this package is the GPU path. ``CPU`` raises, because the CPU path of
HOOMD-blue is not reproduced here (and no CPU fallback is allowed to stand in for the GPU path).
"""

from . import _lib


class Device:
    def __init__(self, notice_level=2):
        self.notice_level = notice_level


class GPU(Device):
    """hoomd.device.GPU (hoomd/device.py): select the (single) MI355X."""

    def __init__(self, gpu_ids=None, num_cpu_threads=None, communicator=None, msg_file=None,
                 notice_level=2):
        super().__init__(notice_level)
        L = _lib.lib()
        n = L.hm_device_count()
        if n < 1:
            raise RuntimeError("hoomd_amd.device.GPU: no GPU is visible to the HIP runtime")
        if gpu_ids not in (None, [0], (0,), 0):
            raise ValueError("hoomd_amd runs on exactly one GPU (gpu_ids=[0])")
        self.gpu_ids = [0]
        self.num_cpu_threads = num_cpu_threads

    @property
    def device(self):
        return _lib.lib().hm_version().decode()

    def enable_profiling(self):
        """Context manager kept for API compatibility; profile with rocprofv3 instead."""
        import contextlib
        return contextlib.nullcontext()


class CPU(Device):
    def __init__(self, *args, **kwargs):
        raise NotImplementedError(
            "hoomd_amd is the GPU path only; use hoomd.device.CPU from HOOMD-blue for CPU runs")


def auto_select(communicator=None, msg_file=None, notice_level=2):
    return GPU(notice_level=notice_level)
