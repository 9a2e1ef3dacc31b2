"""hipEvent timing on a context's stream through libamdhip64 (ctypes)."""
import ctypes


class HipEvents:
    """hipEvent pairs: create(), record(event, stream), elapsed_ms(a, b)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.L.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.L.hipEventDestroy.argtypes = [ctypes.c_void_p]
        self.L.hipStreamWaitEvent.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]

    def create(self):
        e = ctypes.c_void_p()
        if self.L.hipEventCreate(ctypes.byref(e)) != 0:
            raise RuntimeError("hipEventCreate failed")
        return e

    def record(self, e, stream):
        if self.L.hipEventRecord(e, ctypes.c_void_p(stream)) != 0:
            raise RuntimeError("hipEventRecord failed")

    def wait(self, stream, e):
        """Work queued on stream after this call starts after event e."""
        if self.L.hipStreamWaitEvent(ctypes.c_void_p(stream), e, 0) != 0:
            raise RuntimeError("hipStreamWaitEvent failed")

    def elapsed_ms(self, a, b):
        if self.L.hipEventSynchronize(b) != 0:
            raise RuntimeError("hipEventSynchronize failed")
        ms = ctypes.c_float()
        if self.L.hipEventElapsedTime(ctypes.byref(ms), a, b) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return ms.value

    def destroy(self, e):
        self.L.hipEventDestroy(e)


def time_ms(ctx, fn, reps=3, warmup=1, events=None):
    """Mean milliseconds of fn() between hipEvents on ctx's stream."""
    from . import rgrt
    ev = events or HipEvents()
    stream = rgrt.lib().rgrt_stream(ctx.h)
    for _ in range(warmup):
        fn()
    a, b = ev.create(), ev.create()
    ctx.synchronize()
    ev.record(a, stream)
    for _ in range(reps):
        fn()
    ev.record(b, stream)
    ms = ev.elapsed_ms(a, b) / reps
    ev.destroy(a)
    ev.destroy(b)
    return ms


def frames_ms(ctxs, fns, steps, warmup=1, events=None):
    """Mean milliseconds per frame of `steps` frames, frame k enqueued by
    fns[k % len(fns)] on ctxs[k % len(fns)] (frames in flight on several
    contexts). hipEvents bracket the frames: every other stream waits for the
    opening event on the first context's stream, and the closing time is the
    latest of the streams' closing events."""
    from . import rgrt
    ev = events or HipEvents()
    n = len(fns)
    used = ctxs[:n]
    streams = [rgrt.lib().rgrt_stream(c.h) for c in used]
    for k in range(warmup):
        fns[k % n]()
    for c in used:
        c.synchronize()
    e0 = ev.create()
    ends = [ev.create() for _ in range(n)]
    ev.record(e0, streams[0])
    for s in streams[1:]:
        ev.wait(s, e0)
    for k in range(steps):
        fns[k % n]()
    for e, s in zip(ends, streams):
        ev.record(e, s)
    ms = max(ev.elapsed_ms(e0, e) for e in ends) / steps
    for c in used:
        c.synchronize()
    for e in [e0] + ends:
        ev.destroy(e)
    return ms


class DeviceBuffer:
    """Device memory from hipMalloc (for data the library reads from device
    pointers, e.g. animated vertices); freed by close()."""

    def __init__(self, nbytes):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        if self.L.hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)) != 0:
            raise RuntimeError("hipMalloc failed")

    def upload(self, array):
        """Copies the bytes of a C-contiguous host array (blocking)."""
        import numpy as np
        a = np.ascontiguousarray(array)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        if self.L.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) != 0:  # hipMemcpyHostToDevice
            raise RuntimeError("hipMemcpy failed")

    @property
    def address(self):
        return self.ptr.value

    def close(self):
        if self.ptr.value:
            self.L.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()
