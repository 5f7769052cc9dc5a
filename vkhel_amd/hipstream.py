"""HIP streams and events for ordering the asynchronous transfer calls
(vkhel_ext_vector_write_async / _read_async) against the library's stream.

Loads the same HIP runtime as libvkhel.so (by soname), not a copy bundled
with another package.
"""
import ctypes

_H = None


def _hip():
    global _H
    if _H is None:
        h = ctypes.CDLL("libamdhip64.so")
        p, pp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)
        h.hipStreamCreateWithFlags.argtypes = [pp, ctypes.c_uint]
        h.hipStreamDestroy.argtypes = [p]
        h.hipStreamSynchronize.argtypes = [p]
        h.hipStreamWaitEvent.argtypes = [p, p, ctypes.c_uint]
        h.hipEventCreate.argtypes = [pp]
        h.hipEventDestroy.argtypes = [p]
        h.hipEventRecord.argtypes = [p, p]
        h.hipEventSynchronize.argtypes = [p]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), p, p]
        _H = h
    return _H


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class Stream:
    """A non-blocking stream at the default priority."""

    def __init__(self):
        self.handle = ctypes.c_void_p()
        _ok(_hip().hipStreamCreateWithFlags(ctypes.byref(self.handle), 1),
            "hipStreamCreateWithFlags")

    def synchronize(self):
        _ok(_hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def wait(self, event):
        """Later work on this stream waits for `event`."""
        _ok(_hip().hipStreamWaitEvent(self.handle, event.handle, 0),
            "hipStreamWaitEvent")

    def close(self):
        if self.handle:
            _ok(_hip().hipStreamDestroy(self.handle), "hipStreamDestroy")
            self.handle = ctypes.c_void_p()


def wait(stream, event):
    """Later work on `stream` (a Stream or a raw hipStream_t such as
    Context.stream) waits for `event`."""
    s = stream.handle if isinstance(stream, Stream) else stream
    _ok(_hip().hipStreamWaitEvent(s, event.handle, 0), "hipStreamWaitEvent")


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        _ok(_hip().hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream):
        """Record on a Stream or a raw hipStream_t (e.g. Context.stream)."""
        s = stream.handle if isinstance(stream, Stream) else stream
        _ok(_hip().hipEventRecord(self.handle, s), "hipEventRecord")

    def synchronize(self):
        _ok(_hip().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, later):
        later.synchronize()
        out = ctypes.c_float()
        _ok(_hip().hipEventElapsedTime(ctypes.byref(out), self.handle,
                                       later.handle), "hipEventElapsedTime")
        return float(out.value)

    def close(self):
        if self.handle:
            _ok(_hip().hipEventDestroy(self.handle), "hipEventDestroy")
            self.handle = ctypes.c_void_p()
