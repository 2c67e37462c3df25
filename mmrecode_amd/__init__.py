"""mmrecode_amd: MI355X (gfx950) pixel reconstruction behind the C ABI of include/mmr_recon.h.

This module loads the in-tree libmmr_recon.so (built by mmrecode_amd/build.py) and wraps its
context, device memory and submit / wait calls. There is no CPU fallback: if the library is
missing or no HIP device is present, the calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libmmr_recon.so")
_LIB = None


class ReconError(RuntimeError):
    """A non-OK mmr_status, with the library's last error message."""

    def __init__(self, status, message):
        super().__init__(f"mmr_status {status}: {message}")
        self.status = status
        self.message = message


def load_library():
    """Load libmmr_recon.so and declare every entry point of include/mmr_recon.h."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBRARY_PATH):
            raise ImportError(f"{LIBRARY_PATH} is not built; run `python -m mmrecode_amd.build`")
        lib = C.CDLL(LIBRARY_PATH)
        for name, restype, argtypes in abi.ENTRY_POINTS:
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _LIB = lib
    return _LIB


def last_error():
    lib = load_library()
    n = lib.mmr_recon_last_error_message(None, 0)
    buf = C.create_string_buffer(n)
    lib.mmr_recon_last_error_message(buf, n)
    return buf.value.decode(errors="replace")


def check(status):
    if status != abi.STATUS_OK:
        raise ReconError(status, last_error())


class DeviceBuffer:
    """Device allocation owned by a Context (freed on free() or context close)."""

    def __init__(self, ctx, ptr, nbytes):
        self.ctx = ctx
        self.ptr = ptr
        self.nbytes = nbytes

    def free(self):
        if self.ptr:
            self.ctx._free(self)

    def __int__(self):
        return self.ptr


class Context:
    """One mmr_recon_context: a device, its stream and in-flight tickets."""

    def __init__(self, device=0, queue_depth=0, hbm_budget_bytes=0, persistent_groups=0):
        self.lib = load_library()
        cfg = abi.new(abi.Config, device_ordinal=device, queue_depth=queue_depth,
                      hbm_budget_bytes=hbm_budget_bytes, persistent_groups=persistent_groups)
        handle = C.c_void_p()
        check(self.lib.mmr_recon_context_create(C.byref(cfg), C.byref(handle)))
        self.handle = handle
        self._buffers = {}

    def close(self):
        if self.handle:
            for buf in list(self._buffers.values()):
                self._free(buf)
            self.lib.mmr_recon_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- memory
    def alloc(self, nbytes):
        p = C.c_void_p()
        check(self.lib.mmr_recon_device_alloc(self.handle, max(int(nbytes), 1), C.byref(p)))
        buf = DeviceBuffer(self, p.value, int(nbytes))
        self._buffers[p.value] = buf
        return buf

    def _free(self, buf):
        if buf.ptr in self._buffers:
            del self._buffers[buf.ptr]
            check(self.lib.mmr_recon_device_free(self.handle, C.c_void_p(buf.ptr)))
        buf.ptr = None

    def upload(self, array):
        a = np.ascontiguousarray(array)
        buf = self.alloc(a.nbytes)
        if a.nbytes:
            check(self.lib.mmr_recon_copy_to_device(self.handle, C.c_void_p(buf.ptr), a.ctypes.data, a.nbytes))
        return buf

    def write(self, buf, array, offset=0):
        a = np.ascontiguousarray(array)
        check(self.lib.mmr_recon_copy_to_device(self.handle, C.c_void_p(buf.ptr + offset), a.ctypes.data, a.nbytes))

    def download(self, buf, dtype=np.uint8, shape=None, offset=0, nbytes=None):
        dtype = np.dtype(dtype)
        n = buf.nbytes - offset if nbytes is None else nbytes
        out = np.empty(n // dtype.itemsize, dtype)
        check(self.lib.mmr_recon_copy_to_host(self.handle, out.ctypes.data, C.c_void_p(buf.ptr + offset), out.nbytes))
        return out if shape is None else out.reshape(shape)

    # ---- submission
    def _submit(self, fn, struct_type, items, flags):
        arr = (struct_type * len(items))(*items)
        ticket = C.c_uint64()
        check(fn(self.handle, arr, len(items), flags, C.byref(ticket)))
        return ticket.value

    def submit_mjpeg(self, frames, flags=0):
        return self._submit(self.lib.mmr_recon_mjpeg_submit, abi.MjpegFrame, frames, flags)

    def submit_mpeg2(self, pictures, flags=0):
        return self._submit(self.lib.mmr_recon_mpeg2_submit, abi.Mpeg2Picture, pictures, flags)

    def reconstruct_mpeg2_noalloc(self, pictures, flags=0):
        """mmr_recon_mpeg2_reconstruct_noalloc: the pictures reconstructed synchronously inside the
        library's fixed device scratch (no device allocation); raises ReconError on a non-OK status."""
        arr = (abi.Mpeg2Picture * len(pictures))(*pictures)
        check(self.lib.mmr_recon_mpeg2_reconstruct_noalloc(self.handle, arr, len(pictures), flags))

    def submit_dv(self, frames, flags=0):
        return self._submit(self.lib.mmr_recon_dv_submit, abi.DvFrame, frames, flags)

    def submit_h264(self, pictures, flags=0):
        return self._submit(self.lib.mmr_recon_h264_submit, abi.H264Picture, pictures, flags)

    def submit_yuv_to_rgb(self, jobs, flags=0):
        return self._submit(self.lib.mmr_recon_yuv_to_rgb_submit, abi.YuvToRgbJob, jobs, flags)

    def submit_composite(self, jobs, flags=0):
        return self._submit(self.lib.mmr_recon_composite_submit, abi.CompositeJob, jobs, flags)

    def submit_yuv_composite(self, conversions, composites, flags=0):
        """mmr_recon_yuv_composite_submit: conversion i then composite i (over conversion i's RGBA
        output) in one pass."""
        assert len(conversions) == len(composites)
        conv = (abi.YuvToRgbJob * len(conversions))(*conversions)
        comp = (abi.CompositeJob * len(composites))(*composites)
        ticket = C.c_uint64()
        check(self.lib.mmr_recon_yuv_composite_submit(self.handle, conv, comp, len(conversions), flags, C.byref(ticket)))
        return ticket.value

    def submit_linear_composite(self, jobs, flags=0):
        return self._submit(self.lib.mmr_recon_linear_composite_submit, abi.LinearCompositeJob, jobs, flags)

    def submit_linear_convert(self, jobs, flags=0):
        return self._submit(self.lib.mmr_recon_linear_convert_submit, abi.LinearConvertJob, jobs, flags)

    def submit_yuv_blend(self, jobs, flags=0):
        return self._submit(self.lib.mmr_recon_yuv_blend_submit, abi.YuvBlendJob, jobs, flags)

    def submit_overlay_prepare(self, jobs, flags=0):
        return self._submit(self.lib.mmr_recon_overlay_prepare_submit, abi.OverlayPrepareJob, jobs, flags)

    def submit_scale(self, jobs, flags=0):
        return self._submit(self.lib.mmr_recon_scale_submit, abi.ScaleJob, jobs, flags)

    # ---- tuning
    TUNING_FIELDS = [f for f, _ in abi.Tuning._fields_ if f != "struct_size"]

    def get_tuning(self):
        """The context's launch shapes as a dict (mmr_recon_get_tuning)."""
        t = abi.new(abi.Tuning)
        check(self.lib.mmr_recon_get_tuning(self.handle, C.byref(t)))
        return {f: getattr(t, f) for f in self.TUNING_FIELDS}

    def set_tuning(self, **fields):
        """Replaces the given launch-shape fields, keeping the others (mmr_recon_set_tuning)."""
        cur = self.get_tuning()
        unknown = set(fields) - set(cur)
        if unknown:
            raise ValueError(f"unknown tuning fields {sorted(unknown)}")
        cur.update(fields)
        t = abi.new(abi.Tuning, **cur)
        check(self.lib.mmr_recon_set_tuning(self.handle, C.byref(t)))

    def autotune(self, codec, items, flags=0, repeats=3):
        """Times every candidate launch shape of `codec` on a batch and keeps the fastest
        (mmr_recon_autotune). Returns (best tuning dict, [(tuning dict, kernel_ms), ...])."""
        struct = Server.STRUCTS[codec]
        arr = (struct * len(items))(*items)
        best = abi.new(abi.Tuning)
        trials = (abi.TuningTrial * 64)()
        n = C.c_uint32()
        check(self.lib.mmr_recon_autotune(self.handle, codec, C.addressof(arr), len(items), flags, repeats, C.byref(best),
                                          trials, 64, C.byref(n)))
        as_dict = lambda t: {f: getattr(t, f) for f in self.TUNING_FIELDS}  # noqa: E731
        return as_dict(best), [(as_dict(tr.tuning), tr.kernel_ms) for tr in trials[: n.value]]

    def prepare(self, codec, items, flags=0):
        """A Prepared batch of one codec's descriptors (uploaded once, launches captured in a graph)."""
        return Prepared(self, codec, items, flags)

    def wait(self, ticket):
        check(self.lib.mmr_recon_wait(self.handle, ticket))

    def poll(self, ticket):
        done = C.c_int32()
        check(self.lib.mmr_recon_poll(self.handle, ticket, C.byref(done)))
        return bool(done.value)

    def synchronize(self):
        check(self.lib.mmr_recon_synchronize(self.handle))

    def kernel_ms(self, ticket):
        ms = C.c_double()
        check(self.lib.mmr_recon_last_kernel_ms(self.handle, ticket, C.byref(ms)))
        return ms.value

    def kernel_phases_ms(self, ticket):
        """Device milliseconds of each launch phase of a ticket (mmr_recon_kernel_phases_ms)."""
        out = (C.c_double * abi.MAX_PHASES)()
        n = C.c_uint32()
        check(self.lib.mmr_recon_kernel_phases_ms(self.handle, ticket, out, abi.MAX_PHASES, C.byref(n)))
        return [out[i] for i in range(min(n.value, abi.MAX_PHASES))]


class FrameCache:
    """mmr_recon_frame_cache_*: the library-owned resident frame cache (DeviceResourceCache,
    crates/render/src/device_cache.rs:53-287). Keys and descriptors are abi.FrameKey /
    abi.FrameDescriptor or keyword dicts of their fields."""

    def __init__(self, ctx, byte_budget):
        self.ctx = ctx
        h = C.c_void_p()
        check(ctx.lib.mmr_recon_frame_cache_create(ctx.handle, byte_budget, C.byref(h)))
        self.handle = h

    @staticmethod
    def _key(k):
        return k if isinstance(k, abi.FrameKey) else abi.FrameKey(**k)

    @staticmethod
    def _desc(d):
        return d if isinstance(d, abi.FrameDescriptor) else abi.FrameDescriptor(**d)

    def begin_frame(self):
        check(self.ctx.lib.mmr_recon_frame_cache_begin_frame(self.handle))

    def retain(self, key, descriptor, nbytes):
        """(device pointer, abi.CACHE_INSERTED or abi.CACHE_REUSED)."""
        p, st = C.c_void_p(), C.c_uint32()
        check(self.ctx.lib.mmr_recon_frame_cache_retain(self.handle, C.byref(self._key(key)), C.byref(self._desc(descriptor)),
                                                        nbytes, C.byref(p), C.byref(st)))
        return p.value, st.value

    def get(self, key, descriptor):
        p = C.c_void_p()
        check(self.ctx.lib.mmr_recon_frame_cache_get(self.handle, C.byref(self._key(key)), C.byref(self._desc(descriptor)),
                                                     C.byref(p)))
        return p.value

    def remove(self, key):
        r = C.c_int32()
        check(self.ctx.lib.mmr_recon_frame_cache_remove(self.handle, C.byref(self._key(key)), C.byref(r)))
        return bool(r.value)

    def release_idle(self, max_idle_generations):
        n = C.c_uint64()
        check(self.ctx.lib.mmr_recon_frame_cache_release_idle(self.handle, max_idle_generations, C.byref(n)))
        return n.value

    def clear(self):
        check(self.ctx.lib.mmr_recon_frame_cache_clear(self.handle))

    def stats(self):
        st = abi.new(abi.FrameCacheStats)
        check(self.ctx.lib.mmr_recon_frame_cache_stats_get(self.handle, C.byref(st)))
        return {n: getattr(st, n) for n, _ in abi.FrameCacheStats._fields_ if n != "struct_size"}

    def close(self):
        if self.handle:
            self.ctx.lib.mmr_recon_frame_cache_destroy(self.handle)
            self.handle = None


class Prepared:
    """mmr_recon_prepare / mmr_recon_prepared_launch: a batch whose job records are uploaded once and
    whose kernel launches are replayed from a HIP graph."""

    def __init__(self, ctx, codec, items, flags=0):
        self.ctx = ctx
        struct = Server.STRUCTS[codec]
        arr = (struct * len(items))(*items)
        h = C.c_void_p()
        check(ctx.lib.mmr_recon_prepare(ctx.handle, codec, C.addressof(arr), len(items), flags, C.byref(h)))
        self.handle = h

    def launch(self):
        ticket = C.c_uint64()
        check(self.ctx.lib.mmr_recon_prepared_launch(self.ctx.handle, self.handle, C.byref(ticket)))
        return ticket.value

    def close(self):
        if self.handle:
            self.ctx.lib.mmr_recon_prepared_destroy(self.ctx.handle, self.handle)
            self.handle = None


class Server:
    """Continuous-batching request loop of a Context (mmr_recon_server_*)."""

    STRUCTS = {abi.CODEC_MJPEG: abi.MjpegFrame, abi.CODEC_MPEG2: abi.Mpeg2Picture, abi.CODEC_DV: abi.DvFrame,
               abi.CODEC_H264: abi.H264Picture}

    def __init__(self, ctx, hbm_budget_bytes=0, slots=0, max_batch_frames=0, max_frames_per_request_per_batch=0,
                 max_pending=0):
        self.ctx = ctx
        self.lib = ctx.lib
        cfg = abi.new(abi.ServerConfig, hbm_budget_bytes=hbm_budget_bytes, slots=slots,
                      max_batch_frames=max_batch_frames,
                      max_frames_per_request_per_batch=max_frames_per_request_per_batch, max_pending=max_pending)
        h = C.c_void_p()
        check(self.lib.mmr_recon_server_create(ctx.handle, C.byref(cfg), C.byref(h)))
        self.handle = h

    def enqueue(self, codec, frames, convert=None, composite=None, hbm_bytes=0, tag=0, stream_id=0):
        """Enqueue one request; returns its id (raises ReconError, status BUSY when the queue is full)."""
        n = len(frames)
        farr = (self.STRUCTS[codec] * n)(*frames)
        carr = (abi.YuvToRgbJob * n)(*convert) if convert else None
        parr = (abi.CompositeJob * n)(*composite) if composite else None
        req = abi.new(abi.Request, codec=codec, stream_id=stream_id, frame_count=n, frames=C.addressof(farr),
                      convert=C.addressof(carr) if carr is not None else None,
                      composite=C.addressof(parr) if parr is not None else None, hbm_bytes=hbm_bytes, user_tag=tag)
        rid = C.c_uint64()
        check(self.lib.mmr_recon_server_enqueue(self.handle, C.byref(req), C.byref(rid)))
        return rid.value

    def poll(self, capacity=4096):
        out = (abi.FrameCompletion * capacity)()
        n = C.c_uint32()
        check(self.lib.mmr_recon_server_poll(self.handle, out, capacity, C.byref(n)))
        return [(c.request_id, c.user_tag, c.frame_index, c.status, c.latency_ms) for c in out[: n.value]]

    def drain(self):
        check(self.lib.mmr_recon_server_drain(self.handle))

    def stats(self):
        st = abi.new(abi.ServerStats)
        check(self.lib.mmr_recon_server_stats_get(self.handle, C.byref(st)))
        return {f: getattr(st, f) for f, _ in abi.ServerStats._fields_ if f != "struct_size"}

    def close(self):
        if self.handle:
            self.lib.mmr_recon_server_destroy(self.handle)
            self.handle = None


def implicit_weights(poc0, poc1, current_poc, long_term0=False, long_term1=False):
    """(w0, w1) of implicit bi-prediction, computed by the library's host code."""
    lib = load_library()
    w0, w1 = C.c_int32(), C.c_int32()
    check(lib.mmr_recon_h264_implicit_weights(poc0, poc1, current_poc, int(long_term0), int(long_term1),
                                              C.byref(w0), C.byref(w1)))
    return w0.value, w1.value
