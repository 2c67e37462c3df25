"""Checks that run in a child process on the model library (tests/test_served_replay.py).

python tests/served_replay/checks.py planted     -- descriptors with planted mistakes (a level buffer
    one record short, a freed output plane, output planes too small, a DV level buffer cut short)
    are each reported by the kernel address models, and correct submissions are not
python tests/served_replay/checks.py exceptions  -- with the address space limited, a submission
    whose job table cannot be allocated returns MMR_STATUS_INTERNAL_ERROR with a message instead
    of letting std::bad_alloc escape through the C ABI
python tests/served_replay/checks.py cache       -- the resident frame cache's LRU, generation
    protection, budget and stream-ordered release
Prints one JSON line.
"""
import ctypes as C
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from replay import load_model_library, violations  # noqa: E402


def planted():
    lib = load_model_library()
    import numpy as np

    import mmrecode_amd
    from mmrecode_amd import abi
    from mmrecode_amd import workloads as wl

    out = {}
    ctx = mmrecode_amd.Context(device=0)

    def run(name, submit):
        before, _ = violations(lib)
        status = 0
        try:
            ctx.wait(submit())
        except mmrecode_amd.ReconError as e:
            status = e.status
        ctx.synchronize()
        n, msgs = violations(lib)
        out[name] = {"new": n - before, "status": status, "last": msgs[-1] if n > before and msgs else ""}

    # MPEG-2: a GOP reconstructed correctly, then its P picture with a level buffer one record short
    units = wl.mpeg2_gop(seed=3, width=176, height=144, gop="IPB")
    dev = []
    for unit in units:
        extra = {}
        for tag in ("fwd", "bwd"):
            if unit.meta[tag] is not None:
                for n in ("y", "cb", "cr"):
                    extra[f"{tag}_{n}"] = dev[unit.meta[tag]].ptr(n)
        dev.append(wl.DeviceUnit(ctx, unit, extra))
    run("mpeg2_ok", lambda: ctx.submit_mpeg2([d.struct for d in dev], abi.SUBMIT_PER_FRAME_LAUNCH))
    p = units[1]
    short = ctx.upload(p.inputs["levels"][:-1])
    s = p.fill({**{k: b.ptr for k, b in dev[1].bufs.items()}, "levels": short.ptr,
                **{f"fwd_{n}": dev[0].ptr(n) for n in ("y", "cb", "cr")}})
    run("mpeg2_levels_short", lambda: ctx.submit_mpeg2([s]))
    # a freed output plane
    victim = dev[2].bufs["cr"]
    s2 = dev[2].struct
    victim.free()
    run("mpeg2_plane_freed", lambda: ctx.submit_mpeg2([s2], abi.SUBMIT_PER_FRAME_LAUNCH))

    # H.264: planes of a smaller picture
    pic = wl.h264_picture(5, 0, 8, 6, intra_fraction=0.0)
    refs = [dict(zip(("y", "cb", "cr"), wl.h264_reference_planes(5, 8, 6, i))) for i in range(4)]
    rdev = [{n: ctx.upload(a).ptr for n, a in r.items()} for r in refs]
    hd = wl.DeviceUnit(ctx, pic, wl.h264_ref_ptrs(pic, rdev, "dev"))
    run("h264_ok", lambda: ctx.submit_h264([hd.struct]))
    small = {n: ctx.alloc(b.nbytes // 2) for n, b in hd.bufs.items() if n in ("y", "cb", "cr")}
    hs = pic.fill({**{k: b.ptr for k, b in hd.bufs.items()}, **{n: b.ptr for n, b in small.items()},
                   **wl.h264_ref_ptrs(pic, rdev, "dev")})
    run("h264_planes_small", lambda: ctx.submit_h264([hs]))

    # DV: levels one block short
    f = wl.dv_frame(9, 0, system=abi.DV_SYSTEM_525_60)
    dd = wl.DeviceUnit(ctx, f)
    run("dv_ok", lambda: ctx.submit_dv([dd.struct]))
    cut = ctx.upload(f.inputs["levels"][:-1])
    ds = f.fill({**{k: b.ptr for k, b in dd.bufs.items()}, "levels": cut.ptr})
    run("dv_levels_short", lambda: ctx.submit_dv([ds]))
    ctx.close()
    _ = np
    return out


def exceptions():
    import resource

    lib = load_model_library()
    import mmrecode_amd
    from mmrecode_amd import abi

    ctx = mmrecode_amd.Context(device=0)
    soft, hard = resource.getrlimit(resource.RLIMIT_AS)
    resource.setrlimit(resource.RLIMIT_AS, (12 << 30, hard))
    frames = (abi.MjpegFrame * 1)()
    ticket = C.c_uint64()
    # 2^27 frames: the job table alone is about 16 GiB, beyond the address-space limit
    status = lib.mmr_recon_mjpeg_submit(ctx.handle, frames, 1 << 27, 0, C.byref(ticket))
    message = mmrecode_amd.last_error()
    resource.setrlimit(resource.RLIMIT_AS, (soft, hard))
    # the library is still usable afterwards
    d = ctx.alloc(16)
    ctx.write(d, bytes(range(16)))
    ok = bytes(ctx.download(d)) == bytes(range(16))
    ctx.close()
    return {"status": status, "message": message, "usable_after": ok}


def noalloc():
    """mmr_recon_mpeg2_reconstruct_noalloc: no device / pinned allocation and no event during two
    calls (the second smaller), no access outside live allocations, an oversized call rejected."""
    from replay import model_stats

    lib = load_model_library()
    import mmrecode_amd
    from mmrecode_amd import abi
    from mmrecode_amd import workloads as wl

    ctx = mmrecode_amd.Context(device=0)
    sets = []
    for seed, (w, h, gop) in ((41, (720, 480, "IPBBPBB")), (42, (176, 144, "IPB"))):
        units = wl.mpeg2_gop(seed=seed, gop_index=seed, width=w, height=h, gop=gop, field_fraction=0.3)
        dev = []
        for u in units:
            extra = {"forward_source": -1 if u.meta["fwd"] is None else u.meta["fwd"],
                     "backward_source": -1 if u.meta["bwd"] is None else u.meta["bwd"]}
            for tag in ("fwd", "bwd"):
                if u.meta[tag] is not None:
                    for n in ("y", "cb", "cr"):
                        extra[f"{tag}_{n}"] = dev[u.meta[tag]].ptr(n)
            dev.append(wl.DeviceUnit(ctx, u, extra))
        sets.append(dev)
    before = model_stats(lib)
    for dev in sets:
        ctx.reconstruct_mpeg2_noalloc([d.struct for d in dev])
    ctx.reconstruct_mpeg2_noalloc([d.struct for d in sets[1]], abi.SUBMIT_PER_FRAME_LAUNCH)
    calls = model_stats(lib)
    # the oversized case of tests/test_zz_no_alloc.py: 120 1080p descriptors on the buffers of a
    # 64x32 picture (120 x 8160 macroblocks x 96 B of block order > 64 MiB) must be rejected before
    # any kernel runs, or those kernels would read far past the small buffers
    u = wl.mpeg2_gop(seed=43, gop_index=3, width=64, height=32, gop="I")[0]
    d = wl.DeviceUnit(ctx, u)
    after = model_stats(lib)
    big = abi.new(abi.Mpeg2Picture, width=1920, height=1080, picture_type=abi.MPEG2_PICTURE_I, matrices=d.ptr("matrices"),
                  macroblocks=d.ptr("mbs"), levels=d.ptr("levels"), forward_source=-1, backward_source=-1)
    for k, n in enumerate(("y", "cb", "cr")):
        big.planes[k] = d.ptr(n)
    status = lib.mmr_recon_mpeg2_reconstruct_noalloc(ctx.handle, (abi.Mpeg2Picture * 120)(*([big] * 120)), 120, 0)
    message = mmrecode_amd.last_error()
    ran = model_stats(lib)["kernels"] - after["kernels"]
    n, msgs = violations(lib)
    ctx.close()
    return {"allocations": calls["allocations"] - before["allocations"], "events": calls["events"] - before["events"],
            "kernels": calls["kernels"] - before["kernels"], "violations": n, "messages": msgs[:5],
            "oversized_status": status, "oversized_message": message, "oversized_kernels": ran}


def cache():
    """The resident frame cache on the model library: the reference's DeviceResourceCache tests
    (crates/render/src/device_cache.rs:316-392) with device allocations as the resources, plus
    budget errors, remove / clear accounting, and an eviction while a kernel that reads the evicted
    storage is still queued (the release is stream-ordered, so the kernel's reads stay inside a live
    allocation)."""
    lib = load_model_library()
    import numpy as np

    import mmrecode_amd
    from mmrecode_amd import abi

    ctx = mmrecode_amd.Context(device=0)
    out = {}

    def handle(owner):   # device_cache.rs:300-314: MmfxCanvas, revision 1, time-invariant, 2x2, rgba8
        return dict(namespace_id=3, owner=owner, revision=1, local_frame=-1, width=2, height=2, variant=0)

    rgba8_2x2 = dict(width=2, height=2, format=1, color=1, alpha_mode=1)

    def status_of(fn):
        try:
            fn()
            return 0, ""
        except mmrecode_amd.ReconError as e:
            return e.status, e.message

    # stable_key_reuses_one_backend_resource
    c = mmrecode_amd.FrameCache(ctx, 64)
    p1, s1 = c.retain(handle(1), rgba8_2x2, 16)
    p2, s2 = c.retain(handle(1), rgba8_2x2, 16)
    st = c.stats()
    out["reuse"] = {"inserted": s1 == abi.CACHE_INSERTED, "reused": s2 == abi.CACHE_REUSED, "same": p1 == p2 and p1 is not None,
                    "resources": st["resources"], "reuses": st["reuses"]}
    c.close()

    # evicts_oldest_resource_but_protects_current_graph_generation
    c = mmrecode_amd.FrameCache(ctx, 32)
    c.retain(handle(1), rgba8_2x2, 16)
    c.retain(handle(2), rgba8_2x2, 16)
    c.begin_frame()
    second = c.get(handle(2), rgba8_2x2)
    c.retain(handle(3), rgba8_2x2, 16)
    out["evict"] = {"second_before": second is not None, "first": c.get(handle(1), rgba8_2x2),
                    "second": c.get(handle(2), rgba8_2x2) is not None, "third": c.get(handle(3), rgba8_2x2) is not None,
                    "evictions": c.stats()["evictions"], "retained": c.stats()["retained_bytes"]}
    c.close()

    # refuses_to_evict_a_resource_used_by_the_current_graph
    c = mmrecode_amd.FrameCache(ctx, 16)
    c.retain(handle(1), rgba8_2x2, 16)
    status, message = status_of(lambda: c.retain(handle(2), rgba8_2x2, 16))
    out["protect"] = {"status": status, "message": message, "resources": c.stats()["resources"],
                      "retained": c.stats()["retained_bytes"]}
    c.close()

    # rejects_descriptor_collisions (a foreign backend has no counterpart: caches are per context)
    c = mmrecode_amd.FrameCache(ctx, 64)
    c.retain(handle(1), rgba8_2x2, 16)
    status, message = status_of(lambda: c.get(handle(1), dict(rgba8_2x2, width=4, height=4)))
    status2, _ = status_of(lambda: c.retain(handle(1), dict(rgba8_2x2, width=4, height=4), 16))
    out["collision"] = {"get_status": status, "message": message, "retain_status": status2}
    c.close()

    # releases_only_idle_generations
    c = mmrecode_amd.FrameCache(ctx, 64)
    c.retain(handle(1), rgba8_2x2, 16)
    c.begin_frame()
    c.retain(handle(2), rgba8_2x2, 16)
    c.begin_frame()
    c.get(handle(2), rgba8_2x2)
    released = c.release_idle(1)
    out["idle"] = {"released": released, "first": c.get(handle(1), rgba8_2x2), "second": c.get(handle(2), rgba8_2x2) is not None}
    c.close()

    # budget errors and accounting
    status_zero, _ = status_of(lambda: mmrecode_amd.FrameCache(ctx, 0))
    c = mmrecode_amd.FrameCache(ctx, 64)
    status_big, _ = status_of(lambda: c.retain(handle(9), rgba8_2x2, 65))
    status_empty, _ = status_of(lambda: c.retain(handle(9), rgba8_2x2, 0))
    c.retain(handle(1), rgba8_2x2, 16)
    c.retain(handle(2), rgba8_2x2, 32)
    removed = c.remove(handle(1))
    removed_again = c.remove(handle(1))
    after_remove = c.stats()["retained_bytes"]
    c.clear()
    out["budget"] = {"zero_budget": status_zero, "over_budget": status_big, "zero_bytes": status_empty, "removed": removed,
                     "removed_again": removed_again, "after_remove": after_remove, "after_clear": c.stats()["retained_bytes"],
                     "resources_after_clear": c.stats()["resources"]}
    c.close()

    # eviction while a queued kernel still reads the storage
    before, _ = violations(lib)
    W, H = 64, 16
    c = mmrecode_amd.FrameCache(ctx, W * H * 8)
    key = dict(namespace_id=1, owner=7, revision=1, local_frame=0, width=W, height=H, variant=0)
    desc = dict(width=W, height=H, format=2, color=2, alpha_mode=2)
    surface, _ = c.retain(key, desc, W * H * 8)
    ctx.write(mmrecode_amd.DeviceBuffer(ctx, surface, W * H * 8), np.zeros(W * H * 4, np.uint16))
    rgba = ctx.alloc(W * H * 4)
    job = abi.new(abi.LinearConvertJob, direction=abi.LINEAR_TO_SRGBA8, width=W, height=H, linear=surface, linear_stride=W,
                  rgba_stride=W * 4, rgba=rgba.ptr)
    ticket = ctx.submit_linear_convert([job])
    c.begin_frame()
    c.retain(dict(key, owner=8), desc, W * H * 8)   # evicts the surface the queued kernel reads
    evictions = c.stats()["evictions"]
    ctx.wait(ticket)
    after, msgs = violations(lib)
    out["ordered_release"] = {"evictions": evictions, "new_violations": after - before, "messages": msgs[-3:]}
    c.close()
    rgba.free()
    ctx.close()
    return out


if __name__ == "__main__":
    what = sys.argv[1]
    print(json.dumps({"planted": planted, "exceptions": exceptions, "noalloc": noalloc, "cache": cache}[what]()))
