"""The served path on the CPU: the library's real host code (submission, tickets, the
continuous-batching server) on a host model of the HIP runtime with address models of every kernel
(tests/served_replay/). Each check runs in a child process, which loads the model library instead
of libmmr_recon.so."""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "served_replay")


def _run(args, timeout):
    sys.path.insert(0, HERE)
    import build

    build.build()
    r = subprocess.run([sys.executable, os.path.join(HERE, args[0]), *args[1:]], capture_output=True, text=True, timeout=timeout)
    assert r.stdout.strip(), r.stderr[-3000:]
    return r.returncode, json.loads(r.stdout.strip().splitlines()[-1]), r.stderr


def test_model_reports_planted_mistakes():
    rc, out, err = _run(["checks.py", "planted"], 600)
    assert rc == 0, err[-3000:]
    for ok in ("mpeg2_ok", "h264_ok", "dv_ok"):
        assert out[ok]["new"] == 0, (ok, out[ok])
    assert out["mpeg2_levels_short"]["new"] >= 1 and "mpeg2 levels" in out["mpeg2_levels_short"]["last"]
    assert out["mpeg2_plane_freed"]["new"] >= 1 and "already freed" in out["mpeg2_plane_freed"]["last"]
    assert out["h264_planes_small"]["new"] >= 1 and "overrun" in out["h264_planes_small"]["last"]
    assert out["dv_levels_short"]["new"] >= 1 and "dv levels" in out["dv_levels_short"]["last"]


def test_exception_becomes_internal_error_status():
    """crates/capi/src/lib.rs:572-588 wraps every export in catch_unwind; here an allocation failure
    inside a submit call returns MMR_STATUS_INTERNAL_ERROR (100) with a message."""
    rc, out, err = _run(["checks.py", "exceptions"], 300)
    assert rc == 0, err[-3000:]
    assert out["status"] == 100, out
    assert "out of host memory" in out["message"]
    assert out["usable_after"]


@pytest.mark.timeout(900)
def test_served_mix_replay_full_shape():
    """bench.py's served trace at full shapes (64 streams; MJPEG 640x480, MPEG-2 1080p GOP-15, DV,
    H.264 720p / 1080p / 2160p; ring slots; BT.709 RGBA + 3-layer composite) through the server
    and the plain path: every kernel access inside a live allocation, every frame completed OK."""
    rc, out, err = _run(["replay.py", "--rates", "300", "--seconds", "0.4", "--plain-seconds", "2"], 900)
    assert out["violations"] == 0, out["messages"]
    run = out["runs"][0]
    assert run["requests"] > 50 and run["served_frames"] == run["frames"] and run["bad_status"] == 0, run
    assert rc == 0


def test_noalloc_entry_point_allocates_nothing():
    """mmr_recon_mpeg2_reconstruct_noalloc makes no device, pinned or event allocation over three
    calls (a 7-picture 720x480 GOP, then a smaller 3-picture one, then that one per picture), every
    kernel access stays inside live memory, and a submission beyond its scratch is rejected before
    any kernel runs."""
    rc, out, err = _run(["checks.py", "noalloc"], 300)
    assert rc == 0, err[-3000:]
    assert out["allocations"] == 0 and out["events"] == 0, out
    assert out["kernels"] > 0 and out["violations"] == 0, out
    assert out["oversized_status"] == 1 and out["oversized_kernels"] == 0, out
    assert "MMR_RECON_NOALLOC_SCRATCH_BYTES" in out["oversized_message"]


def test_frame_cache_follows_device_resource_cache():
    """mmr_recon_frame_cache_* against the reference's DeviceResourceCache tests
    (crates/render/src/device_cache.rs:316-392) on the model library, plus budget errors,
    accounting, and a stream-ordered release under a queued kernel (checks.py cache)."""
    rc, out, err = _run(["checks.py", "cache"], 300)
    assert rc == 0, err[-3000:]
    assert out["reuse"] == {"inserted": True, "reused": True, "same": True, "resources": 1, "reuses": 1}
    assert out["evict"] == {"second_before": True, "first": None, "second": True, "third": True, "evictions": 1, "retained": 32}
    assert out["protect"]["status"] == 4 and "used in generation" in out["protect"]["message"]
    assert out["protect"]["resources"] == 1 and out["protect"]["retained"] == 16
    assert out["collision"]["get_status"] == 4 and out["collision"]["retain_status"] == 4
    assert out["idle"] == {"released": 1, "first": None, "second": True}
    b = out["budget"]
    assert b["zero_budget"] == 2 and b["over_budget"] == 2 and b["zero_bytes"] == 2
    assert b["removed"] and not b["removed_again"] and b["after_remove"] == 32
    assert b["after_clear"] == 0 and b["resources_after_clear"] == 0
    assert out["ordered_release"]["evictions"] == 1 and out["ordered_release"]["new_violations"] == 0, out["ordered_release"]
