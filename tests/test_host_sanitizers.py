"""CPU: the host code of libvsgm (every vsgm_amd/csrc/host/*.cc except the
CLI mains) built with g++ under AddressSanitizer (leak detection on) and
UndefinedBehaviorSanitizer, running tests/sanitize/host_checks.cc: planner,
graph IO, SNAP ingestion, views, packing and the C-ABI's host calls; and
the pm_pipeline hand-off under ThreadSanitizer (tests/sanitize/queue_checks.cc)."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
AMD = os.path.join(ROOT, "vsgm_amd")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    hip_lib = os.path.join(AMD, "lib", "libvsgm_hip.so")
    if not os.path.exists(hip_lib):
        pytest.skip("libvsgm_hip.so is not built")
    srcs = sorted(s for s in glob.glob(os.path.join(AMD, "csrc", "host", "*.cc")) if not s.endswith("_main.cc"))
    srcs.append(os.path.join(ROOT, "tests", "sanitize", "host_checks.cc"))
    flags = ["-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
             "-fopenmp"]
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(AMD, "csrc", "host"), "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
    objs, jobs = [], []
    for src in srcs:  # one compiler per source, in parallel
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        objs.append(obj)
        jobs.append(subprocess.Popen(["g++"] + flags + inc + ["-c", "-o", obj, src], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                     text=True))
    for job in jobs:
        _, err = job.communicate(timeout=600)
        assert job.returncode == 0, err[-3000:]
    exe = str(tmp_path / "host_checks")
    link = subprocess.run(["g++"] + flags + ["-o", exe] + objs + ["-L" + os.path.join(AMD, "lib"), "-lvsgm_hip",
                                                                   "-Wl,-rpath," + os.path.join(AMD, "lib")],
                          capture_output=True, text=True, timeout=600)
    assert link.returncode == 0, link.stderr[-3000:]
    work = tmp_path / "work"
    work.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe, str(work)], capture_output=True, text=True, timeout=600, env=env)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert "host checks passed" in run.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pipeline_hand_off_is_race_free_under_tsan(tmp_path):
    """The producer/consumer hand-off of pm_pipeline (BoundedQueue and the
    holder protocol, tests/sanitize/queue_checks.cc) under ThreadSanitizer,
    including a failing consumer and close() waking a blocked push."""
    exe = str(tmp_path / "queue_checks")
    build = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread", "-I" + os.path.join(AMD, "csrc", "host"),
                            "-o", exe, os.path.join(ROOT, "tests", "sanitize", "queue_checks.cc")],
                           capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-3000:]
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    assert run.returncode == 0, run.stderr[-4000:]
    assert "queue checks passed" in run.stdout
