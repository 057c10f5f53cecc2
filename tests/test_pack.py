"""The level-set packing of the block ILU factors, checked on the host: every
level of several operators is packed as the device upload packs it and the
ring solve is replayed with the device kernels' data contract (whole steps
and half steps), against plain forward/backward substitution on the block
factors; a ring of half the packed size must fail the same replay.  The
replay is tests/cpp/pack_check.cpp, built with the library.  CPU only."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "pack_check")


def test_ring_replay_matches_substitution():
    assert os.path.exists(CHECK), "tests/cpp/pack_check is missing: run __graft_entry__.build()"
    out = subprocess.run([CHECK], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "control detected" in out.stdout
    assert out.stdout.count(" ok\n") >= 11


SAN = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def test_setup_and_packing_under_sanitizers():
    """The same replay (host setup, ILUT, partition, level-set packing) built
    with AddressSanitizer and UBSan: no out-of-bounds access, use after free
    or undefined behaviour on any of the operators, the control included."""
    exe = os.path.join(ROOT, "tests", "cpp", "pack_check_san")
    assert os.path.exists(exe), "tests/cpp/pack_check_san is missing: run __graft_entry__.build()"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=SAN)
    print(out.stdout[-3000:], out.stderr[-3000:])
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in out.stderr and "runtime error" not in out.stderr
    assert "control detected" in out.stdout


def test_thread_pool_rethrows_worker_exceptions():
    """parallel_for (the setup's thread pool) rethrows a worker's
    std::bad_alloc on the calling thread, where the setup maps it to
    PGS_ERROR_OOM, instead of terminating the process."""
    exe = os.path.join(ROOT, "tests", "cpp", "parallel_for_check_san")
    assert os.path.exists(exe), "tests/cpp/parallel_for_check_san is missing: run __graft_entry__.build()"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, env=SAN)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count(" ok") == 5
