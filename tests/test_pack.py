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
