"""The engine's libstdc++-heap emulation (falconnpp_amd/csrc/heap.h) against
std::priority_queue itself on tie-heavy push/pop sequences (CPU only)."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))


def test_heap_emulation_matches_std_priority_queue():
    src = os.path.join(HERE, "cpp", "heap_check.cpp")
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "heap_check")
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-o", exe, src])
        out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr
        assert out.stdout.startswith("OK")
