"""bench.py's contract on the CPU: one GPU only, and the algorithmic byte
model of the roofline (every operand streamed once)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_byte_model():
    kb = bench.kernel_bytes_per_voxel
    h = 33 / 64   # half-spectrum points per voxel at Nz = 64
    # Newton kernel: E 6, S 6, eps_p 5, CRSS 12, R 9, the 15-entry K_c tile read,
    # ORI (4 B) and PH (1 B), S and P written
    assert kb("nr", 64, 64, 64, False, False) == 525
    assert kb("nr", 64, 64, 64, False, True) == 525 + 12 * 8   # the backward CRSS too
    assert abs(kb("plane_fwd", 64, 64, 64, False, False) - (48 + 96 * h)) < 1e-12
    assert abs(kb("xgreen", 64, 64, 64, False, False) - (2 * 96 * h + 2 * 48 * h)) < 1e-12
    assert abs(kb("plane_inv", 64, 64, 64, False, False) - (96 * h + 96)) < 1e-12
    # FP32 spectra halve the spectrum terms only
    assert abs(kb("plane_inv", 128, 128, 128, True, False) - (48 * 65 / 128 + 96)) < 1e-12
    assert kb("nr", 128, 128, 128, True, False) == 525
