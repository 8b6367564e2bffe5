"""bench.py's command-line contract and its CPU baseline leg, on the CPU."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_cpu_baseline_reports_the_fields_the_line_carries():
    sys.path.insert(0, str(ROOT))
    import bench
    from luma_amd import setups

    case = setups.taylor_green((16, 12, 10), u0=0.05, tau=0.55, dims=3)
    cpu = bench.cpu_baseline(case, target_s=0.2)
    assert cpu["unit"] == "MLUPS" and cpu["kind"] == "port" and cpu["value"] > 0
    assert cpu["cores"] >= 1 and "steps" in cpu["sample"]
    assert set(cpu["host"]) == {"nproc", "affinity", "cpu_model", "OMP_NUM_THREADS"}
    assert "5 warm-up steps" in cpu["sample"] and cpu["luma_active_cell_ops_per_time"] > 0


def test_kernel_source_hash_follows_the_kernel_sources():
    sys.path.insert(0, str(ROOT))
    import bench

    h = bench.kernel_source_hash()
    assert len(h) == 16 and h == bench.kernel_source_hash()
    assert all((ROOT / rel).exists() for rel in bench.KERNEL_SOURCES)
