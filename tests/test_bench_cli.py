"""dedisp_bench: the reference's benchmark tool (bin/benchmark/bench.hpp) as a
command-line front end over the C++ plan classes (DedispPlan, TDDPlan,
FDDGPUPlan)."""
import re
import subprocess
from pathlib import Path

import pytest


def _exe(dd):
    exe = Path(dd.__file__).parent / "bin" / "dedisp_bench"
    assert exe.exists(), "build the library first"
    return exe


def test_bench_cli_options_without_device(dd):
    exe = _exe(dd)
    r = subprocess.run([str(exe), "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "-n [ntrials]" in r.stdout and "-m [plan]" in r.stdout
    # unknown plan, and a DM range whose end is below its start: refused while parsing
    assert subprocess.run([str(exe), "-m", "xyz"], capture_output=True, timeout=60).returncode == 1
    r = subprocess.run([str(exe), "-r", "5,2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Error parsing DM range" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("plan", ["dedisp", "tdd", "fdd"])
@pytest.mark.parametrize("device", [False, True])
def test_bench_cli_runs(dd, plan, device):
    exe = _exe(dd)
    args = [str(exe), "-m", plan, "-s", "60000", "-n", "32", "-c", "256", "-i", "2", "-q"]
    if device:
        args.append("-d")
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Benchmark finished" in r.stdout
    rates = [float(m) for m in re.findall(r"\(([0-9.e+]+) DM\*samples/s\)", r.stdout)]
    assert len(rates) == 2 and all(x > 0 for x in rates)
    assert "accumulated for 2 iteration(s)" in r.stdout


@pytest.mark.gpu
def test_bench_cli_generated_dm_list(dd):
    r = subprocess.run([str(_exe(dd)), "-m", "tdd", "-s", "50000", "-c", "128", "-r", "0,40"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Generating optimal DM trials" in r.stdout
    m = re.search(r"Computing (\d+) DMs from 0\.000000 to ([0-9.]+)", r.stdout)
    assert m and int(m.group(1)) > 1 and float(m.group(2)) >= 40.0
