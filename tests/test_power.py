"""Power-sampling wrapper (convfpga_amd/power.py), the counterpart of the
reference's scripts/power/*.sh: the CLI runs a child, samples until it
exits, writes the CSV and returns the child's status.  CPU tests drive it
with a synthetic reader; the GPU test reads the real board through amdsmi."""
import csv
import sys

import pytest

from convfpga_amd import power


class FakeReader:
    def __init__(self):
        self.e = 100.0

    def power_w(self):
        return 250.0

    def energy_j(self):
        self.e += 5.0
        return self.e


def test_cli_runs_child_and_writes_csv(tmp_path):
    out = tmp_path / "p.csv"
    rc = power.main(["-o", str(out), "-i", "0.005", "--", sys.executable, "-c",
                     "import time, sys; time.sleep(0.2); sys.exit(3)"], reader=FakeReader())
    assert rc == 3  # the child's status
    rows = list(csv.reader(open(out)))
    assert rows[0] == ["time_s", "power_W"]
    assert len(rows) > 5 and all(float(r[1]) == 250.0 for r in rows[1:])
    ts = [float(r[0]) for r in rows[1:]]
    assert ts == sorted(ts) and ts[-1] < 5.0


def test_monitor_summary_uses_energy_counter():
    m = power.PowerMonitor(FakeReader(), 0.005).start()
    s = m.stop()
    assert s["energy_J"] == pytest.approx(5.0)
    assert s["mean_W"] == 250.0 and s["max_W"] == 250.0
    assert s["energy_mean_W"] == pytest.approx(5.0 / s["seconds"])


def test_cli_without_command_is_an_error():
    with pytest.raises(SystemExit):
        power.main([], reader=FakeReader())


@pytest.mark.gpu
def test_reads_board_power_on_gpu():
    import torch
    torch.zeros(1, device="cuda")
    try:
        rd = power.SmiReader(power.torch_device_bdf(0))
    except power.PowerUnavailable as e:
        pytest.skip("amdsmi cannot read this board: %s" % e)
    try:
        w = rd.power_w()
        assert w is not None and 10.0 < w < 3000.0, w
        e0 = rd.energy_j()
        x = torch.randn(1 << 24, device="cuda")
        for _ in range(200):
            x = x * 1.0001
        torch.cuda.synchronize()
        e1 = rd.energy_j()
        if e0 is not None:
            assert e1 >= e0
    finally:
        rd.close()
