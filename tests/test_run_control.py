"""The external stop file and the time budget of jqmc_amd.checkpoint.RunControl (no GPU)."""

import time

from jqmc_amd.checkpoint import RunControl


def test_stop_file_and_time_budget(tmp_path):
    f = tmp_path / "external_control_mcmc.toml"
    rc = RunControl(str(f), max_time=1e9)
    assert f.read_text() == "[external_control]\nstop = false\n"
    assert not rc.should_stop()
    f.write_text("[external_control]\nstop = true\n")
    assert rc.should_stop() and "stop = true" in rc.reason
    rc.close()
    assert not f.exists()
    rc = RunControl(str(f), max_time=0.01)
    time.sleep(0.02)
    assert rc.should_stop() and "max_time" in rc.reason
    rc.close()


def test_unreadable_stop_file_does_not_stop(tmp_path):
    f = tmp_path / "c.toml"
    rc = RunControl(str(f), max_time=1e9)
    f.write_text("this is [ not toml")
    assert not rc.should_stop()
    rc.close()
