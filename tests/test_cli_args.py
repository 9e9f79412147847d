"""Argument handling of the device CLI (no GPU needed)."""
import argparse

import pytest


def test_extents_and_failure_path(monkeypatch, capsys):
    from ym_amd import cli

    assert cli._extents("2,4,6,8") == (2, 4, 6, 8)
    with pytest.raises(argparse.ArgumentTypeError):
        cli._extents("2,4,6")

    def fail(**kw):
        raise cli.GateFailure("gate")

    monkeypatch.setattr(cli, "su3_ensemble", fail)
    assert cli.main(["su3-ensemble", "--master-seed", "0x10"]) == 1
    assert "status=failed" in capsys.readouterr().err
    with pytest.raises(SystemExit):
        cli.main(["no-such-command"])
