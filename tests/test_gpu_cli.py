"""`python -m ym_amd su3-ensemble | su2-ensemble`: the device runs of the
reference binaries su3_ensemble_reference.rs and su2_ensemble_reference.rs
pass their gates and print the reference's report keys."""
import pytest

pytestmark = pytest.mark.gpu

SU3_KEYS = ["sampler_id", "proposal_id", "rng_id", "stream_id", "master_seed", "chain_id", "derived_seed", "lattice",
            "beta", "proposal_angle", "angle_grid_magnitude_bits", "maximum_reference_sweeps",
            "thermalization_sweeps", "measurements", "sweeps_between_measurements", "autocorrelation_max_lag",
            "total_sweeps", "metropolis_proposals", "metropolis_accepts", "metropolis_acceptance", "mean_plaquette",
            "variance", "autocorrelation", "positive_window_lag", "tau_int", "effective_sample_size",
            "standard_error", "final_link_unitarity_residual", "final_link_determinant_residual", "status"]


def parse(out):
    return dict(line.split("=", 1) for line in out.strip().splitlines())


@pytest.mark.parametrize("precision", ["64", "32"])
def test_su3_ensemble_cli(ym, capsys, precision):
    from ym_amd import cli

    assert cli.main(["su3-ensemble", "--precision", precision]) == 0
    kv = parse(capsys.readouterr().out)
    for k in SU3_KEYS:
        assert k in kv, k
    assert kv["status"] == "passed" and kv["lattice"] == "2,2,2,2"
    assert float(kv["metropolis_acceptance"]) == 1.0  # beta = 0 accepts everything
    assert int(kv["total_sweeps"]) == 4 + 16
    assert int(kv["metropolis_proposals"]) == 20 * 8 * 4 * 16
    assert kv["derived_seed"] == f"{ym.derive_su3_chain_seed(0x3C00198212230101, 0):016x}"
    assert len(kv["measurements"].split(",")) == 16


def test_su3_ensemble_cli_beta6(ym, capsys):
    from ym_amd import cli

    assert cli.main(["su3-ensemble", "--extents", "4,4,4,4", "--beta", "6.0", "--measurements", "8"]) == 0
    kv = parse(capsys.readouterr().out)
    assert kv["status"] == "passed" and 0.3 < float(kv["metropolis_acceptance"]) < 1.0


def test_su2_ensemble_cli(ym, capsys):
    from ym_amd import cli

    assert cli.main(["su2-ensemble"]) == 0
    kv = parse(capsys.readouterr().out)
    assert kv["status"] == "passed"
    for k in ("derived_seed", "beta0_mean_plaquette", "beta0_tau_int", "beta0_effective_sample_size",
              "heatbath_acceptance", "overrelaxation_action_residual", "final_link_unit_residual"):
        assert k in kv, k
    assert abs(float(kv["beta0_mean_plaquette"])) <= 0.08
    assert float(kv["overrelaxation_action_residual"]) <= 2e-10
