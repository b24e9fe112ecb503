"""CPU checks of the workload runner's command line (cpfft_amd/run.py): a
step range outside the workload is refused before any GPU work."""
import pytest

from cpfft_amd import run


@pytest.mark.parametrize("args", [["--first", "5", "--last", "2"],
                                  ["--last", "241"],
                                  ["--first", "0"],
                                  ["--group", "0"]])
def test_bad_step_ranges_are_refused(args):
    with pytest.raises(SystemExit) as e:
        run.main(["--workload", "c3_128"] + args)
    assert e.value.code == 2  # argparse's usage error
