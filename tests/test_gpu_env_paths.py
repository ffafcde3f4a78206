"""Code paths that environment variables select, against the oracle: each
setting runs tests/env_paths_check.py in a process of its own, because the
library reads the variables once per process."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

SETTINGS = {
    "early_aggregate_ordinary_walk_two_level_fold": {"RSORT_AMD_EARLY_AGG_TILES": "1",
                                                     "RSORT_AMD_LB_WIDE_TILES": "0",
                                                     "RSORT_AMD_HIST_DIRECT": "0"},
    "global_tile_ticket": {"RSORT_AMD_XCD_TICKETS": "0"},
    "small_sorts_on_onesweep": {"RSORT_AMD_SMALL_LDS": "0"},
    "fill_kernel_clears": {"RSORT_AMD_HIST_CLEAR": "0"},
    "histogram_claims_late": {"RSORT_AMD_HIST_CLAIM": "late"},
}


@pytest.mark.parametrize("name", sorted(SETTINGS))
def test_env_selected_paths(name):
    here = os.path.dirname(os.path.abspath(__file__))
    env = {k: v for k, v in os.environ.items() if not k.startswith("RSORT_AMD_") or k == "RSORT_AMD_LIB"}
    env.update(SETTINGS[name])
    r = subprocess.run([sys.executable, "-u", os.path.join(here, "env_paths_check.py")], env=env,
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0 and "env path parity ok" in r.stdout
