"""Writes tests/golden/sedov_analytic_gamma1.4.txt from the analytic Sedov
profile the reference ships for its Sedov simulation
(Flash-X/source/Simulation/SimulationMain/Sedov/sedovSolutionModLong.dat:
gamma = 1.4, E = 1, rho_up = 1, t = 1, spherical blast, shock at
r_0 = 1.03277747; columns r, v, rho, P).  Every 10th of its 10 000 rows
is kept, and the shock radius row is kept exactly.

Run in a container that holds the reference:
    python tests/golden/make_sedov_profile.py /root/reference
"""
import sys
from pathlib import Path

import numpy as np

src = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / \
    "Flash-X/source/Simulation/SimulationMain/Sedov/sedovSolutionModLong.dat"
rows = np.loadtxt(src, comments="#")
keep = np.unique(np.concatenate([np.arange(0, len(rows), 10), [len(rows) - 1]]))
out = Path(__file__).with_name("sedov_analytic_gamma1.4.txt")
np.savetxt(out, rows[keep], fmt="%.8e",
           header="analytic Sedov profile, gamma 1.4, E = 1, rho_up = 1, t = 1, r_shock = 1.03277747\n"
                  "subsampled (every 10th row) from Flash-X/source/Simulation/SimulationMain/Sedov/"
                  "sedovSolutionModLong.dat by tests/golden/make_sedov_profile.py\ncolumns: r v rho P")
print(f"{out}: {len(keep)} rows")
