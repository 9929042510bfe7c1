"""Fixtures from the reference's own test data (run in a container that
holds the reference at /root/reference):
- regression_logs.json: the regression logs that the reference's rtests
  compare against (programs/*/tests/*_rtest.log, written by
  output_regression_log, src/m_output.f90:788-842): per log, its header
  line and its data lines as text.  The tests format the numbers of every
  line again and compare the text.
- air_chemistry_v1.txt: the input data file of the reference's chemistry
  rtests (programs/standard_2d/tests/air_chemistry_v1.txt: reaction list,
  transport and rate tables), copied unchanged.
- td_air_siglo_swarm.txt: the old-style transport data of the reference's
  standard rtests (programs/standard_2d/tests/td_air_siglo_swarm.txt; the
  3D tests hold the same file), copied unchanged.
- air_chemistry_small_v1.txt: the input data file of the reference's 3D
  chemistry rtests (programs/standard_3d/tests/air_chemistry_small_v1.txt),
  copied unchanged."""
import glob
import json
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

logs = {}
for path in sorted(glob.glob(os.path.join(REF, "programs", "*", "tests", "*_rtest.log"))):
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f if ln.strip()]
    logs[os.path.relpath(path, REF)] = {"header": lines[0], "lines": lines[1:]}
with open(os.path.join(HERE, "regression_logs.json"), "w") as f:
    json.dump(logs, f, indent=0)
print(len(logs), "logs,", sum(len(v["lines"]) for v in logs.values()), "lines")
shutil.copyfile(os.path.join(REF, "programs", "standard_2d", "tests", "air_chemistry_v1.txt"),
                os.path.join(HERE, "air_chemistry_v1.txt"))
shutil.copyfile(os.path.join(REF, "programs", "standard_2d", "tests", "td_air_siglo_swarm.txt"),
                os.path.join(HERE, "td_air_siglo_swarm.txt"))
shutil.copyfile(os.path.join(REF, "programs", "standard_3d", "tests", "air_chemistry_small_v1.txt"),
                os.path.join(HERE, "air_chemistry_small_v1.txt"))
