"""The Python example of README.md, run as written (in a temporary
directory): a 64^3 DOPRI853 run with the HDF5 output and dense outputs,
then the same state continued by the GSL RK8PD stepper."""
import os
import re

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_readme_python_example(tmp_path, monkeypatch):
    text = open(os.path.join(ROOT, "README.md")).read()
    blocks = re.findall(r"```python\n(.*?)```", text, re.S)
    assert len(blocks) == 1
    monkeypatch.chdir(tmp_path)
    ns = {}
    exec(compile(blocks[0], "README.md", "exec"), ns)
    gw, st, gsl = ns["gw"], ns["st"], ns["gsl"]
    assert gw.ndim == 2 and gw.shape[1] == 64 and gw.shape[0] > 4
    assert st["status"] == 0 and st["intervals"] == 2
    assert gsl.t == pytest.approx(4.0)
    assert (tmp_path / "run.h5").exists()
