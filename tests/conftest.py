import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "deepaw-half_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libhalf.so")


def golden(name):
    with open(os.path.join(GOLDEN, name + ".json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def inputs_dir(tmp_path_factory):
    """Directory with the seeded synthetic inputs (si2 from the committed
    fixtures, larger workloads regenerated by the seeded generator)."""
    return str(tmp_path_factory.mktemp("inputs"))


def workload_files(name, inputs_dir):
    from half_amd import synth
    if name == "si2":
        return (os.path.join(GOLDEN, "inputs", "si2.CHGCAR"), os.path.join(GOLDEN, "inputs", "si2.POTCAR"))
    chg, pot, _ = synth.make_inputs(name, inputs_dir)
    return chg, pot


def require_same_inputs(g, chg, pot):
    from half_amd import synth
    if synth.sha256(chg) != g["chgcar_sha256"] or synth.sha256(pot) != g["potcar_sha256"]:
        pytest.fail("regenerated inputs differ from the ones the goldens were made from")


def psi_block(n, nb=3):
    """Deterministic state block used by the golden apply images."""
    ig = np.arange(1, n + 1)[:, None]
    ib = np.arange(1, nb + 1)[None, :]
    return np.cos(0.37 * ig + ib) + 1j * np.sin(0.11 * ig * ib)


def pytest_runtest_logstart(nodeid, location):
    print(f"\n[start] {nodeid}", flush=True)


def pytest_runtest_logreport(report):
    # failure details right away, so they survive a run that ends abruptly
    if report.failed:
        text = str(report.longrepr)
        print(f"[failed] {report.nodeid} ({report.when}):\n" + text[-1500:], flush=True)
