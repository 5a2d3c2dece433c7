import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dm-annihilation-cxx_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X)")


# Figures that tests measure at full size (error against the reference,
# times), printed as one block at the end of the session so they appear
# with the pass/fail summary even under -q.
MEASURED = []


def record(name: str, **figures):
    MEASURED.append((name, figures))


def progress(msg: str):
    """A step of a long test finished: print it and append it to
    gpurun_out/test_progress.log, so a run whose output is captured still
    shows that it advances."""
    print(msg, flush=True)
    d = os.path.join(ROOT, "gpurun_out")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "test_progress.log"), "a") as fh:
        fh.write(msg + "\n")


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    if not MEASURED:
        return
    terminalreporter.write_line("measured:")
    for name, figs in MEASURED:
        parts = []
        for k, v in figs.items():
            parts.append(f"{k}={v:.3g}" if isinstance(v, float) else f"{k}={v}")
        terminalreporter.write_line(f"  {name}: " + " ".join(parts))


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        subprocess.run(["make", "-C", ORACLE, "-s", "liboracle.so"], check=True)
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def pw():
    import pywrapper
    pywrapper.load_library()
    return pywrapper


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def rel_err_report(gpu_map, ref_map, floor=1e-6):
    """Max per-pixel relative error over pixels above floor*max, and the
    relative error of the total."""
    ref = np.asarray(ref_map, dtype=np.float64)
    got = np.asarray(gpu_map, dtype=np.float64)
    mx = np.abs(ref).max() if ref.size else 0.0
    sel = np.abs(ref) > floor * mx
    if sel.any():
        per = np.abs(got[sel] - ref[sel]) / np.abs(ref[sel])
        max_rel = float(per.max())
    else:
        max_rel = 0.0
    # pixels below the floor must still be small in absolute terms
    below = float(np.abs(got[~sel] - ref[~sel]).max() / mx) if (~sel).any() and mx > 0 else 0.0
    tot_ref = ref.sum()
    tot_rel = float(abs(got.sum() - tot_ref) / abs(tot_ref)) if tot_ref != 0 else float(abs(got.sum()))
    # pixels the reference leaves empty (no footprint covers them) must stay
    # exactly empty: a deposit at a wrong address shows up here even when the
    # contribution counts and the total still match
    zero_mismatch = int(np.count_nonzero((ref == 0) & (got != 0)))
    return {"max_rel": max_rel, "total_rel": tot_rel, "below_floor_abs": below, "n_sel": int(sel.sum()),
            "floor": floor, "zero_mismatch": zero_mismatch}


def assert_map(rep, tol, ttol=None):
    """The map bar: per-pixel relative error <= tol above floor*max, the same
    relative bar in absolute terms (tol * floor * max) below it, no deposit
    into a pixel the reference leaves empty, and (ttol) the total."""
    assert rep["max_rel"] <= tol, rep
    assert rep["below_floor_abs"] <= tol * rep["floor"], rep
    assert rep["zero_mismatch"] == 0, rep
    if ttol is not None:
        assert rep["total_rel"] <= ttol, rep
