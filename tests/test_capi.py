"""The C ABI library (CPU checks: no compute call needs a device here).

* libloomgpu.so loads and exports every function include/loomgpu.h declares;
* a plan round-trips through the C++ parser and its DAG kernels compile with
  hiprtc for gfx950 (lg_plan_check needs no device);
* malformed plans are rejected with a message, not a crash.
"""
import ctypes
import os
import re

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "include", "loomgpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lg_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from loom_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.EXPORTS, "ctypes prototype missing for " + n


def test_plan_compiles_without_device():
    from loom_amd import workloads, plan, prover
    prog, _ = workloads.plonk_instance(5, 0)
    w = plan.Plan(prog, 8).words()
    assert prover.check_plan(w) > 0


def test_plan_with_lookups_compiles_without_device():
    """Multi-module plan with range lookups: multiplicity and cyclic logup
    steps, exposed totals, the range module."""
    from loom_amd import workloads, plan, prover
    prog, _ = workloads.uneven_range_instance([4, 6], 1)
    w = plan.Plan(prog, 8).words()
    assert prover.check_plan(w) > 0


def test_plan_rejects_garbage():
    from loom_amd import prover, _lib
    with pytest.raises(_lib.LoomGpuError):
        prover.check_plan(np.array([1, 2, 3], dtype=np.uint32))
    from loom_amd import workloads, plan
    prog, _ = workloads.plonk_instance(4, 0)
    w = plan.Plan(prog, 8).words()
    with pytest.raises(_lib.LoomGpuError):
        prover.check_plan(w[:-3])
