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


def _plan_offsets(w):
    """Word offsets of the fields of a plan stream (the layout loom_amd/plan.py
    writes), for planting bad indexes."""
    w = [int(x) for x in w]
    i = 3
    off = {"cols": [], "perms": [], "rounds": [], "group_cols": [], "group_hdr": [], "shifts": []}
    nm = w[i]; i += 1
    off["modules"] = i
    i += 3 * nm
    nc = w[i]; i += 1
    for _ in range(nc):
        off["cols"].append(i)
        i += 5
    ni = w[i]; i += 1 + ni
    npm = w[i]; i += 1
    for _ in range(npm):
        off["perms"].append(i)
        i += 2

    def dag(i):
        n = w[i]
        return i + 2 + 4 * n

    nr = w[i]; i += 1
    for _ in range(nr):
        off["rounds"].append(i)
        ns = w[i + 3]
        i += 4
        for _ in range(ns):
            no = w[i + 2]
            i += 3 + no
            npar = w[i]
            i += 1 + npar
            nin = w[i]
            i += 1
            for _ in range(nin):
                i = dag(i)
    nx = w[i]; i += 1
    for _ in range(nx):
        ln = w[i + 1]
        i += 2 + (ln + 3) // 4
    npub = w[i]; i += 1
    for _ in range(npub):
        i += 2
        for _ in range(2):
            i += 1 + (w[i] + 3) // 4
        i += 1 + w[i]
    nt = w[i]; i += 1
    for _ in range(nt):
        ng = w[i + 1]
        i += 2
        for _ in range(ng):
            off["group_hdr"].append(i)
            nb = w[i + 2]
            off["group_cols"] += list(range(i + 3, i + 3 + nb))
            i += 3 + nb
            ne = w[i]
            off["group_cols"] += list(range(i + 1, i + 1 + ne))
            i += 1 + ne
            for _ in range(nb + ne):
                k = w[i]
                off["shifts"] += list(range(i + 1, i + 1 + k))
                i += 1 + k
    for _ in range(nm):
        i = dag(i)
    nbus = w[i]; i += 1
    off["buses"] = []
    for _ in range(nbus):
        off["buses"] += list(range(i + 1, i + 1 + w[i]))
        i += 1 + w[i]
    assert i == len(w)
    return off


def _mutations():
    from loom_amd import workloads, plan
    prog, _ = workloads.plonk_instance(4, 0)
    w = plan.Plan(prog, 8).words()
    o = _plan_offsets(w)
    SRC_SIGMA, SRC_LAGRANGE = 3, 1
    sig = [c for c in o["cols"] if w[c + 2] == SRC_SIGMA]
    lag = [c for c in o["cols"] if w[c + 2] == SRC_LAGRANGE]
    assert sig and lag and o["group_cols"] and o["shifts"] and o["rounds"]
    cases = {
        "group column index": (o["group_cols"][0], 100000),
        "group module": (o["group_hdr"][0], 7),
        "group size": (o["group_hdr"][0] + 1, 1 << 10),
        "opening shift": (o["shifts"][0], 1 << 20),
        "round tree": (o["rounds"][0] + 1, 9),
        "round challenge column": (o["rounds"][0], 100000),
        "sigma permutation": (sig[0] + 4, 5),
        "sigma chunk": (sig[0] + 3, 60),
        "Lagrange position": (lag[0] + 3, 1 << 20),
        "column field": (o["cols"][0], 2),
    }
    out = {}
    for name, (at, val) in cases.items():
        bad = w.copy()
        bad[at] = val
        out[name] = bad
    return w, out


def test_malformed_plans_are_rejected():
    """Every index of the plan is checked by the parser: lg_plan_check and
    lg_plan_bytes_per_instance (the admission estimate, which writes per-column
    flags) return an error for each planted bad index instead of touching
    memory out of bounds."""
    import ctypes
    from loom_amd import prover, serve, _lib
    good, bad = _mutations()
    assert prover.check_plan(good) > 0
    assert serve.plan_bytes_per_instance(good) > 0
    for name, w in bad.items():
        with pytest.raises(_lib.LoomGpuError, match="plan"):
            prover.check_plan(w)
        v = ctypes.c_uint64()
        ww = np.ascontiguousarray(w, dtype=np.uint32)
        rc = _lib.lib().lg_plan_bytes_per_instance(ww.ctypes.data_as(_lib.u32p), ww.size, ctypes.byref(v))
        assert rc == -1, name
        assert "plan" in _lib.lib().lg_last_error().decode(), name


def test_host_addresses_follow_replaced_arrays():
    """The Python binding caches an instance's input addresses; replacing an
    input array (not only writing into it) must give the new array's address,
    and the cached arrays stay referenced while their addresses are in use."""
    from loom_amd import workloads, prover
    it = prover.Prepared(*workloads.plonk_instance(5, 1), 8)
    cols, perms = it.host_addresses()
    assert cols == [c.ctypes.data for c in it.columns]
    assert perms == [q.ctypes.data for q in it.perms]
    assert it.host_addresses() == (cols, perms)
    it.perms = [q.copy() for q in it.perms]
    cols2, perms2 = it.host_addresses()
    assert cols2 == cols
    assert perms2 == [q.ctypes.data for q in it.perms] and perms2 != perms
    it.columns[0] = it.columns[0].copy()
    assert it.host_addresses()[0][0] == it.columns[0].ctypes.data
