"""CPU-side checks of the C ABI library and the host logic (no GPU compute)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import magforce
from magforce import configs


def test_library_exports_every_declared_symbol():
    L = magforce.lib()
    declared = magforce.header_functions()
    assert len(declared) >= 30
    missing = [f for f in declared if not hasattr(L, f)]
    assert missing == []


def test_abi_version_and_status_strings():
    L = magforce.lib()
    assert L.mf_abi_version() == 1
    assert L.mf_status_string(magforce.MF_OK) == b"ok"
    assert b"argument" in L.mf_status_string(magforce.MF_ERR_INVALID_ARG)
    assert L.mf_status_string(12345) == b"unknown status"


def test_null_arguments_are_rejected_without_a_device():
    L = magforce.lib()
    assert L.mf_ctx_create(0, None) == magforce.MF_ERR_INVALID_ARG
    assert L.mf_field_allpairs(None, 0, None, None, None, 0, None, None, None, None, None, None,
                               1.0, None, None, None, None) == magforce.MF_ERR_INVALID_ARG
    assert L.mf_plan_field(None, None, None, None, None, None, None, None) == magforce.MF_ERR_INVALID_ARG
    assert L.mf_ctx_destroy(None) == magforce.MF_OK
    assert L.mf_plan_destroy(None) == magforce.MF_OK
    assert L.mf_plan_size(None) == -1


def test_struct_layouts_match_header(tmp_path):
    """sizeof and every member offset of the header's structs, as the C
    compiler lays them out, equal the ctypes mirror's."""
    mirror = {"mf_plan_opts": magforce.PlanOpts, "mf_solve_opts": magforce.SolveOpts,
              "mf_solve_stats": magforce.SolveStats, "mf_plan_info": magforce.PlanInfo}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "mf_capi.h"', 'int main(void) {']
    for cname, py in mirror.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.dirname(magforce.HEADER_PATH)
    subprocess.run(["gcc", "-std=c11", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    seen = 0
    for line in out.splitlines():
        cname, field, value = line.split()
        py = mirror[cname]
        want = ctypes.sizeof(py) if field == "sizeof" else getattr(py, field).offset
        assert int(value) == want, (cname, field)
        seen += 1
    assert seen == sum(len(p._fields_) + 1 for p in mirror.values())


def test_enum_constants_match_header():
    """Every MF_ enumerator of the header has the same value in the mirror."""
    import re
    text = open(magforce.HEADER_PATH).read()
    found = dict((k, int(v)) for k, v in re.findall(r"\b(MF_[A-Z0-9_]+)\s*=\s*(\d+)", text))
    mirror = {"MF_MODE_": "MODE_", "MF_M2L_": "M2L_", "MF_SOLVE_": "SOLVE_", "MF_ERR_": "MF_ERR_",
              "MF_OK": "MF_OK"}
    checked = 0
    for name, value in found.items():
        for cpref, ppref in mirror.items():
            if name.startswith(cpref):
                py = ppref + name[len(cpref):]
                assert getattr(magforce, py) == value, name
                checked += 1
    assert checked >= 13


def test_splitmix64_known_answers():
    # SplitMix64 with state 0: published first outputs of the reference generator.
    z = configs.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    u = configs.uniform(0, 1000)
    assert u.min() >= 0.0 and u.max() < 1.0


def test_workload_sizes_and_determinism():
    a = configs.lattice8k()
    b = configs.lattice8k()
    assert a.n == 8192 and np.array_equal(a.pos, b.pos)
    assert a.kappa == pytest.approx(0.5)
    # jitter within +-0.1 of the lattice
    sites = np.stack(np.meshgrid(np.arange(16), np.arange(16), np.arange(32), indexing="ij"),
                     -1).reshape(-1, 3)
    assert np.abs(a.pos - sites).max() <= 0.1 + 1e-6
    c = configs.cylinder262k()
    assert c.n == 262144
    r = np.sqrt(np.round(c.pos[:, 0]) ** 2 + np.round(c.pos[:, 1]) ** 2)
    assert r.max() <= 32.0
    p = configs.pool1m()
    assert p.n == 1153433
    assert p.pos[1048576:, 2].min() >= 16.0


def test_droplets_workload():
    ds = configs.droplets96()
    assert len(ds) == 96
    ns = np.array([d.n for d in ds])
    assert ns.min() >= 1500 and ns.max() <= 270000
    for d in ds[:5]:
        assert 0.5 <= d.chi <= 3.0
        assert abs(np.linalg.norm(d.hconst) - 1.0) < 1e-12


def test_example_binary_built():
    assert os.path.exists(magforce.EXAMPLE_PATH)
    assert os.access(magforce.EXAMPLE_PATH, os.X_OK)


def test_alloc_hook_installs_and_clears_without_a_device():
    """The fault-injection hook is process state: it can be set and removed
    with no device (its effect is tested on the GPU)."""
    import magforce
    seen = []
    magforce.set_alloc_hook(lambda nbytes: seen.append(nbytes) or False)
    magforce.set_alloc_hook(None)
    assert seen == []
