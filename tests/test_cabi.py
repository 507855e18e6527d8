"""The C-ABI libraries load and export what the headers declare (CPU only, no
compute calls)."""
import os
import re
import subprocess

import pytest

import axhel_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_c_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(axhel_hip_\w+)\s*\(", text))
    return names


def exported(path, demangle=False):
    args = ["nm", "-D", "--defined-only"] + (["-C"] if demangle else []) + [path]
    return subprocess.run(args, capture_output=True, text=True, check=True).stdout


def test_library_loads_and_exports_every_declared_symbol():
    lib = axhel_amd.lib()
    names = declared_c_functions()
    assert len(names) >= 25
    syms = exported(axhel_amd.LIB_PATH)
    for name in sorted(names):
        assert re.search(rf"\bT {name}$", syms, flags=re.M), name
        assert hasattr(lib, name)


def test_python_binding_covers_the_header():
    assert declared_c_functions() == set(axhel_amd.SIGNATURES)


def test_cxx_dropin_exports_reference_api():
    syms = exported(axhel_amd.CXX_LIB_PATH, demangle=True)
    for sig in [
        "unipi::axhel::EltwiseAddMod(unsigned long*, unsigned long const*, unsigned long const*, unsigned long, unsigned long)",
        "unipi::axhel::EltwiseAddMod(unsigned long*, unsigned long const*, unsigned long, unsigned long, unsigned long)",
        "unipi::axhel::EltwiseSubMod(unsigned long*, unsigned long const*, unsigned long const*, unsigned long, unsigned long)",
        "unipi::axhel::EltwiseSubMod(unsigned long*, unsigned long const*, unsigned long, unsigned long, unsigned long)",
        "unipi::axhel::EltwiseMulMod(unsigned long*, unsigned long const*, unsigned long const*, unsigned long, unsigned long, unsigned long)",
        "unipi::axhel::EltwiseFMAMod(unsigned long*, unsigned long const*, unsigned long, unsigned long const*, unsigned long, unsigned long, unsigned long)",
        "unipi::axhel::EltwiseReduceMod(unsigned long*, unsigned long const*, unsigned long, unsigned long, unsigned long, unsigned long)",
        "unipi::axhel::NTT::NTT(unsigned long, unsigned long, unsigned long)",
        "unipi::axhel::NTT::NTT(unsigned long, unsigned long, unipi::axhel::NTTMultiplyOperand const*, unipi::axhel::NTTMultiplyOperand const*, unipi::axhel::NTTMultiplyOperand)",
        "unipi::axhel::NTT::ComputeForward(unsigned long*, unsigned long const*, unsigned long, unsigned long) const",
        "unipi::axhel::NTT::ComputeInverse(unsigned long*, unsigned long const*, unsigned long, unsigned long) const",
        "unipi::axhel::NTT::Degree() const",
        "unipi::axhel::NTT::Modulus() const",
    ]:
        assert sig in syms, sig


def test_kernels_are_built_for_gfx950():
    data = open(axhel_amd.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_configs_are_listed_without_a_gpu():
    for logn in range(1, 21):
        n = 1 << logn
        assert axhel_amd.config_count(n) >= 1
        c = axhel_amd.config_get(n, 0)
        assert c["log_elems"] in (1, 2, 3, 4)
    assert axhel_amd.config_count(3) == 0
    assert axhel_amd.config_count(1 << 21) == 0


def test_rns_create_rejects_bad_arguments():
    """Invalid degrees are refused before any root search (degree 0 would
    divide by zero there); a modulus without a 2N-th root and an empty
    modulus list are refused.  Every case returns before any HIP call."""
    from axhel_amd import params
    q = params.rns_primes(1 << 10, 1)[0]
    for n in (0, 1, 3, 1000, 1 << 21):
        with pytest.raises(axhel_amd.AxhelError):
            axhel_amd.RNS(n, [q])
    with pytest.raises(axhel_amd.AxhelError):
        axhel_amd.RNS(1 << 10, [97])  # 97 - 1 is not a multiple of 2048
    with pytest.raises(axhel_amd.AxhelError):
        axhel_amd.RNS(1 << 10, [])


def test_tuning_table_sets_batch_and_single_call_configurations():
    """The shipped table's columns (batch, single call, and optionally small
    launches, which default to the single-call entry) are what the library
    reports; an explicit selection applies to single calls and small
    launches too, and -1 restores the table's choices.  Host-side state
    only, no compute."""
    path = os.path.join(ROOT, "axhel_amd", "ntt_tuning_gfx950.txt")
    table = {}
    for line in open(path):
        parts = line.split()
        if parts and not line.startswith("#"):
            table[int(parts[0])] = [int(x) for x in parts[1:]]
    assert all(len(v) in (2, 3) for v in table.values())
    for logn, (batch, single, *small) in table.items():
        n = 1 << logn
        assert axhel_amd.config_selected(n) == batch and axhel_amd.config_single(n) == single, logn
        assert axhel_amd.config_small(n) == (small[0] if small else single), logn
    n = 1 << 16
    try:
        axhel_amd.config_select(n, 2)
        assert axhel_amd.config_selected(n) == 2 and axhel_amd.config_single(n) == 2
        assert axhel_amd.config_small(n) == 2
    finally:
        axhel_amd.config_select(n, -1)
    assert axhel_amd.config_selected(n) == table[16][0] and axhel_amd.config_single(n) == table[16][1]
    assert axhel_amd.config_small(n) == table[16][-1]
    # a degree the table does not list: the first configuration for all three
    assert 9 not in table
    assert axhel_amd.config_selected(1 << 9) == 0 and axhel_amd.config_single(1 << 9) == 0
    assert axhel_amd.config_small(1 << 9) == 0


def test_missing_library_fails_loudly():
    """Without the built HIP library every entry point raises ImportError
    before computing anything: there is no CPU fallback.  Run in a child
    process so this process keeps its loaded library."""
    code = "\n".join([
        "import numpy as np, axhel_amd as ax",
        "ax.LIB_PATH, ax._lib = '/nonexistent/libaxhel_hip.so', None",
        "a = np.arange(8, dtype=np.uint64); r = np.zeros_like(a)",
        "calls = [lambda: ax.NTT(8, 17, 3), lambda: ax.EltwiseAddMod(r, a, a, 8, 17),",
        "         lambda: ax.EltwiseMulMod(r, a, a, 8, 17, 1), lambda: ax.RNS(8, [17])]",
        "for c in calls:",
        "    try:",
        "        c()",
        "    except ImportError as e:",
        "        assert 'is not built' in str(e)",
        "    else:",
        "        raise SystemExit('a call ran without the library')",
        "assert r.tolist() == [0] * 8",
        "print('all raised')",
    ])
    out = subprocess.run(["python", "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "all raised" in out.stdout, out.stdout + out.stderr
