"""The unipi::axhel C++ drop-in (libaxhel.so) driven by a C++ program with
host std::vector buffers, checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import axhel_amd as ax
from oracle import oracle as O
from tests.helpers import ntt_params

pytestmark = pytest.mark.gpu


CMAKE_CHECK_PATH = os.path.join(ax.LIB_DIR, "axhel_cxx_check_cmake")


@pytest.mark.parametrize("n,bits", [(4096, 50), (8, 62), (65536, 55)])
@pytest.mark.parametrize("exe", ["g++", "cmake"])
def test_cxx_dropin_program(n, bits, exe):
    """exe: the program linked by hand-written g++ flags, or built by CMake
    through find_package(AXHEL) and AXHEL::axhel (axhel_amd/cmake/example)."""
    q, psi = ntt_params(n, bits)
    mt = O.MT19937_64(0)
    a = mt.random_vector(n, q)
    b = mt.random_vector(n, q)
    stdin = f"{n} {q} {psi}\n" + " ".join(map(str, a.tolist())) + "\n" + " ".join(map(str, b.tolist())) + "\n"
    path = ax.CXX_CHECK_PATH if exe == "g++" else CMAKE_CHECK_PATH
    r = subprocess.run([path], input=stdin, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, *v = line.split()
        out[k] = np.array([int(x) for x in v], dtype=np.uint64)
    t = O.Tables(n, q, psi)
    assert out["kat"].tolist() == [1]
    fl = O.compute_forward(a, t, 1, 4)
    fn = O.compute_forward(a, t, 1, 1)
    assert np.array_equal(out["fwd_lazy"], fl)
    assert np.array_equal(out["fwd"], fn)
    assert np.array_equal(out["inv_lazy"], O.compute_inverse(O.reduce_mod(fl, q, 4, 2), t, 2, 2))
    assert np.array_equal(out["inv"], a)
    assert out["oop"].tolist() == [1] and out["ext"].tolist() == [1]
    assert np.array_equal(out["add"], O.add_mod(a, b, q))
    assert np.array_equal(out["sub"], O.sub_mod(a, b, q))
    assert np.array_equal(out["mul"], O.mul_mod(a, b, q, 1))
    assert np.array_equal(out["fma"], O.fma_mod(a, int(b[0]), b, q, 1))
    assert np.array_equal(out["reduce"], O.reduce_mod(b, q, q, 1))
