"""The staged front end on the CPU: a staged program of this project's own
(tests/frontend/staged_program.cpp) is extracted (forks, loops recovered
from repeated statements) and printed; the printed function, compiled with
g++ against a host-only runtime header, must compute what the program
computes when its loops run directly (restated in Python below)."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FE = os.path.join(ROOT, "easy_graphit_amd", "frontend")
HERE = os.path.join(ROOT, "tests", "frontend")


def expected(n, m):
    i = 0
    while i < n:
        m = m + i if i % 3 == 0 else m - 1
        for _ in range(2):
            m = int((m * 2) - 1000003 * int((m * 2) / 1000003))  # C remainder (truncating)
        i += 1
    k = int(m - 7 * int(m / 7))
    while k > 0:
        m += k
        k -= 1
    return -m if m < 0 else m


@pytest.fixture(scope="module")
def staged(tmp_path_factory):
    d = tmp_path_factory.mktemp("fe")
    gen = str(d / "gen")
    srcs = [os.path.join(FE, "src", "staging.cpp"), os.path.join(FE, "src", "graphit.cpp")]
    subprocess.check_call(["g++", "-std=c++17", "-O0", f"-I{os.path.join(FE, 'include')}",
                           os.path.join(HERE, "staged_program.cpp"), *srcs, "-o", gen])
    text = subprocess.check_output([gen], text=True)
    src = d / "program.cpp"
    src.write_text(text)
    lib = str(d / "libprogram.so")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-shared", "-fPIC", f"-I{HERE}", str(src), "-o", lib])
    L = C.CDLL(lib)
    L.gi_gen_global.restype = C.c_void_p
    L.gi_gen_global.argtypes = [C.c_char_p]
    return text, L


def test_loops_are_recovered(staged):
    text, _ = staged
    body = text[text.index('extern "C" int program'):]
    # the outer loop, the counted inner loop once per outcome of the if before
    # it (what follows a fork is staged once per outcome), the loop whose
    # condition declares a variable (while (true) ... break)
    assert body.count("while (") == 4
    assert "goto" not in text and "break;" in body
    assert body.count("if (") >= 3


@pytest.mark.parametrize("n,m", [(0, 5), (1, -3), (4, 17), (9, 123456), (25, -999)])
def test_printed_program_computes_what_the_program_computes(staged, n, m):
    _, L = staged
    assert L.program(C.c_int(n), C.c_int(m)) == 0
    got = C.c_int.from_address(L.gi_gen_global(b"result")).value
    assert got == expected(n, m)
