"""The C-ABI library, checked without a GPU: it loads, exports every symbol
include/*.h declares, its flop counts equal the oracle's restatement of the
reference's closed forms, and the fatal argument checks of HCORE_dgemm
(reference hcore_dgemm.c:91-95, 743-746, 924-927) behave as in the reference
(message + exit(-1)) or, under HCORE_ON_ERROR=return, record a status."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT


def test_library_exports_every_header_symbol():
    import hcore_amd
    L = hcore_amd.lib()
    names = hcore_amd.header_functions()
    assert "HCORE_dgemm" in names and "HCORE_hip_plan_execute" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    for sym in ("use_scratch", "use_trmm", "gemm_print_index", "sgemm_use_trmm",
                "gemmcd_use_scratch", "hc_nelm_limit"):
        assert hasattr(L, sym), sym


def test_flop_counter_layout_matches_c():
    import ctypes
    import shutil
    import tempfile
    import hcore_amd
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    src = ('#include <stdio.h>\n#include <stddef.h>\n#include "flop_util_structs.h"\n'
           'int main(){printf("%zu %zu %zu\\n", sizeof(flop_counter), '
           'offsetof(flop_counter, update), offsetof(flop_counter, syrk));return 0;}\n')
    d = tempfile.mkdtemp()
    with open(os.path.join(d, "t.c"), "w") as f:
        f.write(src)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), os.path.join(d, "t.c"), "-o",
                    os.path.join(d, "t")], check=True)
    out = subprocess.run([os.path.join(d, "t")], capture_output=True, text=True).stdout.split()
    F = hcore_amd.FlopCounter
    assert int(out[0]) == ctypes.sizeof(F)
    assert int(out[1]) == F.update.offset and int(out[2]) == F.syrk.offset


def test_flop_counts_match_oracle(orc):
    import hcore_amd
    for op, args in [('q', (1024, 187)), ('q', (500, 64)), ('q', (1024, 17)), ('m', (80, 107, 1024)),
                     ('s', (187,)), ('o', (1024, 90, 187, 1)), ('o', (90, 1024, 187, 2)),
                     ('o', (500, 33, 64, 1)), ('o', (33, 500, 64, 2)), ('c', (512,)),
                     ('t', (512, 64, 1)), ('r', (64, 64, 2))]:
        a = list(args) + [0] * (4 - len(args))
        assert hcore_amd.flop_counts(op, *a) == orc.flop_counts(op, *a), (op, args)


_ERR_SCRIPT = r"""
import sys, numpy as np
sys.path.insert(0, %r)
import hcore_amd
M = 16
rng = np.random.default_rng(0)
AU, AV, BU, BV, CU, CV = (rng.standard_normal((M, 4)) for _ in range(6))
case = sys.argv[1]
if case == "zero":
    CU = CU[:, :0]; CV = CV[:, :0]; maxrk = M
elif case == "sum":
    maxrk = 3
else:
    maxrk = 6
try:
    hcore_amd.dgemm(-1.0, AU, AV, BU, BV, 1.0, CU, CV, maxrk=maxrk)
except Exception as e:
    print("EXC", e)
print("STATUS", hcore_amd.last_status())
"""


@pytest.mark.parametrize("case,status,msg", [("zero", 1, "should not be zero"),
                                             ("sum", 2, "Sum of ranks"),
                                             ("k", 3, "Increase maxrank")])
def test_fatal_argument_checks(case, status, msg):
    script = _ERR_SCRIPT % ROOT
    env = dict(os.environ)
    env.pop("HCORE_ON_ERROR", None)
    r = subprocess.run([sys.executable, "-c", script, case], capture_output=True, text=True,
                       env=env, timeout=120)
    assert r.returncode == 255, r.stdout + r.stderr  # exit(-1)
    assert msg in r.stderr
    env["HCORE_ON_ERROR"] = "return"
    r = subprocess.run([sys.executable, "-c", script, case], capture_output=True, text=True,
                       env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ("STATUS %d" % status) in r.stdout
    assert msg in r.stderr
