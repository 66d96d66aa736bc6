"""The C-ABI driven from a C program on the GPU, as a cgo binding calls it:
PackUint32 of the reference's golden inputs gives the golden bytes,
UnpackUint32 and GetUint32 give the inputs back, PackUint64 of 1e12 + {0, 1,
2} takes 19 bytes and round-trips (tests/c_abi/abi_check.c)."""
import os
import subprocess
import sys

import pytest

from tests.test_capi_cpu import ROOT, abi_check_binary

pytestmark = pytest.mark.gpu


def test_c_caller_on_device():
    print("C caller of the C-ABI on device 0", flush=True)
    sys.stdout.flush()
    r = subprocess.run([abi_check_binary(), os.path.join(ROOT, "tests", "golden"), "--gpu"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host + device): 0 failures" in r.stdout
