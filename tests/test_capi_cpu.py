"""CPU-only checks of the C-ABI library: it loads, exports every symbol the
header declares, and its host-side entry points (Header, BlockLength,
MaxBlockLength: a few-byte header read each, no device work) agree with the
oracle. No compute entry point is called without a GPU: the one device
entry point called here returns on its argument checks, before any device
work."""
import os
import re
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(utl_[a-z0-9_]+)\s*\(", text))
    return names


@pytest.fixture(scope="module")
def lib():
    from utlpfor_amd import _lib
    return _lib.load()


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 30
    from utlpfor_amd import _lib
    assert names <= set(_lib.SIGNATURES), names - set(_lib.SIGNATURES)
    for n in names:
        assert hasattr(lib, n), n


def _blocks(orc):
    rng = np.random.default_rng(11)
    out = []
    for i in range(60):
        n = int(rng.integers(1, 129))
        v = rng.integers(0, 1 << int(rng.integers(0, 33)), n, dtype=np.uint64).astype(np.uint32)
        out.append(orc.PackUint32(i % 16, v)[0])
        out.append(orc.PackUint64(i % 16, rng.integers(0, 1 << 63, n, dtype=np.uint64))[0])
    return out


def test_host_entry_points_match_oracle(lib, orc):
    from utlpfor_amd import codec
    for flag in range(64):
        assert codec.MaxBlockLength32(flag) == orc.MaxBlockLength32(flag)
        assert codec.MaxBlockLength64(flag) == orc.MaxBlockLength64(flag)
    rng = np.random.default_rng(3)
    samples = _blocks(orc)
    for _ in range(400):
        samples.append(rng.integers(0, 256, int(rng.integers(0, 12)), dtype=np.uint8).tobytes())
    for blk in samples:
        for cut in {len(blk), min(len(blk), 5), min(len(blk), 3)}:
            b = blk[:cut]
            want = orc.Header(b)
            got = codec.Header(b)
            if want[-1] == 0:
                assert got[:-1] == want[:-1] and got[-1] is None
            else:
                assert got[-1].code == want[-1]
            wl, we = orc.BlockLength(b)
            gl, ge = codec.BlockLength(b)
            assert (ge.code if ge else 0) == we and (gl == wl or we)


def test_walk_rejects_bad_arguments_before_device_work(lib):
    """utl_index_walk checks its arguments before it touches a device: a
    missing required pointer, block starts without meta words, an unaligned
    source and more than 2^32 - 1 streams are UTL_ERR_ARGUMENT. The fake
    pointers are never dereferenced."""
    fake = 0x100000  # 16-byte aligned
    err = 100  # UTL_ERR_ARGUMENT
    assert lib.utl_index_walk(None, None, fake, 1, None, None, None, fake, fake) == err
    assert lib.utl_index_walk(None, fake, fake, 1, None, None, fake, fake, fake) == err
    assert lib.utl_index_walk(None, fake + 4, fake, 1, None, None, None, fake, fake) == err
    assert lib.utl_index_walk(None, fake, fake, 1 << 32, None, None, None, fake, fake) == err


def test_get_dir_rejects_bad_arguments_before_device_work(lib):
    """utl_pos_dir_build and utl_get_batch_dir check their pointers before
    they touch a device (UTL_ERR_ARGUMENT); the fake pointers are never
    dereferenced."""
    fake = 0x100000
    err = 100  # UTL_ERR_ARGUMENT
    assert lib.utl_pos_dir_build(None, None, fake, 4, fake, 1) == err
    assert lib.utl_pos_dir_build(None, fake, None, 4, fake, 1) == err
    assert lib.utl_pos_dir_build(None, fake, fake, 4, None, 1) == err
    assert lib.utl_pos_dir_build(None, fake, fake, 0, fake, 1) == err
    assert lib.utl_get_batch_dir(None, 0, fake, fake, fake, 4, None, 1, fake, 1, fake, fake) == err
    assert lib.utl_get_batch_dir(None, 0, fake, fake, fake, 4, fake, 0, fake, 1, fake, fake) == err
    assert lib.utl_get_batch_dir(None, 0, fake, fake, fake, 4, fake, 1, None, 1, fake, fake) == err


def test_no_cpu_fallback_without_library(tmp_path, monkeypatch):
    from utlpfor_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError):
        _lib.load()


def abi_check_binary():
    """tests/c_abi/abi_check, built by build() (make) when it is missing."""
    import subprocess
    path = os.path.join(ROOT, "tests", "c_abi", "abi_check")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", os.path.dirname(path)], check=True, capture_output=True)
    return path


def test_c_caller_host_entry_points():
    """The header compiles as C99 with warnings as errors and a C program
    linked against the library gets the reference's MaxBlockLength known
    answers and the golden blocks' Header and BlockLength (host reads)."""
    import subprocess
    r = subprocess.run([abi_check_binary(), os.path.join(ROOT, "tests", "golden")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


def _c_prototypes():
    """name -> parameter count of every function declared in include/*.h."""
    protos = {}
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", fn)).read(), flags=re.S)
        for m in re.finditer(r"\b(utl_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
            params = m.group(2).strip()
            protos[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return protos


def _call_args(src, start):
    """Top-level argument count of the call whose '(' is at src[start]."""
    depth, n, empty = 0, 1, True
    for ch in src[start:]:
        if ch in "([{":
            depth += 1
            if depth == 1:
                continue
        elif ch in ")]}":
            depth -= 1
            if depth == 0:
                return 0 if empty else n
        elif ch == "," and depth == 1:
            n += 1
        if depth >= 1 and not ch.isspace():
            empty = False
    raise AssertionError("unbalanced call")


def test_go_binding_matches_the_header():
    """The cgo package in go/rocm (no Go toolchain here, nor on the GPU
    machines, to build it) calls only functions the header declares, each
    with the declared number of arguments, and uses only macros the header
    defines."""
    import glob
    go = "\n".join(open(p).read() for p in sorted(glob.glob(os.path.join(ROOT, "go", "rocm", "*.go")))
                   if not p.endswith("_test.go"))
    assert "import \"C\"" in go
    protos = _c_prototypes()
    calls = [(m.group(1), _call_args(go, m.end() - 1)) for m in re.finditer(r"\bC\.(utl_[a-z0-9_]+)\(", go)]
    assert len(calls) >= 10
    for name, nargs in calls:
        assert name in protos, name
        assert nargs == protos[name], (name, nargs, protos[name])
    header = open(os.path.join(ROOT, "include", "utlpfor_gpu.h")).read()
    for macro in set(re.findall(r"\bC\.(UTL_[A-Z0-9_]+)", go)):
        assert re.search(rf"#define {macro}\b", header), macro
