"""The C ABI: symbols, struct layout and error handling (no GPU needed).

Every entry point declared in include/*.h must be exported by the built
library; the ctypes mirrors must match the C layout (checked against a
probe compiled with gcc from the header); invalid descriptors must be
rejected with SSMX_EINVAL before any HIP call is made.
"""

import ctypes
import glob
import os
import re
import subprocess

import pytest

from superspinsim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(ssmx_\w+)\s*\(",
                            text, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        from superspinsim_amd.build import build_library
        build_library()
    return _lib.load()


def test_every_declared_symbol_is_exported(lib):
    declared = _declared()
    assert "ssmx_propagate" in declared
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.EXPORTS)


def test_struct_layout_matches_header(tmp_path):
    probe = tmp_path / "probe.c"
    probe.write_text(r'''
#include <stddef.h>
#include <stdio.h>
#include "ssmx.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n",
         sizeof(ssmx_propagate_desc),
         offsetof(ssmx_propagate_desc, coef_inst_stride),
         offsetof(ssmx_propagate_desc, h), offsetof(ssmx_propagate_desc, dt),
         sizeof(ssmx_segment), sizeof(ssmx_wave),
         offsetof(ssmx_wave, seg_count),
         offsetof(ssmx_propagate_desc, ckpt_every),
         sizeof(ssmx_adjoint_desc),
         offsetof(ssmx_adjoint_desc, grad_gen),
         offsetof(ssmx_adjoint_desc, coef_inst_stride),
         offsetof(ssmx_adjoint_desc, dt), sizeof(ssmx_dense_desc),
         offsetof(ssmx_dense_desc, coef_inst_stride),
         offsetof(ssmx_dense_desc, scale));
  return 0;
}
''')
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o",
                    str(exe), str(probe)], check=True)
    got = [int(v) for v in subprocess.run(
        [str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    D = _lib.PropagateDesc
    A = _lib.AdjointDesc
    mine = [ctypes.sizeof(D), D.coef_inst_stride.offset, D.h.offset,
            D.dt.offset, ctypes.sizeof(_lib.Segment),
            ctypes.sizeof(_lib.Wave), _lib.Wave.seg_count.offset,
            D.ckpt_every.offset, ctypes.sizeof(A), A.grad_gen.offset,
            A.coef_inst_stride.offset, A.dt.offset,
            ctypes.sizeof(_lib.DenseDesc),
            _lib.DenseDesc.coef_inst_stride.offset,
            _lib.DenseDesc.scale.offset]
    assert got == mine


def _desc(**kw):
    base = dict(n=4, n_inst=1, n_gen=1, slots=2, n_exp=1, n_node=1, fine=1,
                n_out=1, taylor=4, substeps=1, n_obs=1,
                flags=_lib.SSMX_READ_FULL, first_out=0, partner=1,
                coef_inst_stride=0, h=1e-11, t0=0.0, dt=1e-11)
    base.update(kw)
    return _lib.PropagateDesc(**base)


@pytest.mark.parametrize("bad", [
    dict(n=0), dict(n=513), dict(n=100, slots=33),
    dict(n=441, n_gen=8, slots=14), dict(n_inst=0), dict(n_gen=9), dict(slots=0),
    dict(n_exp=7), dict(n_node=4), dict(taylor=0), dict(taylor=41),
    dict(substeps=3), dict(h=0.0), dict(first_out=-1), dict(partner=2),
    dict(ckpt_every=-1),
    dict(flags=0, n_obs=3),
    dict(flags=_lib.SSMX_READ_FULL | _lib.SSMX_READ_INTEGRAL),
    dict(n_exp=6, n_node=3, n_gen=8, flags=_lib.SSMX_ROTATING |
         _lib.SSMX_READ_FULL),
])
def test_invalid_descriptor_is_rejected(lib, bad):
    d = _desc(**bad)
    dummy = ctypes.c_void_p(8)
    rc = lib.ssmx_propagate(ctypes.byref(d), *([dummy] * 13))
    assert rc == -1
    assert lib.ssmx_last_error().decode()


def test_null_pointers_are_rejected(lib):
    d = _desc()
    rc = lib.ssmx_propagate(ctypes.byref(d), *([None] * 13))
    assert rc == -1
    assert b"null" in lib.ssmx_last_error()
    rc = lib.ssmx_propagate(None, *([None] * 13))
    assert rc == -1


def _adesc(**kw):
    base = dict(n=4, n_inst=1, n_gen=2, slots=2, slots_t=3, n_exp=2,
                n_node=2, fine=1, n_out=1, taylor=8, taylor_adj=6, n_obs=1,
                flags=_lib.SSMX_COEF_WAVE, first_out=0, partner=1,
                ckpt_every=10, n_theta=1, theta_gen=0, grad_gen=1,
                reserved=0, coef_inst_stride=0, h=1e-11, t0=0.0, dt=1e-11)
    base.update(kw)
    return _lib.AdjointDesc(**base)


def test_adjoint_workspace_size(lib):
    d = _adesc(n_inst=3, ckpt_every=7, n_exp=2)
    out = ctypes.c_uint64(0)
    assert lib.ssmx_adjoint_workspace_bytes(ctypes.byref(d),
                                            ctypes.byref(out)) == 0
    assert out.value == 3 * 2 * 7 * 2 * 64 * 8
    # n > 64: the inputs of every exponential of one segment, stride NP
    d = _adesc(n=441, n_inst=3, ckpt_every=7, n_exp=2, grad_gen=-1,
               reserved=2)
    assert lib.ssmx_adjoint_workspace_bytes(ctypes.byref(d),
                                            ctypes.byref(out)) == 0
    assert out.value == 3 * 7 * 2 * 448 * 8


@pytest.mark.parametrize("bad", [
    dict(n=513), dict(n=100, grad_gen=1), dict(n=100, grad_gen=-1,
                                                  partner=0),
    dict(n=100, grad_gen=-1, reserved=3), dict(reserved=1),
    dict(n_inst=0), dict(slots=33), dict(slots_t=0),
    dict(taylor_adj=11), dict(taylor_adj=0), dict(ckpt_every=0),
    dict(n_theta=17), dict(theta_gen=2), dict(theta_gen=-1),
    dict(grad_gen=2), dict(flags=_lib.SSMX_ROTATING),
    dict(flags=0, grad_gen=1), dict(n_obs=3), dict(n_out=0),
    dict(n_node=4), dict(n_exp=7),
])
def test_invalid_adjoint_descriptor_is_rejected(lib, bad):
    d = _adesc(**bad)
    dummy = ctypes.c_void_p(8)
    rc = lib.ssmx_adjoint(ctypes.byref(d), *([dummy] * 17))
    assert rc == -1
    assert lib.ssmx_last_error().decode()
    out = ctypes.c_uint64(0)
    assert lib.ssmx_adjoint_workspace_bytes(ctypes.byref(d),
                                            ctypes.byref(out)) == -1


def test_adjoint_null_pointers_are_rejected(lib):
    d = _adesc()
    rc = lib.ssmx_adjoint(ctypes.byref(d), *([None] * 17))
    assert rc == -1
    assert b"null" in lib.ssmx_last_error()


def test_version(lib):
    assert lib.ssmx_version() >= 100


def test_one_hip_runtime_whatever_the_import_order():
    """Loading the library before anything imports torch must not map a
    second HIP runtime next to torch's (launches would find no device)."""
    import sys
    code = ("import re, sys; sys.path.insert(0, %r)\n"
            "from superspinsim_amd import _lib\n_lib.load()\nimport torch\n"
            "maps = open('/proc/self/maps').read()\n"
            "print(sorted(set(re.findall(r'\\S*libamdhip64\\S*', maps))))"
            % ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True,
                         text=True, check=True).stdout.strip()
    assert out.count("libamdhip64") == 1, out


def _ddesc(**kw):
    base = dict(n=4, n_inst=1, n_gen=1, n_exp=1, n_node=1, fine=1, n_out=1,
                squarings=70, flags=0, n_doubles=0, n_singles=0, reserved=0,
                coef_inst_stride=0, h=1e-11, scale=4.0 ** 35)
    base.update(kw)
    return _lib.DenseDesc(**base)


@pytest.mark.parametrize("bad", [
    dict(n=0), dict(n=65), dict(n_inst=0), dict(n_gen=9), dict(n_exp=7),
    dict(n_node=4), dict(fine=0), dict(n_out=-1), dict(squarings=-1),
    dict(scale=0.0), dict(flags=4), dict(n_doubles=1),
    dict(flags=_lib.SSMX_ROTATING, n_doubles=2, n_singles=1),
    dict(reserved=1), dict(h=0.0), dict(coef_inst_stride=-1),
])
def test_invalid_dense_descriptor_is_rejected(lib, bad):
    d = _ddesc(**bad)
    dummy = ctypes.c_void_p(8)
    rc = lib.ssmx_propagate_dense(ctypes.byref(d), *([dummy] * 7))
    assert rc == -1
    assert lib.ssmx_last_error().decode()


def test_dense_null_pointers_are_rejected(lib):
    d = _ddesc()
    assert lib.ssmx_propagate_dense(ctypes.byref(d), *([None] * 7)) == -1
    assert b"null" in lib.ssmx_last_error()


def _cdesc(**kw):
    base = dict(n=9, n_inst=2, n_chunks=3, chunk_out=4, n_out=10, pack=7,
                n_obs=1, reserved=0)
    base.update(kw)
    return _lib.ChunkDesc(**base)


@pytest.mark.parametrize("bad", [
    dict(n=0), dict(n=65), dict(n_inst=0), dict(n_chunks=0),
    dict(chunk_out=0), dict(n_out=13), dict(n_out=8), dict(pack=0),
    dict(pack=8), dict(n_obs=3), dict(n_obs=-1), dict(reserved=1),
])
def test_invalid_chunk_descriptor_is_rejected(lib, bad):
    d = _cdesc(**bad)
    dummy = ctypes.c_void_p(8)
    assert lib.ssmx_chunk_combine(ctypes.byref(d), *([dummy] * 5)) == -1
    assert lib.ssmx_last_error().decode()


def test_chunk_null_pointers_are_rejected(lib):
    d = _cdesc()
    assert lib.ssmx_chunk_combine(ctypes.byref(d), *([None] * 5)) == -1
    assert b"null" in lib.ssmx_last_error()
