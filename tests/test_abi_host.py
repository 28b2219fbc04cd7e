"""CPU checks of the C ABI library and of the host-side plan flattening.
No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from _fixtures import linear_rotor, ocs_fixture
from oracle.cartens import OracleCarTens

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    src = open(os.path.join(ROOT, "include", "richmol_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rm_[a-z_0-9]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    from richmol_amd import _build, _lib
    if not os.path.exists(_build.LIB_PATH):
        _build.build()
    return _lib.load()


def test_library_exports_every_header_symbol(lib):
    from richmol_amd import _lib
    names = _header_functions()
    assert len(names) >= 9
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    src = open(os.path.join(ROOT, "include", "richmol_amd.h")).read()
    assert lib.rm_version() == int(re.search(r"#define RM_ABI_VERSION (\d+)", src).group(1))


def test_plan_create_rejects_bad_indices(lib):
    from richmol_amd import _lib
    i64, i32, dbl = ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    row_ptr = np.array([0, 1, 2], dtype=np.int64)
    col = np.array([0, 5], dtype=np.int32)           # 5 >= n = 2
    pair_ptr = np.array([0, 1, 2], dtype=np.int64)
    pair_term = np.array([0, 0], dtype=np.int32)
    kval = np.zeros(4)
    term_ptr = np.array([0, 1], dtype=np.int64)
    term_field = np.array([0], dtype=np.int32)
    mval = np.zeros(2)
    term_leaf = np.array([0], dtype=np.int32)
    h = ctypes.c_void_p()
    rc = lib.rm_plan_create(2, 2, _lib.ptr(row_ptr, i64), _lib.ptr(col, i32), _lib.ptr(pair_ptr, i64),
                            _lib.ptr(pair_term, i32), _lib.ptr(kval, dbl), 1, _lib.ptr(term_ptr, i64),
                            _lib.ptr(term_field, i32), _lib.ptr(mval, dbl), _lib.ptr(term_leaf, i32),
                            1, 1, ctypes.byref(h))
    assert rc == -1
    assert b"col_idx out of range" in lib.rm_last_error()
    assert not h.value
    rc = lib.rm_plan_create(2, 2, _lib.ptr(row_ptr, i64), _lib.ptr(np.array([0, 1], np.int32), i32),
                            _lib.ptr(pair_ptr, i64), _lib.ptr(pair_term, i32), _lib.ptr(kval, dbl), 1,
                            _lib.ptr(term_ptr, i64), _lib.ptr(term_field, i32), _lib.ptr(mval, dbl),
                            _lib.ptr(term_leaf, i32), 40, 1, ctypes.byref(h))
    assert rc == -1 and b"nfield" in lib.rm_last_error()


def test_entry_points_reject_null_plan(lib):
    assert lib.rm_step(None, None, None, 1, 1, 0.0, -1.0, None, 0, 100, 1e-15, None, None, None) == -1
    assert b"NULL plan" in lib.rm_last_error()
    assert lib.rm_matvec(None, None, None, 1, 1, None) == -1
    assert lib.rm_expect(None, None, 1, 1, None, None) == -1
    assert lib.rm_plan_set_field(None, None, None, None) == -1


def _host_matrix(arr, fprod, thresh):
    """Evaluate the flattened plan on the host: MF_t, then A_e, as dense N x N."""
    n = arr["n"]
    mf = np.zeros(arr["nterm"], dtype=np.complex128)
    for t in range(arr["nterm"]):
        q0, q1 = arr["term_ptr"][t], arr["term_ptr"][t + 1]
        v = np.sum(fprod[arr["term_field"][q0:q1]] * arr["term_mval"][q0:q1])
        thr = thresh[arr["term_leaf"][t]]
        mf[t] = 0 if (thr > 0 and abs(v) < thr) else v
    A = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        for e in range(arr["row_ptr"][i], arr["row_ptr"][i + 1]):
            p0, p1 = arr["pair_ptr"][e], arr["pair_ptr"][e + 1]
            A[i, arr["col_idx"][e]] = np.sum(arr["pair_kval"][p0:p1] * mf[arr["pair_term"][p0:p1]])
    return A


@pytest.mark.parametrize("F,thr", [([0, 0, 1.2e9], None), ([3e8, -2e8, 9e8], None), ([3e8, 0, 9e8], 3e16)])
def test_flattened_plan_equals_full_operator(F, thr):
    """flatten_operator + the device formulas (evaluated on the host) rebuild
    the reference's full-form matrix of the field-contracted operator."""
    from richmol_amd.plan import DevicePlan, flatten_operator
    h0, H, cos2, bas = linear_rotor(5)
    H.field(F, thresh=thr)
    S = h0 + H
    arr = flatten_operator(S)
    fprod, thresh = DevicePlan.field_arrays(type("P", (), {"nfield": arr["nfield"], "nleaf": arr["nleaf"],
                                                            "slots": arr["slots"]})(), S)
    A = _host_matrix(arr, fprod, thresh)
    O = OracleCarTens.from_attrs(h0) + _oracle_field(H, F, thr)
    ref = O.tomat(form="full", repres="dense")
    assert np.abs(A - ref).max() <= 1e-13 * np.abs(ref).max()


def _oracle_field(H, F, thr):
    O = OracleCarTens.from_attrs(H)
    O.field(F, thresh=thr)
    return O


def test_flattened_plan_ocs_fixture():
    from richmol_amd.plan import DevicePlan, flatten_operator
    from richmol_amd import units
    h0, a, d = ocs_fixture()
    H = a * (-0.5 * units.AUpol_x_Vm_to_invcm())
    H.field([0, 0, 3e9])
    arr = flatten_operator(H)
    assert arr["n"] == 16
    fprod, thresh = DevicePlan.field_arrays(type("P", (), {"nfield": arr["nfield"], "nleaf": arr["nleaf"],
                                                            "slots": arr["slots"]})(), H)
    A = _host_matrix(arr, fprod, thresh)
    O = OracleCarTens.from_attrs(a) * (-0.5 * units.AUpol_x_Vm_to_invcm())
    O.field([0, 0, 3e9])
    ref = O.tomat(form="full", repres="dense")
    assert np.abs(A - ref).max() <= 1e-13 * np.abs(ref).max()


def test_job_arrays_keep_the_rm_job_layout(lib):
    """Job records concatenated for rm_batch_propagate keep rm_job's padding
    (numpy's structured promotion packs them to 96 bytes, which shifts every
    pointer after `len` -- the C side would read garbage device addresses);
    a packed array is refused in Python, and the library refuses any record
    size other than sizeof(rm_job) before reading a job."""
    from richmol_amd import serving
    recs = [np.zeros(1, dtype=serving.JOB) for _ in range(3)]
    for i, r in enumerate(recs):
        r["fprod_dev"] = 0x1000 * (i + 1)
        r["nsteps"] = 7 + i
    jobs = serving.job_array(recs)
    assert jobs.dtype == serving.JOB and jobs.dtype.itemsize == ctypes.sizeof(serving._lib.RmJob)
    raw = jobs.tobytes()
    for i in range(3):
        assert int.from_bytes(raw[104 * i + 40:104 * i + 48], "little") == 0x1000 * (i + 1)
        assert int.from_bytes(raw[104 * i + 56:104 * i + 60], "little") == 7 + i
    packed = np.concatenate(recs)
    with pytest.raises(TypeError):
        serving.batch_propagate(lib, None, packed, 12, 100, 1e-15, 0, 0)
    h = ctypes.c_void_p()
    assert lib.rm_batch_create(ctypes.byref(h)) == 0
    assert lib.rm_batch_propagate(h, jobs.ctypes.data, 3, 96, 12, 100, 1e-15, 0, None) == -1
    assert b"bytes" in lib.rm_last_error()
    assert lib.rm_batch_destroy(h) == 0


def test_rm_job_layout_matches_header(tmp_path):
    """The ctypes mirror of rm_job (used to fill job arrays from numpy) has the
    C struct's field offsets and size (compiled with gcc from the header)."""
    import subprocess
    from richmol_amd._lib import RmJob
    names = [n for n, _ in RmJob._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'  printf("{n} %zu\\n", offsetof(rm_job, {n}));\n' for n in names)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "richmol_amd.h"\n'
                   'int main(void) {\n' + body + '  printf("sizeof %zu\\n", sizeof(rm_job));\n'
                   '  return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for n in names:
        assert int(out[n]) == getattr(RmJob, n).offset, n
    assert int(out["sizeof"]) == ctypes.sizeof(RmJob)
