"""C-ABI library checks that need no GPU: it loads, and it exports every
function include/spira_amd.h declares (and the header's constants agree with
the Python mirror).  No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

import spira_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spira_amd.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(spira_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def _defines():
    out = {}
    for m in re.finditer(r"#define\s+(SPIRA_\w+)\s+\(?(-?\d+)u?\)?", open(HEADER).read()):
        out[m.group(1)] = int(m.group(2))
    return out


def test_header_declares_the_boundary():
    names = _declared_functions()
    for must in ["spira_dev_create", "spira_dcsr_upload_f64", "spira_dcsr_download_f64", "spira_lock_f64",
                 "spira_spmv_f64"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = spira_amd.lib()
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_exported_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", spira_amd.lib_path()], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in _declared_functions():
        assert n in exported, n


def test_constants_match_python_mirror():
    d = _defines()
    assert d["SPIRA_OK"] == spira_amd.OK
    assert d["SPIRA_ERR_INVALID_ARGUMENT"] == spira_amd.ERR_INVALID_ARGUMENT
    assert d["SPIRA_ERR_LOGIC"] == spira_amd.ERR_LOGIC
    assert d["SPIRA_ERR_OUT_OF_RANGE"] == spira_amd.ERR_OUT_OF_RANGE
    assert d["SPIRA_ERR_UNSUPPORTED"] == spira_amd.ERR_UNSUPPORTED
    assert d["SPIRA_X_DEVICE"] == spira_amd.X_DEVICE and d["SPIRA_Y_DEVICE"] == spira_amd.Y_DEVICE
    assert d["SPIRA_IN_DEVICE"] == spira_amd.IN_DEVICE and d["SPIRA_NO_SYNC"] == spira_amd.NO_SYNC
    assert d["SPIRA_SPMV_N_VARIANTS"] == spira_amd.N_VARIANTS
    assert d["SPIRA_NOALLOC_SCRATCH_BYTES"] == spira_amd.NOALLOC_SCRATCH_BYTES
    for name, v in spira_amd.VARIANTS.items():
        key = "SPIRA_SPMV_" + name.upper().replace("STREAM1K", "STREAM1K")
        assert d[key] == v, key
    assert d["SPIRA_BATCH_XLDS_FLOATS"] == spira_amd.BATCH_XLDS_FLOATS
    assert d["SPIRA_BATCH_XLDS_TWO_PASS_COLS"] == spira_amd.BATCH_XLDS_TWO_PASS_COLS
    assert d["SPIRA_BATCH_MAX_PLOG2"] == spira_amd.BATCH_MAX_PLOG2
    assert d["SPIRA_BATCH_STREAM"] == spira_amd.BATCH_STREAM


def test_batch_schedule_struct_layout():
    """spira_batch_sched is {u8 plog2, u8 xlds, u16 flags, u32 item_nnz}:
    8 bytes, matched by the Python record type."""
    src = open(HEADER).read()
    body = re.search(r"typedef struct spira_batch_sched \{(.*?)\} spira_batch_sched;", src, re.S).group(1)
    fields = re.findall(r"(uint\d+_t)\s+(\w+);", body)
    assert fields == [("uint8_t", "plog2"), ("uint8_t", "xlds"), ("uint16_t", "flags"), ("uint32_t", "item_nnz")]
    dt = spira_amd.BATCH_SCHED_DTYPE
    assert dt.itemsize == 8 and list(dt.names) == [f for _, f in fields]
    assert [dt.fields[n][1] for n in dt.names] == [0, 1, 2, 4]


def test_null_arguments_are_rejected_without_a_device():
    lib = spira_amd.lib()
    assert lib.spira_dev_create(None, 0, None) == spira_amd.ERR_INVALID_ARGUMENT
    assert lib.spira_spmv_f64(None, None, 0, None, 0, -1, 0) == spira_amd.ERR_INVALID_ARGUMENT
    assert lib.spira_lock_f64(None, None, None, None, None, 0, None) == spira_amd.ERR_INVALID_ARGUMENT
    assert lib.spira_dcsr_info(None, None, None, None, None) == spira_amd.ERR_INVALID_ARGUMENT


def test_device_create_fails_cleanly_without_gpu():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = ctypes.c_void_p()
    rc = spira_amd.lib().spira_dev_create(ctypes.byref(h), 0, None)
    assert rc in (spira_amd.ERR_INVALID_ARGUMENT, spira_amd.ERR_DEVICE)
    assert spira_amd.lib().spira_dev_last_error(None)
    assert not h.value


def test_python_mirror_validates_before_device_calls():
    with pytest.raises(spira_amd.InvalidArgument):
        spira_amd.ParallelMatrix.__init__(object.__new__(spira_amd.ParallelMatrix), 4, 4, 0)
