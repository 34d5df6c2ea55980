"""The C-ABI library: loads, exports every declared symbol, and reports errors without a GPU."""

import ctypes
import re
from pathlib import Path

import numpy as np

from andfn_amd import _lib
from andfn_amd.dtypes import default_constants, make_element_dtype, make_fracture_dtype

HEADER = Path(__file__).resolve().parents[1] / "include" / "andfn_mi.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(andfn_mi_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared()
    assert "andfn_mi_solve" in names and "andfn_mi_last_error" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), (set(names) ^ set(_lib.SIGNATURES))


def test_options_and_info_layout_match_header():
    text = HEADER.read_text()

    def fields(struct):
        body = re.search(r"typedef struct " + struct + r" \{(.*?)\} " + struct, text, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        return re.findall(r"\b(?:const\s+)?(?:int32_t|int64_t|double)\s*\*?\s*([a-z_0-9]+);", body)

    assert fields("andfn_mi_options") == [f for f, _ in _lib.Options._fields_]
    assert fields("andfn_mi_info") == [f for f, _ in _lib.Info._fields_]
    assert fields("andfn_mi_objective") == [f for f, _ in _lib.Objective._fields_]
    assert fields("andfn_mi_adjoint_options") == [f for f, _ in _lib.AdjointOptions._fields_]
    assert fields("andfn_mi_adjoint_info") == [f for f, _ in _lib.AdjointInfo._fields_]


def test_no_device_or_bad_input_is_an_error_not_a_crash():
    lib = _lib.load()
    if lib.andfn_mi_device_count() > 0:
        return
    fa = np.zeros(1, dtype=make_fracture_dtype())
    ea = np.zeros(1, dtype=make_element_dtype())
    c = default_constants()
    rc = lib.andfn_mi_solve(_lib.ptr(fa), 1, 500, _lib.ptr(ea), 1, 200, 50, _lib.ptr(c), None, None, None)
    assert rc in (_lib.ENODEV, _lib.EINVAL)
    assert len(lib.andfn_mi_last_error()) > 0
    rc = lib.andfn_mi_solve(None, 1, 500, None, 1, 200, 50, None, None, None, None)
    assert rc == _lib.EINVAL
    for rc in (lib.andfn_mi_adjoint(None, None, None, None, None, None), lib.andfn_mi_set_parameters(None, None, None),
               lib.andfn_mi_state_heads(None, 0, None, None, None), lib.andfn_mi_state_discharges(None, None, None),
               lib.andfn_mi_state_heads_noalloc(None, 0, None, None, None),
               lib.andfn_mi_state_omega(None, 0, None, None, None),
               lib.andfn_mi_state_w(None, 0, None, None, -1, 1, None),
               lib.andfn_mi_state_streamlines(None, 0, None, None, None, None, 0.01, 0, 10, None, None, None),
               lib.andfn_mi_state_streamlines_fetch(None, None, None, None, None),
               lib.andfn_mi_w(None, 1, 500, None, 1, 200, 0, None, None, -1, 1, None, 0)):
        assert rc == _lib.EINVAL
