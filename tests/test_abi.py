"""CPU checks of the C ABI: the library loads, exports every declared symbol, and the
MEX type checks raise the reference's error identifiers (no GPU call is made)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mbrec.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mbrec_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import mbrec
    L = ctypes.CDLL(mbrec.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from mbrec._lib import SIGNATURES
    assert sorted(n for n, _, _ in SIGNATURES) == declared_functions()


def test_version():
    import mbrec
    assert mbrec.lib().mbrec_version() >= 1


@pytest.mark.parametrize("fn,args,pos,eid,msg", [
    ("forwardDelayAndSumMex", lambda: (np.ones((1, 4), np.float32), np.ones((4, 2)), np.int32(8), 2.0), 0,
     "forwardMex:notDouble", "image must be type double."),
    ("forwardDelayAndSumMex", lambda: (np.ones((1, 4)), np.ones((4, 2)), 8.0, 2.0), 2,
     "forwardMex:notInt32", "numTimesteps must be type int32."),
    ("transposeDelayAndSumMex", lambda: (np.ones((1, 8, 2)), np.ones((4, 2), np.int32), 2.0), 1,
     "forwardMex:notDouble", "travelTimes must be type double."),
    ("forwardMex", lambda: (np.ones((1, 4)), np.ones((4, 2)), np.ones((4, 2)), np.ones((3, 1)), np.int32(8),
                            np.ones((2, 1), np.int32), 2.0), 1,
     "forwardMex:notInt32", "sirIndices must be type int32."),
    ("transposeMex", lambda: (np.ones((1, 8, 2)), np.ones((4, 2), np.int32), np.ones((4, 2)), np.ones((3, 1)),
                              np.ones((2, 1)), 2.0), 4,
     "forwardMex:notInt32", "sirTimestepsStartAndEnd must be type int32."),
    ("transposeMex", lambda: (np.ones((1, 8, 2)), np.ones((4, 2), np.int32), np.ones((4, 2)), np.ones((3, 1)),
                              np.ones((2, 1), np.int32), np.float32(2.0)), 5,
     "forwardMex:notDouble", "circularInterpolationRadiusInSamples must be type double."),
])
def test_mex_type_errors_use_reference_ids(fn, args, pos, eid, msg):
    from mbrec import mex
    with pytest.raises(mex.MexError) as ei:
        getattr(mex, fn)(*args())
    assert ei.value.identifier == eid
    assert msg in str(ei.value)


def test_plan_without_device_fails_loudly():
    """On a machine without a GPU the device API reports an error instead of computing."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import mbrec
    from mbrec import _lib
    with pytest.raises(_lib.MbrecError):
        _lib.check(mbrec.lib().mbrec_init(0))
