"""The KSP facade's option handling (CPU only): tolerances as PETSc reads them
(``rtol = 0`` honoured, ``None`` keeps the current value, negative values rejected) and
the field-split check that only the reference's elasticity split
(``mfe_elasticity_example_1.py:246-258``: stress | displacement + rotation, additive) is
accepted."""

import numpy as np
import pytest

from nfempy_amd.ksp import IS, KSP


def test_set_tolerances_keeps_zero_and_defaults():
    k = KSP().create()
    k.setTolerances(rtol=0.0, atol=1e-12, divtol=5000, max_it=20000)
    assert k.getTolerances() == (0.0, 1e-12, 5000.0, 20000)
    k.setTolerances(max_it=7)
    assert k.getTolerances() == (0.0, 1e-12, 5000.0, 7)
    with pytest.raises(ValueError):
        k.setTolerances(rtol=-1.0)
    with pytest.raises(ValueError):
        k.setTolerances(max_it=0)


class _Sys:
    field_bounds = [0, 10, 13, 16]


def _split(k, a, b, kind="additive"):
    k.getPC().setType("fieldsplit")
    k.getPC().setFieldSplitIS(("s", IS().createGeneral(a)), ("u", IS().createGeneral(b)))
    k.getPC().setFieldSplitType(kind)


def test_field_split_check():
    k = KSP()
    _split(k, np.arange(10), np.arange(10, 16))
    k._check_split(_Sys())
    k = KSP()
    _split(k, np.arange(13), np.arange(13, 16))
    with pytest.raises(NotImplementedError):
        k._check_split(_Sys())
    k = KSP()
    _split(k, np.arange(10), np.arange(10, 16), kind="multiplicative")
    with pytest.raises(NotImplementedError):
        k._check_split(_Sys())
    k = KSP()
    k.getPC().setType("fieldsplit")
    with pytest.raises(NotImplementedError):
        k._check_split(object())
