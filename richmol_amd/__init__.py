"""richmol_amd: the field-driven rotational TDSE step of richmol on MI355X.

Public interface (mirrors richmol's):
  richmol_amd.tdse.TDSE         time grid, initial states, update(H, vecs, ...)
  richmol_amd.field.CarTens     tensor operator container (field, vec, tomat, +, *)
  richmol_amd.rot               rigid-rotor operator builders (H0, tensors, cos^2)
  richmol_amd.trove             reader of richmol-format (.rchm) tensor files
  richmol_amd.units             unit factors (convert_units)
"""
from . import units  # noqa: F401
from .field import CarTens  # noqa: F401
from .tdse import TDSE  # noqa: F401

__all__ = ["CarTens", "TDSE", "units"]
