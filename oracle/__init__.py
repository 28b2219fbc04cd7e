"""FP64 CPU oracle for the field-driven rotational TDSE step.

TEST INFRASTRUCTURE ONLY.  This package is the checker: a plain numpy/scipy
restatement of the reference algorithm (richmol/tdse.py, richmol/field.py).
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only to compute expected values.  The product
package ``richmol_amd`` never imports it and never falls back to it.

Modules
-------
units    frozen CODATA 2022 constants and the unit factors of
         richmol/convert_units.py (identical literals to richmol_amd.units).
cartens  block tensor container: field(), vec(), tomat()/full_form(),
         mul(), add_cartens()  (richmol/field.py:449-1301).
tdse     TDSE (time_grid, init_state, update) and the in-house Lanczos
         expv (richmol/tdse.py:11-486).

Pinning (what fixes each function independently of retyping it): see the
tests ``tests/test_oracle_*.py``; every function listed above is checked
against a golden vector from the reference tree, a closed form, or a brute
force dense computation.  Parity unpinned: none of the functions above; the
Expokit path (``propag='external'``) is not restated.
"""
