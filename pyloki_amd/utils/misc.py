"""Physical constants and small helpers shared by the host-side planning code.

The two constants are the values pyloki uses (``src/pyloki/utils/misc.py:37-39``);
``C_VAL`` is the CODATA speed of light that astropy reports, and ``FLOAT_EPSILON``
is the guard added before truncating a floating grid position to an index.
The gravitational constant and solar mass are the astropy (CODATA 2018 / IAU 2015)
values that pyloki reads through ``astropy.constants`` (``config.py:233-236,474-481``).
"""

from __future__ import annotations

C_VAL = 299792458.0
FLOAT_EPSILON = 1e-6

G_CONST = 6.6743e-11
M_SUN = 1.988409870698051e30
