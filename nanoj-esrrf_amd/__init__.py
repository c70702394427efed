"""eSRRF reconstruction on MI355X (gfx950): HIP kernels behind a C ABI.

The package directory is named ``nanoj-esrrf_amd``; load it with :func:`load` from
``__graft_entry__`` (or put the directory on ``sys.path`` and import ``_capi`` /
``liveSRRF_CL`` directly).

* ``libesrrf.so``      -- HIP kernels + engine (include/esrrf.h), built by build_lib.py
* ``liveSRRF_CL``      -- host mirror of the reference's LiveSRRF_CL on that ABI
* ``sweep``            -- the parameter sweep of gui/ParametersSweep_.java (fused on device)
* ``rse``              -- SQUIRREL RSE of the AVG image and its gradient (ErrorMapLiveSRRF)
* ``squirrel``         -- SQUIRREL error map / RSE / RSP of any (ref, SR) pair, QnR map
* ``rolling``          -- rolling analysis of long stacks (gui/LiveSRRF_optimised_.java)
* ``drift``            -- drift of SR windows for the vibration correction (XYShiftCalculator)
* ``synth``            -- seeded synthetic blinking stacks for the workloads
"""
from . import _capi  # noqa: F401
from .drift import XYShiftCalculator  # noqa: F401
from .liveSRRF_CL import Device, Engine, LiveSRRF_CL  # noqa: F401
from .rolling import LiveSRRF_optimised, Rolling  # noqa: F401
from .rse import RSE  # noqa: F401
from .squirrel import FRC, ErrorMapLiveSRRF, qnr_map, remove_nans  # noqa: F401
from .sweep import ParametersSweep, odd_even, sweep_axes  # noqa: F401
