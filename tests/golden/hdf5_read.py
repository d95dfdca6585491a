"""The package's HDF5 parser (strata_fdtd_amd/hdf5_in.py), under the name the
fixture scripts and tests beside it use."""

from strata_fdtd_amd.hdf5_in import *  # noqa: F401,F403
from strata_fdtd_amd.hdf5_in import UNDEF, _u  # noqa: F401
