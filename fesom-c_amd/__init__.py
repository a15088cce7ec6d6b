"""fesomc_amd: MI355X-native FESOM-C internal time step (host-side mirror).

The package directory is ``fesom-c_amd/``; load it with :func:`load` from
``__graft_entry__`` (or any importlib loader registering it as
``fesomc_amd``).  Submodules:

* ``meshio``    -- FESOM-C mesh/forcing file formats.
* ``runconfig`` -- ``fv_run.dat`` reader (FVCOM-style ``KEY = value``).
* ``meshgen``   -- seeded synthetic channel meshes for the benchmark configs.
* ``gpu``       -- ctypes binding of ``libfesomc_gpu.so`` (the C-ABI).
"""
