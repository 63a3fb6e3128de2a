"""MI355X-native multigrid pressure solve and field stencils of the SARAS solver.

The compute path is the gfx950 HIP library behind include/saras_mg.h (saras_amd/csrc); the Python
modules are thin bindings (``_lib``, ``mg``) plus the synthetic workloads of the benchmark."""
