"""hpmp: MI355X residual/relaxation path of hp-multiphysics (tri_hp / tri_hp_ins).

Import after putting ``hp-multiphysics_amd/`` on ``sys.path``; the compiled
library is ``hp-multiphysics_amd/lib/libhpgpu.so`` (see build.py).
"""
from . import mesh, workloads  # noqa: F401
from .input_map import InputMap  # noqa: F401
