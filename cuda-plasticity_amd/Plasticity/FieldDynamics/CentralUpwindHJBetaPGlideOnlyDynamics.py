"""'lvp' glide-only BetaP dynamics
(Plasticity/FieldDynamics/CentralUpwindHJBetaPGlideOnlyDynamics.py:10-262):
the climb velocity is projected with q (rhov / rhorho), using the code's
plus sign (:119-121, 193-195)."""
from Plasticity.FieldDynamics._gpu_betap import GpuGlideOnlyDynamics


class BetaPDynamics(GpuGlideOnlyDynamics):
    pass
