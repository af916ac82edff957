"""2D-module twin of the glide + climb dynamics
(Plasticity/FieldDynamics/CentralUpwindHJBetaPDynamics.py:10-143).  The GPU
class handles 2D and 3D grids alike (the reference's 2D module asserts in
3D, CentralUpwindHJ.py:181-185)."""
from Plasticity.FieldDynamics._gpu_betap import GpuBetaPDynamics


class BetaPDynamics(GpuBetaPDynamics):
    pass
