"""The woJzz variant of the 2D glide + climb dynamics
(Plasticity/FieldDynamics/CentralUpwindHJBetaPDynamics_woJzz.py:10-135):
climb suppression with tr(sigma)/2 along (ux_yy - uy_xy, uy_xx - ux_yx,
-(ux_zx + uy_zy)), its J term on H_xx and H_yy only, and beta_zz held fixed
(rhs_zz = 0).  2D only, as in the reference; runs on the GPU stepper."""
from Plasticity.amd import lib as L
from Plasticity.FieldDynamics._gpu_betap import GpuBetaPDynamics


class BetaPDynamics(GpuBetaPDynamics):
    dynamics_kind = L.GLIDE_CLIMB_WOJZZ

    def _plan(self, state, **extra):
        if len(state.gridShape) != 2:
            raise NotImplementedError("the woJzz dynamics is 2D only")
        return GpuBetaPDynamics._plan(self, state, **extra)
