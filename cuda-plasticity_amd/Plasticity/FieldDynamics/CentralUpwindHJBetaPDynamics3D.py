"""Glide + climb BetaP dynamics with climb suppression Lambda
(Plasticity/FieldDynamics/CentralUpwindHJBetaPDynamics3D.py:10-221): Lambda
= 0 is 'gcd' (equal glide/climb mobility), Lambda = 1 is 'mdp' (glide only
by stress projection), 0 < Lambda < 1 is partial climb."""
from Plasticity.FieldDynamics._gpu_betap import GpuBetaPDynamics


class BetaPDynamics(GpuBetaPDynamics):
    pass
