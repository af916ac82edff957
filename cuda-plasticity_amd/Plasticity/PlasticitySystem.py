"""PlasticitySystem (Plasticity/PlasticitySystem.py:3-21): holds the state,
mover, dynamics and observers; Run delegates to the mover."""


class PlasticitySystem:
    def __init__(self, gridShape, state, mover, dynamics, observers=None):
        self.gridShape = gridShape
        self.state = state
        self.mover = mover
        self.dynamics = dynamics
        self.observers = observers if observers is not None else []

    def Run(self, startTime=0.0, endTime=10.0, fixedTimeStep=None):
        self.mover.Run(self.state, self.observers, self.dynamics, startTime, endTime, fixedTimeStep=fixedTimeStep)
