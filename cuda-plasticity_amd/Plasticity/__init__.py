"""MI355X-native host layer for the continuum-dislocation-dynamics stepper.

Module names mirror the reference package ``Plasticity`` (PlasticitySystem,
FieldMovers.FieldMover, FieldDynamics.*, PlasticityStates.PlasticityState,
Observers.Observer) so driver scripts written against the reference's
stepping API keep their imports.  All stepping runs through the HIP C ABI
(``Plasticity.amd``); there is no CPU fallback.
"""
