"""create_rate_subs for MI355X: species/reaction-rate (wdot) evaluation.

Mechanism parsing and the generator CLI restate kyleniemeyer/create_rate_subs;
the ``hip`` emitter writes mechanism-specific CDNA4 kernels behind the C ABI
of ``include/crs_hip.h``; ``runtime.load`` binds a built library to torch
device tensors.
"""

__all__ = ['chem', 'mech_interpret', 'utils', 'create_rate_subs',
           'hip_emitter', 'build', 'runtime', 'synth', 'states', 'workloads',
           'metrics', 'cost']
