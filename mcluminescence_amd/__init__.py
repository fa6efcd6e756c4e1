"""mcluminescence_amd — MI355X-native trap/charge-kinetics Monte Carlo.

Drop-in for the reference's src/class simulation path (engine.py, tl_trap_lab.py,
simulate.py, optimizer.py) with the step loop running as hand-written HIP
kernels for gfx950 behind the C ABI in include/mcl.h.
"""
__version__ = "0.1.0"
