"""minibwa_amd -- MI355X (gfx950) seeding + extension hot path for minibwa.

``minibwa_amd.mbg`` binds the C ABI of ``libmbg.so`` (include/mbg.h); ``minibwa_amd.build`` compiles
it.  Nothing here falls back to the CPU.
"""
__all__ = ["mbg", "build"]
