"""gsmfretda_amd -- MI355X (gfx950) Monte-Carlo PDA hot path of gSMFRETda.

pda    : ctypes host interface over libgsmfretda_pda.so (the C ABI of
         include/gsmfretda_pda.h), mirroring the reference's `class mc`
synth  : deterministic synthetic bursts in the reference's HDF5 layout
build  : in-tree hipcc build of the shared library
"""
__version__ = "0.1.0"
