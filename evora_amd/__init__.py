"""evora_amd -- evora's hydro step (MUSCL-Hancock + HLL/HLLC) on MI355X.

The compute path is the HIP library `libevora_hip.so` behind the C-ABI in
include/evora_hip.h; `evora_amd.hydro` is its Python host mirror.  The
modules params, ics and dump are host-side helpers (parameters, initial
conditions, the reference dump format) and need no GPU.
"""
