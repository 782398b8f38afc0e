"""MI355X-native D3Q15 collide-stream path of LAMBReX behind a C ABI.

Modules:
  lbx        -- ctypes binding of include/lbx.h (liblbx.so)
  workloads  -- seeded synthetic initial states for the named configurations
The HIP library and the C++ AmrSim facade are built by lambrex_amd/Makefile
(__graft_entry__.build() runs it).
"""
