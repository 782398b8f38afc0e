"""MI355X-native D3Q15 collide-stream path of LAMBReX behind a C ABI.

Modules:
  lbx        -- ctypes binding of include/lbx.h (liblbx.so)
  workloads  -- seeded synthetic initial states for the named configurations
  build      -- compiles liblbx.so and the C++ AmrSim facade for gfx950
"""
