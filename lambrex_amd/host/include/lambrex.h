// lambrex.h -- process set-up of the reference (include/lambrex.h:6-7).
#pragma once

#include "AmrSim.h"

// Creates the device context (GPU from $LBX_DEVICE, default 0) that every
// AmrSim uses; replaces amrex::Initialize (src/lambrex.cpp:4-9).  Calling it
// twice is harmless.  Throws lambrex::Error when no GPU can be used.
void lambrexInit();
// Releases the device context (src/lambrex.cpp:11-14).  AmrSim objects must
// be destroyed first.
void lambrexFinalise();
