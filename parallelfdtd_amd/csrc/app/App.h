// App.h -- the header name the reference's callers include for the application class
// (src/App.h; src/main.cpp:27, matlab/mex_includes.h:5): FDTD::App and the types its callers
// use (Source, Receiver, SrcType, InputType, UpdateType, the callback typedefs), declared in
// fdtd_app.h, plus the logger and conversions of global_includes.h. HIP-free: a plain C++
// compiler builds callers against it; they link libpfdtd_app.so.
#pragma once

#include "global_includes.h"
#include "fdtd_app.h"
