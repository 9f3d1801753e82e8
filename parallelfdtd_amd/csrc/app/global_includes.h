// global_includes.h -- the reference's shared header (src/global_includes.h): the logger and
// the reflection <-> admittance conversions the callers and the material handler use.
#pragma once

#define APP_EXIT_PAUSE 0

#include "logger.h"

// reflection2Admitance / admitance2Reflection, src/global_includes.h:30-38: (1 - R) / (1 + R),
// an involution, so the inverse is the same map.
inline float reflection2Admitance(float coef) { return (1.f - coef) / (1.f + coef); }
inline float admitance2Reflection(float coef) { return reflection2Admitance(coef); }
