// lin_remap.hpp -- the vertical remap (src/mod_remap.f90:27-356) as
// operations of the linearised step.
#pragma once

#include "lin_ops.hpp"

template <class E>
void lin_remap_program(E &) {}
