// Round launchers with base-field outputs.
#include "ops.hpp"

EFFSC_INSTANTIATE_OPS(gl::Fp, gl::Fp)
