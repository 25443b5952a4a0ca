// Round launchers with F_p^3 outputs (base-field or F_p^3 inputs).
#include "ops.hpp"

EFFSC_INSTANTIATE_OPS(gl::Fp, gl::Fp3)
EFFSC_INSTANTIATE_OPS(gl::Fp3, gl::Fp3)
