// Round launchers with F_p^2 outputs (base-field or F_p^2 inputs).
#include "ops.hpp"

EFFSC_INSTANTIATE_OPS(gl::Fp, gl::Fp2)
EFFSC_INSTANTIATE_OPS(gl::Fp2, gl::Fp2)
