"""qphix-codegen_amd: MI355X-native (gfx950 HIP) Wilson-type Dirac operators
with the semantics of the kernels qphix-codegen generates, behind a C ABI
(include/qphix_amd.h).  This package is the Python host side: ctypes
binding, operator classes mirroring QPhiX's, and synthetic lattices."""
from ._lib import (BF16, F16, F32, F64, FIELD_CLOVER, FIELD_FULLCLOVER, FIELD_GAUGE,  # noqa: F401
                   FIELD_SPINOR, LIB_PATH, QPhiXAMDError, QPhiXLayout, lib)
from .dslash import (PREC_NAME, TORCH_DTYPE, ClovDslash, Dslash, Field, Geometry,  # noqa: F401
                     QPhiXSpinor, TMClovDslash, TMDslash, face_finish, face_pack, qphix_layout)
from . import lattice  # noqa: F401
from .solver import (EvenOddCloverOperator, InvCG, clover_force, clover_from_links, clover_invert,  # noqa: F401
                     reunitarize, tmclover_blocks)
from .autograd import (DslashFunction, PseudofermionAction, dslash_adjoint, dslash_force,  # noqa: F401
                       su3_force_dslash, su3_force_pseudofermion)
