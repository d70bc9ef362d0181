// rekey / retypeval (sb_convert.hip).
#pragma once

#include "sb_ops.h"

namespace sb {

// rekey(A, K2; check) (blocksparse/conversions.jl:327-342, coo/conversions.jl:41-53,
// aliased/conversions.jl:76-93): a copy declaring key type K2 (SB_KEY_*)
TP tensor_rekey(Ctx*, const Tensor& A, int key_code, bool check);
// retypeval(A::COOTensor, T2) (coo/conversions.jl:65-82): a copy declaring value type T2 (SB_VAL_*)
TP coo_retypeval(Ctx*, const Tensor& A, int val_code);

}  // namespace sb
