// stage_kernels.hip -- the stage launcher of one (dim, Np) pair: every RHS / fused-stage
// kernel variant of that pair.  Compiled once per pair with -DPRANDTL_STAGE_DIM=d
// -DPRANDTL_STAGE_NP=n (prandtl_amd/build.py), so the variants build in parallel.
#include "op_internal.hpp"

#if !defined(PRANDTL_STAGE_DIM) || !defined(PRANDTL_STAGE_NP)
#error "compile with -DPRANDTL_STAGE_DIM=<2|3> -DPRANDTL_STAGE_NP=<2..5>"
#endif

namespace prandtl {

template <int DIM, int NP>
int dispatch_visc(prandtl_op_s *op, const StageSpec &S, hipStream_t st) {
    if (op->opts.precision == PRANDTL_MIXED)
        return op->phys.viscous ? dispatch_surf<DIM, NP, true, true>(op, S, st)
                                : dispatch_surf<DIM, NP, false, true>(op, S, st);
    return op->phys.viscous ? dispatch_surf<DIM, NP, true, false>(op, S, st)
                            : dispatch_surf<DIM, NP, false, false>(op, S, st);
}

template int dispatch_visc<PRANDTL_STAGE_DIM, PRANDTL_STAGE_NP>(prandtl_op_s *, const StageSpec &, hipStream_t);

}  // namespace prandtl
