// Entry points declared in deco_hip.h whose implementation has not landed yet.
#include "common.hpp"

using namespace deco;

#define DECO_STUB(sig) \
  sig { return guard([&] { fail(DECO_ERR_UNSUPPORTED, "%s is not implemented", __func__); }); }

extern "C" {
DECO_STUB(int deco_spgemm(const deco_csr*, const deco_csr*, deco_csr**))
DECO_STUB(int deco_spgemm_tril_pattern(const deco_csr*, deco_csr**))
DECO_STUB(int deco_cpr_create(const deco_cpr_params*, deco_csr*, double*, deco_prec**))
DECO_STUB(int deco_fsai_create(const deco_cpr_params*, const deco_csr*, deco_prec**))
DECO_STUB(int deco_jacobi_create(const deco_cpr_params*, const deco_csr*, deco_prec**))
DECO_STUB(int deco_amg_create(const deco_cpr_params*, const deco_csr*, deco_prec**))
DECO_STUB(int deco_identity_create(int, deco_prec**))
DECO_STUB(int deco_prec_destroy(deco_prec*))
DECO_STUB(int deco_prec_apply(deco_prec*, const double*, double*))
DECO_STUB(int deco_prec_apply_transpose(deco_prec*, const double*, double*))
DECO_STUB(int deco_prec_size(const deco_prec*, int*))
DECO_STUB(int deco_prec_set_precision(deco_prec*, int))
DECO_STUB(int deco_cpr_pressure_matrix(const deco_prec*, const deco_csr**))
DECO_STUB(int deco_cpr_fsai(const deco_prec*, const deco_prec**))
DECO_STUB(int deco_cpr_amg(const deco_prec*, const deco_prec**))
DECO_STUB(int deco_cpr_left_transform(const deco_prec*, double*))
DECO_STUB(int deco_fsai_factors(const deco_prec*, const deco_csr**, const deco_csr**))
DECO_STUB(int deco_amg_num_levels(const deco_prec*, int*))
DECO_STUB(int deco_amg_level(const deco_prec*, int, const deco_csr**, const deco_csr**, const deco_csr**,
                             const double**, const int**))
DECO_STUB(int deco_amg_coarse_inverse(const deco_prec*, int*, const double**))
DECO_STUB(int deco_solve(const deco_sol_params*, const deco_csr*, deco_prec*, const double*, double*,
                         deco_sol_info*))
DECO_STUB(int deco_solve_transpose(const deco_sol_params*, deco_csr*, deco_prec*, const double*, double*,
                                   deco_sol_info*))
DECO_STUB(int deco_grad_A_pattern(const deco_csr*, const double*, const double*, double*))
DECO_STUB(int deco_cpr_transform_transpose(const deco_prec*, const double*, double*))
}
