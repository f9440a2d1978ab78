// MOM6-variant evaluator: single tree traversal lists and per-point sums
// (SINGLE_TREE_TRAVERSAL, tree_module.f90:401-487; mom6_fast_sum_module.f90:33-192).
#include "common.h"
#include "devmath.h"
#include "plan.h"

using namespace sal;

extern "C" {

int sal_stt(sal_geom g, double theta, const double* x, const double* y, const double* z, int64_t n_own,
            int64_t* n_pc, int64_t* n_pp) {
    return guard([&] { SAL_REQUIRE(false, SAL_ERR_STATE, "sal_stt: single tree traversal not available in this build"); });
}

int sal_stt_export(sal_geom g, int* pc_target, int* pc_source, int* pp_target, int* pp_source) {
    return guard([&] { SAL_REQUIRE(false, SAL_ERR_STATE, "sal_stt_export: no lists"); });
}

int sal_mom6_proxy_source(sal_geom g, const double* area, const double* ssh, int degree, double* w) {
    return guard([&] { SAL_REQUIRE(false, SAL_ERR_STATE, "sal_mom6_proxy_source not available in this build"); });
}

int sal_mom6_sums(sal_geom g, const double* area, const double* ssh, int degree, int transposed, double* gx,
                  double* gy, int accumulate) {
    return guard([&] { SAL_REQUIRE(false, SAL_ERR_STATE, "sal_mom6_sums not available in this build"); });
}

}  // extern "C"
