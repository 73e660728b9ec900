// pdlp_gpu.hpp -- the GPU PDLP entry point (reference: include/pdlp_gpu.hpp:
// 8-26), served by the MI355X engine.  Same contract as the reference: same
// inputs, same PdlpResult, iterations count attempts, x / y in unscaled PDLP-
// form coordinates, cooperative cancellation through stop_flag (polled by the
// host loop), std::runtime_error on a device failure.  The build gate the
// reference spells SOVEREIGN_WITH_CUDA is SUPREMA_WITH_HIP here.
#pragma once
#include <atomic>
#include <vector>

#include "pdlp.hpp"
#include "suprema_pdlp.hpp"

#define SUPREMA_WITH_HIP 1

inline PdlpResult solve_pdhg_gpu(const RangedLP& unscaled_form, const RangedLP& scaled_form,
                                 const std::vector<double>& Dr, const std::vector<double>& Dc,
                                 double eta, int max_iterations, int check_every, double tol,
                                 std::atomic<bool>* stop_flag) {
    return suprema::solve_pdhg_gpu_as<PdlpResult>(unscaled_form, scaled_form, Dr, Dc, eta,
                                                  max_iterations, check_every, tol, stop_flag);
}
