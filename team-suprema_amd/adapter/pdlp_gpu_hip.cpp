// pdlp_gpu_hip.cpp -- drop-in replacement for the reference's
// gpu/pdlp_gpu_solve.cu: defines solve_pdhg_gpu() (reference:
// include/pdlp_gpu.hpp:22-25) on top of the MI355X engine's C ABI.  Compile it
// against the reference's own headers with the GPU gate renamed to
// SUPREMA_WITH_HIP and defined, e.g.
//   g++ -std=c++17 -DSUPREMA_WITH_HIP=1 -I<reference>/include -I<this repo>/include \
//       -c pdlp_gpu_hip.cpp
// (unrenamed headers: add -DSOVEREIGN_WITH_CUDA=SUPREMA_WITH_HIP) and link
// libsuprema_pdlp.so.  The reference's types are used as they are.
#include "pdlp_gpu.hpp"
#include "suprema_pdlp.hpp"

PdlpResult solve_pdhg_gpu(const RangedLP& unscaled_form, const RangedLP& scaled_form,
                          const std::vector<double>& Dr, const std::vector<double>& Dc, double eta,
                          int max_iterations, int check_every, double tol,
                          std::atomic<bool>* stop_flag) {
    return suprema::solve_pdhg_gpu_as<PdlpResult>(unscaled_form, scaled_form, Dr, Dc, eta,
                                                  max_iterations, check_every, tol, stop_flag);
}
