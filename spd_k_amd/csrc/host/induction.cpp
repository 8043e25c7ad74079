// Induction integrator (src/induction_ader.hpp).
#include <cstdio>
#include "driver.hpp"
namespace spdk_host {
int run_induction(RunConfig& cfg, const std::string& outdir, int warmup, spdk_run_result* res) {
    (void)cfg; (void)outdir; (void)warmup; (void)res;
    std::fprintf(stderr, "ERROR: induction system not built yet\n");
    return SPDK_E_CONFIG;
}
}  // namespace spdk_host
