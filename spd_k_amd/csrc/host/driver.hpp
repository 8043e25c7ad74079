// Shared pieces of the host integrators.
#pragma once
#include <string>
#include <vector>
#include "../../../include/spdk.h"
#include "setup.hpp"

namespace spdk_host {

void check(int rc, const char* what);

// Device allocations of one run, zero-initialised on the context's stream.
struct DeviceArena {
    std::vector<void*> ptrs;
    double* alloc(spdk_ctx* ctx, size_t count);
    void release();
    ~DeviceArena() { release(); }
};

size_t sd_count(const spdk_sd_shape& s);
size_t fv_count(const spdk_fv_shape& s);
spdk_ctx* make_context(const RunConfig& cfg, const Operators& op, const Axis* ax[3]);
void fill_kernel_timing(spdk_ctx* ctx, spdk_run_result* res);

int run_hydro(RunConfig& cfg, const std::string& outdir, int warmup, spdk_run_result* res);
int run_induction(RunConfig& cfg, const std::string& outdir, int warmup, spdk_run_result* res);

}  // namespace spdk_host
